// The reference's three C++ entry points (fherma.h:46-48), as wrappers over
// the C ABI: same signatures and types, errors thrown as std::runtime_error.
#include "../../include/fherma.h"

#include "../../include/polymult.h"

void* fherma_init(const fherma::Point& pt) {
    pm_state* s = nullptr;
    if (pm_init(pt.N, pt.W, pt.L, pt.q.data.data(), pt.q.data.size(), &s) != 0)
        throw std::runtime_error(pm_last_error());
    return s;
}

fherma::Outputs fherma_run(void* state, const fherma::Inputs& in) {
    auto* s = static_cast<pm_state*>(state);
    pm_info info{};
    if (pm_query(s, &info) != 0) throw std::runtime_error(pm_last_error());
    const size_t words = (size_t)info.N * info.L;
    if (in.a.data.size() != words || in.b.data.size() != words)
        throw std::runtime_error("operand size does not match N x L");
    fherma::Outputs out;
    out.c.shape = {static_cast<int64_t>(info.N), static_cast<int64_t>(info.L)};
    out.c.data.resize(words);
    if (pm_run(s, in.a.data.data(), in.b.data.data(), out.c.data.data()) != 0)
        throw std::runtime_error(pm_last_error());
    return out;
}

void fherma_free(void* state) { pm_free(static_cast<pm_state*>(state)); }
