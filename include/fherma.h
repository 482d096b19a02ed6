// fherma.h -- the three C++ entry points the reference's harness binds, with
// the same types, so a harness built against the reference's fherma.h links
// against libpolymult unchanged.
//
// The layout is fixed by the reference interface (fherma.h:13-48 there):
// Tensor<T> is {shape, data} with row-major data; Point carries N, W, L and q
// (L limbs); Inputs carries a and b, Outputs carries c, each N x L u32. The
// names, member order and signatures must stay identical for ABI identity.
#pragma once

#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

namespace fherma {

template <class T>
struct Tensor {
    std::vector<int64_t> shape;
    std::vector<T> data;  // row-major

    int64_t count() const {
        int64_t n = 1;
        for (auto d : shape) n *= d;
        return n;
    }
};

struct Point {
    uint32_t N = 0;
    uint32_t W = 0;
    uint32_t L = 0;
    Tensor<uint32_t> q;  // L limbs
};

struct Inputs {
    Tensor<uint32_t> a;  // N x L
    Tensor<uint32_t> b;  // N x L
};

struct Outputs {
    Tensor<uint32_t> c;  // N x L
};

}  // namespace fherma

// Throw std::runtime_error on failure, as the reference does.
void* fherma_init(const fherma::Point& p);
fherma::Outputs fherma_run(void* state, const fherma::Inputs& in);
void fherma_free(void* state);
