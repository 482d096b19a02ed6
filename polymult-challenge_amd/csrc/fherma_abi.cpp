// The reference's three C++ entry points (fherma.h:46-48), as wrappers over
// the C ABI: same signatures and types, errors thrown as std::runtime_error.
//
// fherma_run must return a fresh N x L std::vector, and value-initialising
// 14 MB costs milliseconds of page faults. A helper thread therefore builds
// the next result vector between calls; the call only takes it. The helper
// never sees operand data.
#include <condition_variable>
#include <memory>
#include <mutex>
#include <thread>

#include "../../include/fherma.h"
#include "../../include/polymult.h"

namespace {

class NextResult {
  public:
    explicit NextResult(size_t words) : words_(words), th_([this] { loop(); }) { ask(); }
    ~NextResult() {
        {
            std::lock_guard<std::mutex> lk(m_);
            stop_ = true;
        }
        cv_.notify_all();
        th_.join();
    }
    std::vector<uint32_t> take() {
        std::unique_lock<std::mutex> lk(m_);
        cv_.wait(lk, [&] { return ready_; });
        ready_ = false;
        return std::move(buf_);
    }
    void ask() {
        {
            std::lock_guard<std::mutex> lk(m_);
            wanted_ = true;
        }
        cv_.notify_all();
    }

  private:
    void loop() {
        std::unique_lock<std::mutex> lk(m_);
        for (;;) {
            cv_.wait(lk, [&] { return stop_ || wanted_; });
            if (stop_) return;
            wanted_ = false;
            lk.unlock();
            std::vector<uint32_t> v(words_);
            lk.lock();
            buf_ = std::move(v);
            ready_ = true;
            cv_.notify_all();
        }
    }
    size_t words_;
    std::mutex m_;
    std::condition_variable cv_;
    bool wanted_ = false, ready_ = false, stop_ = false;
    std::vector<uint32_t> buf_;
    std::thread th_;
};

struct FhermaState {
    pm_state* s = nullptr;
    uint32_t N = 0, L = 0;
    std::unique_ptr<NextResult> next;
    ~FhermaState() {
        next.reset();  // stop the helper before the state goes
        pm_free(s);
    }
};

}  // namespace

void* fherma_init(const fherma::Point& pt) {
    pm_state* s = nullptr;
    if (pm_init(pt.N, pt.W, pt.L, pt.q.data.data(), pt.q.data.size(), &s) != 0)
        throw std::runtime_error(pm_last_error());
    auto F = std::make_unique<FhermaState>();  // owns s from here on
    F->s = s;
    F->N = pt.N;
    F->L = pt.L;
    F->next = std::make_unique<NextResult>((size_t)pt.N * pt.L);
    return F.release();
}

fherma::Outputs fherma_run(void* state, const fherma::Inputs& in) {
    auto* F = static_cast<FhermaState*>(state);
    const size_t words = (size_t)F->N * F->L;
    if (in.a.data.size() != words || in.b.data.size() != words)
        throw std::runtime_error("operand size does not match N x L");
    fherma::Outputs out;
    out.c.shape = {static_cast<int64_t>(F->N), static_cast<int64_t>(F->L)};
    out.c.data = F->next->take();
    const int rc = pm_run(F->s, in.a.data.data(), in.b.data.data(), out.c.data.data());
    F->next->ask();
    if (rc != 0) throw std::runtime_error(pm_last_error());
    return out;
}

void fherma_free(void* state) { delete static_cast<FhermaState*>(state); }
