// The reference's three C++ entry points (fherma.h:46-48), as wrappers over
// the C ABI: same signatures and types, errors thrown as std::runtime_error.
//
// fherma_run must return a fresh N x L std::vector, and value-initialising
// 14 MB costs milliseconds of page faults. So each call hands back a vector
// built ahead of time: after a multiply, the call orders the next one from a
// worker thread (a promise in the worker's queue) and the following call
// collects it through the matching future. The worker never sees operand data.
#include <condition_variable>
#include <deque>
#include <future>
#include <memory>
#include <mutex>
#include <thread>

#include "../../include/fherma.h"
#include "../../include/polymult.h"

namespace {

using Result = std::vector<uint32_t>;

class ResultSupply {
  public:
    explicit ResultSupply(size_t words) : words_(words), worker_([this] { serve(); }) { pending_ = order(); }
    ~ResultSupply() {
        {
            std::lock_guard<std::mutex> g(mu_);
            closing_ = true;
        }
        wake_.notify_one();
        worker_.join();
    }
    // The vector ordered last; the caller orders the next one with refill().
    // If building it failed, the error reaches the caller and a new order
    // keeps the next call going.
    Result collect() {
        try {
            return pending_.get();
        } catch (...) {
            pending_ = order();
            throw;
        }
    }
    void refill() { pending_ = order(); }

  private:
    std::future<Result> order() {
        std::promise<Result> pr;
        std::future<Result> f = pr.get_future();
        {
            std::lock_guard<std::mutex> g(mu_);
            orders_.push_back(std::move(pr));
        }
        wake_.notify_one();
        return f;
    }
    void serve() {
        for (;;) {
            std::promise<Result> pr;
            {
                std::unique_lock<std::mutex> g(mu_);
                wake_.wait(g, [&] { return closing_ || !orders_.empty(); });
                if (orders_.empty()) return;  // closing, nothing left to fill
                pr = std::move(orders_.front());
                orders_.pop_front();
            }
            try {
                pr.set_value(Result(words_));
            } catch (...) {
                pr.set_exception(std::current_exception());  // bad_alloc reaches the caller
            }
        }
    }
    const size_t words_;
    std::mutex mu_;
    std::condition_variable wake_;
    std::deque<std::promise<Result>> orders_;
    bool closing_ = false;
    std::future<Result> pending_;
    std::thread worker_;
};

struct FhermaState {
    pm_state* s = nullptr;
    uint32_t N = 0, L = 0;
    std::unique_ptr<ResultSupply> next;
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
    F->next = std::make_unique<ResultSupply>((size_t)pt.N * pt.L);
    return F.release();
}

fherma::Outputs fherma_run(void* state, const fherma::Inputs& in) {
    auto* F = static_cast<FhermaState*>(state);
    const size_t words = (size_t)F->N * F->L;
    if (in.a.data.size() != words || in.b.data.size() != words)
        throw std::runtime_error("operand size does not match N x L");
    fherma::Outputs out;
    out.c.shape = {static_cast<int64_t>(F->N), static_cast<int64_t>(F->L)};
    out.c.data = F->next->collect();
    const int rc = pm_run(F->s, in.a.data.data(), in.b.data.data(), out.c.data.data());
    F->next->refill();
    if (rc != 0) throw std::runtime_error(pm_last_error());
    return out;
}

void fherma_free(void* state) { delete static_cast<FhermaState*>(state); }
