// Host-side helpers shared by the PCIe pipeline and the serving loop.
#pragma once

#include <atomic>
#include <condition_variable>
#include <cstddef>
#include <cstdint>
#include <functional>
#include <mutex>
#include <thread>
#include <vector>

#if defined(__x86_64__)
#include <immintrin.h>
#endif

namespace pm {

inline void relax() {
#if defined(__x86_64__)
    _mm_pause();
#endif
}

// memcpy with non-temporal stores (AVX2 where the CPU has it).
void copy_stream(void* dst, const void* src, size_t n);
// CPUs this process may run on.
int host_cpus();

// A team of host threads that runs fn(0..n-1) with the calling thread taking
// part. Between calls the workers sleep; while a multiply is in progress
// (hot()) they spin so a copy job starts within microseconds.
class Team {
  public:
    explicit Team(int workers) {
        for (int i = 0; i < workers; ++i) threads_.emplace_back([this] { loop(); });
    }
    ~Team() {
        {
            std::lock_guard<std::mutex> lk(m_);
            stop_ = true;
        }
        cv_.notify_all();
        for (auto& t : threads_) t.join();
    }
    int size() const { return (int)threads_.size() + 1; }
    void hot(bool on) {
        {
            std::lock_guard<std::mutex> lk(m_);
            hot_ = on;
        }
        if (on) cv_.notify_all();
    }
    void run(int n, const std::function<void(int)>& fn) {
        Job job{&fn, n};
        current_.store(&job);
        epoch_.fetch_add(1);
        if (sleeping_.load() > 0) {
            // taking the lock orders the wake-up after any sleeper's predicate check
            { std::lock_guard<std::mutex> lk(m_); }
            cv_.notify_all();
        }
        work(job);
        while (job.done.load() < n) relax();
        current_.store(nullptr);
        while (busy_.load() > 0) relax();  // nobody still holds &job
    }

  private:
    struct Job {
        const std::function<void(int)>* fn;
        int total;
        std::atomic<int> next{0};
        std::atomic<int> done{0};
    };
    static void work(Job& j) {
        for (int i; (i = j.next.fetch_add(1)) < j.total;) {
            (*j.fn)(i);
            j.done.fetch_add(1);
        }
    }
    void loop() {
        uint64_t seen = epoch_.load();
        for (;;) {
            int idle = 0;
            while (epoch_.load() == seen) {
                relax();
                if (++idle > 4096) {
                    std::unique_lock<std::mutex> lk(m_);
                    if (stop_) return;
                    if (!hot_ && epoch_.load() == seen) {
                        sleeping_.fetch_add(1);
                        cv_.wait(lk, [&] { return stop_ || hot_ || epoch_.load() != seen; });
                        sleeping_.fetch_sub(1);
                        if (stop_) return;
                    }
                    idle = 0;
                }
            }
            seen = epoch_.load();
            busy_.fetch_add(1);
            if (Job* j = current_.load()) work(*j);
            busy_.fetch_sub(1);
        }
    }
    std::vector<std::thread> threads_;
    std::atomic<Job*> current_{nullptr};
    std::atomic<uint64_t> epoch_{0};
    std::atomic<int> busy_{0}, sleeping_{0};
    std::mutex m_;
    std::condition_variable cv_;
    bool hot_ = false, stop_ = false;
};

}  // namespace pm
