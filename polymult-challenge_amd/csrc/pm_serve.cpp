// Serving loop: multiply requests for a fixed catalog of points arrive one at a
// time from host buffers, and a scheduler thread groups whatever is waiting
// into a batch (continuous batching). A request is admitted into a batch only
// once its HBM footprint (operands, result, residues) and its pinned staging
// fit the server's budgets; requests are admitted in arrival order. Each batch
// runs as
//   stage (host team: pageable -> pinned) -> one H2D per request ->
//   five phase launches over every request's tiles -> one D2H per request ->
//   scatter (host team: pinned -> the caller's buffer),
// on three streams (in, and two for compute and out) with up to `max_inflight`
// batches at once, so one batch's transfers overlap another's compute. A batch
// with at most 2 MiB of operands skips the copies: its kernels read the pinned
// staging and write the results there. Latency is measured from a request's
// arrival to the end of its scatter.
#include <hip/hip_runtime.h>
#include <sys/resource.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <random>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../../include/polymult.h"
#include "pm_batch.h"
#include "pm_host.h"

namespace pm {
namespace {

using Clock = std::chrono::steady_clock;
double now_s() { return std::chrono::duration<double>(Clock::now().time_since_epoch()).count(); }
size_t round256(size_t n) { return (n + 255) & ~size_t(255); }

// Nearest-rank percentile of a sorted, non-empty sample.
double percentile(const std::vector<double>& v, double p) {
    size_t i = (size_t)std::ceil(p * (double)v.size());
    i = i ? i - 1 : 0;
    return v[std::min(v.size() - 1, i)];
}

// First-fit allocator over [0, size) in 256-byte granules; free ranges are
// kept sorted and merged on release.
class Arena {
  public:
    void reset(size_t size) {
        free_.clear();
        if (size) free_[0] = size;
        size_ = size;
        used_ = peak_ = 0;
    }
    bool alloc(size_t n, size_t& off) {
        n = round256(n);
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second < n) continue;
            off = it->first;
            const size_t rest = it->second - n;
            free_.erase(it);
            if (rest) free_[off + n] = rest;
            used_ += n;
            peak_ = std::max(peak_, used_);
            return true;
        }
        return false;
    }
    void release(size_t off, size_t n) {
        n = round256(n);
        used_ -= n;
        auto it = free_.emplace(off, n).first;
        auto nx = std::next(it);
        if (nx != free_.end() && it->first + it->second == nx->first) {
            it->second += nx->second;
            free_.erase(nx);
        }
        if (it != free_.begin()) {
            auto pv = std::prev(it);
            if (pv->first + pv->second == it->first) {
                pv->second += it->second;
                free_.erase(it);
            }
        }
    }
    size_t size() const { return size_; }
    size_t used() const { return used_; }
    size_t peak() const { return peak_; }
    void clear_peak() { peak_ = used_; }

  private:
    std::map<size_t, size_t> free_;
    size_t size_ = 0, used_ = 0, peak_ = 0;
};

struct Request {
    uint64_t id = 0;
    uint32_t point = 0;
    const uint32_t* a = nullptr;
    const uint32_t* b = nullptr;
    uint32_t* c = nullptr;
    double t_arrive = 0, t_take = 0, t_done = 0;
    size_t dev_off = 0, pin_off = 0;
    int state = 0;  // 0 queued, 1 in a batch, 2 done, -1 failed
    std::string error;
};

// A piece of a host copy, so the team can split large operands.
struct Piece {
    void* dst;
    const void* src;
    size_t n;
};
constexpr size_t kPiece = 2u << 20;

void add_pieces(std::vector<Piece>& v, void* dst, const void* src, size_t n, size_t piece = kPiece) {
    for (size_t o = 0; o < n; o += piece)
        v.push_back({static_cast<char*>(dst) + o, static_cast<const char*>(src) + o, std::min(piece, n - o)});
}

// The host-to-device copy of a batch's first request goes in chunks of these
// sizes (the last repeats): the copy engine starts on the first MiB while the
// team stages the next chunk, instead of waiting for the whole batch to be
// staged; large chunks after that keep the engine near its large-copy rate.
// Each later request goes as one copy: its staging already overlaps the
// copies of the requests before it.
constexpr size_t kCopyChunks[] = {1u << 20, 3u << 20, 16u << 20};

// The results go back the same way: the first request's in chunks of these
// sizes (the last repeats) closed by a 0.5 MiB tail, each later request in
// one copy, with an event after each of the first request's chunks and after
// about every MiB of later results. The completion thread scatters each marked
// run into the callers' buffers while the copies after it are still running,
// so only the tail's scatter follows the last copy.
constexpr size_t kOutChunks[] = {1u << 20, 5u << 20};
constexpr size_t kOutTail = 512u << 10;
constexpr size_t kOutMark = 1u << 20;

// Batches with at most this many bytes of operands skip the copy engines: the
// kernels read the operands from the pinned staging arena and the CRT writes
// the results into it. A lone N=8192/W=438 request then waits for no
// copy-engine or cross-stream hand-off.
constexpr size_t kZeroCopyBytes = 2u << 20;

}  // namespace

struct Server {
    Catalog cat;
    BatchConfig cfg;
    std::vector<size_t> io_bytes, res_bytes, dev_need, pin_need;  // per point
    void* d_arena = nullptr;
    unsigned char* h_arena = nullptr;
    Arena dev_arena, pin_arena;
    size_t max_batch = 64;
    size_t max_batch_bytes = 0;
    // Copies in on one stream; consecutive batches compute, then copy out, on
    // alternate streams, so one batch's kernels overlap the other's copies.
    // Three streams, not four: a process has four hardware queues by default,
    // shared by every stream it has made, and once other streams have come and
    // gone a fourth server stream could share the input stream's queue and
    // serialize the two copy directions (capacity fell from 10,000 to 6,200
    // requests/s that way).
    hipStream_t s_in = nullptr, s_comp[2] = {nullptr, nullptr};
    uint64_t batch_seq = 0;  // scheduler thread only
    std::unique_ptr<Team> stage_team, scatter_team;

    struct Slot {
        Runner run;
        std::vector<Request*> reqs;
        hipEvent_t ev_in = nullptr, ev_out = nullptr;
        double t_take = 0, t_staged = 0, t_issued = 0;
        // Results leave in marked runs of copies: mark k's event follows its
        // copies, and its scatter pieces are out[out_end[k - 1] .. out_end[k]).
        std::vector<hipEvent_t> mark_ev;
        std::vector<size_t> out_end;
        std::vector<Piece> out;
        size_t nmark = 0;
    };
    std::vector<std::unique_ptr<Slot>> slots;

    // Everything below is guarded by m.
    std::mutex m;
    std::condition_variable cv_work;    // scheduler: work_gen moved
    std::condition_variable cv_flight;  // completer: a batch was issued, or the scheduler ended
    std::condition_variable cv_done;    // waiters: a request finished
    std::deque<Slot*> inflight;         // issued batches, oldest first
    std::vector<Slot*> idle;
    std::deque<Request*> queue;
    std::unordered_map<uint64_t, std::unique_ptr<Request>> reqs;
    uint64_t next_id = 1;
    uint64_t work_gen = 0;  // bumped by every submission, release and stop
    bool stop = false, sched_exit = false;
    std::string broken;
    uint64_t last_deferred = 0;

    // statistics since the last reset
    std::vector<double> lat_ms;
    uint64_t n_submitted = 0, n_completed = 0, n_failed = 0, n_batches = 0, n_batched = 0, n_deferred = 0;
    double s_stage = 0, s_issue = 0, s_gpu = 0, s_scatter = 0, t_reset = 0;

    std::thread sched, completer;

    // PM_SERVE_TRACE=<prefix>: per-batch and per-request timelines written
    // as <prefix>.batches.csv and <prefix>.requests.csv when the server ends.
    std::string trace_path;
    struct BatchRec {
        double take, staged, issued, seen, done;
        uint32_t nreq;
        uint64_t bytes;
    };
    struct ReqRec {
        double arrive, take, done;
        uint32_t point;
    };
    std::vector<BatchRec> trace_b;
    std::vector<ReqRec> trace_r;
    void write_trace();

    void create(uint32_t npoints, const uint32_t* N, const uint32_t* W, const uint32_t* q, const pm_server_config* c);
    ~Server();

    uint64_t submit(uint32_t point, const uint32_t* a, const uint32_t* b, uint32_t* c, double t_arrive);
    // 0 done, 1 not yet; latency from arrival to done, queue_ms from arrival
    // to admission into a batch
    int wait(uint64_t id, double timeout_s, double* latency_ms, double* queue_ms = nullptr);
    void stats(pm_server_stats* out, bool reset);

  private:
    void schedule_loop();
    void complete_loop();
    std::vector<Request*> take_locked();
    void launch(Slot& S);
    void retire(Slot& S, hipError_t status, double t_seen);
    void fail_batch(Slot& S, const std::string& why);
};

void Server::create(uint32_t npoints, const uint32_t* N, const uint32_t* W, const uint32_t* q,
                    const pm_server_config* c) {
    pm_server_config C{};
    if (c) C = *c;
    const size_t dev_budget = C.device_budget ? C.device_budget : (size_t)4 << 30;
    const size_t pin_budget = C.pinned_budget ? C.pinned_budget : (size_t)1 << 30;
    max_batch = C.max_batch ? C.max_batch : 64;
    max_batch_bytes = C.max_batch_bytes ? C.max_batch_bytes : (size_t)64 << 20;
    const int ninflight = C.max_inflight ? (int)C.max_inflight : 3;
    if (ninflight > 8) throw std::runtime_error("max_inflight above 8");
    cfg.col_tc = C.col_tile ? C.col_tile : 32;
    if (cfg.col_tc != 8 && cfg.col_tc != 16 && cfg.col_tc != 32) throw std::runtime_error("col_tile must be 8, 16 or 32");

    cat.build(npoints, N, W, q);
    for (uint32_t i = 0; i < npoints; ++i) {
        io_bytes.push_back(round256(cat.io_words(i) * 4));
        res_bytes.push_back(round256((size_t)cat.plans[i].K * cat.plans[i].N * 4));
        dev_need.push_back(2 * io_bytes[i] + 2 * res_bytes[i]);  // a|b (c over a), R0, R1
        pin_need.push_back(2 * io_bytes[i]);                      // a|b in, c back over a
    }
    hip_check(hipMalloc(&d_arena, dev_budget), "hipMalloc request arena");
    hip_check(hipHostMalloc(reinterpret_cast<void**>(&h_arena), pin_budget, hipHostMallocDefault),
              "hipHostMalloc staging arena");
    // The first CPU touch of a pinned page faults; do it here, not while a
    // batch is staging.
    std::memset(h_arena, 0, pin_budget);
    dev_arena.reset(dev_budget);
    pin_arena.reset(pin_budget);
    hip_check(hipStreamCreateWithFlags(&s_in, hipStreamNonBlocking), "stream");
    hip_check(hipStreamCreateWithFlags(&s_comp[0], hipStreamNonBlocking), "stream");
    hip_check(hipStreamCreateWithFlags(&s_comp[1], hipStreamNonBlocking), "stream");
    for (int i = 0; i < ninflight; ++i) {
        auto S = std::make_unique<Slot>();
        S->run.init();
        hip_check(hipEventCreateWithFlags(&S->ev_in, hipEventDisableTiming), "event");
        hip_check(hipEventCreateWithFlags(&S->ev_out, hipEventDisableTiming), "event");
        idle.push_back(S.get());
        slots.push_back(std::move(S));
    }
    // Every point must plan on this device (LDS) before a request can arrive;
    // the largest tile count of one product sizes the slots' tile buffers.
    size_t max_tiles = 0;
    {
        Runner& R = slots[0]->run;
        for (uint32_t i = 0; i < npoints; ++i) {
            R.prods.assign(1, BatchProduct{});
            R.prods[0].D = cat.dev[i];
            R.plan(cfg);
            max_tiles = std::max(max_tiles, R.tasks.size());
        }
        R.prods.clear();
        R.tasks.clear();
    }
    for (auto& S : slots) S->run.reserve(max_batch, max_batch * max_tiles);
    // host threads: half stage inputs (with the scheduler), half scatter
    // results (with the completer)
    const int threads = C.host_threads ? (int)C.host_threads : std::max(2, std::min(8, host_cpus() - 2));
    stage_team = std::make_unique<Team>(std::max(0, threads / 2 - 1));
    scatter_team = std::make_unique<Team>(std::max(0, threads - threads / 2 - 1));
    t_reset = now_s();
    if (const char* tp = std::getenv("PM_SERVE_TRACE")) trace_path = tp;
    sched = std::thread([this] { schedule_loop(); });
    completer = std::thread([this] { complete_loop(); });
}

Server::~Server() {
    {
        std::lock_guard<std::mutex> lk(m);
        stop = true;
        ++work_gen;
    }
    cv_work.notify_all();
    if (sched.joinable()) sched.join();
    if (completer.joinable()) completer.join();
    if (!trace_path.empty()) write_trace();
    for (hipStream_t s : {s_in, s_comp[0], s_comp[1]})
        if (s) (void)hipStreamSynchronize(s);
    for (auto& S : slots) {
        for (hipEvent_t e : {S->ev_in, S->ev_out})
            if (e) (void)hipEventDestroy(e);
        for (hipEvent_t e : S->mark_ev) (void)hipEventDestroy(e);
    }
    slots.clear();
    for (hipStream_t s : {s_in, s_comp[0], s_comp[1]})
        if (s) (void)hipStreamDestroy(s);
    if (d_arena) (void)hipFree(d_arena);
    if (h_arena) (void)hipHostFree(h_arena);
}

uint64_t Server::submit(uint32_t point, const uint32_t* a, const uint32_t* b, uint32_t* c, double t_arrive) {
    if (point >= cat.plans.size()) throw std::runtime_error("point index out of range");
    if (!a || !b || !c) throw std::runtime_error("null operand pointer");
    auto r = std::make_unique<Request>();
    r->point = point;
    r->a = a;
    r->b = b;
    r->c = c;
    r->t_arrive = t_arrive > 0 ? t_arrive : now_s();
    uint64_t id;
    {
        std::lock_guard<std::mutex> lk(m);
        if (!broken.empty()) throw std::runtime_error("server stopped after an error: " + broken);
        if (stop) throw std::runtime_error("server is shutting down");
        if (dev_need[point] > dev_arena.size())
            throw std::runtime_error("request needs " + std::to_string(dev_need[point]) +
                                     " bytes of device memory, above the server's budget of " +
                                     std::to_string(dev_arena.size()));
        if (pin_need[point] > pin_arena.size())
            throw std::runtime_error("request needs " + std::to_string(pin_need[point]) +
                                     " bytes of pinned staging, above the server's budget of " +
                                     std::to_string(pin_arena.size()));
        id = r->id = next_id++;
        queue.push_back(r.get());
        reqs.emplace(id, std::move(r));
        ++n_submitted;
        ++work_gen;
    }
    cv_work.notify_one();
    return id;
}

int Server::wait(uint64_t id, double timeout_s, double* latency_ms, double* queue_ms) {
    std::unique_lock<std::mutex> lk(m);
    auto it = reqs.find(id);
    if (it == reqs.end()) throw std::runtime_error("unknown ticket");
    Request* r = it->second.get();
    auto ready = [&] { return r->state == 2 || r->state < 0; };
    if (!ready()) {
        if (timeout_s <= 0) return 1;
        if (!cv_done.wait_for(lk, std::chrono::duration<double>(timeout_s), ready)) return 1;
    }
    if (latency_ms) *latency_ms = 1e3 * (r->t_done - r->t_arrive);
    if (queue_ms) *queue_ms = 1e3 * (r->t_take - r->t_arrive);
    if (r->state < 0) {
        const std::string e = r->error;
        reqs.erase(it);
        throw std::runtime_error(e);
    }
    reqs.erase(it);
    return 0;
}

void Server::stats(pm_server_stats* out, bool reset) {
    std::lock_guard<std::mutex> lk(m);
    *out = pm_server_stats{};
    out->submitted = n_submitted;
    out->completed = n_completed;
    out->failed = n_failed;
    out->batches = n_batches;
    out->deferred = n_deferred;
    out->mean_batch = n_batches ? (double)n_batched / n_batches : 0.0;
    if (!lat_ms.empty()) {
        std::vector<double> v = lat_ms;
        std::sort(v.begin(), v.end());
        out->p50_ms = percentile(v, 0.50);
        out->p99_ms = percentile(v, 0.99);
        out->max_ms = v.back();
        double s = 0;
        for (double x : v) s += x;
        out->mean_ms = s / v.size();
    }
    if (n_batches) {
        out->stage_ms = 1e3 * s_stage / n_batches;
        out->issue_ms = 1e3 * s_issue / n_batches;
        out->gpu_ms = 1e3 * s_gpu / n_batches;
        out->scatter_ms = 1e3 * s_scatter / n_batches;
    }
    const double wall = std::max(1e-9, now_s() - t_reset);
    out->sched_busy = (s_stage + s_issue) / wall;
    out->completer_busy = s_scatter / wall;
    out->device_budget = dev_arena.size();
    out->device_peak = dev_arena.peak();
    out->pinned_budget = pin_arena.size();
    out->pinned_peak = pin_arena.peak();
    if (reset) {
        lat_ms.clear();
        n_submitted = n_completed = n_failed = n_batches = n_batched = n_deferred = 0;
        s_stage = s_issue = s_gpu = s_scatter = 0;
        t_reset = now_s();
        dev_arena.clear_peak();
        pin_arena.clear_peak();
    }
}

// Admits requests from the head of the queue while they fit both budgets.
// Strict arrival order: a head that does not fit holds back the ones behind
// it (no starvation of large requests). Called with m held.
std::vector<Request*> Server::take_locked() {
    std::vector<Request*> batch;
    size_t bytes = 0;
    while (!queue.empty() && batch.size() < max_batch) {
        Request* r = queue.front();
        // a large backlog leaves as several pipelined batches, not one
        if (!batch.empty() && bytes + pin_need[r->point] > max_batch_bytes) break;
        size_t doff = 0, poff = 0;
        if (!dev_arena.alloc(dev_need[r->point], doff)) {
            if (r->id != last_deferred) ++n_deferred, last_deferred = r->id;
            break;
        }
        if (!pin_arena.alloc(pin_need[r->point], poff)) {
            dev_arena.release(doff, dev_need[r->point]);
            if (r->id != last_deferred) ++n_deferred, last_deferred = r->id;
            break;
        }
        r->dev_off = doff;
        r->pin_off = poff;
        r->state = 1;
        r->t_take = now_s();
        queue.pop_front();
        batch.push_back(r);
        bytes += pin_need[r->point];
    }
    return batch;
}

void Server::launch(Slot& S) {
    unsigned char* dev = static_cast<unsigned char*>(d_arena);
    Runner& R = S.run;
    size_t in_bytes = 0;
    for (const Request* r : S.reqs) in_bytes += 2 * cat.io_words(r->point) * 4;
    const bool zero_copy = in_bytes <= kZeroCopyBytes;
    R.prods.assign(S.reqs.size(), BatchProduct{});
    for (size_t i = 0; i < S.reqs.size(); ++i) {
        const Request* r = S.reqs[i];
        BatchProduct& P = R.prods[i];
        P.D = cat.dev[r->point];
        uint32_t* base = reinterpret_cast<uint32_t*>(dev + r->dev_off);
        // a | padding | b, in device memory or, for a zero-copy batch, in the
        // pinned arena (the same layout); the result goes over a: the CRT
        // phase runs after every RNS tile has read a
        uint32_t* io = zero_copy ? reinterpret_cast<uint32_t*>(h_arena + r->pin_off) : base;
        P.a = io;
        P.b = io + io_bytes[r->point] / 4;
        P.c = io;
        P.R0 = base + 2 * io_bytes[r->point] / 4;
        P.R1 = P.R0 + res_bytes[r->point] / 4;
    }
    R.plan(cfg);
    // A zero-copy batch whose groups all run the single-product kernels (plan
    // and pointers as kernel arguments) reads no tables and puts nothing on
    // the input stream, so it needs no cross-stream hand-off either (each
    // costs 10-20 us of host time).
    bool all_single = R.cfg.blocks_per_cu <= 0;
    for (int g : R.single) all_single = all_single && g >= 0;
    const bool via_input_stream = !(zero_copy && all_single);
    // The tile lists go first on the input stream, ahead of the operands: on
    // the compute stream their small copy would wait behind the copy engines'
    // queue of later batches' operands. The slot's tile buffers are free: a
    // slot is reused only after its batch retired.
    if (via_input_stream) R.upload(s_in, true);
    // Stage each request's a and b into pinned memory and copy them in, chunk
    // by chunk: a request's a|b region (a, its padding to 256 bytes, b) is
    // contiguous in pinned and in device memory alike.
    std::vector<Piece> pieces;
    const int T = stage_team->size();
    if (zero_copy) {
        // stage every operand in one team job; the kernels read them in place
        for (Request* r : S.reqs) {
            const size_t n = cat.io_words(r->point) * 4, ib = io_bytes[r->point];
            unsigned char* h = h_arena + r->pin_off;
            const size_t piece = std::max<size_t>(64u << 10, (2 * n + T - 1) / T);
            add_pieces(pieces, h, r->a, n, piece);
            add_pieces(pieces, h + ib, r->b, n, piece);
        }
        stage_team->run((int)pieces.size(), [&](int i) { copy_stream(pieces[i].dst, pieces[i].src, pieces[i].n); });
    }
    for (Request* r : S.reqs) {
        if (zero_copy) break;
        const size_t n = cat.io_words(r->point) * 4, ib = io_bytes[r->point], span = ib + n;
        unsigned char* h = h_arena + r->pin_off;
        const bool first = r == S.reqs.front();
        size_t k = 0;
        for (size_t lo = 0, hi; lo < span; lo = hi) {
            hi = first ? std::min(span, lo + kCopyChunks[k]) : span;
            if (k + 1 < sizeof(kCopyChunks) / sizeof(kCopyChunks[0])) ++k;
            const size_t piece = std::max<size_t>(64u << 10, (hi - lo + T - 1) / T);
            pieces.clear();
            // the parts of a = [0, n) and b = [ib, ib + n) inside [lo, hi)
            if (lo < n) add_pieces(pieces, h + lo, reinterpret_cast<const char*>(r->a) + lo, std::min(hi, n) - lo, piece);
            if (hi > ib) {
                const size_t b0 = std::max(lo, ib);
                add_pieces(pieces, h + b0, reinterpret_cast<const char*>(r->b) + (b0 - ib), hi - b0, piece);
            }
            stage_team->run((int)pieces.size(), [&](int i) { copy_stream(pieces[i].dst, pieces[i].src, pieces[i].n); });
            hip_check(hipMemcpyAsync(dev + r->dev_off + lo, h + lo, hi - lo, hipMemcpyHostToDevice, s_in), "H2D");
        }
    }
    S.t_staged = now_s();
    hipStream_t sc = s_comp[batch_seq++ % 2];
    if (via_input_stream) {
        hip_check(hipEventRecord(S.ev_in, s_in), "record");
        hip_check(hipStreamWaitEvent(sc, S.ev_in, 0), "wait");
    }
    hip_check(R.launch(sc), "batch launch");
    S.out.clear();
    S.out_end.clear();
    S.nmark = 0;
    if (zero_copy) {
        // the results are in the pinned arena once the batch's kernels end
        const int Ts = scatter_team->size();
        for (Request* r : S.reqs) {
            const size_t n = cat.io_words(r->point) * 4;
            add_pieces(S.out, r->c, h_arena + r->pin_off, n, std::max<size_t>(64u << 10, (n + Ts - 1) / Ts));
        }
        if (S.nmark == S.mark_ev.size()) {
            hipEvent_t e = nullptr;
            hip_check(hipEventCreateWithFlags(&e, hipEventDisableTiming), "event");
            S.mark_ev.push_back(e);
        }
        hip_check(hipEventRecord(S.mark_ev[S.nmark++], sc), "record");
        S.out_end.push_back(S.out.size());
        hip_check(hipEventRecord(S.ev_out, sc), "record");
        S.t_issued = now_s();
        return;
    }
    const int Ts = scatter_team->size();
    size_t pending = 0;
    auto mark = [&] {
        if (S.nmark == S.mark_ev.size()) {
            hipEvent_t e = nullptr;
            hip_check(hipEventCreateWithFlags(&e, hipEventDisableTiming), "event");
            S.mark_ev.push_back(e);
        }
        hip_check(hipEventRecord(S.mark_ev[S.nmark++], sc), "record");
        S.out_end.push_back(S.out.size());
        pending = 0;
    };
    for (Request* r : S.reqs) {
        const size_t n = cat.io_words(r->point) * 4;
        unsigned char* h = h_arena + r->pin_off;
        const bool first = r == S.reqs.front();
        size_t k = 0;
        for (size_t lo = 0, hi; lo < n; lo = hi) {
            hi = n;
            if (first && n - lo > kOutTail) {
                // a chunk of the plan, but never leaving less than the tail
                hi = std::min(lo + kOutChunks[k], n - kOutTail);
                if (k + 1 < sizeof(kOutChunks) / sizeof(kOutChunks[0])) ++k;
            }
            hip_check(hipMemcpyAsync(h + lo, dev + r->dev_off + lo, hi - lo, hipMemcpyDeviceToHost, sc), "D2H");
            add_pieces(S.out, reinterpret_cast<char*>(r->c) + lo, h + lo, hi - lo,
                       std::max<size_t>(64u << 10, (hi - lo + Ts - 1) / Ts));
            pending += hi - lo;
            if (first || pending >= kOutMark) mark();
        }
    }
    if (pending) mark();
    hip_check(hipEventRecord(S.ev_out, sc), "record");
    S.t_issued = now_s();
}

// Called by the completer once the batch's results are in the callers'
// buffers (t_seen: when its last D2H was seen finished), or on a failure.
void Server::retire(Slot& S, hipError_t status, double t_seen) {
    if (status != hipSuccess) {
        fail_batch(S, std::string("batch failed: ") + hipGetErrorString(status));
        return;
    }
    const double t = now_s();
    {
        std::lock_guard<std::mutex> lk(m);
        s_stage += S.t_staged - S.t_take;
        s_issue += S.t_issued - S.t_staged;
        s_gpu += t_seen - S.t_issued;
        s_scatter += t - t_seen;
        if (!trace_path.empty()) {
            uint64_t bytes = 0;
            for (Request* r : S.reqs) {
                bytes += 12 * cat.io_words(r->point);
                trace_r.push_back({r->t_arrive, r->t_take, t, r->point});
            }
            trace_b.push_back({S.t_take, S.t_staged, S.t_issued, t_seen, t, (uint32_t)S.reqs.size(), bytes});
        }
        for (Request* r : S.reqs) {
            r->t_done = t;
            r->state = 2;
            lat_ms.push_back(1e3 * (t - r->t_arrive));
            dev_arena.release(r->dev_off, dev_need[r->point]);
            pin_arena.release(r->pin_off, pin_need[r->point]);
        }
        n_completed += S.reqs.size();
        n_batches += 1;
        n_batched += S.reqs.size();
        ++work_gen;
    }
    S.reqs.clear();
    cv_done.notify_all();
    cv_work.notify_one();
}

void Server::fail_batch(Slot& S, const std::string& why) {
    const double t = now_s();
    {
        std::lock_guard<std::mutex> lk(m);
        for (Request* r : S.reqs) {
            r->t_done = t;
            r->state = -1;
            r->error = why;
            dev_arena.release(r->dev_off, dev_need[r->point]);
            pin_arena.release(r->pin_off, pin_need[r->point]);
        }
        n_failed += S.reqs.size();
        if (broken.empty()) broken = why;
        // nothing queued can run on a device in an unknown state
        for (Request* r : queue) {
            r->t_done = t;
            r->state = -1;
            r->error = why;
            ++n_failed;
        }
        queue.clear();
        ++work_gen;
    }
    S.reqs.clear();
    cv_done.notify_all();
    cv_work.notify_one();
}

void Server::write_trace() {
    if (FILE* f = std::fopen((trace_path + ".batches.csv").c_str(), "w")) {
        std::fprintf(f, "take,staged,issued,seen,done,nreq,bytes\n");
        for (const BatchRec& b : trace_b)
            std::fprintf(f, "%.6f,%.6f,%.6f,%.6f,%.6f,%u,%llu\n", b.take, b.staged, b.issued, b.seen, b.done, b.nreq,
                         (unsigned long long)b.bytes);
        std::fclose(f);
    }
    if (FILE* f = std::fopen((trace_path + ".requests.csv").c_str(), "w")) {
        std::fprintf(f, "arrive,take,done,point\n");
        for (const ReqRec& r : trace_r) std::fprintf(f, "%.6f,%.6f,%.6f,%u\n", r.arrive, r.take, r.done, r.point);
        std::fclose(f);
    }
}

// Forms and issues batches: whenever a slot is idle and the queue head fits
// the budgets, takes everything admissible, stages it and enqueues the
// batch's copies and launches; otherwise sleeps until a submission or a
// release changes that.
void Server::schedule_loop() {
    for (;;) {
        Slot* S = nullptr;
        std::vector<Request*> batch;
        {
            std::unique_lock<std::mutex> lk(m);
            for (;;) {
                if (stop && queue.empty()) {
                    sched_exit = true;
                    lk.unlock();
                    cv_flight.notify_all();
                    return;
                }
                if (!queue.empty() && !idle.empty()) {
                    batch = take_locked();
                    if (!batch.empty()) break;
                }
                const uint64_t seen = work_gen;
                cv_work.wait(lk, [&] { return work_gen != seen; });
            }
            S = idle.back();
            idle.pop_back();
        }
        S->reqs = std::move(batch);
        S->t_take = now_s();
        try {
            launch(*S);
        } catch (const std::exception& e) {
            // drain what was already issued before reusing anything
            for (hipStream_t s : {s_in, s_comp[0], s_comp[1]}) (void)hipStreamSynchronize(s);
            fail_batch(*S, e.what());
            std::lock_guard<std::mutex> lk(m);
            idle.push_back(S);
            continue;
        }
        {
            std::lock_guard<std::mutex> lk(m);
            inflight.push_back(S);
        }
        cv_flight.notify_one();
    }
}

// Retires issued batches in order: scatters each marked run of results once
// its copies have finished, then releases the memory and hands the slot back.
void Server::complete_loop() {
    for (;;) {
        Slot* S = nullptr;
        {
            std::unique_lock<std::mutex> lk(m);
            cv_flight.wait(lk, [&] { return !inflight.empty() || sched_exit; });
            if (inflight.empty()) return;
            S = inflight.front();
        }
        // each marked run of copies, scattered as soon as its event completes
        hipError_t q = hipSuccess;
        double t_seen = 0;
        for (size_t k = 0, p0 = 0; k < S->nmark; ++k) {
            while ((q = hipEventQuery(S->mark_ev[k])) == hipErrorNotReady) relax();
            if (q != hipSuccess) break;
            if (k + 1 == S->nmark) t_seen = now_s();
            const Piece* pc = S->out.data() + p0;
            const size_t p1 = S->out_end[k];
            scatter_team->run((int)(p1 - p0), [&](int i) { copy_stream(pc[i].dst, pc[i].src, pc[i].n); });
            p0 = p1;
        }
        if (q == hipSuccess)
            while ((q = hipEventQuery(S->ev_out)) == hipErrorNotReady) relax();
        retire(*S, q, t_seen);
        {
            std::lock_guard<std::mutex> lk(m);
            inflight.pop_front();
            idle.push_back(S);
            ++work_gen;
        }
        cv_work.notify_one();
    }
}

// ---------------------------------------------------------------------------
// Poisson load driver
// ---------------------------------------------------------------------------
namespace {

// Operand uniform in [0, q): W-bit draws per coefficient, redrawn while >= q
// (q > 2^(W-1), so fewer than half are redrawn; with q = 2^W - delta and a
// small delta, almost none).
std::vector<uint32_t> uniform_operand(std::mt19937_64& g, const HostPlan& P) {
    const uint32_t L = P.L, top = P.W - 32 * (L - 1);  // bits of the top limb, 1..32
    const uint32_t mask = top == 32 ? 0xFFFFFFFFu : ((1u << top) - 1u);
    std::vector<uint32_t> v((size_t)P.N * L);
    for (uint32_t n = 0; n < P.N; ++n) {
        uint32_t* x = v.data() + (size_t)n * L;
        for (bool below = false; !below;) {
            for (uint32_t i = 0; i < L; ++i) x[i] = (uint32_t)g();
            x[L - 1] &= mask;
            int i = (int)L - 1;
            while (i >= 0 && x[i] == P.q[i]) --i;
            below = i >= 0 && x[i] < P.q[i];
        }
    }
    return v;
}

// The operands a drive sends and the products it checks them against. The
// products come from the plain path (pm_init / pm_run: the point's own prime
// set and the single-product kernels), not from the batch phase kernels that
// serve the requests.
struct DriveSet {
    std::vector<std::vector<std::vector<uint32_t>>> ops;   // [point][4]: a0, b0, a1, b1
    std::vector<std::vector<std::vector<uint32_t>>> want;  // [point][2]: a0 b0, a1 b1
};

void plain_products(const HostPlan& P, DriveSet& D, size_t i) {
    pm_state* st = nullptr;
    if (pm_init(P.N, P.W, P.L, P.q.data(), P.L, &st) != 0)
        throw std::runtime_error(std::string("expected products: ") + pm_last_error());
    std::unique_ptr<pm_state, void (*)(pm_state*)> hold(st, pm_free);
    for (int k = 0; k < 2; ++k) {
        std::vector<uint32_t> c(D.ops[i][2 * k].size());
        if (pm_run(st, D.ops[i][2 * k].data(), D.ops[i][2 * k + 1].data(), c.data()) != 0)
            throw std::runtime_error(std::string("expected products: ") + pm_last_error());
        D.want[i].push_back(std::move(c));
    }
}

using Buf = std::unique_ptr<std::vector<uint32_t>>;

// Result buffers per point, shared by the driving thread (takes) and the
// checker (gives back).
struct BufPool {
    std::mutex m;
    std::vector<std::vector<Buf>> free;
    Buf take(size_t point, size_t words) {
        {
            std::lock_guard<std::mutex> lk(m);
            if (!free[point].empty()) {
                Buf b = std::move(free[point].back());
                free[point].pop_back();
                return b;
            }
        }
        return std::make_unique<std::vector<uint32_t>>(words);
    }
    void give(size_t point, Buf b) {
        std::lock_guard<std::mutex> lk(m);
        free[point].push_back(std::move(b));
    }
};

// Compares every finished request's result with its expected product on a
// thread of its own, so the driving thread keeps to its arrival schedule,
// then returns the buffer to the pool.
class Checker {
  public:
    Checker(const DriveSet& D, BufPool& pool) : D_(D), pool_(pool), th_([this] { loop(); }) {}
    ~Checker() { finish(); }
    void push(uint32_t point, int pair, Buf b) {
        {
            std::lock_guard<std::mutex> lk(m_);
            q_.push_back({point, pair, std::move(b)});
        }
        cv_.notify_one();
    }
    // Checks what is queued, then stops the thread.
    void finish() {
        {
            std::lock_guard<std::mutex> lk(m_);
            if (done_) return;
            done_ = true;
        }
        cv_.notify_one();
        th_.join();
    }
    uint64_t checked = 0, mismatches = 0;

  private:
    struct Item {
        uint32_t point;
        int pair;
        Buf buf;
    };
    void loop() {
        for (;;) {
            Item it;
            {
                std::unique_lock<std::mutex> lk(m_);
                cv_.wait(lk, [&] { return done_ || !q_.empty(); });
                if (q_.empty()) return;
                it = std::move(q_.front());
                q_.pop_front();
            }
            const std::vector<uint32_t>& w = D_.want[it.point][it.pair];
            if (it.buf->size() != w.size() || std::memcmp(it.buf->data(), w.data(), w.size() * 4) != 0) ++mismatches;
            ++checked;
            pool_.give(it.point, std::move(it.buf));
        }
    }
    const DriveSet& D_;
    BufPool& pool_;
    std::mutex m_;
    std::condition_variable cv_;
    std::deque<Item> q_;
    bool done_ = false;
    std::thread th_;
};

// Open loop (depth 0): Poisson arrivals at `rate`, latency from each
// scheduled arrival. Closed loop (depth > 0): `depth` requests kept in
// flight, a new one submitted as soon as one completes; the achieved rate is
// then the server's capacity on the mix. Every completed result is compared
// with the plain path's product of the same operands.
void drive(Server& S, double rate, uint32_t depth, double seconds, const double* weights, uint64_t seed,
           pm_serve_report* out) {
    if ((depth == 0 && !(rate > 0)) || !(seconds > 0)) throw std::runtime_error("rate and seconds must be positive");
    const size_t np = S.cat.plans.size();
    std::vector<double> w(weights, weights + np);
    double wsum = 0;
    for (double x : w) {
        if (!(x >= 0)) throw std::runtime_error("mix weights must be non-negative");
        wsum += x;
    }
    if (!(wsum > 0)) throw std::runtime_error("mix weights sum to zero");
    std::mt19937_64 g(seed);
    // two operand pairs per point, alternating over the requests; kept on the
    // heap, so that requests a failed drive could not wait for can still read
    // them (see Drain)
    auto D = std::make_shared<DriveSet>();
    D->ops.resize(np);
    D->want.resize(np);
    for (size_t i = 0; i < np; ++i)
        if (w[i] > 0) {
            for (int k = 0; k < 4; ++k) D->ops[i].push_back(uniform_operand(g, S.cat.plans[i]));
            plain_products(S.cat.plans[i], *D, i);
        }
    // Result buffers, allocated and touched before the clock starts (a first
    // touch inside the run would charge page faults to the latencies); the
    // pool grows only if more requests of a point are outstanding at once.
    auto pool = std::make_shared<BufPool>();
    pool->free.resize(np);
    for (size_t i = 0; i < np; ++i)
        if (w[i] > 0)
            for (int k = (int)std::ceil((depth ? depth : 64) * w[i] / wsum) + 2; k > 0; --k)
                pool->free[i].push_back(std::make_unique<std::vector<uint32_t>>(S.cat.io_words(i), 1u));
    std::discrete_distribution<int> pick(w.begin(), w.end());
    std::exponential_distribution<double> gap(depth ? 1.0 : rate);
    Checker check(*D, *pool);

    struct Out {
        uint64_t id;
        uint32_t point;
        int pair;
        Buf buf;
    };
    std::deque<Out> outstanding;
    // On an error (a request still undone after 60 s, a refused submission)
    // the server may still read the operands and write the result buffers of
    // the requests in flight: wait for each of them before anything is freed,
    // and if one still does not finish, keep its memory for the life of the
    // process rather than free it under the server.
    struct Drain {
        Server& S;
        std::deque<Out>& outstanding;
        std::shared_ptr<DriveSet> D;
        ~Drain() {
            bool leak = false;
            for (Out& o : outstanding) {
                int rc;
                try {
                    rc = S.wait(o.id, 60.0, nullptr, nullptr);
                } catch (...) {
                    rc = -1;
                }
                if (rc == 1) {
                    leak = true;
                    (void)o.buf.release();
                }
            }
            if (leak) (void)new std::shared_ptr<DriveSet>(D);
        }
    } drain{S, outstanding, D};
    std::vector<double> lat, qlat, slat;
    std::vector<std::vector<double>> plat(np);
    uint64_t failed = 0, nreq = 0;
    double max_late = 0;
    auto finished = [&](Out& o, int rc, double ms, double qms) {
        if (rc == 0) {
            lat.push_back(ms);
            qlat.push_back(qms);
            slat.push_back(ms - qms);
            plat[o.point].push_back(ms);
            check.push(o.point, o.pair, std::move(o.buf));
        } else {
            ++failed;
            pool->give(o.point, std::move(o.buf));
        }
    };
    // Requests finish in batch order, so reaping stops at the first one not
    // done (or waits for each, at the end).
    auto reap = [&](bool block) {
        while (!outstanding.empty()) {
            Out& o = outstanding.front();
            double ms = 0, qms = 0;
            int rc;
            try {
                rc = S.wait(o.id, block ? 60.0 : 0.0, &ms, &qms);
            } catch (const std::exception&) {
                rc = -1;
            }
            if (rc == 1) {
                if (block) throw std::runtime_error("request did not complete within 60 s");
                break;
            }
            finished(o, rc, ms, qms);
            outstanding.pop_front();
        }
    };

    pm_server_stats st0;
    S.stats(&st0, true);
    rusage ru0{}, rt0{};
    getrusage(RUSAGE_SELF, &ru0);
    getrusage(RUSAGE_THREAD, &rt0);
    auto submit_one = [&](double t_arrive) {
        const uint32_t p = (uint32_t)pick(g);
        const int pair = (int)(nreq % 2);
        Out o{0, p, pair, pool->take(p, S.cat.io_words(p))};
        const auto& op = D->ops[p];
        o.id = S.submit(p, op[2 * pair].data(), op[2 * pair + 1].data(), o.buf->data(), t_arrive);
        outstanding.push_back(std::move(o));
        ++nreq;
    };
    const double t0 = now_s() + 0.02;
    if (depth > 0) {
        while (now_s() < t0) relax();
        while (now_s() - t0 < seconds) {
            while (outstanding.size() < depth) submit_one(now_s());
            // wait for the oldest, then take whatever else has finished
            double ms = 0, qms = 0;
            Out& f = outstanding.front();
            int rc;
            try {
                rc = S.wait(f.id, 60.0, &ms, &qms);
            } catch (const std::exception&) {
                rc = -1;
            }
            if (rc == 1) throw std::runtime_error("request did not complete within 60 s");
            finished(f, rc, ms, qms);
            outstanding.pop_front();
            reap(false);
        }
    } else {
        double t = 0;
        for (;;) {
            t += gap(g);
            if (t >= seconds) break;
            const double target = t0 + t;
            for (double d; (d = target - now_s()) > 0;) {
                if (d > 300e-6) std::this_thread::sleep_for(std::chrono::duration<double>(d - 200e-6));
                else relax();
            }
            max_late = std::max(max_late, now_s() - target);
            submit_one(target);
            if (outstanding.size() > 8) reap(false);
        }
    }
    reap(true);
    const double t_end = now_s();
    check.finish();
    rusage ru1{}, rt1{};
    getrusage(RUSAGE_SELF, &ru1);
    getrusage(RUSAGE_THREAD, &rt1);
    auto secs = [](const timeval& a, const timeval& b) { return (b.tv_sec - a.tv_sec) + 1e-6 * (b.tv_usec - a.tv_usec); };
    pm_server_stats st;
    S.stats(&st, true);
    *out = pm_serve_report{};
    out->offered_rate = depth ? 0.0 : rate;
    out->seconds = seconds;
    out->requests = nreq;
    out->completed = lat.size();
    out->failed = failed;
    out->achieved_rate = lat.size() / std::max(1e-9, t_end - t0);
    if (!lat.empty()) {
        std::sort(lat.begin(), lat.end());
        std::sort(qlat.begin(), qlat.end());
        std::sort(slat.begin(), slat.end());
        out->queue_p50_ms = percentile(qlat, 0.50);
        out->queue_p99_ms = percentile(qlat, 0.99);
        out->service_p50_ms = percentile(slat, 0.50);
        out->service_p99_ms = percentile(slat, 0.99);
        for (size_t i = 0; i < np && i < 8; ++i) {
            if (plat[i].empty()) continue;
            std::sort(plat[i].begin(), plat[i].end());
            out->point_p50_ms[i] = percentile(plat[i], 0.50);
            out->point_p99_ms[i] = percentile(plat[i], 0.99);
        }
        out->p50_ms = percentile(lat, 0.50);
        out->p99_ms = percentile(lat, 0.99);
        out->max_ms = lat.back();
        double s = 0;
        for (double x : lat) s += x;
        out->mean_ms = s / lat.size();
    }
    out->mean_batch = st.mean_batch;
    out->batches = st.batches;
    out->deferred = st.deferred;
    out->device_peak = st.device_peak;
    out->max_late_ms = 1e3 * max_late;
    out->cpu_util = (secs(ru0.ru_utime, ru1.ru_utime) + secs(ru0.ru_stime, ru1.ru_stime)) / std::max(1e-9, t_end - t0);
    out->ctx_invol = (uint64_t)(ru1.ru_nivcsw - ru0.ru_nivcsw);
    out->driver_ctx_invol = (uint64_t)(rt1.ru_nivcsw - rt0.ru_nivcsw);
    out->stage_ms = st.stage_ms;
    out->issue_ms = st.issue_ms;
    out->gpu_ms = st.gpu_ms;
    out->scatter_ms = st.scatter_ms;
    out->sched_busy = st.sched_busy;
    out->completer_busy = st.completer_busy;
    out->verified = check.checked;
    out->mismatches = check.mismatches;
}

}  // namespace
}  // namespace pm

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
struct pm_server {
    std::unique_ptr<pm::Server> s;
};

namespace {
thread_local std::string g_server_error;
template <class F>
int guarded(F&& body) {
    try {
        return body();
    } catch (const std::exception& e) {
        g_server_error = e.what();
    } catch (...) {
        g_server_error = "unknown error";
    }
    return -1;
}
}  // namespace

extern "C" {

const char* pm_server_last_error(void) { return g_server_error.c_str(); }

int pm_server_create(uint32_t npoints, const uint32_t* N, const uint32_t* W, const uint32_t* q_concat,
                     const pm_server_config* cfg, pm_server** out) {
    return guarded([&] {
        if (!out || !N || !W || !q_concat) throw std::runtime_error("null argument");
        *out = nullptr;
        auto h = std::make_unique<pm_server>();
        h->s = std::make_unique<pm::Server>();
        h->s->create(npoints, N, W, q_concat, cfg);
        *out = h.release();
        return 0;
    });
}

void pm_server_free(pm_server* s) { delete s; }

int pm_server_submit(pm_server* s, uint32_t point, const uint32_t* a, const uint32_t* b, uint32_t* c,
                     uint64_t* ticket) {
    return guarded([&] {
        if (!s || !ticket) throw std::runtime_error("null argument");
        *ticket = s->s->submit(point, a, b, c, 0.0);
        return 0;
    });
}

int pm_server_wait(pm_server* s, uint64_t ticket, double timeout_s, double* latency_ms) {
    return guarded([&] {
        if (!s) throw std::runtime_error("null argument");
        return s->s->wait(ticket, timeout_s, latency_ms);
    });
}

int pm_server_stats_get(pm_server* s, int reset, pm_server_stats* out) {
    return guarded([&] {
        if (!s || !out) throw std::runtime_error("null argument");
        s->s->stats(out, reset != 0);
        return 0;
    });
}

int pm_server_drive(pm_server* s, double rate, double seconds, const double* weights, uint64_t seed,
                    pm_serve_report* out) {
    return guarded([&] {
        if (!s || !weights || !out) throw std::runtime_error("null argument");
        pm::drive(*s->s, rate, 0, seconds, weights, seed, out);
        return 0;
    });
}

int pm_server_saturate(pm_server* s, uint32_t depth, double seconds, const double* weights, uint64_t seed,
                       pm_serve_report* out) {
    return guarded([&] {
        if (!s || !weights || !out) throw std::runtime_error("null argument");
        if (depth == 0 || depth > 4096) throw std::runtime_error("depth must be in 1..4096");
        pm::drive(*s->s, 0.0, depth, seconds, weights, seed, out);
        return 0;
    });
}

}  // extern "C"
