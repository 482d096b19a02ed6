// Host-buffer multiply through PCIe: the drop-in path behind fherma_run.
//
// The call moves 2 N L 4 bytes in and N L 4 bytes out, so everything else is
// arranged to hide behind those transfers:
//   * a is cut into growing contiguous coefficient chunks (1, 3, then 16 MiB)
//     and b goes as one copy: a host worker team stages a chunk from the
//     caller's pageable buffer into pinned memory (non-temporal stores) at
//     about 2-4x the link rate while the copy engine moves the one before;
//     the engines run fastest on large copies, and all of b is staged during
//     a's last copy. The RNS launches follow once every copy is queued (each
//     waits on the device for its own chunk), so no host-side wait sits
//     between staging one chunk and the next;
//   * a's forward column pass and forward rows run once a is in (while b
//     crosses the link); once b is resident the other passes run, the row
//     pass transforming only b's rows before the pointwise product;
//   * the result leaves in chunks over two copy streams: CRT of the first
//     chunk alone, then the rest, device-to-host copies alternating between
//     the streams, and the host team scatters each chunk from pinned staging
//     into the caller's buffer while the next is copied.
// Chunks are contiguous coefficient ranges (the input layout is [N][L]), so
// staging needs no gather and the column NTT pass starts after the last input
// chunk. Column windows are the other option (PM_IN_COLS, PM_OUT_COLS; off by
// default, measured no faster): b staged one window of columns at a time (each
// row's run of CW coefficients), each window's RNS and forward column pass
// running as it lands; the result leaving first window first, the other
// columns' inverse column pass under its copy. Points whose plans are single
// chunks (a at most 1 MiB) skip the copy engines: the kernels read and write
// the pinned staging buffers directly.
//
// PM_PROFILE=1 prints one phase line per call on stderr (the chunked pipeline).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <mutex>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

#if defined(__x86_64__)
#include <immintrin.h>
#endif
#if defined(__linux__)
#include <sched.h>
#endif

#include "pm_host.h"
#include "pm_state.h"

namespace pm {

namespace {

void hip_check(hipError_t e, const char* what) {
    if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

#if defined(__x86_64__)
__attribute__((target("avx2"))) void copy_stream_avx2(unsigned char* d, const unsigned char* s, size_t n) {
    while (n && ((uintptr_t)d & 31)) {
        *d++ = *s++;
        --n;
    }
    size_t i = 0;
    for (; i + 128 <= n; i += 128) {
        const __m256i x0 = _mm256_loadu_si256((const __m256i*)(s + i));
        const __m256i x1 = _mm256_loadu_si256((const __m256i*)(s + i + 32));
        const __m256i x2 = _mm256_loadu_si256((const __m256i*)(s + i + 64));
        const __m256i x3 = _mm256_loadu_si256((const __m256i*)(s + i + 96));
        _mm256_stream_si256((__m256i*)(d + i), x0);
        _mm256_stream_si256((__m256i*)(d + i + 32), x1);
        _mm256_stream_si256((__m256i*)(d + i + 64), x2);
        _mm256_stream_si256((__m256i*)(d + i + 96), x3);
    }
    if (i < n) std::memcpy(d + i, s + i, n - i);
    _mm_sfence();
}
#endif

}  // namespace

// The destination is streamed past the caches (the staging buffers are written
// once and read by DMA; the result is read later by the caller), so a copy
// costs one read and one write of memory rather than three.
void copy_stream(void* dst, const void* src, size_t n) {
#if defined(__x86_64__)
    static const bool avx2 = __builtin_cpu_supports("avx2");
    if (avx2) {
        copy_stream_avx2(static_cast<unsigned char*>(dst), static_cast<const unsigned char*>(src), n);
        return;
    }
#endif
    std::memcpy(dst, src, n);
}

int host_cpus() {
#if defined(__linux__)
    cpu_set_t set;
    if (sched_getaffinity(0, sizeof(set), &set) == 0) return CPU_COUNT(&set);
#endif
    return (int)std::max(1u, std::thread::hardware_concurrency());
}

struct Pipeline {
    Team* team = nullptr;
    hipStream_t s_in = nullptr, s_out = nullptr;
    hipStream_t s_in2 = nullptr;   // second H2D stream: input chunks alternate
    hipStream_t s_out2 = nullptr;  // optional second D2H stream (PM_D2H_STREAMS=2)
    uint32_t* h_in[2] = {nullptr, nullptr};  // pinned staging, [N][L]
    uint32_t* h_out = nullptr;
    struct Chunk {
        uint32_t first, count;
    };
    std::vector<Chunk> cin[2], cout;  // input plans of a and b, output plan
    // Column windows of 2^lcw columns (0: contiguous coefficient chunks): b
    // staged window by window (cin[1] then holds staging positions, one window
    // per chunk), and the result leaving first window first (cout likewise).
    int lcw_in = 0, lcw_out = 0;
    LaunchCfg cfg_in, cfg_out;  // the RNS and CRT tiles sized to fit a window
    std::vector<hipEvent_t> ev_in, ev_crt, ev_out;
    bool profile = false;
    // PM_PROFILE timeline: start of input, last H2D done, last RNS done,
    // NTT passes done, first CRT done, first D2H done, last D2H done
    hipEvent_t tl[3] = {nullptr, nullptr, nullptr};
    hipEvent_t tl_more[4] = {nullptr, nullptr, nullptr, nullptr};
};

namespace {

// Contiguous chunks: the body follows `sizes` (bytes; the last entry repeats),
// then the `tail` sizes close the range; boundaries on multiples of `align`
// coefficients. Small chunks at the ends start the copy engine early and
// shorten the drain; large ones in the middle keep it near its large-transfer
// rate (about 56 GB/s from 4 MB up, 40 GB/s at 1 MB).
std::vector<Pipeline::Chunk> plan_chunks(uint32_t N, size_t bytes_per_coeff, const std::vector<size_t>& sizes,
                                         const std::vector<size_t>& tail_sizes, uint32_t align) {
    auto coeffs = [&](size_t b) {
        const uint32_t c = (uint32_t)std::max<size_t>(align, b / bytes_per_coeff);
        return (c + align - 1) / align * align;
    };
    std::vector<uint32_t> tail;
    uint32_t tail_total = 0;
    for (size_t b : tail_sizes) {
        tail.push_back(coeffs(b));
        tail_total += tail.back();
    }
    while (!tail.empty() && tail_total >= N) {
        tail_total -= tail.front();
        tail.erase(tail.begin());
    }
    const uint32_t body_end = N - tail_total;
    std::vector<Pipeline::Chunk> out;
    auto it = sizes.begin();
    for (uint32_t f = 0; f < body_end;) {
        const uint32_t take = std::min(coeffs(*it), body_end - f);
        out.push_back({f, take});
        f += take;
        if (std::next(it) != sizes.end()) ++it;
    }
    uint32_t f = body_end;
    for (uint32_t c : tail) {
        out.push_back({f, c});
        f += c;
    }
    return out;
}

// "0.5,2,8" (MiB) from the environment ("0": none), or the default.
std::vector<size_t> sizes_from_env(const char* name, std::vector<size_t> dflt) {
    const char* e = std::getenv(name);
    if (!e || !*e) return dflt;
    if (std::strcmp(e, "0") == 0) return {};
    std::vector<size_t> out;
    for (const char* p = e; *p;) {
        char* end = nullptr;
        const double mb = std::strtod(p, &end);
        if (end == p) break;
        if (mb > 0) out.push_back((size_t)(mb * (1 << 20)));
        p = *end == ',' ? end + 1 : end;
    }
    return out.empty() ? dflt : out;
}

void wait_event(hipEvent_t e) {
    for (;;) {
        const hipError_t r = hipEventQuery(e);
        if (r == hipSuccess) return;
        if (r != hipErrorNotReady) hip_check(r, "event");
        relax();
    }
}

}  // namespace

void State::pipeline_init() {
    Pipeline* P = new Pipeline;
    pipe = P;
    P->profile = std::getenv("PM_PROFILE") != nullptr;
    int workers = std::min(15, std::max(1, host_cpus() - 1));
    if (const char* e = std::getenv("PM_HOST_THREADS")) workers = std::max(0, std::atoi(e) - 1);
    P->team = new Team(workers);
    hip_check(hipStreamCreateWithFlags(&P->s_in, hipStreamNonBlocking), "stream");
    // Input chunks alternate over two H2D streams, so two copy engines share
    // the link (measured 3-5% faster per call at N=65536/W=1747 than one);
    // PM_H2D_STREAMS=1 keeps one.
    const char* hs = std::getenv("PM_H2D_STREAMS");
    if (!hs || std::atoi(hs) != 1) hip_check(hipStreamCreateWithFlags(&P->s_in2, hipStreamNonBlocking), "stream");
    hip_check(hipStreamCreateWithFlags(&P->s_out, hipStreamNonBlocking), "stream");
    // Output chunks alternate over two D2H streams (2-3% faster per call at
    // N=65536/W=1747 than one); PM_D2H_STREAMS=1 keeps one.
    const char* ds = std::getenv("PM_D2H_STREAMS");
    if (!ds || std::atoi(ds) != 1) hip_check(hipStreamCreateWithFlags(&P->s_out2, hipStreamNonBlocking), "stream");
    const size_t bytes = coeff_words * 4;
    for (int i = 0; i < 2; ++i) hip_check(hipHostMalloc((void**)&P->h_in[i], bytes, hipHostMallocDefault), "pinned");
    hip_check(hipHostMalloc((void**)&P->h_out, bytes, hipHostMallocDefault), "pinned");
    const size_t bpc = (size_t)host.L * 4;
    const size_t MiB = 1u << 20;
    // Defaults from sweeps at N=65536/W=1747. The copy engine sets the pace:
    // 1 MiB copies move at 40 GB/s, 8 MiB at 54, 64 MiB at 57. a starts small
    // (the engine starts early) and grows; b, staged in full while a's last
    // copy runs, goes as one copy. Against {1, 4, 8} MiB plus a 1 MiB tail
    // for both operands: 1.031-1.036 ms per call against 1.078-1.093 (with
    // the two D2H streams). On the output side, 5 MiB chunks after the first
    // let the scatter into the caller's buffer run under the next chunk's
    // copy: 1.08-1.11 ms per call against 1.13-1.18 ms with one 14 MiB chunk.
    P->cin[0] = plan_chunks(host.N, bpc, sizes_from_env("PM_IN_CHUNKS", {MiB, 3 * MiB, 16 * MiB}),
                            sizes_from_env("PM_IN_TAIL", {}), 128);
    P->cin[1] = plan_chunks(host.N, bpc, sizes_from_env("PM_IN_CHUNKS_B", {64 * MiB}),
                            sizes_from_env("PM_IN_TAIL_B", {}), 128);
    P->cout = plan_chunks(host.N, bpc, sizes_from_env("PM_OUT_CHUNKS", {MiB, 5 * MiB}),
                          sizes_from_env("PM_OUT_TAIL", {MiB / 2}), 128);
    // Column windows, for chunked two-pass plans only (PM_IN_COLS and
    // PM_OUT_COLS: window width in columns, 0 for contiguous chunks).
    const bool chunked = !(P->cin[0].size() == 1 && P->cin[1].size() == 1 && P->cout.size() == 1);
    auto window_log = [&](const char* name, int dflt) {
        const char* e = std::getenv(name);
        const int cols = e ? std::atoi(e) : dflt;
        if (!chunked || dev.npass != 2 || cols < 32 || (cols & (cols - 1)) || cols % cfg.col_tc ||
            cols >= (1 << host.w2))
            return 0;
        int l = 0;
        while ((1 << l) < cols) ++l;
        return l;
    };
    // Both are off by default. At 65536/1747, per call against contiguous
    // chunks (alternating rounds of 100 calls): 64-column output windows
    // 0.984-0.999 ms against 0.990-1.031 on one box, 1.000-1.009 against
    // 1.000-1.008 on another; b in 128-column windows as well 0.987-1.009 ms;
    // b in 64-column windows 1.005-1.008 against 0.987-0.997. The windows
    // take the inverse column pass off the first output chunk's path, but the
    // row pass before it cannot be windowed, the first window is a larger first
    // copy, and b's windows are more and smaller copies.
    const uint32_t rows = 1u << host.w1;
    P->lcw_in = window_log("PM_IN_COLS", 0);
    if (P->lcw_in) {
        P->cfg_in = cfg;
        if (P->cfg_in.rns_tn > (1 << P->lcw_in)) {
            P->cfg_in.rns_tn = 1 << P->lcw_in;
            P->cfg_in.rns_lds = rns_lds_bytes(P->cfg_in.rns_tn, (int)host.L, dev.rns_rt);
        }
        const uint32_t win = rows << P->lcw_in;
        P->cin[1].clear();
        for (uint32_t f = 0; f < host.N; f += win) P->cin[1].push_back({f, win});
    }
    P->lcw_out = cfg.crt_alias ? window_log("PM_OUT_COLS", 0) : 0;
    if (P->lcw_out) {
        P->cfg_out = cfg;
        if (P->cfg_out.crt_tn > (1 << P->lcw_out)) {
            P->cfg_out.crt_tn = 1 << P->lcw_out;
            P->cfg_out.crt_lds = crt_alias_lds_bytes(P->cfg_out.crt_tn, cfg.crt_ys, (int)host.L, dev.crt_ks);
        }
        // the first window alone, then the usual chunks over the rest, each a
        // whole number of window rows (runs of 2^lcw coefficients)
        const uint32_t win = rows << P->lcw_out;
        // (4 MiB chunks and a 1 + 0.5 MiB tail, so the last scatter is short:
        // 5 MiB chunks and a 0.5 MiB tail left it exposed, 1.003-1.019 ms)
        std::vector<size_t> sz = sizes_from_env("PM_OUT_CHUNKS", {MiB, 4 * MiB});
        if (sz.size() > 1) sz.erase(sz.begin());
        P->cout = {{0, win}};
        for (const auto ch : plan_chunks(host.N - win, bpc, sz, sizes_from_env("PM_OUT_TAIL", {MiB, MiB / 2}),
                                         1u << P->lcw_out))
            P->cout.push_back({win + ch.first, ch.count});
    }
    if (P->profile) {
        for (int i = 0; i < 3; ++i) hip_check(hipEventCreate(&P->tl[i]), "event");
        for (auto& e : P->tl_more) hip_check(hipEventCreate(&e), "event");
    }
    auto mk = [](std::vector<hipEvent_t>& v, size_t n) {
        v.resize(n);
        for (auto& e : v) hip_check(hipEventCreateWithFlags(&e, hipEventDisableTiming), "event");
    };
    mk(P->ev_in, P->cin[0].size() + P->cin[1].size());
    mk(P->ev_crt, P->cout.size());
    mk(P->ev_out, P->cout.size());
    if (P->profile)
        std::fprintf(stderr,
                     "pm: N=%u W=%u L=%u K=%d, w1=%d w2=%d, chunks in a %zu / b %zu / out %zu, column windows in %d / "
                     "out %d, %d host threads\n",
                     host.N, host.W, host.L, host.K, host.w1, host.w2, P->cin[0].size(), P->cin[1].size(), P->cout.size(),
                     P->lcw_in ? 1 << P->lcw_in : 0, P->lcw_out ? 1 << P->lcw_out : 0, P->team->size());
}

void State::pipeline_free() {
    Pipeline* P = pipe;
    if (!P) return;
    if (P->s_in) (void)hipStreamSynchronize(P->s_in);
    if (P->s_in2) (void)hipStreamSynchronize(P->s_in2);
    if (P->s_out) (void)hipStreamSynchronize(P->s_out);
    if (P->s_out2) (void)hipStreamSynchronize(P->s_out2);
    for (auto e : P->ev_in) (void)hipEventDestroy(e);
    for (auto e : P->ev_crt) (void)hipEventDestroy(e);
    for (auto e : P->ev_out) (void)hipEventDestroy(e);
    for (auto e : P->tl)
        if (e) (void)hipEventDestroy(e);
    for (auto e : P->tl_more)
        if (e) (void)hipEventDestroy(e);
    for (int i = 0; i < 2; ++i)
        if (P->h_in[i]) (void)hipHostFree(P->h_in[i]);
    if (P->h_out) (void)hipHostFree(P->h_out);
    if (P->s_in) (void)hipStreamDestroy(P->s_in);
    if (P->s_in2) (void)hipStreamDestroy(P->s_in2);
    if (P->s_out) (void)hipStreamDestroy(P->s_out);
    if (P->s_out2) (void)hipStreamDestroy(P->s_out2);
    delete P->team;
    delete P;
    pipe = nullptr;
}

void State::run_host(const uint32_t* a, const uint32_t* b, uint32_t* c) {
    Pipeline& P = *pipe;
    using clock = std::chrono::steady_clock;
    const auto t0 = clock::now();
    const size_t L = host.L;
    const int T = P.team->size();
    const uint32_t* src[2] = {a, b};
    struct Hot {
        Team* t;
        explicit Hot(Team* x) : t(x) { t->hot(true); }
        ~Hot() { t->hot(false); }
    } hot(P.team);

    // Each copy job is split into T pieces of whole coefficients.
    auto copy_job = [&](uint32_t* dst, const uint32_t* from, size_t words) {
        const size_t piece = (words / L + T - 1) / T * L;
        P.team->run(T, [&](int i) {
            const size_t lo = std::min(words, (size_t)i * piece), hi = std::min(words, lo + piece);
            if (hi > lo) copy_stream(dst + lo, from + lo, (hi - lo) * 4);
        });
    };

    // Staging positions [f, f + cnt) of a column-window order (runs of 2^lcw
    // coefficients; run u is row u mod 2^w1 of window u / 2^w1) to or from
    // their places in an [N][L] array, the runs split over the team.
    auto move_runs = [&](uint32_t* to, const uint32_t* from, uint32_t f, uint32_t cnt, int lcw, bool gather) {
        const uint32_t u0 = f >> lcw, nu = cnt >> lcw, rmask = (1u << host.w1) - 1u;
        const size_t run = L << lcw;
        const uint32_t per = (nu + T - 1) / T;
        P.team->run(T, [&](int t) {
            for (uint32_t u = u0 + (uint32_t)t * per, e = std::min(u0 + nu, u + per); u < e; ++u) {
                const size_t pos = (size_t)u * run;
                const size_t n = (((size_t)(u & rmask) << host.w2) + ((size_t)(u >> host.w1) << lcw)) * L;
                if (gather)
                    copy_stream(to + pos, from + n, run * 4);
                else
                    copy_stream(to + n, from + pos, run * 4);
            }
        });
    };

    // ---- small points: one chunk each way, nothing to overlap ----
    // Every plan is a single chunk (a at most 1 MiB), so the pipeline's copy
    // streams and cross-stream events would only add their latency. Both
    // operands are staged in one team job; the RNS reads them from the pinned
    // staging buffers and the CRT writes the result into pinned memory, so no
    // copy engine is involved, and the five kernels run in order on the
    // compute stream. Per call at N=4096/W=256: 0.062 ms, against 0.086-0.089
    // with the copies on the compute stream and 0.107-0.119 through the
    // chunked pipeline.
    if (P.cin[0].size() == 1 && P.cin[1].size() == 1 && P.cout.size() == 1 && !P.profile) {
        const size_t words = (size_t)host.N * L;
        const size_t piece = (2 * host.N + T - 1) / T * L;  // whole coefficients over both operands
        P.team->run(T, [&](int k) {
            for (size_t lo = (size_t)k * piece, hi = std::min(2 * words, lo + piece); lo < hi;) {
                const int op = lo >= words;
                const size_t o = lo - op * words, n = std::min(hi, (op + 1) * words) - lo;
                copy_stream(P.h_in[op] + o, src[op] + o, n * 4);
                lo += n;
            }
        });
        acquire_R(stream);
        hip_check(enqueue_multiply(P.h_in[0], P.h_in[1], P.h_out, stream), "multiply");
        release_R(stream);
        hip_check(hipEventRecord(P.ev_out[0], stream), "record");
        wait_event(P.ev_out[0]);
        copy_job(c, P.h_out, words);
        return;
    }

    // ---- input: stage and H2D per chunk, every copy queued first ----
    // A cross-stream hipStreamWaitEvent costs 11-20 us of host time, about a
    // 1 MiB copy's worth of link time, and it would sit between staging one
    // chunk and the next. So the loop queues only the copies (and a cheap
    // event after each); the waits and the RNS launches follow once the last
    // chunk is staged, while b is still crossing the link.
    acquire_R(stream);  // d_R: after any caller-stream or sharded call still in flight
    if (P.profile) hip_check(hipEventRecord(P.tl[0], P.s_in), "record");
    int i = -1;  // running input chunk index over both operands
    for (int op = 0; op < 2; ++op) {
        for (const auto ch : P.cin[op]) {
            ++i;
            const size_t off = (size_t)ch.first * L, words = (size_t)ch.count * L;
            if (op == 1 && P.lcw_in)
                move_runs(P.h_in[1], b, ch.first, ch.count, P.lcw_in, true);
            else
                copy_job(P.h_in[op] + off, src[op] + off, words);
            hipStream_t sc = (P.s_in2 && (i & 1)) ? P.s_in2 : P.s_in;
            hip_check(hipMemcpyAsync(d_in[op] + off, P.h_in[op] + off, words * 4, hipMemcpyHostToDevice, sc), "H2D");
            hip_check(hipEventRecord(P.ev_in[i], sc), "record");
        }
    }
    const auto t_in = clock::now();
    // ---- RNS per chunk, each after its copy; a's column pass before b's RNS ----
    i = -1;
    for (int op = 0; op < 2; ++op) {
        for (const auto ch : P.cin[op]) {
            ++i;
            hip_check(hipStreamWaitEvent(stream, P.ev_in[i], 0), "wait");
            if (op == 1 && P.lcw_in) {  // b's window: its RNS, then its columns' forward pass
                hip_check(launch_rns(dev, P.cfg_in, d_in[1], nullptr, d_R[1], nullptr, 1, stream, ch.first, ch.count,
                                     P.lcw_in),
                          "rns");
                hip_check(launch_col(dev, cfg, d_R[1], nullptr, 1, false, stream, ch.first >> host.w1,
                                     ch.count >> host.w1),
                          "col b");
                continue;
            }
            hip_check(launch_rns(dev, cfg, d_in[op], nullptr, d_R[op], nullptr, 1, stream, ch.first, ch.count), "rns");
        }
        // a's column pass and forward rows run while b is still crossing the link
        if (op == 0) {
            hip_check(launch_col(dev, cfg, d_R[0], nullptr, 1, false, stream), "col a");
            hip_check(launch_row(dev, cfg, d_R[0], nullptr, ROW_FWD, stream), "rows a");
        }
    }
    if (P.profile) {
        if (P.s_in2) hip_check(hipStreamWaitEvent(P.s_in, P.ev_in[i], 0), "wait");
        hip_check(hipEventRecord(P.tl[1], P.s_in), "record");
        hip_check(hipEventRecord(P.tl[2], stream), "record");
    }

    // ---- the rest of the NTT passes ----
    if (!P.lcw_in) hip_check(launch_col(dev, cfg, d_R[1], nullptr, 1, false, stream), "col b");
    hip_check(launch_row(dev, cfg, d_R[1], d_R[0], ROW_FUSED_PRE, stream), "row");
    const int nout = (int)P.cout.size();
    const uint32_t first_n = P.cout[0].count;
    const int lo = P.lcw_out;
    // with output windows, the first window's columns only (the rest follow
    // its CRT, under its copy)
    hip_check(lo ? launch_col(dev, cfg, d_R[1], nullptr, 1, true, stream, 0, 1u << lo)
                 : launch_col(dev, cfg, d_R[1], nullptr, 1, true, stream),
              "col inv");
    if (P.profile) hip_check(hipEventRecord(P.tl_more[0], stream), "record");

    // ---- output: CRT, then D2H per chunk ----
    // The CRT of the first chunk runs alone so its copy starts early; the rest
    // runs as one full-width launch (small CRT launches leave most CUs idle and
    // would pace the copies). Each D2H waits for the launch that covers it.
    const LaunchCfg& co = lo ? P.cfg_out : cfg;
    hip_check(launch_crt(dev, co, d_R[1], d_out, dev.G, stream, 0, first_n, d_slots, lo), "crt");
    hip_check(hipEventRecord(P.ev_crt[0], stream), "record");
    if (P.profile) hip_check(hipEventRecord(P.tl_more[1], stream), "record");
    if (nout > 1) {
        if (lo)
            hip_check(launch_col(dev, cfg, d_R[1], nullptr, 1, true, stream, 1u << lo, (1u << host.w2) - (1u << lo)),
                      "col inv");
        hip_check(launch_crt(dev, co, d_R[1], d_out, dev.G, stream, first_n, host.N - first_n, d_slots, lo), "crt");
        hip_check(hipEventRecord(P.ev_crt[1], stream), "record");
    }
    release_R(stream);
    // Chunk 0 waits for the first CRT launch; each D2H stream's first chunk
    // after it waits for the second (which follows the first on the compute
    // stream), and the stream's later chunks follow that copy.
    bool waited_rest[2] = {false, false};
    for (int k = 0; k < nout; ++k) {
        const auto ch = P.cout[k];
        const size_t off = (size_t)ch.first * L, words = (size_t)ch.count * L;
        const int si = (P.s_out2 && (k & 1)) ? 1 : 0;
        hipStream_t so = si ? P.s_out2 : P.s_out;
        if (k == 0) {
            hip_check(hipStreamWaitEvent(so, P.ev_crt[0], 0), "wait");
        } else if (!waited_rest[si]) {
            hip_check(hipStreamWaitEvent(so, P.ev_crt[1], 0), "wait");
            waited_rest[si] = true;
        }
        hip_check(hipMemcpyAsync(P.h_out + off, d_out + off, words * 4, hipMemcpyDeviceToHost, so), "D2H");
        hip_check(hipEventRecord(P.ev_out[k], so), "record");
        if (P.profile && k == 0) hip_check(hipEventRecord(P.tl_more[2], so), "record");
    }
    if (P.profile) {
        if (P.s_out2) hip_check(hipStreamWaitEvent(P.s_out, P.ev_out[nout - 1], 0), "wait");
        hip_check(hipEventRecord(P.tl_more[3], P.s_out), "record");
    }
    clock::time_point t_first{};
    for (int k = 0; k < nout; ++k) {
        wait_event(P.ev_out[k]);
        if (k == 0) t_first = clock::now();
        const auto ch = P.cout[k];
        const size_t off = (size_t)ch.first * L;
        if (lo)
            move_runs(c, P.h_out, ch.first, ch.count, lo, false);
        else
            copy_job(c + off, P.h_out + off, (size_t)ch.count * L);
    }
    if (P.profile) {
        auto us = [&](clock::time_point t) { return std::chrono::duration<double, std::micro>(t - t0).count(); };
        std::fprintf(stderr, "pm: inputs staged+submitted %.1f us | first out chunk %.1f | done %.1f\n", us(t_in),
                     us(t_first), us(clock::now()));
        hip_check(hipEventSynchronize(P.tl_more[3]), "sync");  // recorded after the last D2H's own event
        auto el = [&](hipEvent_t e) {
            float ms = 0;
            hip_check(hipEventElapsedTime(&ms, P.tl[0], e), "elapsed");
            return 1000.0 * ms;
        };
        std::fprintf(stderr,
                     "pm: gpu: last H2D %.1f us | last RNS %.1f | NTT passes %.1f | first CRT %.1f | first D2H %.1f | "
                     "last D2H %.1f\n",
                     el(P.tl[1]), el(P.tl[2]), el(P.tl_more[0]), el(P.tl_more[1]), el(P.tl_more[2]), el(P.tl_more[3]));
    }
}

}  // namespace pm

// ---------------------------------------------------------------------------
// Transfer probe (the evidence behind the chunk sizes and thread counts):
// PCIe copy rates from pinned and pageable memory and the host staging rate.
// ---------------------------------------------------------------------------
namespace pm {

void probe_transfers(size_t bytes, int threads, int reps, double* out) {
    using clock = std::chrono::steady_clock;
    auto gbps = [&](double s) { return (double)bytes / s / 1e9; };
    void *d = nullptr, *hp = nullptr, *hq = nullptr;
    hip_check(hipMalloc(&d, bytes), "probe alloc");
    hip_check(hipHostMalloc(&hp, bytes, hipHostMallocDefault), "probe pinned");
    hip_check(hipHostMalloc(&hq, bytes, hipHostMallocDefault), "probe pinned");
    std::vector<unsigned char> pg(bytes, 1), pg2(bytes, 0);
    std::memset(hp, 2, bytes);
    hipStream_t s, s2;
    hip_check(hipStreamCreateWithFlags(&s, hipStreamNonBlocking), "probe stream");
    hip_check(hipStreamCreateWithFlags(&s2, hipStreamNonBlocking), "probe stream");
    auto time_it = [&](const std::function<void()>& f) {
        f();
        double best = 1e30;
        for (int r = 0; r < reps; ++r) {
            const auto t = clock::now();
            f();
            best = std::min(best, std::chrono::duration<double>(clock::now() - t).count());
        }
        return best;
    };
    // 0, 1: pinned H2D and D2H through the copy engine
    out[0] = gbps(time_it([&] {
        hip_check(hipMemcpyAsync(d, hp, bytes, hipMemcpyHostToDevice, s), "h2d");
        hip_check(hipStreamSynchronize(s), "sync");
    }));
    out[1] = gbps(time_it([&] {
        hip_check(hipMemcpyAsync(hq, d, bytes, hipMemcpyDeviceToHost, s), "d2h");
        hip_check(hipStreamSynchronize(s), "sync");
    }));
    // 2, 3: pageable H2D and D2H (the runtime's own staging)
    out[2] = gbps(time_it([&] { hip_check(hipMemcpy(d, pg.data(), bytes, hipMemcpyHostToDevice), "h2d pageable"); }));
    out[3] = gbps(time_it([&] { hip_check(hipMemcpy(pg2.data(), d, bytes, hipMemcpyDeviceToHost), "d2h pageable"); }));
    // 4, 5: host staging copy pageable -> pinned, one thread and `threads` threads
    out[4] = gbps(time_it([&] { copy_stream(hp, pg.data(), bytes); }));
    {
        Team team(std::max(0, threads - 1));
        team.hot(true);
        const int T = team.size();
        const size_t piece = (bytes / 64 + T - 1) / T * 64;
        out[5] = gbps(time_it([&] {
            team.run(T, [&](int i) {
                const size_t lo = std::min(bytes, (size_t)i * piece), hi = std::min(bytes, lo + piece);
                if (hi > lo) copy_stream((char*)hp + lo, pg.data() + lo, hi - lo);
            });
        }));
        // 6: both copy directions at once (pinned), full duplex
        out[6] = gbps(time_it([&] {
            hip_check(hipMemcpyAsync(d, hp, bytes / 2, hipMemcpyHostToDevice, s), "h2d");
            hip_check(hipMemcpyAsync(hq, (char*)d + bytes / 2, bytes / 2, hipMemcpyDeviceToHost, s2), "d2h");
            hip_check(hipStreamSynchronize(s), "sync");
            hip_check(hipStreamSynchronize(s2), "sync");
        }));
        team.hot(false);
    }
    (void)hipStreamDestroy(s);
    (void)hipStreamDestroy(s2);
    (void)hipFree(d);
    (void)hipHostFree(hp);
    (void)hipHostFree(hq);
}

}  // namespace pm
