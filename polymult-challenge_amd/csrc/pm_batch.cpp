// Batched form: many products of uneven N and W in one graph launch.
//
// Every product draws its K primes from one shared set (p = 1 mod 2^17), so a
// single twiddle table serves all of them; the RNS and CRT tables stay per
// product. The host builds one tile list per phase (RNS tiles, forward column
// tiles, fused row tiles, inverse column tiles, CRT tiles) over every product
// of the batch, and the batch runs as five launches, one per phase, each over
// all products' tiles, so even a batch of small products fills the device.
// Tiles reach the blocks either through the hardware dispatcher (one block per
// tile) or through a persistent grid that claims them from an atomic counter.
// The launches are captured once in a hipGraph and replayed per batch. An
// autotuner times the graph over the column-tile width, the 4-step split of
// odd logN and the tile hand-out mode.
#include "pm_batch.h"

#include <algorithm>
#include <cstring>
#include <memory>

#include "../../include/polymult.h"

namespace pm {
namespace {

int line_threads(int nb) { return nb >= 4 ? 1 << (nb - 4) : 1; }

template <class T>
void ensure(T*& p, size_t& cap, size_t n, const char* what) {
    if (n <= cap) return;
    if (p) hip_check(hipFree(p), "hipFree");
    p = nullptr;
    cap = std::max<size_t>(n, cap * 2);
    hip_check(hipMalloc(&p, cap * sizeof(T)), what);
}

template <class T>
void ensure_pinned(T*& p, size_t& cap, size_t n) {
    if (n <= cap) return;
    if (p) hip_check(hipHostFree(p), "hipHostFree");
    p = nullptr;
    cap = std::max<size_t>(n, cap * 2);
    hip_check(hipHostMalloc(reinterpret_cast<void**>(&p), cap * sizeof(T), hipHostMallocDefault), "hipHostMalloc");
}

}  // namespace

// ---------------------------------------------------------------------------
// Catalog
// ---------------------------------------------------------------------------
void Catalog::build(uint32_t count, const uint32_t* N, const uint32_t* W, const uint32_t* q_concat) {
    if (count == 0) throw std::runtime_error("empty batch");
    // Plan every point on a generous shared prime set, then keep the K the
    // largest point needs.
    shared_primes(S, 300);
    plans.resize(count);
    size_t qoff = 0;
    int kmax = 0;
    for (uint32_t i = 0; i < count; ++i) {
        const uint32_t L = (W[i] + 31) / 32;
        plan_point_shared(plans[i], N[i], W[i], L, q_concat + qoff, L, S);
        qoff += L;
        build_tables(plans[i], false);
        kmax = std::max(kmax, plans[i].K);
    }
    S.p.resize(kmax);
    S.pinv.resize(kmax);
    S.psi.resize(kmax);

    std::vector<uint32_t> tw;
    build_shared_twiddles(S, kmax, tw);
    const size_t pbytes = (size_t)kmax * 4, twoff = (2 * pbytes + 255) & ~size_t(255);
    hip_check(hipMalloc(&d_shared, twoff + tw.size() * 4), "hipMalloc shared");
    auto* shared8 = static_cast<unsigned char*>(d_shared);
    hip_check(hipMemcpy(shared8, S.p.data(), pbytes, hipMemcpyHostToDevice), "upload");
    hip_check(hipMemcpy(shared8 + pbytes, S.pinv.data(), pbytes, hipMemcpyHostToDevice), "upload");
    hip_check(hipMemcpy(shared8 + twoff, tw.data(), tw.size() * 4, hipMemcpyHostToDevice), "upload");
    bytes = twoff + tw.size() * 4;

    std::vector<unsigned char> blob;
    auto add = [&](const void* d, size_t n) {
        const size_t off = (blob.size() + 255) & ~size_t(255);
        blob.resize(off + n);
        if (n) std::memcpy(blob.data() + off, d, n);
        return off;
    };
    struct Off {
        size_t rf, rb, cf, G, Gp, C0, T, q;
    };
    std::vector<Off> offs(count);
    for (uint32_t i = 0; i < count; ++i) {
        const HostPlan& H = plans[i];
        offs[i] = {add(H.rns_frag.data(), H.rns_frag.size()), add(H.rns_bias.data(), H.rns_bias.size() * 8),
                   add(H.crt_frag.data(), H.crt_frag.size()),  add(H.G.data(), H.G.size() * 4),
                   add(H.Gplain.data(), H.Gplain.size() * 4),  add(H.C0.data(), H.C0.size() * 4),
                   add(H.T.data(), H.T.size() * 4),            add(H.q.data(), H.q.size() * 4)};
    }
    hip_check(hipMalloc(&d_tables, blob.size()), "hipMalloc tables");
    hip_check(hipMemcpy(d_tables, blob.data(), blob.size(), hipMemcpyHostToDevice), "upload tables");
    bytes += blob.size();

    auto at = [&](size_t off) { return static_cast<unsigned char*>(d_tables) + off; };
    dev.assign(count, DevPlan{});
    for (uint32_t i = 0; i < count; ++i) {
        const HostPlan& H = plans[i];
        DevPlan& D = dev[i];
        D.N = H.N;
        D.logN = H.logN;
        D.w1 = H.w1;
        D.w2 = H.w2;
        D.npass = 2;
        D.pw[0] = D.w1;
        D.pw[1] = D.w2;
        D.K = H.K;
        D.L = (int)H.L;
        D.W = (int)H.W;
        D.delta = H.delta;
        D.p = reinterpret_cast<const uint32_t*>(shared8);
        D.pinv = reinterpret_cast<const uint32_t*>(shared8 + pbytes);
        D.tw = reinterpret_cast<const uint32_t*>(shared8 + twoff);
        D.tw_stride = kSharedN;
        D.rns_frag = reinterpret_cast<const v4i*>(at(offs[i].rf));
        D.rns_bias = reinterpret_cast<const uint64_t*>(at(offs[i].rb));
        D.rns_rt = H.rns_rt;
        D.rns_ks = H.rns_ks;
        D.magic_L = H.L > 1 ? (uint32_t)((0x100000000ull + H.L - 1) / H.L) : 0u;
        D.crt_frag = reinterpret_cast<const v4i*>(at(offs[i].cf));
        D.crt_rt = H.crt_rt;
        D.crt_ks = H.crt_ks;
        D.crt_ks_full = H.crt_ks;
        D.G = reinterpret_cast<const uint2*>(at(offs[i].G));
        D.Gplain = reinterpret_cast<const uint2*>(at(offs[i].Gp));
        D.C0 = reinterpret_cast<const uint32_t*>(at(offs[i].C0));
        D.F0 = H.F0;
        D.T = reinterpret_cast<const uint32_t*>(at(offs[i].T));
        D.q = reinterpret_cast<const uint32_t*>(at(offs[i].q));
    }
}

Catalog::~Catalog() {
    if (d_shared) (void)hipFree(d_shared);
    if (d_tables) (void)hipFree(d_tables);
}

// ---------------------------------------------------------------------------
// Runner
// ---------------------------------------------------------------------------
void Runner::init() {
    int dev = 0;
    hip_check(hipGetDevice(&dev), "hipGetDevice");
    hip_check(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev), "attribute");
    hip_check(hipDeviceGetAttribute(&max_lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev), "attribute");
}

void Runner::plan(const BatchConfig& c) {
    cfg = c;
    tasks.clear();
    for (size_t& l : lds) l = 0;
    for (BatchProduct& B : prods) {
        DevPlan& D = B.D;
        const int logN = D.logN;
        D.w1 = (logN + (c.ceil_split ? 1 : 0)) / 2;
        D.w2 = logN - D.w1;
        D.npass = 2;
        D.pw[0] = D.w1;
        D.pw[1] = D.w2;
        const uint32_t nrows = 1u << D.w1, ncols = 1u << D.w2;
        const int Tc = line_threads(D.w1), Tr = line_threads(D.w2);
        B.col_tc = (int)std::min<uint32_t>({(uint32_t)c.col_tc, ncols, (uint32_t)(kBatchThreads[PH_COLF] / Tc)});
        B.row_rpb = (int)std::min<uint32_t>(nrows, (uint32_t)(kBatchThreads[PH_ROW] / Tr));
        B.crt_ys = 32 * D.crt_ks + 16;
        lds[PH_RNS] = std::max(lds[PH_RNS], rns_lds_bytes(kBatchRnsTile, D.L, D.rns_rt));
        lds[PH_CRT] = std::max(lds[PH_CRT], crt_lds_bytes(kBatchCrtTile, B.crt_ys, D.L, D.crt_ks));
        const size_t col_l = col_lds_bytes(D.w1, B.col_tc);
        lds[PH_COLF] = lds[PH_COLI] = std::max(lds[PH_COLF], col_l);
        const size_t row_l = D.w2 > 4 ? (size_t)B.row_rpb * (((size_t)1 << D.w2) + 16) * 4 : 0;
        lds[PH_ROW] = std::max(lds[PH_ROW], row_l);
    }
    for (size_t l : lds)
        if (l > (size_t)max_lds) throw std::runtime_error("batch product too large for the device's LDS");
    // Groups of consecutive products; inside a group, phase-major tile lists
    // with the products in order.
    const uint32_t P = (uint32_t)prods.size();
    const size_t cap = c.group_mb > 0 ? (size_t)c.group_mb << 20 : ~size_t(0);
    off.clear();
    single.clear();
    single_cfg.clear();
    for (uint32_t g0 = 0; g0 < P;) {
        uint32_t g1 = g0;
        size_t bytes = 0;
        do {
            bytes += 8 * (size_t)prods[g1].D.K * prods[g1].D.N;
            ++g1;
        } while (g1 < P && bytes + 8 * (size_t)prods[g1].D.K * prods[g1].D.N <= cap);
        if (g1 == g0 + 1 && c.single_kernels) {
            single.push_back((int)g0);
            single_cfg.push_back(choose_launch(prods[g0].D, std::min<size_t>((size_t)max_lds, 64 * 1024), (size_t)max_lds));
            const LaunchCfg& L = single_cfg.back();
            if (L.crt_lds > (size_t)max_lds || L.rns_lds > (size_t)max_lds)
                throw std::runtime_error("batch product too large for the device's LDS");
        } else {
            single.push_back(-1);
            single_cfg.push_back(LaunchCfg{});
        }
        off.push_back((uint32_t)tasks.size());  // RNS
        for (uint32_t i = g0; i < g1; ++i)
            for (uint32_t op = 0; op < 2; ++op)
                for (uint32_t n = 0; n < prods[i].D.N; n += kBatchRnsTile) tasks.push_back({i, n, op});
        off.push_back((uint32_t)tasks.size());  // forward columns
        for (uint32_t i = g0; i < g1; ++i) {
            const BatchProduct& B = prods[i];
            if (!B.D.w1) continue;
            for (uint32_t op = 0; op < 2; ++op)
                for (int j = 0; j < B.D.K; ++j)
                    for (uint32_t c0 = 0; c0 < (1u << B.D.w2); c0 += B.col_tc)
                        tasks.push_back({i, c0, (uint32_t)j | (op << 16)});
        }
        off.push_back((uint32_t)tasks.size());  // fused rows
        for (uint32_t i = g0; i < g1; ++i) {
            const BatchProduct& B = prods[i];
            for (int j = 0; j < B.D.K; ++j)
                for (uint32_t r0 = 0; r0 < (1u << B.D.w1); r0 += B.row_rpb) tasks.push_back({i, r0, (uint32_t)j});
        }
        off.push_back((uint32_t)tasks.size());  // inverse columns
        for (uint32_t i = g0; i < g1; ++i) {
            const BatchProduct& B = prods[i];
            if (!B.D.w1) continue;
            for (int j = 0; j < B.D.K; ++j)
                for (uint32_t c0 = 0; c0 < (1u << B.D.w2); c0 += B.col_tc) tasks.push_back({i, c0, (uint32_t)j});
        }
        off.push_back((uint32_t)tasks.size());  // CRT
        for (uint32_t i = g0; i < g1; ++i)
            for (uint32_t n = 0; n < prods[i].D.N; n += kBatchCrtTile) tasks.push_back({i, n, 0});
        g0 = g1;
    }
    off.push_back((uint32_t)tasks.size());
}

void Runner::upload(hipStream_t s, bool pinned) {
    ensure(d_prods, prod_cap, prods.size(), "hipMalloc products");
    ensure(d_tasks, task_cap, tasks.size(), "hipMalloc tiles");
    ensure(d_ctr, ctr_cap, std::max<size_t>(1, off.size() - 1), "hipMalloc counters");
    const void* hp = prods.data();
    const void* ht = tasks.data();
    if (pinned) {
        ensure_pinned(h_prods, h_prod_cap, prods.size());
        ensure_pinned(h_tasks, h_task_cap, tasks.size());
        std::memcpy(h_prods, prods.data(), prods.size() * sizeof(BatchProduct));
        std::memcpy(h_tasks, tasks.data(), tasks.size() * sizeof(BatchTask));
        hp = h_prods;
        ht = h_tasks;
    }
    hip_check(hipMemcpyAsync(d_prods, hp, prods.size() * sizeof(BatchProduct), hipMemcpyHostToDevice, s),
              "upload products");
    hip_check(hipMemcpyAsync(d_tasks, ht, tasks.size() * sizeof(BatchTask), hipMemcpyHostToDevice, s),
              "upload tiles");
}

void Runner::reserve(size_t nprods, size_t ntasks) {
    ensure(d_ctr, ctr_cap, PH_COUNT * nprods, "hipMalloc counters");
    ensure(d_prods, prod_cap, nprods, "hipMalloc products");
    ensure(d_tasks, task_cap, ntasks, "hipMalloc tiles");
    ensure_pinned(h_prods, h_prod_cap, nprods);
    ensure_pinned(h_tasks, h_task_cap, ntasks);
    std::memset(static_cast<void*>(h_prods), 0, h_prod_cap * sizeof(BatchProduct));
    std::memset(static_cast<void*>(h_tasks), 0, h_task_cap * sizeof(BatchTask));
}

hipError_t Runner::launch_reset(hipStream_t s) const {
    if (cfg.blocks_per_cu <= 0) return hipSuccess;
    return launch_batch_reset(d_ctr, (uint32_t)(off.size() - 1), s);
}

hipError_t Runner::launch_group(size_t g, hipStream_t s) const {
    const bool persistent = cfg.blocks_per_cu > 0;
    const int grid = std::max(1, cus * cfg.blocks_per_cu);
    if (single[g] >= 0) {
        // the plain multiply's sequence
        const BatchProduct& B = prods[single[g]];
        const LaunchCfg& C = single_cfg[g];
        hipError_t e;
        if ((e = launch_rns(B.D, C, B.a, B.b, B.R0, B.R1, 2, s)) != hipSuccess) return e;
        if ((e = launch_col(B.D, C, B.R0, B.R1, 2, false, s)) != hipSuccess) return e;
        if ((e = launch_row(B.D, C, B.R1, B.R0, ROW_FUSED, s)) != hipSuccess) return e;
        if ((e = launch_col(B.D, C, B.R1, nullptr, 1, true, s)) != hipSuccess) return e;
        return launch_crt(B.D, C, B.R1, B.c, B.D.G, s);
    }
    for (int ph = 0; ph < PH_COUNT; ++ph) {
        const size_t k = g * PH_COUNT + ph;
        const hipError_t e = launch_batch_phase(ph, d_prods, d_tasks + off[k], off[k + 1] - off[k],
                                                persistent ? d_ctr + k : nullptr, grid, lds[ph], s);
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

hipError_t Runner::launch(hipStream_t s) const {
    hipError_t e = launch_reset(s);
    for (size_t g = 0; e == hipSuccess && g < groups(); ++g) e = launch_group(g, s);
    return e;
}

Runner::~Runner() {
    if (d_prods) (void)hipFree(d_prods);
    if (d_tasks) (void)hipFree(d_tasks);
    if (d_ctr) (void)hipFree(d_ctr);
    if (h_prods) (void)hipHostFree(h_prods);
    if (h_tasks) (void)hipHostFree(h_tasks);
}

// ---------------------------------------------------------------------------
// Batch: a fixed set of products, one hipGraph replay per launch
// ---------------------------------------------------------------------------
struct Batch {
    Catalog cat;
    Runner run;
    uint32_t* d_R = nullptr;
    hipStream_t stream = nullptr;
    static constexpr int kMaxBranches = 3;
    hipStream_t side[kMaxBranches - 1] = {};  // further branches, used only inside the capture
    hipEvent_t fork = nullptr, join[kMaxBranches - 1] = {};
    hipEvent_t launched = nullptr;  // recorded after each graph launch, on its stream
    bool any_launch = false;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;

    ~Batch() {
        if (stream) (void)hipStreamSynchronize(stream);
        drop_graph();
        if (d_R) (void)hipFree(d_R);
        if (launched) (void)hipEventSynchronize(launched);
        if (fork) (void)hipEventDestroy(fork);
        if (launched) (void)hipEventDestroy(launched);
        for (auto e : join)
            if (e) (void)hipEventDestroy(e);
        for (auto t : side)
            if (t) (void)hipStreamDestroy(t);
        if (stream) (void)hipStreamDestroy(stream);
    }
    void drop_graph() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        exec = nullptr;
        graph = nullptr;
    }
    void configure(const BatchConfig& c) {
        // launches still in flight (on any stream) read the buffers the new
        // plan may reallocate
        hip_check(hipDeviceSynchronize(), "synchronize");
        run.plan(c);
        run.upload(stream);
        hip_check(hipStreamSynchronize(stream), "upload");
        drop_graph();
    }
    void set_operands(const uint32_t* const* a, const uint32_t* const* b, uint32_t* const* c) {
        // every pointer is checked before any is taken, so a refused call
        // leaves the batch's pointers (host and device copies) in step
        for (size_t i = 0; i < run.prods.size(); ++i) {
            if (!a[i] || !b[i] || !c[i]) throw std::runtime_error("null operand pointer");
            if ((uintptr_t)a[i] % 16 || (uintptr_t)b[i] % 16)
                throw std::runtime_error("batch operands must be 16-byte aligned");
        }
        bool changed = false;
        for (size_t i = 0; i < run.prods.size(); ++i) {
            BatchProduct& P = run.prods[i];
            if (P.a != a[i] || P.b != b[i] || P.c != c[i]) {
                P.a = a[i];
                P.b = b[i];
                P.c = c[i];
                changed = true;
            }
        }
        if (!changed) return;
        // The last launch, on whatever stream it went, may still read d_prods
        // and run its graph: the upload waits for it on the device.
        if (any_launch) hip_check(hipStreamWaitEvent(stream, launched, 0), "wait");
        hip_check(hipMemcpyAsync(run.d_prods, run.prods.data(), run.prods.size() * sizeof(BatchProduct),
                                 hipMemcpyHostToDevice, stream),
                  "upload products");
        // The phase kernels read the pointers from d_prods, but a group of one
        // product was recorded with them as kernel arguments: capture again,
        // once the graph in flight has finished.
        for (int g : run.single)
            if (g >= 0) {
                if (any_launch) hip_check(hipEventSynchronize(launched), "wait");
                drop_graph();
                break;
            }
    }
    void enqueue(hipStream_t st) {
        if (!exec) {
            hip_check(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal), "capture");
            hipError_t e = run.launch_reset(stream);
            const int nb = std::max(1, std::min(run.cfg.branches, kMaxBranches));
            if (nb > 1) {
                // groups go round-robin over nb chains that fork and join inside the graph
                if (e == hipSuccess) e = hipEventRecord(fork, stream);
                for (int k = 0; e == hipSuccess && k < nb - 1; ++k) e = hipStreamWaitEvent(side[k], fork, 0);
                for (size_t g = 0; e == hipSuccess && g < run.groups(); ++g)
                    e = run.launch_group(g, g % nb ? side[g % nb - 1] : stream);
                for (int k = 0; e == hipSuccess && k < nb - 1; ++k) {
                    e = hipEventRecord(join[k], side[k]);
                    if (e == hipSuccess) e = hipStreamWaitEvent(stream, join[k], 0);
                }
            } else {
                for (size_t g = 0; e == hipSuccess && g < run.groups(); ++g) e = run.launch_group(g, stream);
            }
            hipGraph_t g = nullptr;
            const hipError_t e2 = hipStreamEndCapture(stream, &g);
            hip_check(e, "batch launch");
            hip_check(e2, "end capture");
            graph = g;
            hip_check(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0), "graph instantiate");
        }
        hip_check(hipGraphLaunch(exec, st), "graph launch");
        hip_check(hipEventRecord(launched, st), "record");
        any_launch = true;
    }
    float time_ms(int iters) {
        hipEvent_t e0, e1;
        hip_check(hipEventCreate(&e0), "event");
        hip_check(hipEventCreate(&e1), "event");
        enqueue(stream);
        hip_check(hipStreamSynchronize(stream), "sync");
        hip_check(hipEventRecord(e0, stream), "record");
        for (int i = 0; i < iters; ++i) enqueue(stream);
        hip_check(hipEventRecord(e1, stream), "record");
        hip_check(hipEventSynchronize(e1), "sync");
        float ms = 0;
        hip_check(hipEventElapsedTime(&ms, e0, e1), "elapsed");
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        return ms / std::max(1, iters);
    }
};

void batch_build(Batch& B, uint32_t count, const uint32_t* N, const uint32_t* W, const uint32_t* q_concat) {
    B.cat.build(count, N, W, q_concat);
    B.run.init();
    hip_check(hipStreamCreateWithFlags(&B.stream, hipStreamNonBlocking), "stream");
    for (auto& t : B.side) hip_check(hipStreamCreateWithFlags(&t, hipStreamNonBlocking), "stream");
    hip_check(hipEventCreateWithFlags(&B.fork, hipEventDisableTiming), "event");
    hip_check(hipEventCreateWithFlags(&B.launched, hipEventDisableTiming), "event");
    for (auto& e : B.join) hip_check(hipEventCreateWithFlags(&e, hipEventDisableTiming), "event");
    size_t rwords = 0;
    for (size_t i = 0; i < count; ++i) rwords += B.cat.residue_words(i);
    hip_check(hipMalloc(&B.d_R, rwords * 4), "hipMalloc residues");
    B.run.prods.assign(count, BatchProduct{});
    size_t roff = 0;
    for (uint32_t i = 0; i < count; ++i) {
        BatchProduct& P = B.run.prods[i];
        P.D = B.cat.dev[i];
        const size_t kn = (size_t)B.cat.plans[i].K * B.cat.plans[i].N;
        P.R0 = B.d_R + roff;
        P.R1 = B.d_R + roff + kn;
        roff += 2 * kn;
    }
    B.configure(BatchConfig{});
}

}  // namespace pm

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
struct pm_batch {
    std::unique_ptr<pm::Batch> b;
};

namespace {
thread_local std::string g_batch_error;
template <class F>
int guarded(F&& body) {
    try {
        body();
        return 0;
    } catch (const std::exception& e) {
        g_batch_error = e.what();
    } catch (...) {
        g_batch_error = "unknown error";
    }
    return -1;
}
}  // namespace

extern "C" {

const char* pm_batch_last_error(void) { return g_batch_error.c_str(); }

int pm_batch_create(uint32_t count, const uint32_t* N, const uint32_t* W, const uint32_t* q_concat, pm_batch** out) {
    return guarded([&] {
        if (!out || !N || !W || !q_concat) throw std::runtime_error("null argument");
        *out = nullptr;
        auto h = std::make_unique<pm_batch>();
        h->b = std::make_unique<pm::Batch>();
        pm::batch_build(*h->b, count, N, W, q_concat);
        *out = h.release();
    });
}

void pm_batch_free(pm_batch* b) { delete b; }

int pm_batch_product_info(const pm_batch* b, uint32_t i, pm_info* out) {
    return guarded([&] {
        if (!b || !out || i >= b->b->cat.plans.size()) throw std::runtime_error("bad product index");
        const pm::HostPlan& H = b->b->cat.plans[i];
        const pm::DevPlan& D = b->b->run.prods[i].D;
        out->N = H.N;
        out->W = H.W;
        out->L = H.L;
        out->K = (uint32_t)H.K;
        out->delta = H.delta;
        out->w1 = (uint32_t)D.w1;
        out->w2 = (uint32_t)D.w2;
        out->device_bytes = b->b->cat.bytes;
    });
}

int pm_batch_run(pm_batch* b, const uint32_t* const* d_a, const uint32_t* const* d_b, uint32_t* const* d_c,
                 void* stream) {
    return guarded([&] {
        if (!b || !d_a || !d_b || !d_c) throw std::runtime_error("null argument");
        pm::Batch& B = *b->b;
        B.set_operands(d_a, d_b, d_c);
        hipStream_t st = stream ? static_cast<hipStream_t>(stream) : B.stream;
        if (st != B.stream) {
            hipEvent_t e;
            pm::hip_check(hipEventCreateWithFlags(&e, hipEventDisableTiming), "event");
            pm::hip_check(hipEventRecord(e, B.stream), "record");
            pm::hip_check(hipStreamWaitEvent(st, e, 0), "wait");
            (void)hipEventDestroy(e);
        }
        B.enqueue(st);
    });
}

int pm_batch_sync(pm_batch* b) {
    return guarded([&] {
        pm::hip_check(hipDeviceSynchronize(), "synchronize");
    });
}

int pm_batch_time(pm_batch* b, const uint32_t* const* d_a, const uint32_t* const* d_b, uint32_t* const* d_c,
                  int iters, float* ms) {
    return guarded([&] {
        if (!b || !d_a || !d_b || !d_c || !ms) throw std::runtime_error("null argument");
        pm::Batch& B = *b->b;
        B.set_operands(d_a, d_b, d_c);
        *ms = B.time_ms(iters);
    });
}

int pm_batch_configure(pm_batch* b, int col_tile, int ceil_split, int blocks_per_cu, int group_mb, int branches) {
    return guarded([&] {
        if (!b) throw std::runtime_error("null argument");
        if (col_tile != 8 && col_tile != 16 && col_tile != 32) throw std::runtime_error("col_tile must be 8, 16 or 32");
        if (ceil_split < 0 || ceil_split > 1 || blocks_per_cu < 0 || blocks_per_cu > 8 || group_mb < 0 ||
            branches < 1 || branches > pm::Batch::kMaxBranches)
            throw std::runtime_error("batch configuration out of range");
        pm::BatchConfig c;
        c.col_tc = col_tile;
        c.ceil_split = ceil_split;
        c.blocks_per_cu = blocks_per_cu;
        c.group_mb = group_mb;
        c.branches = branches;
        b->b->configure(c);
    });
}

int pm_batch_autotune(pm_batch* b, const uint32_t* const* d_a, const uint32_t* const* d_b, uint32_t* const* d_c,
                      int iters, int* best) {
    return guarded([&] {
        if (!b || !d_a || !d_b || !d_c || !best) throw std::runtime_error("null argument");
        pm::Batch& B = *b->b;
        B.set_operands(d_a, d_b, d_c);
        pm::BatchConfig bestc = B.run.cfg;
        float bestms = 1e30f;
        for (int tc : {8, 16, 32})
            for (int split = 0; split < 2; ++split)
                for (int bpc : {0, 2})
                    for (int gmb : {0, 48, 96, 192})
                        for (int br : {1, 2, 3}) {
                    pm::BatchConfig c;
                    c.col_tc = tc;
                    c.ceil_split = split;
                    c.blocks_per_cu = bpc;
                    c.group_mb = gmb;
                    c.branches = br;
                    try {
                        B.configure(c);
                    } catch (const std::runtime_error&) {
                        continue;  // e.g. LDS: not a candidate for these products
                    }
                    const float ms = B.time_ms(iters);
                    if (ms < bestms) {
                        bestms = ms;
                        bestc = c;
                    }
                }
        B.configure(bestc);
        if (best) {
            best[0] = bestc.col_tc;
            best[1] = bestc.ceil_split;
            best[2] = bestc.blocks_per_cu;
            best[3] = (int)(bestms * 1000.0f);  // microseconds
            best[4] = bestc.group_mb;
            best[5] = bestc.branches;
        }
    });
}

}  // extern "C"
