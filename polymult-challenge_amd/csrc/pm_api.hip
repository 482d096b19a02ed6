// Host runtime and C ABI: plan upload, the device-resident multiply, the
// host-buffer multiply (chunked PCIe pipeline in pm_pipeline.cpp), stage
// entry points and device-buffer helpers.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/polymult.h"
#include "pm_kernels.h"
#include "pm_plan.h"
#include "pm_state.h"

namespace {

thread_local std::string g_error;

void hip_check(hipError_t e, const char* what) {
    if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

template <class F>
int guarded(F&& body) {
    try {
        body();
        return 0;
    } catch (const std::exception& e) {
        g_error = e.what();
    } catch (...) {
        g_error = "unknown error";
    }
    return -1;
}

// One device allocation for every constant table; returns the offset of each.
struct TableUploader {
    std::vector<unsigned char> blob;
    size_t add(const void* data, size_t bytes) {
        const size_t off = (blob.size() + 255) & ~size_t(255);
        blob.resize(off + bytes);
        if (bytes) std::memcpy(blob.data() + off, data, bytes);
        return off;
    }
};

}  // namespace

// ---------------------------------------------------------------------------
// Residues of the allocation-free device call (pm_run_device_scratch): one
// array per GPU, in the code object, so the call needs no device allocation.
// Each GPU holds it from the library's load until the process ends. In host
// code `scratch` names a host-side stand-in, never the device memory: the
// device address comes from hipGetSymbolAddress, looked up once per device
// (at pm_init, with that device current) and kept in g_scratch.
// ---------------------------------------------------------------------------
constexpr size_t kScratchWords = (size_t)24 << 20;  // 96 MiB: both residue arrays, 2 K N words
static __device__ __align__(256) uint32_t scratch[kScratchWords];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
struct ScratchSlot {
    bool looked_up = false;
    uint32_t* ptr = nullptr;
    hipError_t err = hipSuccess;
};
ScratchSlot g_scratch[kMaxDevices];

// The device address of `device`'s scratch array (the device must be
// current), or nullptr with *err set.
uint32_t* scratch_address(int device, hipError_t* err) {
    if (device < 0 || device >= kMaxDevices) {
        *err = hipErrorInvalidDevice;
        return nullptr;
    }
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    ScratchSlot& S = g_scratch[device];
    if (!S.looked_up) {
        void* p = nullptr;
        S.err = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        S.ptr = S.err == hipSuccess ? static_cast<uint32_t*>(p) : nullptr;
        S.looked_up = true;
    }
    *err = S.err;
    return S.ptr;
}

}  // namespace

namespace pm {

void State::upload_plan() {
    const HostPlan& H = host;
    TableUploader U;
    const size_t o_p = U.add(H.p.data(), H.p.size() * 4);
    const size_t o_pinv = U.add(H.pinv.data(), H.pinv.size() * 4);
    const size_t o_tw = U.add(H.tw.data(), H.tw.size() * 4);
    const size_t o_rf = U.add(H.rns_frag.data(), H.rns_frag.size());
    const size_t o_rb = U.add(H.rns_bias.data(), H.rns_bias.size() * 8);
    const size_t o_cf = U.add(H.crt_frag.data(), H.crt_frag.size());
    const size_t o_G = U.add(H.G.data(), H.G.size() * 4);
    const size_t o_Gp = U.add(H.Gplain.data(), H.Gplain.size() * 4);
    const size_t o_C0 = U.add(H.C0.data(), H.C0.size() * 4);
    const size_t o_T = U.add(H.T.data(), H.T.size() * 4);
    const size_t o_q = U.add(H.q.data(), H.q.size() * 4);
    hip_check(hipMalloc(&d_tables, U.blob.size()), "hipMalloc tables");
    hip_check(hipMemcpy(d_tables, U.blob.data(), U.blob.size(), hipMemcpyHostToDevice), "upload tables");
    table_bytes = U.blob.size();
    auto at = [&](size_t off) { return static_cast<unsigned char*>(d_tables) + off; };

    DevPlan& D = dev;
    D.N = H.N;
    D.logN = H.logN;
    D.w1 = H.w1;
    D.w2 = H.w2;
    D.npass = H.npass;
    for (int i = 0; i < 4; ++i) D.pw[i] = H.pw[i];
    D.K = H.K;
    D.L = (int)H.L;
    D.W = (int)H.W;
    D.delta = H.delta;
    D.p = reinterpret_cast<const uint32_t*>(at(o_p));
    D.pinv = reinterpret_cast<const uint32_t*>(at(o_pinv));
    D.tw = reinterpret_cast<const uint32_t*>(at(o_tw));
    D.tw_stride = H.N;
    D.rns_frag = reinterpret_cast<const v4i*>(at(o_rf));
    D.rns_bias = reinterpret_cast<const uint64_t*>(at(o_rb));
    D.rns_rt = H.rns_rt;
    D.rns_ks = H.rns_ks;
    D.magic_L = H.L > 1 ? (uint32_t)((0x100000000ull + H.L - 1) / H.L) : 0u;
    D.crt_frag = reinterpret_cast<const v4i*>(at(o_cf));
    D.crt_rt = H.crt_rt;
    D.crt_ks = H.crt_ks;
    D.crt_ks_full = H.crt_ks;
    D.G = reinterpret_cast<const uint2*>(at(o_G));
    D.Gplain = reinterpret_cast<const uint2*>(at(o_Gp));
    D.C0 = reinterpret_cast<const uint32_t*>(at(o_C0));
    D.F0 = H.F0;
    D.T = reinterpret_cast<const uint32_t*>(at(o_T));
    D.q = reinterpret_cast<const uint32_t*>(at(o_q));
}

void State::init(uint32_t N, uint32_t W, uint32_t L, const uint32_t* q, size_t q_limbs) {
    plan_point(host, N, W, L, q, q_limbs);
    build_tables(host);
    hip_check(hipGetDevice(&device), "hipGetDevice");
    d_scratch = scratch_address(device, &scratch_err);  // checked by pm_run_device_scratch
    int max_lds = 0;
    hip_check(hipDeviceGetAttribute(&max_lds, hipDeviceAttributeMaxSharedMemoryPerBlock, device), "device attribute");
    hip_check(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking), "stream");
    hip_check(hipEventCreateWithFlags(&ev_R, hipEventDisableTiming), "event");
    if (const char* e = std::getenv("PM_LANES")) lanes = std::min(kMaxLanes, std::max(1, std::atoi(e)));
    // The lane streams are created right after the state's stream (their
    // residues only on first use): streams share the process's few hardware
    // queues, and streams created back to back land on different ones, so
    // the lanes really run side by side.
    for (int L = 1; L < lanes; ++L)
        hip_check(hipStreamCreateWithFlags(&lane_stream[L], hipStreamNonBlocking), "stream");
    upload_plan();
    cfg = choose_launch(dev, std::min<size_t>((size_t)max_lds, 64 * 1024), (size_t)max_lds);
    if (cfg.crt_lds > (size_t)max_lds || cfg.rns_lds > (size_t)max_lds ||
        crt_combine_lds_bytes(cfg.crt_tn, (int)L, cfg.combine_direct) > (size_t)max_lds)
        throw std::runtime_error("point too large for this device's LDS");
    coeff_words = (size_t)N * L;
    const size_t rbytes = (size_t)host.K * N * 4;
    for (int i = 0; i < 2; ++i) {
        hip_check(hipMalloc(&d_R[i], rbytes), "hipMalloc residues");
        hip_check(hipMalloc(&d_in[i], coeff_words * 4), "hipMalloc operands");
    }
    hip_check(hipMalloc(&d_out, coeff_words * 4), "hipMalloc result");
    device_bytes = table_bytes + 2 * rbytes + 3 * coeff_words * 4;
    if (cfg.crt_split) {
        hip_check(hipMalloc(&d_slots, slot_bytes()), "hipMalloc CRT slots");
        device_bytes += slot_bytes();
    }
    pipeline_init();
    // Warm-up on zero operands: loads every code object, touches every buffer.
    hip_check(hipMemsetAsync(d_in[0], 0, coeff_words * 4, stream), "memset");
    hip_check(hipMemsetAsync(d_in[1], 0, coeff_words * 4, stream), "memset");
    hip_check(enqueue_multiply(d_in[0], d_in[1], d_out, stream), "warm-up multiply");
    hip_check(hipStreamSynchronize(stream), "warm-up");
}

void State::acquire_R(hipStream_t st) {
    if (R_used) hip_check(hipStreamWaitEvent(st, ev_R, 0), "wait");
}

void State::release_R(hipStream_t st) {
    hip_check(hipEventRecord(ev_R, st), "record");
    R_used = true;
}

hipError_t State::enqueue_multiply(const uint32_t* a, const uint32_t* b, uint32_t* c, hipStream_t st) {
    return enqueue_multiply_with(a, b, c, st, d_R[0], d_R[1], d_slots);
}

hipError_t State::enqueue_multiply_with(const uint32_t* a, const uint32_t* b, uint32_t* c, hipStream_t st,
                                        uint32_t* R0, uint32_t* R1, long long* slots) {
    hipError_t e;
    if ((e = launch_rns(dev, cfg, a, b, R0, R1, 2, st)) != hipSuccess) return e;
    if ((e = launch_col(dev, cfg, R0, R1, 2, false, st)) != hipSuccess) return e;
    if ((e = launch_row(dev, cfg, R1, R0, ROW_FUSED, st)) != hipSuccess) return e;
    if ((e = launch_col(dev, cfg, R1, nullptr, 1, true, st)) != hipSuccess) return e;
    return launch_crt(dev, cfg, R1, c, dev.G, st, 0, 0, slots);
}

void State::enqueue_lanes(const uint32_t* a, const uint32_t* b, uint32_t* c) {
    const int L = next_lane;
    next_lane = (next_lane + 1) % lanes;
    if (!ev_crt[L]) {
        const size_t rbytes = (size_t)host.K * host.N * 4;
        if (L == 0) {
            lane_stream[0] = stream;
            lane_R[0][0] = d_R[0];
            lane_R[0][1] = d_R[1];
            lane_slots[0] = d_slots;
        } else {
            for (int i = 0; i < 2; ++i) hip_check(hipMalloc(&lane_R[L][i], rbytes), "hipMalloc residues");
            device_bytes += 2 * rbytes;
            if (cfg.crt_split) {
                hip_check(hipMalloc(&lane_slots[L], slot_bytes()), "hipMalloc CRT slots");
                device_bytes += slot_bytes();
            }
        }
        hip_check(hipEventCreateWithFlags(&ev_crt[L], hipEventDisableTiming), "event");
    }
    hipStream_t st = lane_stream[L];
    uint32_t* R0 = lane_R[L][0];
    uint32_t* R1 = lane_R[L][1];
    if (L == 0) acquire_R(st);  // lane 0 works in d_R
    wait_for_inputs(st, L, a, b);
    hip_check(launch_rns(dev, cfg, a, b, R0, R1, 2, st), "rns");
    hip_check(launch_col(dev, cfg, R0, R1, 2, false, st), "col");
    hip_check(launch_row(dev, cfg, R1, R0, ROW_FUSED, st), "row");
    hip_check(launch_col(dev, cfg, R1, nullptr, 1, true, st), "col inv");
    if (prev_lane >= 0 && prev_lane != L) hip_check(hipStreamWaitEvent(st, ev_crt[prev_lane], 0), "wait");
    hip_check(launch_crt(dev, cfg, R1, c, dev.G, st, 0, 0, lane_slots[L]), "crt");
    note_output(lane_written[L], ev_crt[L], c);  // before ev_crt[L] moves on to this call
    hip_check(hipEventRecord(ev_crt[L], st), "record");
    if (L == 0) release_R(st);
    prev_lane = L;
}

namespace {
bool overlaps(const State::Range& r, uintptr_t lo, uintptr_t hi) { return r.first < hi && lo < r.second; }
}  // namespace

void State::wait_for_inputs(hipStream_t st, int lane, const uint32_t* a, const uint32_t* b) {
    const uintptr_t n = coeff_words * 4;
    const uintptr_t a0 = (uintptr_t)a, b0 = (uintptr_t)b;
    auto hit = [&](const std::vector<Range>& list) {
        for (const Range& r : list)
            if (overlaps(r, a0, a0 + n) || overlaps(r, b0, b0 + n)) return true;
        return false;
    };
    for (int L = 0; L < kMaxLanes; ++L)
        if (L != lane && ev_crt[L] && hit(lane_written[L])) hip_check(hipStreamWaitEvent(st, ev_crt[L], 0), "wait");
    // lane 0 and the default-residue calls already wait on ev_R (acquire_R)
    if (lane > 0 && R_used && hit(R_written)) hip_check(hipStreamWaitEvent(st, ev_R, 0), "wait");
}

void State::note_output(std::vector<Range>& list, hipEvent_t ev_before, const uint32_t* c) {
    const Range r{(uintptr_t)c, (uintptr_t)c + coeff_words * 4};
    if (std::find(list.begin(), list.end(), r) != list.end()) return;  // the usual case: one output buffer reused
    // ev_before still marks the end of the previous call that this list
    // covers: once it has completed, none of the listed outputs is pending.
    if (!list.empty() && hipEventQuery(ev_before) == hipSuccess) list.clear();
    list.push_back(r);
}

void State::sync_all() {
    hip_check(hipStreamSynchronize(stream), "synchronize");
    for (int L = 1; L < kMaxLanes; ++L)
        if (lane_stream[L]) hip_check(hipStreamSynchronize(lane_stream[L]), "synchronize");
    if (R_used) hip_check(hipEventSynchronize(ev_R), "synchronize");
    for (auto& w : lane_written) w.clear();
    R_written.clear();
}

std::vector<uint32_t> State::shard_bounds(int shards) const {
    const uint32_t K = (uint32_t)dev.K, groups = (K + 7) / 8;
    if (shards < 1 || (uint32_t)shards > groups)
        throw std::runtime_error("shards must be between 1 and " + std::to_string(groups) +
                                 " (whole groups of 8 primes) at this point");
    std::vector<uint32_t> b(shards + 1);
    // larger shards first: shard 0 is a largest one
    for (int g = 0; g <= shards; ++g)
        b[g] = std::min(K, 8 * (uint32_t)(((uint64_t)groups * g + shards - 1) / shards));
    return b;
}

DevPlan State::shard_view(uint32_t j0, uint32_t cnt) const {
    DevPlan V = dev;
    V.K = (int)cnt;
    V.p += j0;
    V.pinv += j0;
    V.tw += (size_t)j0 * dev.tw_stride;
    V.rns_frag += (size_t)(j0 / 8) * dev.rns_ks * 64;  // row tile = 8 primes
    V.rns_bias += j0;
    V.rns_rt = (4 * (int)cnt + 31) / 32;
    V.crt_frag += (size_t)(j0 / 8) * 64;  // k-step = 8 primes; rows keep the full stride
    V.crt_ks = ((int)cnt + 7) / 8;
    V.crt_ks_full = dev.crt_ks_full;
    V.G += j0;
    V.Gplain += j0;
    return V;
}

// What one GPU of a prime-sharded multiply runs: its primes through RNS, the
// three NTT passes and the CRT's linear part.
hipError_t State::enqueue_shard(const DevPlan& V, const LaunchCfg& C, uint32_t j0, const uint32_t* a,
                                const uint32_t* b, long long* part, hipStream_t st) {
    uint32_t* R0 = d_R[0] + (size_t)j0 * dev.N;
    uint32_t* R1 = d_R[1] + (size_t)j0 * dev.N;
    hipError_t e;
    if ((e = launch_rns(V, C, a, b, R0, R1, 2, st)) != hipSuccess) return e;
    if ((e = launch_col(V, C, R0, R1, 2, false, st)) != hipSuccess) return e;
    if ((e = launch_row(V, C, R1, R0, ROW_FUSED, st)) != hipSuccess) return e;
    if ((e = launch_col(V, C, R1, nullptr, 1, true, st)) != hipSuccess) return e;
    return launch_crt_partial(V, C, R1, part, st);
}

void State::enqueue_sharded(const uint32_t* a, const uint32_t* b, uint32_t* c, int shards, hipStream_t st) {
    const std::vector<uint32_t> bd = shard_bounds(shards);
    const size_t stride = (size_t)(dev.L + 3) * dev.N;  // int64 slots per shard
    const size_t need = (size_t)shards * stride * 8;
    acquire_R(st);
    if (need > part_bytes) {
        // the partials are reallocated: every earlier use has to be done
        if (R_used) hip_check(hipEventSynchronize(ev_R), "sync");
        hip_check(hipStreamSynchronize(st), "sync");
        if (d_part) hip_check(hipFree(d_part), "hipFree");
        d_part = nullptr;
        part_bytes = 0;
        hip_check(hipMalloc(&d_part, need), "hipMalloc shard partials");
        part_bytes = need;
    }
    int max_lds = 0;
    hip_check(hipDeviceGetAttribute(&max_lds, hipDeviceAttributeMaxSharedMemoryPerBlock, device), "attribute");
    for (int g = 0; g < shards; ++g) {
        const DevPlan V = shard_view(bd[g], bd[g + 1] - bd[g]);
        const LaunchCfg C = choose_launch(V, std::min<size_t>((size_t)max_lds, 64 * 1024), (size_t)max_lds);
        hip_check(enqueue_shard(V, C, bd[g], a, b, d_part + (size_t)g * stride, st), "shard");
    }
    hip_check(launch_crt_combine(dev, cfg, d_part, shards, stride, c, st), "combine");
    release_R(st);
}

State::~State() {
    if (R_used) (void)hipEventSynchronize(ev_R);
    if (stream) (void)hipStreamSynchronize(stream);
    for (int L = 1; L < kMaxLanes; ++L)
        if (lane_stream[L]) (void)hipStreamSynchronize(lane_stream[L]);
    for (auto& e : ev_crt)
        if (e) (void)hipEventDestroy(e);
    for (int L = 1; L < kMaxLanes; ++L) {
        for (auto& r : lane_R[L])
            if (r) (void)hipFree(r);
        if (lane_slots[L]) (void)hipFree(lane_slots[L]);
        if (lane_stream[L]) (void)hipStreamDestroy(lane_stream[L]);
    }
    if (d_part) (void)hipFree(d_part);
    if (d_slots) (void)hipFree(d_slots);
    pipeline_free();
    for (int i = 0; i < 2; ++i) {
        if (d_R[i]) (void)hipFree(d_R[i]);
        if (d_in[i]) (void)hipFree(d_in[i]);
    }
    if (d_out) (void)hipFree(d_out);
    if (d_tables) (void)hipFree(d_tables);
    if (ev_R) (void)hipEventDestroy(ev_R);
    if (stream) (void)hipStreamDestroy(stream);
}

}  // namespace pm

struct pm_state {
    std::unique_ptr<pm::State> s;
};

extern "C" {

const char* pm_last_error(void) { return g_error.c_str(); }

int pm_point_info(uint32_t N, uint32_t W, uint32_t L, const uint32_t* q, size_t q_limbs, uint32_t* K,
                  uint64_t* delta, uint32_t* primes, size_t cap) {
    return guarded([&] {
        pm::HostPlan P;
        pm::plan_point(P, N, W, L, q, q_limbs);
        if (K) *K = (uint32_t)P.K;
        if (delta) *delta = P.delta;
        if (primes)
            for (size_t i = 0; i < cap && i < P.p.size(); ++i) primes[i] = P.p[i];
    });
}

int pm_init(uint32_t N, uint32_t W, uint32_t L, const uint32_t* q, size_t q_limbs, pm_state** out) {
    return guarded([&] {
        if (!out) throw std::runtime_error("null output handle");
        *out = nullptr;
        auto h = std::make_unique<pm_state>();
        h->s = std::make_unique<pm::State>();
        h->s->init(N, W, L, q, q_limbs);
        *out = h.release();
    });
}

void pm_free(pm_state* s) { delete s; }

int pm_query(const pm_state* s, pm_info* out) {
    return guarded([&] {
        if (!s || !out) throw std::runtime_error("null argument");
        const pm::HostPlan& H = s->s->host;
        out->N = H.N;
        out->W = H.W;
        out->L = H.L;
        out->K = (uint32_t)H.K;
        out->delta = H.delta;
        out->w1 = (uint32_t)H.w1;
        out->w2 = (uint32_t)H.w2;
        out->device_bytes = s->s->device_bytes;
    });
}

int pm_run(pm_state* s, const uint32_t* a, const uint32_t* b, uint32_t* c) {
    return guarded([&] {
        if (!s || !a || !b || !c) throw std::runtime_error("null argument");
        s->s->run_host(a, b, c);
    });
}

int pm_run_device(pm_state* s, const uint32_t* d_a, const uint32_t* d_b, uint32_t* d_c, void* stream) {
    return guarded([&] {
        if (!s || !d_a || !d_b || !d_c) throw std::runtime_error("null argument");
        if (stream) {
            hipStream_t st = static_cast<hipStream_t>(stream);
            pm::State& S = *s->s;
            S.acquire_R(st);
            S.wait_for_inputs(st, -1, d_a, d_b);
            hip_check(S.enqueue_multiply(d_a, d_b, d_c, st), "multiply");
            if (S.R_used) S.note_output(S.R_written, S.ev_R, d_c);  // before ev_R moves on
            else S.R_written.assign(1, {(uintptr_t)d_c, (uintptr_t)d_c + S.coeff_words * 4});
            S.release_R(st);
        } else
            s->s->enqueue_lanes(d_a, d_b, d_c);
    });
}

int pm_run_device_scratch(pm_state* s, const uint32_t* d_a, const uint32_t* d_b, uint32_t* d_c, void* stream) {
    return guarded([&] {
        if (!s || !d_a || !d_b || !d_c) throw std::runtime_error("null argument");
        pm::State& S = *s->s;
        const size_t kn = (size_t)S.host.K * S.host.N;
        // residues of a and b, then (points with the split CRT) its int64 slots
        const size_t slot_words = S.cfg.crt_split ? S.slot_bytes() / 4 : 0;
        if (2 * kn + slot_words > kScratchWords)
            throw std::runtime_error("point too large for the allocation-free call: its residues need " +
                                     std::to_string((2 * kn + slot_words) * 4) + " bytes, the scratch array holds " +
                                     std::to_string(kScratchWords * 4) + " (2 K N <= " +
                                     std::to_string(kScratchWords) + " words)");
        if (!S.d_scratch)
            throw std::runtime_error(std::string("scratch array unavailable on this device: ") +
                                     hipGetErrorString(S.scratch_err));
        hipStream_t st = stream ? static_cast<hipStream_t>(stream) : S.stream;
        long long* slots = slot_words ? reinterpret_cast<long long*>(S.d_scratch + 2 * kn) : nullptr;
        hip_check(S.enqueue_multiply_with(d_a, d_b, d_c, st, S.d_scratch, S.d_scratch + kn, slots), "multiply");
    });
}

int pm_sync(pm_state* s) {
    return guarded([&] { s->s->sync_all(); });
}

int pm_dev_alloc(size_t bytes, void** out) {
    return guarded([&] { hip_check(hipMalloc(out, bytes ? bytes : 1), "hipMalloc"); });
}
int pm_dev_free(void* p) {
    return guarded([&] { hip_check(hipFree(p), "hipFree"); });
}
int pm_copy_to_device(void* dst, const void* src, size_t bytes) {
    return guarded([&] { hip_check(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice), "copy to device"); });
}
int pm_stream_create(void** out) {
    return guarded([&] {
        if (!out) throw std::runtime_error("null argument");
        hipStream_t s = nullptr;
        hip_check(hipStreamCreateWithFlags(&s, hipStreamNonBlocking), "stream");
        *out = s;
    });
}
int pm_stream_sync(void* stream) {
    return guarded([&] { hip_check(hipStreamSynchronize(static_cast<hipStream_t>(stream)), "synchronize"); });
}
int pm_stream_destroy(void* stream) {
    return guarded([&] {
        if (stream) hip_check(hipStreamDestroy(static_cast<hipStream_t>(stream)), "stream");
    });
}
int pm_copy_to_host(void* dst, const void* src, size_t bytes) {
    return guarded([&] { hip_check(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost), "copy to host"); });
}

int pm_time_device(pm_state* h, const uint32_t* d_a, const uint32_t* d_b, uint32_t* d_c, int iters, float* ms) {
    return guarded([&] {
        pm::State& S = *h->s;
        const pm::DevPlan& D = S.dev;
        const pm::LaunchCfg& C = S.cfg;
        hipStream_t st = S.stream;
        const int n = iters > 0 ? iters : 1;
        hipEvent_t e0, e1;
        hip_check(hipEventCreate(&e0), "event");
        hip_check(hipEventCreate(&e1), "event");
        S.acquire_R(st);
        // Average of n launches enqueued back to back between two events: an
        // event between every pair of kernels would add its own gap to each.
        auto time_runs = [&](const std::function<void()>& launch) {
            launch();  // warm
            hip_check(hipEventRecord(e0, st), "record");
            for (int i = 0; i < n; ++i) launch();
            hip_check(hipEventRecord(e1, st), "record");
            hip_check(hipEventSynchronize(e1), "sync");
            float t = 0;
            hip_check(hipEventElapsedTime(&t, e0, e1), "elapsed");
            return t / n;
        };
        // Each kernel alone, repeated on the residues the chain leaves (every
        // pass keeps them canonical, so repeats are valid inputs).
        ms[1] = time_runs([&] { hip_check(pm::launch_rns(D, C, d_a, d_b, S.d_R[0], S.d_R[1], 2, st), "rns"); });
        ms[2] = time_runs([&] { hip_check(pm::launch_col(D, C, S.d_R[0], S.d_R[1], 2, false, st), "col"); });
        ms[3] = time_runs([&] { hip_check(pm::launch_row(D, C, S.d_R[1], S.d_R[0], pm::ROW_FUSED, st), "row"); });
        ms[4] = time_runs([&] { hip_check(pm::launch_col(D, C, S.d_R[1], nullptr, 1, true, st), "col inv"); });
        ms[5] = time_runs([&] { hip_check(pm::launch_crt(D, C, S.d_R[1], d_c, D.G, st, 0, 0, S.d_slots), "crt"); });
        // Whole multiplies last, so d_c ends holding a * b.
        ms[0] = time_runs([&] { hip_check(S.enqueue_multiply(d_a, d_b, d_c, st), "multiply"); });
        S.release_R(st);
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
    });
}

int pm_shard_bounds(pm_state* h, int shards, uint32_t* bounds) {
    return guarded([&] {
        if (!h || !bounds) throw std::runtime_error("null argument");
        const std::vector<uint32_t> b = h->s->shard_bounds(shards);
        std::memcpy(bounds, b.data(), b.size() * 4);
    });
}

int pm_run_device_sharded(pm_state* h, const uint32_t* d_a, const uint32_t* d_b, uint32_t* d_c, int shards,
                          void* stream) {
    return guarded([&] {
        pm::State& S = *h->s;
        S.enqueue_sharded(d_a, d_b, d_c, shards, stream ? static_cast<hipStream_t>(stream) : S.stream);
    });
}

int pm_time_shard(pm_state* h, const uint32_t* d_a, const uint32_t* d_b, uint32_t* d_c, int shards, int iters,
                  float* ms) {
    return guarded([&] {
        pm::State& S = *h->s;
        hipStream_t st = S.stream;
        S.enqueue_sharded(d_a, d_b, d_c, shards, st);  // sizes the partials, warms up; orders after earlier uses
        hip_check(hipStreamSynchronize(st), "sync");
        const std::vector<uint32_t> bd = S.shard_bounds(shards);
        int max_lds = 0;
        hip_check(hipDeviceGetAttribute(&max_lds, hipDeviceAttributeMaxSharedMemoryPerBlock, S.device), "attribute");
        // shard 0 is a largest one
        const pm::DevPlan V = S.shard_view(bd[0], bd[1] - bd[0]);
        const pm::LaunchCfg C = pm::choose_launch(V, std::min<size_t>((size_t)max_lds, 64 * 1024), (size_t)max_lds);
        const size_t stride = (size_t)(S.dev.L + 3) * S.dev.N;
        uint32_t* R0 = S.d_R[0];
        uint32_t* R1 = S.d_R[1];
        constexpr int NE = 8;
        hipEvent_t ev[NE];
        for (auto& e : ev) hip_check(hipEventCreate(&e), "event");
        double acc[NE] = {0};
        for (int it = 0; it < iters; ++it) {
            hip_check(hipEventRecord(ev[0], st), "record");
            S.enqueue_sharded(d_a, d_b, d_c, shards, st);
            hip_check(hipEventRecord(ev[1], st), "record");
            hip_check(pm::launch_rns(V, C, d_a, d_b, R0, R1, 2, st), "rns");
            hip_check(hipEventRecord(ev[2], st), "record");
            hip_check(pm::launch_col(V, C, R0, R1, 2, false, st), "col");
            hip_check(hipEventRecord(ev[3], st), "record");
            hip_check(pm::launch_row(V, C, R1, R0, pm::ROW_FUSED, st), "row");
            hip_check(hipEventRecord(ev[4], st), "record");
            hip_check(pm::launch_col(V, C, R1, nullptr, 1, true, st), "col inv");
            hip_check(hipEventRecord(ev[5], st), "record");
            hip_check(pm::launch_crt_partial(V, C, R1, S.d_part, st), "crt partial");
            hip_check(hipEventRecord(ev[6], st), "record");
            hip_check(pm::launch_crt_combine(S.dev, S.cfg, S.d_part, shards, stride, d_c, st), "combine");
            S.release_R(st);
            hip_check(hipEventRecord(ev[7], st), "record");
            hip_check(hipEventSynchronize(ev[7]), "sync");
            for (int k = 1; k < NE; ++k) {
                float t = 0;
                hip_check(hipEventElapsedTime(&t, ev[k - 1], ev[k]), "elapsed");
                acc[k - 1] += t;
            }
        }
        for (auto& e : ev) (void)hipEventDestroy(e);
        for (int k = 0; k < NE - 1; ++k) ms[k] = (float)(acc[k] / (iters > 0 ? iters : 1));
    });
}

int pm_probe_transfers(size_t bytes, int threads, int reps, double* out) {
    return guarded([&] { pm::probe_transfers(bytes, threads, reps, out); });
}

int pm_probe_copy(size_t bytes, int iters, double* gbps) {
    return guarded([&] {
        if (!gbps || bytes == 0 || bytes % 16 || iters < 1) throw std::runtime_error("bad copy probe arguments");
        void *src = nullptr, *dst = nullptr;
        hipEvent_t e0 = nullptr, e1 = nullptr;
        int dev = 0, cus = 0;
        hip_check(hipGetDevice(&dev), "hipGetDevice");
        hip_check(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev), "attribute");
        // 4 blocks of 256 threads per CU: the fastest of 4, 8, 16 and 32 per CU
        // over 1 GiB (5.66 TB/s against 4.55-4.78 TB/s for 8 and more)
        const int grid = cus * 4;
        try {
            hip_check(hipMalloc(&src, bytes), "hipMalloc");
            hip_check(hipMalloc(&dst, bytes), "hipMalloc");
            hip_check(hipMemset(src, 1, bytes), "hipMemset");
            hip_check(hipEventCreate(&e0), "event");
            hip_check(hipEventCreate(&e1), "event");
            hip_check(pm::launch_copy_probe(src, dst, bytes, grid, nullptr), "copy");
            hip_check(hipEventRecord(e0, nullptr), "record");
            for (int i = 0; i < iters; ++i) hip_check(pm::launch_copy_probe(src, dst, bytes, grid, nullptr), "copy");
            hip_check(hipEventRecord(e1, nullptr), "record");
            hip_check(hipEventSynchronize(e1), "sync");
            float ms = 0;
            hip_check(hipEventElapsedTime(&ms, e0, e1), "elapsed");
            *gbps = 2.0 * (double)bytes * iters / (ms * 1e-3) / 1e9;  // read + write
        } catch (...) {
            if (e0) (void)hipEventDestroy(e0);
            if (e1) (void)hipEventDestroy(e1);
            if (src) (void)hipFree(src);
            if (dst) (void)hipFree(dst);
            throw;
        }
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        (void)hipFree(src);
        (void)hipFree(dst);
    });
}

int pm_probe_valu(int iters, double* out) {
    return guarded([&] {
        if (!out || iters < 1) throw std::runtime_error("bad VALU probe arguments");
        int dev = 0, cus = 0;
        hip_check(hipGetDevice(&dev), "hipGetDevice");
        hip_check(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev), "attribute");
        const int grid = cus * 8;  // 8 blocks of 4 waves per CU: 8 waves per SIMD
        uint32_t* d = nullptr;
        hipEvent_t e0 = nullptr, e1 = nullptr;
        hipStream_t st = nullptr;
        auto release = [&] {
            if (e0) (void)hipEventDestroy(e0);
            if (e1) (void)hipEventDestroy(e1);
            if (st) (void)hipStreamDestroy(st);
            if (d) (void)hipFree(d);
        };
        try {
            hip_check(hipMalloc(&d, (size_t)grid * 256 * 4), "hipMalloc");
            hip_check(hipStreamCreateWithFlags(&st, hipStreamNonBlocking), "stream");
            hip_check(hipEventCreate(&e0), "event");
            hip_check(hipEventCreate(&e1), "event");
            for (int op = 0; op < 4; ++op) {
                hip_check(pm::launch_valu_probe(op, d, grid, iters, st), "probe");  // warm
                constexpr int reps = 5;
                hip_check(hipEventRecord(e0, st), "record");
                for (int r = 0; r < reps; ++r) hip_check(pm::launch_valu_probe(op, d, grid, iters, st), "probe");
                hip_check(hipEventRecord(e1, st), "record");
                hip_check(hipEventSynchronize(e1), "sync");
                float ms = 0;
                hip_check(hipEventElapsedTime(&ms, e0, e1), "elapsed");
                const double lane_ops = (double)grid * 256 * iters * pm::kValuProbeOpsPerIter * reps;
                out[op] = lane_ops / (ms * 1e-3) / 1e12;
            }
        } catch (...) {
            release();
            throw;
        }
        release();
    });
}

int pm_stage_rns(pm_state* h, const uint32_t* d_a, uint32_t* d_R) {
    return guarded([&] {
        pm::State& S = *h->s;
        hip_check(pm::launch_rns(S.dev, S.cfg, d_a, nullptr, d_R, nullptr, 1, S.stream), "rns");
        hip_check(hipStreamSynchronize(S.stream), "sync");
    });
}

int pm_stage_ntt_fwd(pm_state* h, uint32_t* d_R) {
    return guarded([&] {
        pm::State& S = *h->s;
        hip_check(pm::launch_col(S.dev, S.cfg, d_R, nullptr, 1, false, S.stream), "col");
        hip_check(pm::launch_row(S.dev, S.cfg, d_R, nullptr, pm::ROW_FWD, S.stream), "row");
        hip_check(hipStreamSynchronize(S.stream), "sync");
    });
}

int pm_stage_ntt_inv(pm_state* h, uint32_t* d_R) {
    return guarded([&] {
        pm::State& S = *h->s;
        hip_check(pm::launch_row(S.dev, S.cfg, d_R, nullptr, pm::ROW_INV, S.stream), "row");
        hip_check(pm::launch_col(S.dev, S.cfg, d_R, nullptr, 1, true, S.stream), "col");
        hip_check(hipStreamSynchronize(S.stream), "sync");
    });
}

int pm_stage_crt(pm_state* h, const uint32_t* d_X, uint32_t* d_c) {
    return guarded([&] {
        pm::State& S = *h->s;
        hip_check(pm::launch_crt(S.dev, S.cfg, d_X, d_c, S.dev.Gplain, S.stream, 0, 0, S.d_slots), "crt");
        hip_check(hipStreamSynchronize(S.stream), "sync");
    });
}

}  // extern "C"
