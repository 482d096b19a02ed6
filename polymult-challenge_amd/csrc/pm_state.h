// Per-point state behind the C ABI.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <utility>
#include <vector>

#include "pm_kernels.h"
#include "pm_plan.h"

namespace pm {

struct Pipeline;  // pm_pipeline.cpp: pinned staging, streams, host workers

struct State {
    HostPlan host;
    DevPlan dev{};
    LaunchCfg cfg;
    int device = 0;
    hipStream_t stream = nullptr;  // compute stream
    void* d_tables = nullptr;
    size_t table_bytes = 0;
    uint32_t* d_R[2] = {nullptr, nullptr};   // residues [K][N] of a and b
    uint32_t* d_in[2] = {nullptr, nullptr};  // operands [N][L] for host calls
    uint32_t* d_out = nullptr;               // result [N][L] for host calls
    size_t coeff_words = 0;                  // N * L
    size_t device_bytes = 0;
    Pipeline* pipe = nullptr;
    long long* d_part = nullptr;  // prime-shard CRT partials, grown on demand
    // Points with the split CRT (cfg.crt_split): its int64 slots, (L + 3) N,
    // for the default residues (d_slots) and for each lane.
    long long* d_slots = nullptr;
    size_t slot_bytes() const { return (size_t)(host.L + 3) * host.N * 8; }
    // This device's scratch array (the allocation-free call's residues), or
    // nullptr and the lookup's error.
    uint32_t* d_scratch = nullptr;
    hipError_t scratch_err = hipSuccess;
    size_t part_bytes = 0;
    // d_R and d_part serve the caller-stream calls, lane 0, the host path, the
    // sharded path and the timing helpers, on whatever stream each runs. Every
    // use waits on the device for the one before (acquire_R) and marks its
    // own end (release_R).
    hipEvent_t ev_R = nullptr;
    bool R_used = false;
    void acquire_R(hipStream_t st);
    void release_R(hipStream_t st);

    void init(uint32_t N, uint32_t W, uint32_t L, const uint32_t* q, size_t q_limbs);
    void upload_plan();
    // The whole multiply on device buffers, enqueued on st.
    hipError_t enqueue_multiply(const uint32_t* a, const uint32_t* b, uint32_t* c, hipStream_t st);
    hipError_t enqueue_multiply_with(const uint32_t* a, const uint32_t* b, uint32_t* c, hipStream_t st,
                                     uint32_t* R0, uint32_t* R1, long long* slots);

    // Device calls without a caller stream rotate over `lanes` lanes (lane 0
    // is the state's stream; each lane has its own stream and residues), so
    // one call's latency-bound kernels overlap the next call's. A call's CRT
    // waits for the previous call's CRT, so results land in call order.
    // PM_LANES (1..kMaxLanes) sets the count. The lane streams are created
    // with the state's stream, their residues on first use.
    static constexpr int kMaxLanes = 4;
    int lanes = 3;
    hipStream_t lane_stream[kMaxLanes] = {};
    uint32_t* lane_R[kMaxLanes][2] = {};
    long long* lane_slots[kMaxLanes] = {};
    hipEvent_t ev_crt[kMaxLanes] = {};
    int prev_lane = -1;  // lane of the last call, -1 before the first
    int next_lane = 0;
    void enqueue_lanes(const uint32_t* a, const uint32_t* b, uint32_t* c);
    void sync_all();

    // Read after write across lanes and streams: the outputs written by calls
    // that may still be in flight, per event that orders them (a lane's
    // ev_crt, or ev_R for the calls that use the default residues). A call
    // whose a or b overlaps one of them first waits for that event, so a
    // result can feed the next call (x = a b, then y = x b) without pm_sync.
    using Range = std::pair<uintptr_t, uintptr_t>;  // [first byte, end)
    std::vector<Range> lane_written[kMaxLanes];
    std::vector<Range> R_written;
    void wait_for_inputs(hipStream_t st, int lane, const uint32_t* a, const uint32_t* b);
    // Adds c's range to `list`; call before `ev_before` is recorded again.
    void note_output(std::vector<Range>& list, hipEvent_t ev_before, const uint32_t* c);

    // Prime shards (the multi-GPU split, run here one shard after another):
    // boundaries j_0 = 0 < j_1 < ... < j_G = K on whole MFMA k-steps (8
    // primes), a shard's view of the plan, and the sharded multiply.
    std::vector<uint32_t> shard_bounds(int shards) const;
    DevPlan shard_view(uint32_t j0, uint32_t cnt) const;
    hipError_t enqueue_shard(const DevPlan& V, const LaunchCfg& C, uint32_t j0, const uint32_t* a, const uint32_t* b,
                             long long* part, hipStream_t st);
    void enqueue_sharded(const uint32_t* a, const uint32_t* b, uint32_t* c, int shards, hipStream_t st);

    void pipeline_init();
    void pipeline_free();
    // c = a b from host buffers through the PCIe pipeline; blocking.
    void run_host(const uint32_t* a, const uint32_t* b, uint32_t* c);

    ~State();
};

// Measures, for `bytes` per transfer (best of `reps`), in GB/s:
// out[0] pinned H2D, out[1] pinned D2H, out[2] pageable H2D, out[3] pageable
// D2H, out[4] host staging copy on one thread, out[5] on `threads` threads,
// out[6] H2D and D2H of bytes/2 each at once (aggregate).
void probe_transfers(size_t bytes, int threads, int reps, double* out);

}  // namespace pm
