// Device view of a plan and the kernel launchers (pm_kernels.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace pm {

typedef int v4i __attribute__((ext_vector_type(4)));

// Everything a kernel needs, passed by value.
struct DevPlan {
    uint32_t N;
    // NTT passes: pw[0 .. npass-1] stages each (sum logN, each at most 8);
    // the last is the row pass. w1 = pw[0], w2 = pw[npass-1] (two passes up to
    // N = 2^16: the 4-step's columns and rows).
    int logN, w1, w2, K, L, W;
    int npass;
    int pw[4];
    uint64_t delta;
    const uint32_t* p;     // [K]
    const uint32_t* pinv;  // [K]
    const uint32_t* tw;    // [K][tw_stride] Shoup companions w' of w = psi^bitrev(k), k < N
    uint32_t tw_stride;    // N for a single point; 65536 for the shared batch table
    const v4i* rns_frag;   // [rns_rt][rns_ks][64]
    const uint64_t* rns_bias;
    int rns_rt, rns_ks;
    uint32_t magic_L;      // ceil(2^32 / L) for L >= 2: idx / L = umulhi(idx, magic_L) for idx < 2^20
    const v4i* crt_frag;   // [crt_rt][crt_ks_full][64]
    int crt_rt, crt_ks;    // crt_ks: k-steps (8 primes each) this plan covers
    int crt_ks_full;       // k-steps per row tile of the table (crt_ks unless a prime shard)
    const uint2* G;        // [K]
    const uint2* Gplain;   // [K]
    const uint32_t* C0;    // [L + 2]
    int64_t F0;
    const uint32_t* T;     // [K + 1][L]
    const uint32_t* q;     // [L]
};

// Launch geometry chosen once per plan.
struct LaunchCfg {
    int rns_tn = 128;   // coefficients per RNS block (32, 64 or 128)
    int rns_threads = 256;  // threads per RNS block (256 or 512)
    int rns_stride = 0; // LDS bytes per coefficient row
    size_t rns_lds = 0;
    int crt_tn = 64;    // coefficients per CRT block
    int crt_ys = 0;     // LDS bytes per coefficient of y bytes
    size_t crt_lds = 0;
    // The CRT block of y bytes and slots does not fit the device's LDS: the
    // CRT runs as its linear part (slots to a global buffer of (L + 3) N
    // int64) and the finish, which reads them from global memory when they do
    // not fit the LDS either (combine_direct).
    bool crt_split = false;
    bool combine_direct = false;
    bool crt_alias = false;  // k_crt_alias: slots over the y bytes (32 coefficients, crt_rt <= 8)
    int col_tc = 32;    // columns per column-pass block
    int row_rpb = 16;   // rows per row-pass block
};

enum RowMode { ROW_FUSED = 0, ROW_FWD = 1, ROW_INV = 2, ROW_FUSED_PRE = 3 };

// Chooses block shapes for a plan: tiles sized to `max_lds`, the CRT split
// when its block exceeds `device_lds` (the device's LDS per block).
LaunchCfg choose_launch(const DevPlan& D, size_t max_lds, size_t device_lds);

// R_i[j][n] = a_i[n] mod p_j for nops (1 or 2) operands, for the coefficients
// n in [first, first + count) (count 0: to the end). a_i and R_i are the
// whole arrays; `first` must be a multiple of 128. lcw > 0: a_i is staged in
// column windows of 2^lcw columns (two-pass plans; staging position i holds
// coefficient n = r 2^w2 + w 2^lcw + c for i = (w 2^w1 + r) 2^lcw + c), and
// [first, first + count) are staging positions, `first` on a window.
hipError_t launch_rns(const DevPlan& D, const LaunchCfg& C, const uint32_t* a0, const uint32_t* a1,
                      uint32_t* R0, uint32_t* R1, int nops, hipStream_t s, uint32_t first = 0, uint32_t count = 0,
                      int lcw = 0);
// The strided passes (every pass but the row pass: stages 0 .. logN - w2 - 1,
// or their inverse in reverse order) on nops residue arrays.
// col_count > 0 (two-pass plans): only columns [col_first, col_first +
// col_count) of the 2^w2, both multiples of the column tile.
hipError_t launch_col(const DevPlan& D, const LaunchCfg& C, uint32_t* R0, uint32_t* R1, int nops, bool inverse,
                      hipStream_t s, uint32_t col_first = 0, uint32_t col_count = 0);
// Row pass. ROW_FUSED: RB <- inv(fwd(RB) * fwd(RA)) per row (RA untouched,
// the product carries 2^-32). ROW_FUSED_PRE: the same with RA already through
// ROW_FWD (so a's rows can run before b is in). ROW_FWD / ROW_INV: RB only.
hipError_t launch_row(const DevPlan& D, const LaunchCfg& C, uint32_t* RB, const uint32_t* RA, RowMode mode,
                      hipStream_t s);
// out[n][L] = CRT of R[j][n] (scaled by G) mod q, for n in [first, first + count).
// With C.crt_split, `part` is the (L + 3) N int64 slot buffer. lcw > 0 (the
// aliased CRT only): out is in column-window order as launch_rns's input, and
// [first, first + count) are its positions.
hipError_t launch_crt(const DevPlan& D, const LaunchCfg& C, const uint32_t* R, uint32_t* out, const uint2* G,
                      hipStream_t s, uint32_t first = 0, uint32_t count = 0, long long* part = nullptr, int lcw = 0);

// Prime shards: D a shard's view (its primes, whole k-steps). P[slot][n]
// int64, L + 3 slots; the combine sums `shards` such arrays (shard_stride
// apart) with the full plan D and finishes the CRT into out[N][L].
hipError_t launch_crt_partial(const DevPlan& D, const LaunchCfg& C, const uint32_t* R, long long* P, hipStream_t s);
hipError_t launch_crt_combine(const DevPlan& D, const LaunchCfg& C, const long long* P, int shards, size_t shard_stride,
                              uint32_t* out, hipStream_t s);

// VALU issue-rate probe: `grid` blocks of 256 threads, each thread 8 chains
// of 128 instructions per iteration of kind op (0 v_add_u32, 1 v_mul_lo_u32,
// 2 v_mul_hi_u32, 3 v_mad_u64_u32); out[grid * 256] keeps the results.
constexpr int kValuProbeOpsPerIter = 128;
hipError_t launch_valu_probe(int op, uint32_t* out, int grid, int iters, hipStream_t s);

// dst = src, `bytes` a multiple of 16, on `grid` blocks of 256 threads.
hipError_t launch_copy_probe(const void* src, void* dst, size_t bytes, int grid, hipStream_t s);

// ---- batched products: one launch per phase over every product's tiles ----
enum BatchPhase { PH_RNS = 0, PH_COLF = 1, PH_ROW = 2, PH_COLI = 3, PH_CRT = 4, PH_COUNT = 5 };

// One product of a batch as the phase kernels see it.
struct BatchProduct {
    DevPlan D;
    const uint32_t* a;
    const uint32_t* b;
    uint32_t* c;
    uint32_t* R0;
    uint32_t* R1;
    int col_tc, row_rpb, crt_ys;
};

// A tile of one phase: product and tile coordinates
//   RNS  x = first coefficient (128 per tile), y = operand (0 a, 1 b)
//   COLF x = first column, y = prime | operand << 16
//   ROW  x = first row, y = prime
//   COLI x = first column, y = prime
//   CRT  x = first coefficient (32 per tile)
struct BatchTask {
    uint32_t prod, x, y;
};

// Block size of each phase kernel (as the plain kernels: 8 waves for RNS
// and the column passes, 4 for the rows and the CRT).
constexpr int kBatchThreads[PH_COUNT] = {512, 512, 256, 512, 256};
constexpr int kBatchRnsTile = 128;

// LDS of one RNS block of tn coefficients: their 4L bytes each, 32 bytes of
// zeros read by the last k-step, and 16 bytes of per-prime constants
// {p, pinv, bias lo, bias hi} for each of the 8 rns_rt primes of the tiles.
// LDS of one column block (w1 > 4 only): the [2^w1][tc] transpose tile and
// the first 2^w1 twiddle pairs of the prime.
inline size_t col_lds_bytes(int w1, int tc) { return w1 > 4 ? ((size_t)1 << w1) * ((size_t)tc * 4 + 8) : 0; }
inline size_t rns_lds_bytes(int tn, int L, int rns_rt) { return (size_t)tn * 4 * L + 32 + (size_t)128 * rns_rt; }
// LDS of one CRT block of tn coefficients: their y bytes (ys per coefficient),
// the L + 3 int64 slots of each, and 16 bytes of per-prime constants
// {G lo, G hi, p, 0} for each of the 8 crt_ks primes of the k-steps.
inline size_t crt_lds_bytes(int tn, int ys, int L, int crt_ks) {
    return (size_t)tn * ys + (size_t)tn * (L + 3) * 8 + (size_t)128 * crt_ks;
}
// The CRT's linear part alone (y bytes and constants) and the finish alone
// (slots unless direct, then the L + 2 limbs of each coefficient).
inline size_t crt_partial_lds_bytes(int tn, int ys, int crt_ks) { return (size_t)tn * ys + (size_t)128 * crt_ks; }
inline size_t crt_combine_lds_bytes(int tn, int L, bool direct) {
    return (size_t)tn * ((direct ? 0 : (size_t)(L + 3) * 8) + (size_t)(L + 2) * 4);
}
constexpr int kBatchCrtTile = 32;
// k_crt_alias (tn coefficients): the y bytes, and over them the slot rows
// [tn][2 (L + 3) + 1] words, then the constants at a 16-byte boundary.
inline size_t crt_alias_kc_off(int tn, int ys, int L) {
    const size_t y = (size_t)tn * ys, slots = (size_t)tn * (2 * (L + 3) + 1) * 4;
    return ((y > slots ? y : slots) + 15) / 16 * 16;
}
inline size_t crt_alias_lds_bytes(int tn, int ys, int L, int crt_ks) {
    return crt_alias_kc_off(tn, ys, L) + (size_t)128 * crt_ks;
}

// Zeroes n words (the persistent tile counters) with a kernel, so the launch
// can sit in a captured graph.
hipError_t launch_batch_reset(uint32_t* ctl, uint32_t nctl, hipStream_t s);
// Runs the ntasks tiles of one phase. ctr == nullptr: one block per tile;
// otherwise `grid` persistent blocks claim tiles from *ctr (zeroed before).
hipError_t launch_batch_phase(int phase, const BatchProduct* prods, const BatchTask* tasks, uint32_t ntasks,
                              uint32_t* ctr, int grid, size_t lds, hipStream_t s);

}  // namespace pm
