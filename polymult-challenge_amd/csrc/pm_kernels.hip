// gfx950 kernels of the negacyclic multiply c = a b mod (X^N + 1, q).
//
//   k_rns   a[N][L] -> R[K][N]: residues as an exact int8 MFMA GEMM over the
//           bytes of each coefficient (v_mfma_i32_32x32x32_i8), one REDC.
//   k_col   NTT stages 0 .. w1-1 (or their inverse) on column tiles.
//   k_row   NTT stages w1 .. logN-1 on rows; the fused mode runs a's and b's
//           forward rows, the pointwise product and b's inverse rows in one
//           pass over the data.
//   k_crt   R[K][N] -> c[N][L]: y_j = r_j G_j, then z = sum_j y_j Q_j and the
//           estimate of v = round(sum_j y_j / p_j) as one int8 MFMA GEMM over
//           the bytes of y, then + T[v], the 2^W = delta fold and one -q.
//
// Residues live as R[j][n] with n = row 2^w2 + col: one contiguous vector per
// prime, rows contiguous, columns strided. Every NTT kernel keeps 16 values
// per thread in registers and does four radix-2 stages there; the two halves
// of an 8-stage pass meet through one LDS transpose.
//
// What each kernel computes is what the reference computes at the same step
// (the values are identical, the method is not):
//   k_rns  = k_decompose (solve.cu:147-202): a_n mod p_j
//   k_col  = k_col<INV> (solve.cu:254-284): NTT stages 0 .. w1-1 per column
//   k_row  = k_row_fwd / k_row_fused (solve.cu:340-378): stages w1 .. logN-1,
//            the pointwise product (Montgomery, 2^-32) and the inverse rows
//   k_crt  = k_crt (solve.cu:380-501): y_j, v, z = sum y_j Q_j + T[v], fold, -q
// The twiddle of CT stage s, block i is psi^bitrev(2^s + i) and the inverse
// uses the mirrored entry; the table entries are the reference's tw(k)
// (solve.cu:947-953), here used as Shoup pairs rather than Montgomery forms
// and stored as the companions alone (see TwComp).
#include "pm_kernels.h"

#include <cstdlib>

#include "pm_arith.h"

namespace pm {
namespace {

typedef uint32_t u32;
typedef uint64_t u64;
typedef int v16i __attribute__((ext_vector_type(16)));

__device__ __forceinline__ v16i mfma_i8(v4i a, v4i b, v16i c) {
    return __builtin_amdgcn_mfma_i32_32x32x32_i8(a, b, c, 0, 0, 0);
}

// ---------------------------------------------------------------------------
// NTT lines
//
// A line is 2^NB consecutive NTT positions of one pass: a column (stride
// ncols) in the column pass, a row in the row pass. T threads share a line,
// each holding E registers. Pass 1 owns positions t + k T and runs the first
// log2(E) stages; pass 2 owns positions t E + k and runs the rest. Stage ls of
// a line with global stage offset s0 and line index `line` uses twiddle
// tw[2^(s0+ls) + line 2^ls + i], i the block index inside the line; the
// inverse uses the mirrored entry tw[2^(s0+ls+1) - 1 - (line 2^ls + i)].
// ---------------------------------------------------------------------------

// Twiddle sources. The tables in global memory hold only the Shoup companion
// w' = floor(w 2^32 / p) of each twiddle w (4 bytes, half the pairs' traffic);
// w = mulhi(w', p) + 1, exact for every w != 0 (all twiddles are powers of
// psi): w' p = w 2^32 - r with r = w 2^32 mod p in [1, p), so w' p lies in
// [(w - 1) 2^32, w 2^32). The column pass stages whole pairs in LDS.
struct TwComp {
    const u32* t;
    u32 p;
    __device__ __forceinline__ TwComp operator+(u32 o) const { return {t + o, p}; }
    __device__ __forceinline__ uint2 operator[](int i) const {
        const u32 c = t[i];
        return make_uint2(__umulhi(c, p) + 1u, c);
    }
};
struct TwPair {
    const uint2* t;
    __device__ __forceinline__ TwPair operator+(u32 o) const { return {t + o}; }
    __device__ __forceinline__ uint2 operator[](int i) const { return t[i]; }
};

// Residues are written once by each pass and read by the next kernel, which
// may run on any XCD. Stores with the sc1 policy (device scope: through the
// XCD's L2 to the last-level cache) leave the L2 no dirty lines to write back
// and the next kernel finds the data in the last-level cache; non-temporal
// stores would bypass that cache too (measured slower for the kernel after).
// At 65536/1747, against plain stores (one box, alternating runs): the RNS
// 25.9-26.3 -> 20.7-21.3 us (+0.7% three-lane throughput on its own), then
// forward columns 30.0 -> 26.0-26.3, inverse columns 16.0 -> 14.5-14.8, fused
// rows 41.2-41.8 -> 40.3-40.6 (+2% more). Non-temporal RNS stores ran 22.5 us
// but slowed the column pass by 1.5 us, -1.5% overall.
constexpr int kResidueStorePolicy = 16;  // gfx950 CPol: sc1

struct ResidueStore {
    __amdgpu_buffer_rsrc_t rsrc;
    // over R[0 .. words): every store of a pass lands inside its [K][N] array
    __device__ __forceinline__ ResidueStore(u32* R, size_t words)
        : rsrc(__builtin_amdgcn_make_buffer_rsrc(R, (short)0, (int)(words * 4), 0x00020000)) {}
    __device__ __forceinline__ void operator()(int byte_off, u32 v) const {
        __builtin_amdgcn_raw_buffer_store_b32(v, rsrc, byte_off, 0, kResidueStorePolicy);
    }
};

// Column windows (two-pass plans). A host call can stage an operand, or take
// its result, one window of CW = 2^lcw columns at a time, so that each
// window's column pass can run as soon as that window is in, or its result can
// leave as soon as that window's inverse column pass is done. Staging position
// i = (w 2^w1 + r) CW + c, with c < CW, row r < 2^w1 and window w, holds
// coefficient n = r 2^w2 + w CW + c: a window is contiguous in staging, and
// each of its rows is one run of CW coefficients of the [N][L] array.
// lcw = 0: no window, n = i.
__device__ __forceinline__ u32 window_pos(const DevPlan& D, u32 i, int lcw) {
    if (lcw == 0) return i;
    const u32 r = (i >> lcw) & ((1u << D.w1) - 1u);
    return (r << D.w2) | ((i >> (lcw + D.w1)) << lcw) | (i & ((1u << lcw) - 1u));
}

template <int NB>
struct Line {
    static constexpr int LE = NB >= 4 ? 4 : NB;
    static constexpr int E = 1 << LE;
    static constexpr int T = (1 << NB) / E;
    static constexpr int R1 = LE;
    static constexpr int R2 = NB - LE;
};

template <int NB, int NA, class TW>
__device__ __forceinline__ void fwd_pass1(u32 (&x)[NA][Line<NB>::E], const TW tab, int s0,
                                          u32 line, u32 p) {
    constexpr int E = Line<NB>::E, R1 = Line<NB>::R1;
#pragma unroll
    for (int st = 0; st < R1; ++st) {
        const int half = E >> (st + 1);
        const bool LAZY = st < R1 - 1;  // the pass's last stage leaves every output reduced
        // the stage's first entry as a pointer, so b becomes an immediate offset
        const TW ts = tab + ((1u << (s0 + st)) + (line << st));
#pragma unroll
        for (int b = 0; b < (1 << st); ++b) {
            const uint2 w = ts[b];
#pragma unroll
            for (int a = 0; a < NA; ++a)
#pragma unroll
                for (int qq = 0; qq < half; ++qq) {
                    // outputs at qq >= half/2 are the next stage's twiddled inputs
                    if (LAZY && qq >= half / 2)
                        ct_butterfly_lazy(x[a][2 * b * half + qq], x[a][2 * b * half + qq + half], w.x, w.y, p);
                    else
                        ct_butterfly(x[a][2 * b * half + qq], x[a][2 * b * half + qq + half], w.x, w.y, p);
                }
        }
    }
}

template <int NB, int NA, class TW>
__device__ __forceinline__ void fwd_pass2(u32 (&x)[NA][Line<NB>::E], const TW tab, int s0,
                                          u32 line, int t, u32 p) {
    constexpr int E = Line<NB>::E, R1 = Line<NB>::R1, R2 = Line<NB>::R2;
#pragma unroll
    for (int st = 0; st < R2; ++st) {
        const int ls = R1 + st;
        const int half = 1 << (R2 - 1 - st);
        const bool LAZY = st < R2 - 1;
        const int nblk = E / (2 * half);
        const TW ts = tab + ((1u << (s0 + ls)) + (line << ls) + (u32)(t * nblk));
#pragma unroll
        for (int b = 0; b < nblk; ++b) {
            const uint2 w = ts[b];
#pragma unroll
            for (int a = 0; a < NA; ++a)
#pragma unroll
                for (int qq = 0; qq < half; ++qq) {
                    // outputs at qq >= half/2 are the next stage's twiddled inputs
                    if (LAZY && qq >= half / 2)
                        ct_butterfly_lazy(x[a][2 * b * half + qq], x[a][2 * b * half + qq + half], w.x, w.y, p);
                    else
                        ct_butterfly(x[a][2 * b * half + qq], x[a][2 * b * half + qq + half], w.x, w.y, p);
                }
        }
    }
}

template <int NB, class TW>
__device__ __forceinline__ void inv_pass2(u32 (&x)[Line<NB>::E], const TW tab, int s0, u32 line,
                                          int t, u32 p) {
    constexpr int E = Line<NB>::E, R1 = Line<NB>::R1, R2 = Line<NB>::R2;
#pragma unroll
    for (int st = R2 - 1; st >= 0; --st) {
        const int ls = R1 + st;
        const int half = 1 << (R2 - 1 - st);
        const int nblk = E / (2 * half);
        // entry (2^(s0+ls+1) - 1 - i) for i = (line << ls) + t nblk + b: the
        // last block's entry as a pointer, b counted back from it
        const TW ts = tab + ((2u << (s0 + ls)) - 1u - ((line << ls) + (u32)(t * nblk + nblk - 1)));
#pragma unroll
        for (int b = 0; b < nblk; ++b) {
            const uint2 m = ts[nblk - 1 - b];
#pragma unroll
            for (int qq = 0; qq < half; ++qq) gs_butterfly(x[2 * b * half + qq], x[2 * b * half + qq + half], m.x, m.y, p);
        }
    }
}

template <int NB, class TW>
__device__ __forceinline__ void inv_pass1(u32 (&x)[Line<NB>::E], const TW tab, int s0, u32 line,
                                          u32 p) {
    constexpr int E = Line<NB>::E, R1 = Line<NB>::R1;
#pragma unroll
    for (int st = R1 - 1; st >= 0; --st) {
        const int half = E >> (st + 1);
        const TW ts = tab + ((2u << (s0 + st)) - 1u - ((line << st) + (u32)((1 << st) - 1)));
#pragma unroll
        for (int b = 0; b < (1 << st); ++b) {
            const uint2 m = ts[(1 << st) - 1 - b];
#pragma unroll
            for (int qq = 0; qq < half; ++qq) gs_butterfly(x[2 * b * half + qq], x[2 * b * half + qq + half], m.x, m.y, p);
        }
    }
}

// ---------------------------------------------------------------------------
// Strided tile: one pass of NB NTT stages, s0 .. s0 + NB - 1, over the lines
// of block `hi` (hi < 2^s0) of prime j: position m of the line for column c
// is hi 2^(NB + rest) + m 2^rest + c, m < 2^NB. The tile is columns [col0,
// col0 + tc) (c < 2^rest), tc T threads (threads from tc T on only take part
// in the barriers). With s0 = 0 (hi = 0, rest = w2) it is the 4-step's column
// pass; N above 2^16 chains several such passes before the row pass.
// LDS tile [2^NB][tc] for the transpose between the two register passes,
// then the 2^NB - 2^R1 twiddle pairs pass 2 reads: stage ls of the line
// needs entries 2^(s0+ls) + hi 2^ls + i, i < 2^ls (the inverse: the mirrored
// line 2^s0 - 1 - hi), staged at 2^ls + i, which is where pass 2 looks with
// s0 = 0 and line 0 (for s0 = 0 the staged copy is the table's first 2^NB
// pairs). They are staged before pass 1, behind the transpose barrier, so
// pass 2 does not wait on global memory stage by stage. Threads sweep
// columns fastest, so every global access of a wave reads or writes whole
// row segments.
// ---------------------------------------------------------------------------
// TC > 0 fixes the tile width at compile time (the plain path's 32 columns):
// the LDS offsets of the unrolled accesses become immediates.
template <int NB, bool INV, int TC = 0>
__device__ __forceinline__ void col_tile(const DevPlan& D, u32* __restrict__ R, int j, int col0, int tc_in, u32* lds,
                                         int tid, int s0, u32 hi, int rest) {
    const int tc = TC > 0 ? TC : tc_in;
    using Ln = Line<NB>;
    constexpr int E = Ln::E, T = Ln::T, R1 = Ln::R1, R2 = Ln::R2;
    const bool on = tid < tc * T;
    const int c = tid % tc, t = tid / tc;
    const u32 p = D.p[j];
    const TwComp tab{D.tw + (size_t)j * D.tw_stride, p};
    // The tile's base is the same for the whole block; each access adds a
    // 32-bit byte offset to it (the prime's vector is below 2^32 bytes), one
    // 64-bit add per access instead of a sign extension, a 64-bit shift and
    // an add. Line position m sits m 2^rest words on; the offsets are split
    // into the thread's part and a uniform step per k.
    unsigned char* base =
        reinterpret_cast<unsigned char*>(R + (size_t)j * D.N + ((size_t)hi << (NB + rest)) + col0);
    const u32 sh = (u32)rest;
    const u32 strided0 = (((u32)t << sh) + (u32)c) << 2;      // positions t + k T
    const u32 blocked0 = (((u32)(t * E) << sh) + (u32)c) << 2;  // positions t E + k
    auto strided = [&](int k) -> u32& { return *reinterpret_cast<u32*>(base + strided0 + ((u32)(k * T) << (sh + 2))); };
    auto blocked = [&](int k) -> u32& { return *reinterpret_cast<u32*>(base + blocked0 + ((u32)k << (sh + 2))); };
    // stores go through a buffer resource (byte offsets from R, same addresses)
    const ResidueStore store(R, (size_t)D.K * D.N);
    const int obase = (int)(((size_t)j * D.N + ((size_t)hi << (NB + rest)) + col0) * 4);
    auto store_strided = [&](int k, u32 v) { store(obase + (int)(strided0 + ((u32)(k * T) << (sh + 2))), v); };
    auto store_blocked = [&](int k, u32 v) { store(obase + (int)(blocked0 + ((u32)k << (sh + 2))), v); };
    u32 x[1][E];
    uint2* ltw = reinterpret_cast<uint2*>(lds + (1 << NB) * tc);
    auto stage_twiddles = [&] {
        // entry 2^ls + i <- table (2^s0 + line) 2^ls + i
        const u32 line = INV ? (1u << s0) - 1u - hi : hi;
        const u32 lift = (1u << s0) + line - 1u;  // 0 for s0 = 0
        if (on)
            for (int i = (1 << R1) + tid; i < (1 << NB); i += tc * T) {
                const int ls = 31 - __clz(i);
                ltw[i] = tab[(u32)i + (lift << ls)];
            }
    };
    if (!INV) {
        if (on) {
#pragma unroll
            for (int k = 0; k < E; ++k) x[0][k] = strided(k);
        }
        if constexpr (R2 > 0) stage_twiddles();  // read after the transpose barrier
        if (on) fwd_pass1<NB, 1>(x, tab, s0, hi, p);
        if constexpr (R2 > 0) {
            if (on) {
#pragma unroll
                for (int k = 0; k < E; ++k) lds[(t + k * T) * tc + c] = x[0][k];
            }
            __syncthreads();
            if (on) {
#pragma unroll
                for (int k = 0; k < E; ++k) x[0][k] = lds[(t * E + k) * tc + c];
                fwd_pass2<NB, 1>(x, TwPair{ltw}, 0, 0, t, p);
#pragma unroll
                for (int k = 0; k < E; ++k) store_blocked(k, x[0][k]);
            }
        } else if (on) {
#pragma unroll
            for (int k = 0; k < E; ++k) store_strided(k, x[0][k]);
        }
    } else {
        if constexpr (R2 > 0) {
            if (on) {
#pragma unroll
                for (int k = 0; k < E; ++k) x[0][k] = blocked(k);
            }
            stage_twiddles();
            __syncthreads();
            if (on) {
                inv_pass2<NB>(x[0], TwPair{ltw}, 0, 0, t, p);
#pragma unroll
                for (int k = 0; k < E; ++k) lds[(t * E + k) * tc + c] = x[0][k];
            }
            __syncthreads();
            if (on) {
#pragma unroll
                for (int k = 0; k < E; ++k) x[0][k] = lds[(t + k * T) * tc + c];
            }
        } else if (on) {
#pragma unroll
            for (int k = 0; k < E; ++k) x[0][k] = strided(k);
        }
        if (on) {
            inv_pass1<NB>(x[0], tab, s0, hi, p);
#pragma unroll
            for (int k = 0; k < E; ++k) store_strided(k, x[0][k]);
        }
    }
}

// Block x = hi 2^tiles_log + tile: 2^tiles_log column tiles of tc per line block.
template <int NB, bool INV, int TC>
__global__ void __launch_bounds__(1024) k_col(DevPlan D, u32* __restrict__ R0, u32* __restrict__ R1, int tc, int s0,
                                              int rest, int tiles_log) {
    extern __shared__ u32 lds[];
    const u32 hi = blockIdx.x >> tiles_log, tile = blockIdx.x & ((1u << tiles_log) - 1u);
    col_tile<NB, INV, TC>(D, blockIdx.z ? R1 : R0, (int)blockIdx.y, (int)(tile * (u32)tc), tc, lds, (int)threadIdx.x,
                          s0, hi, rest);
}

// ---------------------------------------------------------------------------
// Row tile number row0 / rpb of prime j, rpb rows, rpb T threads. The tile
// holds rows in mirrored pairs: its first rpb/2 rows come from the lower half
// of the rows, the other rpb/2 are their mirrors nrows - 1 - r. The inverse
// stages of row r read the twiddles the forward stages of row nrows - 1 - r
// read (the mirrored entries), so a pair shares its twiddle lines in cache
// and the table streams from HBM once instead of twice. Row position e sits
// at e + e/16 in its LDS row (one pad word after every 16) and rows are
// 2^NB + 16 words apart, so both the strided pass-1 accesses (t + k T) and
// the contiguous pass-2 accesses (16 t + k, stride 17 across the threads)
// spread over the banks. The padded index is the thread's base plus a
// constant for every k, so each access takes an immediate offset.
// LDS: rpb (2^NB + 16) words (one row buffer per row).
// ---------------------------------------------------------------------------

template <int NB, int MODE>
__device__ __forceinline__ void row_tile(const DevPlan& D, u32* __restrict__ RB, const u32* __restrict__ RA, int j,
                                         u32 row0, int rpb, u32* lds, int tid) {
    using Ln = Line<NB>;
    constexpr int E = Ln::E, T = Ln::T, R2 = Ln::R2;
    constexpr int RS = (1 << NB) + (R2 > 0 ? 16 : 0);
    const bool on = tid < rpb * T;
    const int r = on ? tid / T : 0, t = tid % T;
    const int s0 = D.logN - NB;  // stages before the row pass
    const u32 nrows = 1u << s0;
    u32 row = row0 + (u32)r;
    if ((rpb & 1) == 0 && (u32)rpb < nrows) {
        const u32 h = (u32)rpb / 2, base = (row0 / (u32)rpb) * h;
        row = (u32)r < h ? base + (u32)r : nrows - 1u - (base + (u32)r - h);
    }
    const u32 p = D.p[j];
    const TwComp tab{D.tw + (size_t)j * D.tw_stride, p};
    const size_t off = (size_t)j * D.N + ((size_t)row << NB);
    u32* rb = RB + off;
    u32* lb = lds + r * RS;
    const ResidueStore store(RB, (size_t)D.K * D.N);  // rb[i] is byte rboff + 4 i of it
    const int rboff = (int)(off * 4);
    // padded LDS index of position t + k T (t < T, T a power of two up to 16:
    // (t + k T) / 16 = k T / 16) and of position 16 t + k (k < 16)
    static_assert(R2 == 0 || (E == 16 && T <= 16), "pad mapping assumes 16 values per thread");
    auto ps = [&](int k) { return t + k * T + ((k * T) >> 4); };
    auto pc = [&](int k) { return t * (E + 1) + k; };

    if constexpr (MODE == ROW_FUSED) {
        const u32* ra = RA + off;
        const u32 pinv = D.pinv[j];
        u32 x[2][E];
        if (on) {
#pragma unroll
            for (int k = 0; k < E; ++k) {
                x[0][k] = ra[t + k * T];
                x[1][k] = rb[t + k * T];
            }
            fwd_pass1<NB, 2>(x, tab, s0, row, p);
        }
        if constexpr (R2 > 0) {
            // a, then b, through the one row buffer: half the LDS of a
            // buffer each, so twice the rows (and waves) fit on a CU
#pragma unroll
            for (int a = 0; a < 2; ++a) {
                if (a) __syncthreads();  // every read of a's transpose is done
                if (on) {
#pragma unroll
                    for (int k = 0; k < E; ++k) lb[ps(k)] = x[a][k];
                }
                __syncthreads();
                if (on) {
#pragma unroll
                    for (int k = 0; k < E; ++k) x[a][k] = lb[pc(k)];
                }
            }
            if (on) fwd_pass2<NB, 2>(x, tab, s0, row, t, p);
        }
        u32 y[E];
        if (on) {
#pragma unroll
            for (int k = 0; k < E; ++k) y[k] = mont_mul(x[0][k], x[1][k], p, pinv);
        }
        if constexpr (R2 > 0) {
            if (on) inv_pass2<NB>(y, tab, s0, row, t, p);
            __syncthreads();
            if (on) {
#pragma unroll
                for (int k = 0; k < E; ++k) lb[pc(k)] = y[k];
            }
            __syncthreads();
            if (on) {
#pragma unroll
                for (int k = 0; k < E; ++k) y[k] = lb[ps(k)];
            }
        }
        if (on) {
            inv_pass1<NB>(y, tab, s0, row, p);
#pragma unroll
            for (int k = 0; k < E; ++k) store(rboff + (int)((t + k * T) * 4), y[k]);
        }
    } else if constexpr (MODE == ROW_FUSED_PRE) {
        // b's forward rows; a's row is already in the NTT domain, each
        // position where ROW_FWD left it (t E + k after a second register
        // pass, t + k T otherwise), which is where this thread's b value sits
        const u32* ra = RA + off;
        const u32 pinv = D.pinv[j];
        u32 x[1][E];
        if (on) {
#pragma unroll
            for (int k = 0; k < E; ++k) x[0][k] = rb[t + k * T];
            fwd_pass1<NB, 1>(x, tab, s0, row, p);
        }
        if constexpr (R2 > 0) {
            if (on) {
#pragma unroll
                for (int k = 0; k < E; ++k) lb[ps(k)] = x[0][k];
            }
            __syncthreads();
            if (on) {
#pragma unroll
                for (int k = 0; k < E; ++k) x[0][k] = lb[pc(k)];
                fwd_pass2<NB, 1>(x, tab, s0, row, t, p);
            }
        }
        u32 y[E];
        if (on) {
#pragma unroll
            for (int k = 0; k < E; ++k) y[k] = mont_mul(R2 > 0 ? ra[t * E + k] : ra[t + k * T], x[0][k], p, pinv);
        }
        if constexpr (R2 > 0) {
            if (on) inv_pass2<NB>(y, tab, s0, row, t, p);
            __syncthreads();
            if (on) {
#pragma unroll
                for (int k = 0; k < E; ++k) lb[pc(k)] = y[k];
            }
            __syncthreads();
            if (on) {
#pragma unroll
                for (int k = 0; k < E; ++k) y[k] = lb[ps(k)];
            }
        }
        if (on) {
            inv_pass1<NB>(y, tab, s0, row, p);
#pragma unroll
            for (int k = 0; k < E; ++k) store(rboff + (int)((t + k * T) * 4), y[k]);
        }
    } else if constexpr (MODE == ROW_FWD) {
        u32 x[1][E];
        if (on) {
#pragma unroll
            for (int k = 0; k < E; ++k) x[0][k] = rb[t + k * T];
            fwd_pass1<NB, 1>(x, tab, s0, row, p);
        }
        if constexpr (R2 > 0) {
            if (on) {
#pragma unroll
                for (int k = 0; k < E; ++k) lb[ps(k)] = x[0][k];
            }
            __syncthreads();
            if (on) {
#pragma unroll
                for (int k = 0; k < E; ++k) x[0][k] = lb[pc(k)];
                fwd_pass2<NB, 1>(x, tab, s0, row, t, p);
#pragma unroll
                for (int k = 0; k < E; ++k) store(rboff + (int)((t * E + k) * 4), x[0][k]);
            }
        } else if (on) {
#pragma unroll
            for (int k = 0; k < E; ++k) store(rboff + (int)((t + k * T) * 4), x[0][k]);
        }
    } else {  // ROW_INV
        u32 y[E];
        if constexpr (R2 > 0) {
            if (on) {
#pragma unroll
                for (int k = 0; k < E; ++k) y[k] = rb[t * E + k];
                inv_pass2<NB>(y, tab, s0, row, t, p);
#pragma unroll
                for (int k = 0; k < E; ++k) lb[pc(k)] = y[k];
            }
            __syncthreads();
            if (on) {
#pragma unroll
                for (int k = 0; k < E; ++k) y[k] = lb[ps(k)];
            }
        } else if (on) {
#pragma unroll
            for (int k = 0; k < E; ++k) y[k] = rb[t + k * T];
        }
        if (on) {
            inv_pass1<NB>(y, tab, s0, row, p);
#pragma unroll
            for (int k = 0; k < E; ++k) store(rboff + (int)((t + k * T) * 4), y[k]);
        }
    }
}

// At most 80 VGPRs (6 waves per SIMD): with one row buffer per row the LDS
// allows that many resident blocks. The fused pass takes 77 with the twiddle
// recovery; held to 7 waves it fits 62 without spills but measured 0.5% lower
// in three-lane throughput at N=65536/W=1747.
template <int NB, int MODE>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(6, 8)))
k_row(DevPlan D, u32* __restrict__ RB, const u32* __restrict__ RA, int rpb) {
    extern __shared__ u32 lds[];
    row_tile<NB, MODE>(D, RB, RA, (int)blockIdx.y, blockIdx.x * (u32)rpb, rpb, lds, (int)threadIdx.x);
}

// ---------------------------------------------------------------------------
// int8 MFMA GEMM rows shared by the RNS and CRT kernels:
//   acc[ct] += sum_ks A(rt, ks) x B(ct, ks)
// A fragments come from a table in lane order and are prefetched eight
// k-steps at a time, so one global latency covers eight MFMAs per column
// tile. MFMA 32x32x32 lane layout: A lane l holds row l%32 and k = 16(l/32)
// + 0..15; B lane l holds column l%32 and the same k; result register r of
// lane l is row 8(r/4) + 4(l/32) + r%4, column l%32.
// ---------------------------------------------------------------------------
template <int NCT, class BF, int PF = 8>
__device__ __forceinline__ void gemm_rows(v16i (&acc)[NCT], const v4i* __restrict__ frag, int KS, BF bfrag) {
    for (int ks0 = 0; ks0 < KS; ks0 += PF) {
        v4i Af[PF];
#pragma unroll
        for (int i = 0; i < PF; ++i) Af[i] = ks0 + i < KS ? frag[(size_t)(ks0 + i) * 64] : (v4i){0, 0, 0, 0};
#pragma unroll
        for (int i = 0; i < PF; ++i) {
            if (ks0 + i < KS) {
#pragma unroll
                for (int ct = 0; ct < NCT; ++ct) acc[ct] = mfma_i8(Af[i], bfrag(ct, ks0 + i), acc[ct]);
            }
        }
    }
}

__device__ __forceinline__ long long slot_value(const v16i& acc, int g) {
    return (long long)acc[4 * g] + (long long)acc[4 * g + 1] * 256 + (long long)acc[4 * g + 2] * 65536 +
           (long long)acc[4 * g + 3] * 16777216;
}

// ---------------------------------------------------------------------------
// RNS decomposition. Block = 4 or 8 waves and TN = 32 NCT consecutive coefficients,
// copied verbatim (4L bytes each) into LDS. The GEMM is
//   C[(j,t)][n] = sum_b D[(j,t)][b] (byte_b(a_n) - 128),
// D the balanced digits of 2^(8b+32) mod p_j; sum_t 256^t C + bias_j is then
// a_n 2^32 (mod p_j), positive and below p 2^32, and one REDC gives a_n mod p_j.
// Columns of the last k-step past 4L read the next coefficient's bytes; the
// matching digits of D are zero.
// |C| <= 4L 128^2. Up to L = 127 (WIDE false), x = C0 + 256 C1 and y = C2 +
// 256 C3 stay below 2^31 in magnitude, and the sum is bias + x + 2^16 y with
// two 64-bit adds; wider points (WIDE) form x and y in 64 bits. Either way
// |x + 2^16 y| <= L 2^40 < p 2^20 (the bias's offset) for L < 2^11, so the
// REDC input is positive and below p 2^32.
// ---------------------------------------------------------------------------
constexpr int kRnsNarrowL = 127;

// Coefficients [n0, n0 + TN) of a (count in all) into residue positions
// [nst, nst + TN) of R (nst = n0 but for a column window).
template <int NCT, bool WIDE = false>
__device__ __forceinline__ void rns_tile(const DevPlan& D, const u32* __restrict__ a, u32* __restrict__ R, u32 n0,
                                         u32 count, int vec, u32* lds32, int tid, int nthreads, u32 nst) {
    constexpr int TN = 32 * NCT;
    const int L = D.L, K = D.K;
    const u32 N = D.N;
    const int cnt = (int)min((u32)TN, count - n0);
    const u32* src = a + (size_t)n0 * L;
    const int words = TN * L;

    // The bytes go into LDS already re-centred (^0x80 per byte), once, rather
    // than at every fragment load (each is read by every row tile).
    constexpr u32 kC = 0x80808080u;
    if (vec && cnt == TN) {
        const uint4* s4 = reinterpret_cast<const uint4*>(src);
        uint4* d4 = reinterpret_cast<uint4*>(lds32);
        for (int i = tid; i < words / 4; i += nthreads) {
            const uint4 v = s4[i];
            d4[i] = make_uint4(v.x ^ kC, v.y ^ kC, v.z ^ kC, v.w ^ kC);
        }
    } else {
        for (int i = tid; i < words; i += nthreads) lds32[i] = (i < cnt * L ? src[i] : 0u) ^ kC;
    }
    if (tid < 8) lds32[words + tid] = 0u;  // read by the last k-step of the last column; digits there are zero
    // Per-prime constants in LDS: the epilogue reads them without waiting on
    // the global-memory counter, which also counts the tile's earlier stores.
    uint4* kc = reinterpret_cast<uint4*>(lds32 + words + 8);  // 16-byte aligned: words is a multiple of 32
    for (int j = tid; j < 8 * D.rns_rt; j += nthreads) {
        uint4 v = make_uint4(0u, 0u, 0u, 0u);
        if (j < K) {
            const u64 bj = D.rns_bias[j];
            v = make_uint4(D.p[j], D.pinv[j], (u32)bj, (u32)(bj >> 32));
        }
        kc[j] = v;
    }
    __syncthreads();

    const int lane = tid & 63, wave = tid >> 6, nw = nthreads >> 6;
    const int half = lane >> 5, col = lane & 31;
    auto bfrag = [&](int ct, int ks) {
        const u32* q = lds32 + (32 * ct + col) * L + 8 * ks + 4 * half;
        const v4i B = {(int)q[0], (int)q[1], (int)q[2], (int)q[3]};
        return B;
    };
    auto reduce = [&](const v16i& acc, int g, u64 bias, u32 p, u32 pinv) {
        if constexpr (WIDE) {
            const long long x = (long long)acc[4 * g] + (long long)acc[4 * g + 1] * 256;
            const long long y = (long long)acc[4 * g + 2] + (long long)acc[4 * g + 3] * 256;
            return redc64(bias + (u64)x + ((u64)y << 16), p, pinv);
        } else {
            const int x = (int)((u32)acc[4 * g] + ((u32)acc[4 * g + 1] << 8));
            const int y = (int)((u32)acc[4 * g + 2] + ((u32)acc[4 * g + 3] << 8));
            return redc64(bias + (u64)(long long)x + ((u64)(long long)y << 16), p, pinv);
        }
    };
    // full tiles store through a buffer resource over R (every offset is in
    // range: j < K, nst + TN <= N)
    const ResidueStore store(R, (size_t)K * N);
    for (int rt = wave; rt < D.rns_rt; rt += nw) {
        v16i acc[NCT];
#pragma unroll
        for (int ct = 0; ct < NCT; ++ct) acc[ct] = (v16i){0};
        gemm_rows<NCT>(acc, D.rns_frag + (size_t)rt * D.rns_ks * 64 + lane, D.rns_ks, bfrag);
#pragma unroll
        for (int g = 0; g < 4; ++g) {
            const int j = 8 * rt + 2 * g + half;
            const uint4 k4 = kc[j];
            if (j >= K) continue;
            const u64 bias = (u64)k4.z | ((u64)k4.w << 32);
            u32* row = R + (size_t)j * N + nst + col;
            if (cnt == TN) {
                const int off = (int)(((size_t)j * N + nst + col) * 4);
#pragma unroll
                for (int ct = 0; ct < NCT; ++ct) store(off + 128 * ct, reduce(acc[ct], g, bias, k4.x, k4.y));
            } else {
#pragma unroll
                for (int ct = 0; ct < NCT; ++ct)
                    if (32 * ct + col < cnt) row[32 * ct] = reduce(acc[ct], g, bias, k4.x, k4.y);
            }
        }
    }
}

template <int NCT, bool WIDE>
__global__ void __launch_bounds__(512) k_rns(DevPlan D, const u32* __restrict__ a0, const u32* __restrict__ a1,
                                              u32* __restrict__ R0, u32* __restrict__ R1, u32 count, int vec, int lcw) {
    extern __shared__ __align__(16) u32 lds32[];
    const u32 n0 = blockIdx.x * (u32)(32 * NCT);
    rns_tile<NCT, WIDE>(D, blockIdx.y ? a1 : a0, blockIdx.y ? R1 : R0, n0, count, vec, lds32, (int)threadIdx.x,
                        (int)blockDim.x, window_pos(D, n0, lcw));
}

// ---------------------------------------------------------------------------
// CRT. Block = 4 waves and TN = 32 NCT coefficients (32 by default).
//   A: y_j = r_j G_j, stored as 4K bytes per coefficient, re-centred (^0x80).
//   B: MFMA over (j, s) against the digit table: slot i of the result is
//      sum_t 256^t C[4i + t]; slots 0..L give z - C0, slots L+1, L+2 give
//      F - F0 with F = sum_j y_j floor(2^48 / p_j).
//   C: one thread per coefficient: v = (F + 2^47) >> 48, carry-propagate
//      z = sum_i P_i 2^(32 i) + C0 + T[v] into L + 2 limbs, fold with
//      2^W = delta until z < 2^W, subtract q once if needed.
//   D: coalesced store of the block's contiguous [TN][L] output.
// A and B are linear in the primes: over a subset of the primes (a prime
// shard, whole MFMA k-steps of 8 primes) they give that shard's share of
// every slot, and the shares add up to the full slots (k_crt_partial,
// k_crt_combine).
// ---------------------------------------------------------------------------

// {G lo, G hi, p, 0} of the primes j < 8 crt_ks (zeros from K on) into LDS,
// so phase A reads them without waiting on the global-memory counter. The
// caller syncs before phase A.
__device__ __forceinline__ void crt_consts(const DevPlan& D, const uint2* __restrict__ G, uint4* kc, int tid,
                                           int nthreads) {
    for (int j = tid; j < 8 * D.crt_ks; j += nthreads) {
        uint4 v = make_uint4(0u, 0u, 0u, 0u);
        if (j < D.K) {
            const uint2 g = G[j];
            v = make_uint4(g.x, g.y, D.p[j], 0u);
        }
        kc[j] = v;
    }
}

// Phase A for coefficients [n0, n0 + cnt): thread (c, r) covers primes j = r,
// r + JS, ... of coefficient c, eight loads in flight per batch (lanes sweep
// c, so loads coalesce). Past cnt and from K on, the residue and the
// constants are zero, and so is y.
template <int TN>
__device__ __forceinline__ void crt_y(const DevPlan& D, const u32* __restrict__ R, const uint4* kc, u32 n0, int cnt,
                                      int ys, unsigned char* Y, int tid, int nthreads) {
    const int K = D.K;
    const u32 N = D.N;
    const int kp = 8 * D.crt_ks;
    const int JS = nthreads / TN;
    const int c = tid % TN, r = tid / TN;
    const bool live = c < cnt;
    const u32* Rc = R + n0 + c;
    for (int j0 = r; j0 < kp; j0 += 8 * JS) {
        u32 v[8];
        uint4 k[8];
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            const int j = j0 + i * JS;
            v[i] = (live && j < K) ? Rc[(size_t)j * N] : 0u;
        }
        // all eight constant reads before the first Y write (the compiler
        // cannot tell the two LDS regions apart and would order them)
#pragma unroll
        for (int i = 0; i < 8; ++i) k[i] = kc[min(j0 + i * JS, kp - 1)];
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            const int j = j0 + i * JS;
            if (j < kp) {
                const u32 y = shoup_mul(v[i], k[i].x, k[i].y, k[i].z);
                *reinterpret_cast<u32*>(Y + c * ys + 4 * j) = y ^ 0x80808080u;
            }
        }
    }
}

// Phase B: put(slot, c, value) for every slot < L + 3 and column c < TN.
template <int NCT, class PUT>
__device__ __forceinline__ void crt_gemm(const DevPlan& D, const unsigned char* Y, int ys, int tid, int nthreads,
                                         PUT put) {
    const int nslot = D.L + 3;
    const int lane = tid & 63, wave = tid >> 6, nw = nthreads >> 6;
    const int half = lane >> 5, col = lane & 31;
    auto bfrag = [&](int ct, int ks) {
        return *reinterpret_cast<const v4i*>(Y + (32 * ct + col) * ys + 32 * ks + 16 * half);
    };
    for (int rt = wave; rt < D.crt_rt; rt += nw) {
        v16i acc[NCT];
#pragma unroll
        for (int ct = 0; ct < NCT; ++ct) acc[ct] = (v16i){0};
        gemm_rows<NCT>(acc, D.crt_frag + (size_t)rt * D.crt_ks_full * 64 + lane, D.crt_ks, bfrag);
#pragma unroll
        for (int g = 0; g < 4; ++g) {
            const int slot = 8 * rt + 2 * g + half;
            if (slot >= nslot) continue;
#pragma unroll
            for (int ct = 0; ct < NCT; ++ct) put(slot, 32 * ct + col, slot_value(acc[ct], g));
        }
    }
}

// Phases C and D for coefficients [n0, n0 + cnt); get(slot, c) returns the
// slot value of coefficient c. Z: LDS [cnt][L + 2] words.
template <class GET>
__device__ __forceinline__ void crt_finish(const DevPlan& D, GET get, u32* __restrict__ out, u32 n0, int cnt, u32* Z,
                                           int tid, int nthreads) {
    const int L = D.L, K = D.K, ZL = L + 2;
    const int W = D.W, wl = W >> 5, wb = W & 31;
    const u32* C0 = D.C0;
    for (int c = tid; c < cnt; c += nthreads) {
        u32* z = Z + (size_t)c * ZL;
        const u64 F = (u64)get(L + 1, c) + ((u64)get(L + 2, c) << 32) + (u64)D.F0;
        // v <= K for every valid input; the clamp only keeps a corrupted sum
        // from indexing past T.
        const u32 v = min((u32)((F + (1ull << 47)) >> 48), (u32)K);
        const u32* Tv = D.T + (size_t)v * L;
        long long carry = 0;
#pragma unroll 8
        for (int i = 0; i < L; ++i) {
            const long long s = carry + get(i, c) + (long long)C0[i] + (long long)Tv[i];
            z[i] = (u32)s;
            carry = s >> 32;  // arithmetic: the running sum may be negative
        }
        {
            long long s = carry + get(L, c) + (long long)C0[L];
            z[L] = (u32)s;
            s = (s >> 32) + (long long)C0[L + 1];
            z[L + 1] = (u32)s;
        }
        for (;;) {
            u64 hi;
            if (wb == 0) {
                hi = (u64)z[wl] | ((u64)z[wl + 1] << 32);
            } else {
                hi = ((u64)z[wl] >> wb) | ((u64)z[wl + 1] << (32 - wb));
                if (wl + 2 < ZL) hi |= (u64)z[wl + 2] << (64 - wb);
            }
            if (hi == 0) break;
            z[wl] = wb ? (z[wl] & ((1u << wb) - 1u)) : 0u;
            for (int i = wl + 1; i < ZL; ++i) z[i] = 0;
            const u64 lo64 = hi * D.delta, hi64 = __umul64hi(hi, D.delta);
            const u32 add[4] = {(u32)lo64, (u32)(lo64 >> 32), (u32)hi64, (u32)(hi64 >> 32)};
            u64 cy = 0;
            for (int i = 0; i < ZL; ++i) {
                const u64 s = (u64)z[i] + (i < 4 ? add[i] : 0u) + cy;
                z[i] = (u32)s;
                cy = s >> 32;
                if (i >= 3 && cy == 0) break;
            }
        }
        bool ge = true;
        for (int i = L - 1; i >= 0; --i) {
            const u32 qi = D.q[i];
            if (z[i] != qi) {
                ge = z[i] > qi;
                break;
            }
        }
        if (ge) {
            long long br = 0;
            for (int i = 0; i < L; ++i) {
                const long long d = (long long)z[i] - D.q[i] - br;
                z[i] = (u32)d;
                br = d < 0;
            }
        }
    }
    __syncthreads();
    // Non-temporal: no kernel reads the result, so it need not displace the
    // residues and twiddles in the caches.
    u32* dst = out + (size_t)n0 * L;
    for (int idx = tid; idx < cnt * L; idx += nthreads) {
        const int c = L == 1 ? idx : (int)__umulhi((u32)idx, D.magic_L);  // idx / L
        __builtin_nontemporal_store(Z[(size_t)c * ZL + (idx - c * L)], &dst[idx]);
    }
}

template <int NCT>
__device__ __forceinline__ void crt_tile(const DevPlan& D, const u32* __restrict__ R, u32* __restrict__ out,
                                         const uint2* __restrict__ G, u32 n0, u32 count, int ys, unsigned char* lds8,
                                         int tid, int nthreads) {
    constexpr int TN = 32 * NCT;
    const int cnt = (int)min((u32)TN, count - n0);
    unsigned char* Y = lds8;  // [TN][ys]
    // slot-major, so lanes (consecutive coefficients) touch consecutive words
    long long* Pp = reinterpret_cast<long long*>(lds8 + (size_t)TN * ys);  // [nslot][TN]
    uint4* kc = reinterpret_cast<uint4*>(lds8 + (size_t)TN * ys + (size_t)(D.L + 3) * TN * 8);  // [8 crt_ks]
    crt_consts(D, G, kc, tid, nthreads);
    __syncthreads();
    crt_y<TN>(D, R, kc, n0, cnt, ys, Y, tid, nthreads);
    __syncthreads();
    crt_gemm<NCT>(D, Y, ys, tid, nthreads, [&](int slot, int c, long long v) { Pp[(size_t)slot * TN + c] = v; });
    __syncthreads();
    crt_finish(D, [&](int slot, int c) { return Pp[(size_t)slot * TN + c]; }, out, n0, cnt,
               reinterpret_cast<u32*>(Y) /* reuses the y bytes */, tid, nthreads);
}

// The finish of one coefficient, in place (the steps of crt_finish): z is
// the coefficient's row of words, slot i (i < L + 3) at z[2 i] (low) and
// z[2 i + 1] (high); limb i (i < L + 2) is written over slot i's low word
// once slot i has been read.
__device__ __forceinline__ void crt_finish_inplace(const DevPlan& D, u32* __restrict__ z) {
    const int L = D.L, K = D.K, ZL = L + 2;
    const int W = D.W, wl = W >> 5, wb = W & 31;
    const u32* C0 = D.C0;
    auto slot = [&](int i) { return (long long)(((u64)z[2 * i + 1] << 32) | z[2 * i]); };
    const u64 F = (u64)slot(L + 1) + ((u64)slot(L + 2) << 32) + (u64)D.F0;
    // v <= K for every valid input; the clamp only keeps a corrupted sum
    // from indexing past T.
    const u32 v = min((u32)((F + (1ull << 47)) >> 48), (u32)K);
    const u32* Tv = D.T + (size_t)v * L;
    long long carry = 0;
#pragma unroll 8
    for (int i = 0; i < L; ++i) {
        const long long s = carry + slot(i) + (long long)C0[i] + (long long)Tv[i];
        z[2 * i] = (u32)s;
        carry = s >> 32;  // arithmetic: the running sum may be negative
    }
    {
        long long s = carry + slot(L) + (long long)C0[L];
        z[2 * L] = (u32)s;
        s = (s >> 32) + (long long)C0[L + 1];
        z[2 * (L + 1)] = (u32)s;
    }
    for (;;) {
        u64 hi;
        if (wb == 0) {
            hi = (u64)z[2 * wl] | ((u64)z[2 * (wl + 1)] << 32);
        } else {
            hi = ((u64)z[2 * wl] >> wb) | ((u64)z[2 * (wl + 1)] << (32 - wb));
            if (wl + 2 < ZL) hi |= (u64)z[2 * (wl + 2)] << (64 - wb);
        }
        if (hi == 0) break;
        z[2 * wl] = wb ? (z[2 * wl] & ((1u << wb) - 1u)) : 0u;
        for (int i = wl + 1; i < ZL; ++i) z[2 * i] = 0;
        const u64 lo64 = hi * D.delta, hi64 = __umul64hi(hi, D.delta);
        const u32 add[4] = {(u32)lo64, (u32)(lo64 >> 32), (u32)hi64, (u32)(hi64 >> 32)};
        u64 cy = 0;
        for (int i = 0; i < ZL; ++i) {
            const u64 s = (u64)z[2 * i] + (i < 4 ? add[i] : 0u) + cy;
            z[2 * i] = (u32)s;
            cy = s >> 32;
            if (i >= 3 && cy == 0) break;
        }
    }
    bool ge = true;
    for (int i = L - 1; i >= 0; --i) {
        const u32 qi = D.q[i], zi = z[2 * i];
        if (zi != qi) {
            ge = zi > qi;
            break;
        }
    }
    if (ge) {
        long long br = 0;
        for (int i = 0; i < L; ++i) {
            const long long d = (long long)z[2 * i] - D.q[i] - br;
            z[2 * i] = (u32)d;
            br = d < 0;
        }
    }
}

// The CRT with the slots over the y bytes: TN = 32 NCT coefficients per
// block of 4 NCT waves, so the eight GEMM row tiles (crt_rt <= 8) make two
// per wave at NCT = 1 and one per wave over two column tiles at NCT = 2.
// Each wave keeps its results (reduced to slot values) in registers until
// every wave has read the y bytes, then writes them as slot rows [TN][S]
// (S = 2 (L + 3) + 1 words, an odd stride, so lanes on consecutive
// coefficients hit distinct banks); the finish writes each limb over its
// slot, and the store reads the limbs in place. The block's LDS is max(y
// bytes, slot rows) + constants instead of their sum (17.8 KB instead of
// 32.6 KB at L = 55, NCT = 1), so more blocks share a CU to hide the residue
// and fragment loads; at NCT = 2 each digit fragment feeds two MFMAs, which
// halves the fragment traffic from L2 per coefficient.
// Four fragments in flight per wave keep NCT = 1 at 64 VGPRs: eight 4-wave
// blocks per CU, 22.8-23.3 us against 25.5-25.7 us with eight in flight at
// six blocks per CU (65536/1747).
constexpr int kCrtAliasPF = 4;
template <int NCT, int MINW>  // NCT 1 or 2; MINW: waves per SIMD the registers must allow
__global__ void __launch_bounds__(NCT == 1 ? 256 : 512) __attribute__((amdgpu_waves_per_eu(MINW, 8)))
k_crt_alias(DevPlan D, const u32* __restrict__ R, u32* __restrict__ out, const uint2* __restrict__ G, u32 count, int ys,
            int kc_off, int lcw) {
    extern __shared__ __align__(16) unsigned char lds8[];
    constexpr int TN = 32 * NCT, RTW = NCT == 1 ? 2 : 1;  // coefficients per block, row tiles per wave
    const int tid = (int)threadIdx.x, nthreads = (int)blockDim.x;
    const u32 n0 = blockIdx.x * (u32)TN;
    const int cnt = (int)min((u32)TN, count - n0);
    const int L = D.L, nslot = L + 3, S = 2 * nslot + 1;
    unsigned char* Y = lds8;                                  // [TN][ys], then the slot rows
    uint4* kc = reinterpret_cast<uint4*>(lds8 + kc_off);      // [8 crt_ks]
    crt_consts(D, G, kc, tid, nthreads);
    __syncthreads();
    crt_y<TN>(D, R, kc, window_pos(D, n0, lcw), cnt, ys, Y, tid, nthreads);
    __syncthreads();
    const int lane = tid & 63, wave = tid >> 6, nw = nthreads >> 6;
    const int half = lane >> 5, col = lane & 31;
    auto bfrag = [&](int ct, int ks) {
        return *reinterpret_cast<const v4i*>(Y + (32 * ct + col) * ys + 32 * ks + 16 * half);
    };
    // each row tile's result, reduced to its lane's four slot values per column tile
    long long sv[RTW][NCT][4];
#pragma unroll
    for (int h = 0; h < RTW; ++h) {
        v16i acc[NCT];
#pragma unroll
        for (int ct = 0; ct < NCT; ++ct) acc[ct] = (v16i){0};
        const int rt = wave + h * nw;
        if (rt < D.crt_rt)
            gemm_rows<NCT, decltype(bfrag), kCrtAliasPF>(acc, D.crt_frag + (size_t)rt * D.crt_ks_full * 64 + lane,
                                                          D.crt_ks, bfrag);
#pragma unroll
        for (int ct = 0; ct < NCT; ++ct)
#pragma unroll
            for (int g = 0; g < 4; ++g) sv[h][ct][g] = slot_value(acc[ct], g);
    }
    __syncthreads();  // every wave has read the y bytes
    u32* sl = reinterpret_cast<u32*>(lds8);
#pragma unroll
    for (int h = 0; h < RTW; ++h) {
        const int rt = wave + h * nw;
        if (rt >= D.crt_rt) continue;
#pragma unroll
        for (int ct = 0; ct < NCT; ++ct)
#pragma unroll
            for (int g = 0; g < 4; ++g) {
                const int slot = 8 * rt + 2 * g + half;
                if (slot >= nslot) continue;
                const int c = 32 * ct + col;
                sl[c * S + 2 * slot] = (u32)sv[h][ct][g];
                sl[c * S + 2 * slot + 1] = (u32)((u64)sv[h][ct][g] >> 32);
            }
    }
    __syncthreads();
    for (int c = tid; c < cnt; c += nthreads) crt_finish_inplace(D, sl + c * S);
    __syncthreads();
    // non-temporal, as in crt_finish: no kernel reads the result (20.4-20.6 ->
    // 18.9-19.1 us at 65536/1747, three-lane throughput unchanged)
    u32* dst = out + (size_t)n0 * L;
    for (int idx = tid; idx < cnt * L; idx += nthreads) {
        const int c = L == 1 ? idx : (int)__umulhi((u32)idx, D.magic_L);  // idx / L
        __builtin_nontemporal_store(sl[c * S + 2 * (idx - c * L)], &dst[idx]);
    }
}

// One prime shard's share of every slot: P[slot][n] (int64), for the
// coefficients [32 NCT blockIdx.x, ...) below count. D is the shard's view.
template <int NCT>
__global__ void __launch_bounds__(256) k_crt_partial(DevPlan D, const u32* __restrict__ R, const uint2* __restrict__ G,
                                                      long long* __restrict__ P, u32 count, int ys) {
    extern __shared__ __align__(16) unsigned char lds8[];
    constexpr int TN = 32 * NCT;
    const u32 n0 = blockIdx.x * (u32)TN;
    const int cnt = (int)min((u32)TN, count - n0);
    const int tid = (int)threadIdx.x;
    uint4* kc = reinterpret_cast<uint4*>(lds8 + (size_t)TN * ys);  // [8 crt_ks]
    crt_consts(D, G, kc, tid, (int)blockDim.x);
    __syncthreads();
    crt_y<TN>(D, R, kc, n0, cnt, ys, lds8, tid, (int)blockDim.x);
    __syncthreads();
    const u32 N = D.N;
    crt_gemm<NCT>(D, lds8, ys, tid, (int)blockDim.x, [&](int slot, int c, long long v) {
        if (c < cnt) P[(size_t)slot * N + n0 + c] = v;
    });
}

// Sums the shards' slots and finishes: c = CRT result mod q. D is the full
// plan. The block first sums its coefficients' slots over the shards into LDS
// (all threads, coalesced), then finishes one coefficient per thread. DIRECT
// (points whose slots do not fit the LDS beside the finish's limbs, W above
// about 9,000 bits): the finishing threads read and sum the slots from global
// memory as they go (consecutive threads, consecutive words).
template <int NCT, bool DIRECT>
__global__ void __launch_bounds__(256) k_crt_combine(DevPlan D, const long long* __restrict__ P, int shards,
                                                      size_t shard_stride, u32* __restrict__ out, u32 count) {
    extern __shared__ __align__(16) unsigned char lds8[];
    constexpr int TN = 32 * NCT;
    const int nslot = D.L + 3;
    const u32 n0 = blockIdx.x * (u32)TN;
    const int cnt = (int)min((u32)TN, count - n0);
    const u32 N = D.N;
    const int tid = (int)threadIdx.x;
    if constexpr (DIRECT) {
        u32* Z = reinterpret_cast<u32*>(lds8);  // [TN][L + 2]
        crt_finish(D,
                   [&](int slot, int c) {
                       long long s = 0;
                       for (int g = 0; g < shards; ++g) s += P[(size_t)g * shard_stride + (size_t)slot * N + n0 + c];
                       return s;
                   },
                   out, n0, cnt, Z, tid, (int)blockDim.x);
    } else {
        long long* S = reinterpret_cast<long long*>(lds8);                   // [nslot][TN]
        u32* Z = reinterpret_cast<u32*>(lds8 + (size_t)nslot * TN * 8);      // [TN][L + 2]
        for (int idx = tid; idx < nslot * TN; idx += (int)blockDim.x) {
            const int slot = idx / TN, c = idx % TN;
            long long s = 0;
            if (c < cnt)
                for (int g = 0; g < shards; ++g) s += P[(size_t)g * shard_stride + (size_t)slot * N + n0 + c];
            S[idx] = s;
        }
        __syncthreads();
        crt_finish(D, [&](int slot, int c) { return S[(size_t)slot * TN + c]; }, out, n0, cnt, Z, tid,
                   (int)blockDim.x);
    }
}

template <int NCT>
__global__ void __launch_bounds__(256) k_crt(DevPlan D, const u32* __restrict__ R, u32* __restrict__ out,
                                              const uint2* __restrict__ G, u32 count, int ys) {
    extern __shared__ __align__(16) unsigned char lds8[];
    crt_tile<NCT>(D, R, out, G, blockIdx.x * (u32)(32 * NCT), count, ys, lds8, (int)threadIdx.x, (int)blockDim.x);
}

// ---------------------------------------------------------------------------
// Batched products (N <= 2^16, two NTT passes): one launch per phase, each
// over that phase's tiles of every product. Inside a phase tiles are independent, so there is no
// cross-block wait and no device-scope fence; the launch boundary orders the
// phases. Each phase kernel has its own register budget (one kernel holding
// every tile type would run at the largest one's). Two ways to hand out
// tiles: the hardware dispatcher (one block per tile, ctr == nullptr) or a
// persistent grid whose blocks claim tiles from an atomic counter.
// ---------------------------------------------------------------------------
template <bool INV>
__device__ __forceinline__ void col_tile_nb(int nb, const DevPlan& D, u32* R, int j, int col0, int tc, u32* lds,
                                            int tid) {
    switch (nb) {
        case 1: col_tile<1, INV>(D, R, j, col0, tc, lds, tid, 0, 0, D.w2); break;
        case 2: col_tile<2, INV>(D, R, j, col0, tc, lds, tid, 0, 0, D.w2); break;
        case 3: col_tile<3, INV>(D, R, j, col0, tc, lds, tid, 0, 0, D.w2); break;
        case 4: col_tile<4, INV>(D, R, j, col0, tc, lds, tid, 0, 0, D.w2); break;
        case 5: col_tile<5, INV>(D, R, j, col0, tc, lds, tid, 0, 0, D.w2); break;
        case 6: col_tile<6, INV>(D, R, j, col0, tc, lds, tid, 0, 0, D.w2); break;
        case 7: col_tile<7, INV>(D, R, j, col0, tc, lds, tid, 0, 0, D.w2); break;
        case 8: col_tile<8, INV>(D, R, j, col0, tc, lds, tid, 0, 0, D.w2); break;
        default: break;
    }
}

__device__ __forceinline__ void row_tile_nb(int nb, const DevPlan& D, u32* RB, const u32* RA, int j, u32 row0, int rpb,
                                            u32* lds, int tid) {
    switch (nb) {
        case 0: row_tile<0, ROW_FUSED>(D, RB, RA, j, row0, rpb, lds, tid); break;
        case 1: row_tile<1, ROW_FUSED>(D, RB, RA, j, row0, rpb, lds, tid); break;
        case 2: row_tile<2, ROW_FUSED>(D, RB, RA, j, row0, rpb, lds, tid); break;
        case 3: row_tile<3, ROW_FUSED>(D, RB, RA, j, row0, rpb, lds, tid); break;
        case 4: row_tile<4, ROW_FUSED>(D, RB, RA, j, row0, rpb, lds, tid); break;
        case 5: row_tile<5, ROW_FUSED>(D, RB, RA, j, row0, rpb, lds, tid); break;
        case 6: row_tile<6, ROW_FUSED>(D, RB, RA, j, row0, rpb, lds, tid); break;
        case 7: row_tile<7, ROW_FUSED>(D, RB, RA, j, row0, rpb, lds, tid); break;
        case 8: row_tile<8, ROW_FUSED>(D, RB, RA, j, row0, rpb, lds, tid); break;
        default: break;
    }
}

template <int PH>
__device__ __forceinline__ void batch_task(const BatchProduct& P, u32 x, u32 y, u32* lds, int tid) {
    // A copy, so the plan's fields are loaded once per tile (uniform, into
    // scalar registers) rather than re-read through the product pointer.
    const DevPlan D = P.D;
    if constexpr (PH == PH_RNS) {
        rns_tile<kBatchRnsTile / 32>(D, y ? P.b : P.a, y ? P.R1 : P.R0, x, D.N, 1, lds, tid, kBatchThreads[PH], x);
    } else if constexpr (PH == PH_COLF) {
        col_tile_nb<false>(D.w1, D, (y >> 16) ? P.R1 : P.R0, (int)(y & 0xFFFF), (int)x, P.col_tc, lds, tid);
    } else if constexpr (PH == PH_ROW) {
        row_tile_nb(D.w2, D, P.R1, P.R0, (int)y, x, P.row_rpb, lds, tid);
    } else if constexpr (PH == PH_COLI) {
        col_tile_nb<true>(D.w1, D, P.R1, (int)y, (int)x, P.col_tc, lds, tid);
    } else {
        crt_tile<kBatchCrtTile / 32>(D, P.R1, P.c, D.G, x, D.N, P.crt_ys, reinterpret_cast<unsigned char*>(lds), tid,
                                     kBatchThreads[PH]);
    }
}

// Plain streaming copy (16 B per lane): the reference point for the HBM
// counters and for the achievable copy bandwidth.
__global__ void __launch_bounds__(256) k_copy_probe(const uint4* __restrict__ src, uint4* __restrict__ dst, size_t n) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

// VALU issue-rate probe: every thread runs 8 independent chains of one
// instruction (0 v_add_u32, 1 v_mul_lo_u32, 2 v_mul_hi_u32, 3
// v_mad_u64_u32), 16 x 8 per loop pass, so enough waves per SIMD hide each
// chain's latency and the time measures issue rate. Written as asm so the
// compiler can neither fold the chains nor change the instruction; they touch
// registers only. out gets one word per thread (keeps the results live).
template <int OP>
__global__ void __launch_bounds__(256) k_valu_probe(u32* __restrict__ out, u32 seed, int iters) {
    const u32 tid = blockIdx.x * blockDim.x + threadIdx.x;
    u32 x[8];
    u64 y[8];
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        x[i] = seed + tid * 8u + (u32)i;
        y[i] = x[i];
    }
    const u32 k = seed | 1u;
    for (int it = 0; it < iters; ++it) {
#pragma unroll
        for (int u = 0; u < 16; ++u)
#pragma unroll
            for (int i = 0; i < 8; ++i) {
                if constexpr (OP == 0) asm volatile("v_add_u32 %0, %1, %0" : "+v"(x[i]) : "v"(k));
                else if constexpr (OP == 1) asm volatile("v_mul_lo_u32 %0, %0, %1" : "+v"(x[i]) : "v"(k));
                else if constexpr (OP == 2) asm volatile("v_mul_hi_u32 %0, %0, %1" : "+v"(x[i]) : "v"(k));
                else asm volatile("v_mad_u64_u32 %0, vcc, %1, %2, %0" : "+v"(y[i]) : "v"(x[i]), "v"(k) : "vcc");
            }
    }
    u32 r = 0;
#pragma unroll
    for (int i = 0; i < 8; ++i) r ^= x[i] ^ (u32)y[i] ^ (u32)(y[i] >> 32);
    out[tid] = r;
}

__global__ void k_batch_reset(u32* ctl, u32 n) {
    for (u32 i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) ctl[i] = 0;
}

template <int PH>
__global__ void __launch_bounds__(kBatchThreads[PH]) k_batch_phase(const BatchProduct* __restrict__ prods,
                                                               const BatchTask* __restrict__ tasks, u32 ntasks,
                                                               u32* ctr) {
    extern __shared__ __align__(16) u32 lds[];
    const int tid = (int)threadIdx.x;
    if (!ctr) {
        const BatchTask tk = tasks[blockIdx.x];
        const u32 pi = __builtin_amdgcn_readfirstlane(tk.prod);
        const u32 x = __builtin_amdgcn_readfirstlane(tk.x), y = __builtin_amdgcn_readfirstlane(tk.y);
        batch_task<PH>(prods[pi], x, y, lds, tid);
        return;
    }
    __shared__ u32 sh_t;
    for (;;) {
        if (tid == 0) sh_t = atomicAdd(ctr, 1u);
        __syncthreads();
        const u32 t = __builtin_amdgcn_readfirstlane(sh_t);
        if (t >= ntasks) break;
        const BatchTask tk = tasks[t];
        const u32 pi = __builtin_amdgcn_readfirstlane(tk.prod);
        const u32 x = __builtin_amdgcn_readfirstlane(tk.x), y = __builtin_amdgcn_readfirstlane(tk.y);
        batch_task<PH>(prods[pi], x, y, lds, tid);
        // the next claim overwrites sh_t and the next tile reuses the LDS
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// template dispatch
// ---------------------------------------------------------------------------
template <bool INV>
hipError_t col_dispatch(int nb, dim3 grid, dim3 block, size_t lds, hipStream_t s, const DevPlan& D, u32* R0, u32* R1,
                        int tc, int s0, int rest, int tiles_log) {
    switch (nb) {
#define PM_COL(n)                                                                                                    \
    case n:                                                                                                          \
        if (tc == 32)                                                                                                \
            hipLaunchKernelGGL((k_col<n, INV, 32>), grid, block, lds, s, D, R0, R1, tc, s0, rest, tiles_log);        \
        else                                                                                                         \
            hipLaunchKernelGGL((k_col<n, INV, 0>), grid, block, lds, s, D, R0, R1, tc, s0, rest, tiles_log);         \
        break;
        PM_COL(1) PM_COL(2) PM_COL(3) PM_COL(4) PM_COL(5) PM_COL(6) PM_COL(7) PM_COL(8)
#undef PM_COL
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

template <int MODE>
hipError_t row_dispatch(int nb, dim3 grid, dim3 block, size_t lds, hipStream_t s, const DevPlan& D, u32* RB,
                        const u32* RA, int rpb) {
    switch (nb) {
#define PM_ROW(n) \
    case n: hipLaunchKernelGGL((k_row<n, MODE>), grid, block, lds, s, D, RB, RA, rpb); break;
        PM_ROW(0) PM_ROW(1) PM_ROW(2) PM_ROW(3) PM_ROW(4) PM_ROW(5) PM_ROW(6) PM_ROW(7) PM_ROW(8)
#undef PM_ROW
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

int line_threads(int nb) { return nb >= 4 ? 1 << (nb - 4) : 1; }

}  // namespace

// The finish over `count` coefficients of `shards` slot arrays (shard_stride
// apart, stride N per slot), DIRECT when the slots do not fit the LDS.
static hipError_t launch_combine(const DevPlan& D, const LaunchCfg& C, const long long* P, int shards,
                                 size_t shard_stride, u32* out, u32 count, hipStream_t s) {
    const dim3 grid((count + C.crt_tn - 1) / C.crt_tn), block(256);
    const size_t lds = crt_combine_lds_bytes(C.crt_tn, D.L, C.combine_direct);
    if (C.combine_direct) {
        if (C.crt_tn != 32) return hipErrorInvalidValue;
        hipLaunchKernelGGL((k_crt_combine<1, true>), grid, block, lds, s, D, P, shards, shard_stride, out, count);
        return hipGetLastError();
    }
    switch (C.crt_tn) {
        case 128: hipLaunchKernelGGL((k_crt_combine<4, false>), grid, block, lds, s, D, P, shards, shard_stride, out, count); break;
        case 64: hipLaunchKernelGGL((k_crt_combine<2, false>), grid, block, lds, s, D, P, shards, shard_stride, out, count); break;
        case 32: hipLaunchKernelGGL((k_crt_combine<1, false>), grid, block, lds, s, D, P, shards, shard_stride, out, count); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

// PM_RNS_TN (32 .. 256) / PM_CRT_TN (32 .. 128) set the coefficients per
// RNS / CRT block, for tuning; the LDS limit still applies.
static int tile_cap(const char* name, int dflt, int maxv) {
    const char* v = std::getenv(name);
    const int t = v ? std::atoi(v) : dflt;
    return (t == 32 || t == 64 || t == 128 || t == 256) && t <= maxv ? t : dflt;
}

LaunchCfg choose_launch(const DevPlan& D, size_t max_lds, size_t device_lds) {
    LaunchCfg C;
    const u32 nrows = 1u << (D.logN - D.w2), ncols = 1u << D.w2;  // rows of the row pass, and their length
    C.rns_stride = 4 * D.L;
    C.rns_tn = D.L > kRnsNarrowL ? 32 : tile_cap("PM_RNS_TN", 128, 256);  // the 64-bit epilogue runs 32
    while (C.rns_tn > 32 && rns_lds_bytes(C.rns_tn, D.L, D.rns_rt) > max_lds) C.rns_tn /= 2;
    C.rns_lds = rns_lds_bytes(C.rns_tn, D.L, D.rns_rt);
    // 4-wave RNS blocks by default: four blocks per CU overlap their tile
    // loads better than two 8-wave blocks (PM_RNS_THREADS=512 for the latter).
    if (const char* v = std::getenv("PM_RNS_THREADS")) C.rns_threads = std::atoi(v) == 512 ? 512 : 256;
    C.crt_ys = 32 * D.crt_ks + 16;
    // 32 coefficients per CRT block by default: half the LDS of 64, so more
    // blocks per CU hide the latency of the residue loads (measured faster at
    // N=65536).
    C.crt_tn = tile_cap("PM_CRT_TN", 32, 128);
    auto crt_need = [&](int tn) { return crt_lds_bytes(tn, C.crt_ys, D.L, D.crt_ks); };
    while (C.crt_tn > 32 && crt_need(C.crt_tn) > max_lds) C.crt_tn /= 2;
    C.crt_lds = crt_need(C.crt_tn);
    // Wide points whose CRT block exceeds the device's LDS split the CRT into
    // the linear part (y and the GEMM, slots to global memory) and the finish.
    C.crt_split = C.crt_lds > device_lds;
    if (C.crt_split) {
        C.crt_tn = 32;
        C.crt_lds = crt_partial_lds_bytes(32, C.crt_ys, D.crt_ks);
    }
    C.combine_direct = crt_combine_lds_bytes(C.crt_tn, D.L, false) > device_lds;
    // 32-coefficient blocks with at most two row tiles per wave put the slots
    // over the y bytes (k_crt_alias); PM_CRT_ALIAS=0 keeps the separate slots.
    // 64 coefficients per block of 8 waves where that fits the LDS budget
    // (measured 20.5-20.7 us against 22.7-22.9 us for 32 at 65536/1747; 128
    // per block, 65 KB of LDS, ran 20.1-20.6 us but cost 2% of three-lane
    // throughput). PM_CRT_ALIAS=32 keeps 32, PM_CRT_ALIAS=0 the separate slots.
    const char* ca = std::getenv("PM_CRT_ALIAS");
    C.crt_alias = !C.crt_split && C.crt_tn == 32 && D.crt_rt <= 8 && !(ca && ca[0] == '0');
    if (C.crt_alias) {
        if (!(ca && std::atoi(ca) == 32) && crt_alias_lds_bytes(64, C.crt_ys, D.L, D.crt_ks) <= max_lds) C.crt_tn = 64;
        C.crt_lds = crt_alias_lds_bytes(C.crt_tn, C.crt_ys, D.L, D.crt_ks);
    }
    C.col_tc = (int)std::min<u32>(ncols, 32);
    C.row_rpb = (int)std::min<u32>(nrows, (u32)(256 / line_threads(D.w2)));
    return C;
}

// A column window's staging range starts on a window and, within the
// launch, keeps every tile inside one run of CW coefficients; the residues of
// staging position `first` then sit first / 2^w1 positions on.
static bool window_ok(const DevPlan& D, int lcw, int tile, u32 first, u32 count) {
    if (lcw == 0) return true;
    const u32 win = 1u << (lcw + D.w1);
    return D.npass == 2 && lcw <= D.w2 && (1 << lcw) >= tile && first % win == 0 && count % (u32)tile == 0;
}
static u32 residue_first(const DevPlan& D, int lcw, u32 first) { return lcw ? first >> D.w1 : first; }

hipError_t launch_rns(const DevPlan& D, const LaunchCfg& C, const u32* a0, const u32* a1, u32* R0, u32* R1, int nops,
                      hipStream_t s, u32 first, u32 count, int lcw) {
    if (count == 0) count = D.N - first;
    if (first + count > D.N || !window_ok(D, lcw, C.rns_tn, first, count)) return hipErrorInvalidValue;
    const dim3 grid((count + C.rns_tn - 1) / C.rns_tn, nops), block(C.rns_threads);
    const size_t L = (size_t)D.L;
    const u32 rfirst = residue_first(D, lcw, first);
    a0 += first * L;
    R0 += rfirst;
    if (nops > 1) {
        a1 += first * L;
        R1 += rfirst;
    }
    const int vec = ((uintptr_t)a0 % 16 == 0) && (nops < 2 || (uintptr_t)a1 % 16 == 0);
    if (D.L > kRnsNarrowL) {  // 64-bit epilogue; such rows fill the LDS at 32 coefficients
        if (C.rns_tn != 32) return hipErrorInvalidValue;
        hipLaunchKernelGGL((k_rns<1, true>), grid, block, C.rns_lds, s, D, a0, a1, R0, R1, count, vec, lcw);
        return hipGetLastError();
    }
    switch (C.rns_tn) {
        case 256: hipLaunchKernelGGL((k_rns<8, false>), grid, block, C.rns_lds, s, D, a0, a1, R0, R1, count, vec, lcw); break;
        case 128: hipLaunchKernelGGL((k_rns<4, false>), grid, block, C.rns_lds, s, D, a0, a1, R0, R1, count, vec, lcw); break;
        case 64: hipLaunchKernelGGL((k_rns<2, false>), grid, block, C.rns_lds, s, D, a0, a1, R0, R1, count, vec, lcw); break;
        case 32: hipLaunchKernelGGL((k_rns<1, false>), grid, block, C.rns_lds, s, D, a0, a1, R0, R1, count, vec, lcw); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_col(const DevPlan& D, const LaunchCfg& C, u32* R0, u32* R1, int nops, bool inverse, hipStream_t s,
                      u32 col_first, u32 col_count) {
    if (col_count) {
        // columns [col_first, col_first + col_count) of a two-pass plan: the
        // tiles of the one strided pass that cover them (its twiddles depend on
        // the stage and the row, not on the column)
        const u32 tc = (u32)std::min<u32>((u32)C.col_tc, 1u << D.w2);
        if (D.npass != 2 || col_first % tc || col_count % tc || col_first + col_count > (1u << D.w2))
            return hipErrorInvalidValue;
        const int nb = D.pw[0], T = line_threads(nb);
        const u32 tiles = col_count / tc;
        int tiles_log = 0;
        while ((1u << tiles_log) < tiles) ++tiles_log;
        R0 += col_first;
        if (nops > 1) R1 += col_first;
        const dim3 grid(tiles, D.K, nops), block(tc * T);
        const size_t lds = col_lds_bytes(nb, (int)tc);
        return inverse ? col_dispatch<true>(nb, grid, block, lds, s, D, R0, R1, (int)tc, 0, D.w2, tiles_log)
                       : col_dispatch<false>(nb, grid, block, lds, s, D, R0, R1, (int)tc, 0, D.w2, tiles_log);
    }
    // the strided passes 0 .. npass - 2 (forward in order, inverse in reverse)
    const int last = D.npass - 1;
    for (int k = 0; k < last; ++k) {
        const int pi = inverse ? last - 1 - k : k;
        const int nb = D.pw[pi];
        if (nb == 0) continue;
        int s0 = 0;
        for (int i = 0; i < pi; ++i) s0 += D.pw[i];
        const int rest = D.logN - s0 - nb;
        const int tc = (int)std::min<u32>((u32)C.col_tc, 1u << rest);
        int tiles_log = 0;
        while ((tc << tiles_log) < (1 << rest)) ++tiles_log;
        const int T = line_threads(nb);
        const dim3 grid((1u << s0) << tiles_log, D.K, nops), block(tc * T);
        const size_t lds = col_lds_bytes(nb, tc);
        const hipError_t e = inverse ? col_dispatch<true>(nb, grid, block, lds, s, D, R0, R1, tc, s0, rest, tiles_log)
                                     : col_dispatch<false>(nb, grid, block, lds, s, D, R0, R1, tc, s0, rest, tiles_log);
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

hipError_t launch_row(const DevPlan& D, const LaunchCfg& C, u32* RB, const u32* RA, RowMode mode, hipStream_t s) {
    const u32 nrows = 1u << (D.logN - D.w2);
    const int T = line_threads(D.w2);
    const dim3 grid(nrows / C.row_rpb, D.K), block(C.row_rpb * T);
    const size_t rs = ((size_t)1 << D.w2) + (D.w2 > 4 ? 16 : 0);
    const size_t lds = D.w2 > 4 ? (size_t)C.row_rpb * rs * 4 : 0;
    switch (mode) {
        case ROW_FUSED: return row_dispatch<ROW_FUSED>(D.w2, grid, block, lds, s, D, RB, RA, C.row_rpb);
        case ROW_FWD: return row_dispatch<ROW_FWD>(D.w2, grid, block, lds, s, D, RB, RA, C.row_rpb);
        case ROW_INV: return row_dispatch<ROW_INV>(D.w2, grid, block, lds, s, D, RB, RA, C.row_rpb);
        case ROW_FUSED_PRE: return row_dispatch<ROW_FUSED_PRE>(D.w2, grid, block, lds, s, D, RB, RA, C.row_rpb);
    }
    return hipErrorInvalidValue;
}

hipError_t launch_crt(const DevPlan& D, const LaunchCfg& C, const u32* R, u32* out, const uint2* G, hipStream_t s,
                      u32 first, u32 count, long long* part, int lcw) {
    if (count == 0) count = D.N - first;
    if (first + count > D.N) return hipErrorInvalidValue;
    if (lcw) {  // column windows: the aliased CRT only
        if (!C.crt_alias || !window_ok(D, lcw, C.crt_tn, first, count)) return hipErrorInvalidValue;
        const int koff = (int)crt_alias_kc_off(C.crt_tn, C.crt_ys, D.L);
        const dim3 grid(count / C.crt_tn);
        R += residue_first(D, lcw, first);
        out += (size_t)first * D.L;
        if (C.crt_tn == 64)
            hipLaunchKernelGGL((k_crt_alias<2, 8>), grid, dim3(512), C.crt_lds, s, D, R, out, G, count, C.crt_ys, koff,
                               lcw);
        else
            hipLaunchKernelGGL((k_crt_alias<1, 8>), grid, dim3(256), C.crt_lds, s, D, R, out, G, count, C.crt_ys, koff,
                               lcw);
        return hipGetLastError();
    }
    const dim3 grid((count + C.crt_tn - 1) / C.crt_tn), block(256);
    if (C.crt_split) {
        // slots of [first, first + count) into part (stride N per slot), then the finish
        if (!part) return hipErrorInvalidValue;
        hipLaunchKernelGGL((k_crt_partial<1>), grid, block, C.crt_lds, s, D, R + first, G, part + first, count,
                           C.crt_ys);
        if (const hipError_t e = hipGetLastError(); e != hipSuccess) return e;
        return launch_combine(D, C, part + first, 1, 0, out + (size_t)first * D.L, count, s);
    }
    R += first;
    out += (size_t)first * D.L;
    if (C.crt_alias) {
        const int koff = (int)crt_alias_kc_off(C.crt_tn, C.crt_ys, D.L);
        if (C.crt_tn == 64)
            hipLaunchKernelGGL((k_crt_alias<2, 8>), grid, dim3(512), C.crt_lds, s, D, R, out, G, count, C.crt_ys, koff, 0);
        else
            hipLaunchKernelGGL((k_crt_alias<1, 8>), grid, block, C.crt_lds, s, D, R, out, G, count, C.crt_ys, koff, 0);
        return hipGetLastError();
    }
    switch (C.crt_tn) {
        case 128: hipLaunchKernelGGL((k_crt<4>), grid, block, C.crt_lds, s, D, R, out, G, count, C.crt_ys); break;
        case 64: hipLaunchKernelGGL((k_crt<2>), grid, block, C.crt_lds, s, D, R, out, G, count, C.crt_ys); break;
        case 32: hipLaunchKernelGGL((k_crt<1>), grid, block, C.crt_lds, s, D, R, out, G, count, C.crt_ys); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_crt_partial(const DevPlan& D, const LaunchCfg& C, const u32* R, long long* P, hipStream_t s) {
    const dim3 grid((D.N + C.crt_tn - 1) / C.crt_tn), block(256);
    const size_t lds = crt_partial_lds_bytes(C.crt_tn, C.crt_ys, D.crt_ks);
    switch (C.crt_tn) {
        case 128: hipLaunchKernelGGL((k_crt_partial<4>), grid, block, lds, s, D, R, D.G, P, D.N, C.crt_ys); break;
        case 64: hipLaunchKernelGGL((k_crt_partial<2>), grid, block, lds, s, D, R, D.G, P, D.N, C.crt_ys); break;
        case 32: hipLaunchKernelGGL((k_crt_partial<1>), grid, block, lds, s, D, R, D.G, P, D.N, C.crt_ys); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_crt_combine(const DevPlan& D, const LaunchCfg& C, const long long* P, int shards, size_t shard_stride,
                              u32* out, hipStream_t s) {
    return launch_combine(D, C, P, shards, shard_stride, out, D.N, s);
}

hipError_t launch_copy_probe(const void* src, void* dst, size_t bytes, int grid, hipStream_t s) {
    if (bytes % 16) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_copy_probe, dim3(grid), dim3(256), 0, s, static_cast<const uint4*>(src),
                       static_cast<uint4*>(dst), bytes / 16);
    return hipGetLastError();
}

hipError_t launch_valu_probe(int op, u32* out, int grid, int iters, hipStream_t s) {
    switch (op) {
        case 0: hipLaunchKernelGGL((k_valu_probe<0>), dim3(grid), dim3(256), 0, s, out, 12345u, iters); break;
        case 1: hipLaunchKernelGGL((k_valu_probe<1>), dim3(grid), dim3(256), 0, s, out, 12345u, iters); break;
        case 2: hipLaunchKernelGGL((k_valu_probe<2>), dim3(grid), dim3(256), 0, s, out, 12345u, iters); break;
        case 3: hipLaunchKernelGGL((k_valu_probe<3>), dim3(grid), dim3(256), 0, s, out, 12345u, iters); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_batch_reset(u32* ctl, u32 nctl, hipStream_t s) {
    hipLaunchKernelGGL(k_batch_reset, dim3((nctl + 255) / 256), dim3(256), 0, s, ctl, nctl);
    return hipGetLastError();
}

hipError_t launch_batch_phase(int phase, const BatchProduct* prods, const BatchTask* tasks, u32 ntasks, u32* ctr,
                              int grid, size_t lds, hipStream_t s) {
    if (ntasks == 0) return hipSuccess;
    const dim3 g(ctr ? (u32)grid : ntasks), b(kBatchThreads[phase]);
    switch (phase) {
#define PM_PH(n) \
    case n: hipLaunchKernelGGL((k_batch_phase<n>), g, b, lds, s, prods, tasks, ntasks, ctr); break;
        PM_PH(PH_RNS) PM_PH(PH_COLF) PM_PH(PH_ROW) PM_PH(PH_COLI) PM_PH(PH_CRT)
#undef PM_PH
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

}  // namespace pm
