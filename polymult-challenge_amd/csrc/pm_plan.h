// Host-side plan for one point (N, W, L, q): validation, RNS primes and every
// constant table the device kernels read. Pure C++, no HIP: the tables are
// uploaded by pm_api.hip.
#pragma once

#include <cstddef>
#include <cstdint>
#include <vector>

namespace pm {

// Limits. N: primes p = 1 (mod 2N) below 2^31 exist only for 2N < 2^31, and
// four NTT passes of at most 8 stages cover 2^30. K < 1024, as the reference
// (solve.cu:911-913), which bounds W (about 15,800 bits at small N).
constexpr int kMaxLogN = 30;
constexpr int kMaxK = 1023;
// The RNS epilogue sums its 4L-term digit products in 32 bits two at a time
// (4L 128^2 257 < 2^31) up to L = 127, W <= 4064; wider points take the 64-bit
// epilogue. The batched and served forms keep to L <= kNarrowL.
constexpr uint32_t kNarrowL = 127;

struct HostPlan {
    uint32_t N = 0, W = 0, L = 0;
    int logN = 0, w1 = 0, w2 = 0;  // 4-step split: nrows = 2^w1, ncols = 2^w2 (two passes)
    // NTT passes of pw[i] stages each, the last the row pass (w1 = pw[0], w2 =
    // pw[npass-1]): two up to N = 2^16, then ceil(logN / 8).
    int npass = 2;
    int pw[4] = {0, 0, 0, 0};
    int K = 0;                      // RNS primes
    uint64_t delta = 0;             // 2^W - q
    std::vector<uint32_t> q;        // [L]
    std::vector<uint32_t> p, pinv;  // [K]; pinv = -p^-1 mod 2^32
    std::vector<uint32_t> psi;      // [K] primitive 2N-th roots

    // NTT: tw[j][k] = w', the Shoup companion floor(w 2^32 / p_j) of
    // w = psi_j^bitrev_logN(k) mod p_j; w = floor(w' p_j / 2^32) + 1. [K][N]
    std::vector<uint32_t> tw;

    // RNS GEMM (int8 MFMA, 32x32x32): A fragments in lane order,
    // [rns_rt][rns_ks][64 lanes][16 bytes]. Row m = 4 j + t holds the
    // balanced base-256 digit t of c_jb = 2^(8b+32) mod p_j (symmetric
    // representative), column k = byte b of the coefficient.
    std::vector<int8_t> rns_frag;
    std::vector<uint64_t> rns_bias;  // [K] 128 sum_b c_jb mod p_j + p_j 2^20
    int rns_rt = 0, rns_ks = 0;

    // CRT GEMM: rows are digit positions of z (4L + 4 of them) and of the
    // v estimate F (6, padded to 8); columns k = 4 j + s index byte s of y_j.
    std::vector<int8_t> crt_frag;    // [crt_rt][crt_ks][64][16]
    int crt_rt = 0, crt_ks = 0;
    std::vector<uint32_t> G;         // [K][2] Shoup pair of N^-1 2^32 (M/p_j)^-1
    std::vector<uint32_t> Gplain;    // [K][2] Shoup pair of (M/p_j)^-1 (stage tests)
    std::vector<uint32_t> C0;        // [L + 2] 128 * 0x01010101 * sum_j Q_j
    int64_t F0 = 0;                  // 128 * 0x01010101 * sum_j floor(2^48 / p_j)
    std::vector<uint32_t> T;         // [K + 1][L] (-v M) mod q
};

// Validates the point and chooses the primes. Throws std::runtime_error with
// the reference's messages (solve.cu:873-913).
void plan_point(HostPlan& P, uint32_t N, uint32_t W, uint32_t L, const uint32_t* q, size_t q_limbs);

// Builds the RNS and CRT tables of a planned point, and its own NTT twiddles
// unless `twiddles` is false (batched products share one table).
void build_tables(HostPlan& P, bool twiddles = true);

// The prime set shared by every product of a batch: p = 1 + k 2^17 < 2^31 for
// descending k, so p = 1 (mod 2N) for every N <= 65536, with psi of order
// 2^17. One twiddle table then serves every N: psi_N = psi^(65536/N) gives
// psi_N^bitrev_logN(k) = psi^bitrev_16(k) for k < N.
struct SharedPrimes {
    std::vector<uint32_t> p, pinv, psi;
};
constexpr uint32_t kSharedN = 65536;
void shared_primes(SharedPrimes& S, int count);
// Plans a product over the shared primes: the same validation and K rule as
// plan_point (the first K primes with sum log2 p >= 2W + logN + 3).
void plan_point_shared(HostPlan& P, uint32_t N, uint32_t W, uint32_t L, const uint32_t* q, size_t q_limbs,
                       const SharedPrimes& S);
// tw[j][k] (Shoup companions, as HostPlan::tw) for j < K, k < 65536 over the shared primes.
void build_shared_twiddles(const SharedPrimes& S, int K, std::vector<uint32_t>& tw);

}  // namespace pm
