/* pm_oracle.h -- CPU restatement of the reference's negacyclic multiply.
 *
 * TEST INFRASTRUCTURE ONLY. This is the checker that the HIP path is compared
 * against; the product never links, loads or calls it. Only tests/, the smoke
 * check and bench.py's CPU-baseline leg may use it.
 *
 * c = a*b mod (X^N + 1, q), W-bit coefficients as L little-endian u32 limbs,
 * computed the way the reference does it (solve.cu:5-15): RNS over K 31-bit
 * primes p = 1 (mod 2N), a merged-psi negacyclic NTT per prime, an exact CRT
 * straight to mod q with the 2^W = delta (mod q) fold.
 *
 * Every intermediate is canonical (values in [0, p) or [0, q)), so each stage
 * can be compared bit for bit with the GPU path's stage dumps.
 */
#ifndef PM_ORACLE_H
#define PM_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct pmo_plan pmo_plan;

/* Validates the point and builds every table (solve.cu:869-1020).
 * Returns 0, or -1 with a message in err (the reference's own wording). */
int pmo_plan_create(uint32_t N, uint32_t W, uint32_t L, const uint32_t* q,
                    pmo_plan** out, char* err, size_t errlen);
void pmo_plan_free(pmo_plan* P);

/* K, delta = 2^W - q, and optionally the primes [K] and roots psi [K]. */
int pmo_plan_info(const pmo_plan* P, uint32_t* K, uint64_t* delta,
                  uint32_t* primes, uint32_t* psi);

/* R[j][n] = a_n mod p_j. a is [N][L]; R is [K][N]. (solve.cu:147-202) */
void pmo_rns(const pmo_plan* P, const uint32_t* a, uint32_t* R);
/* In-place forward negacyclic NTT of every prime row of R [K][N]:
 * Cooley-Tukey, twiddle tw(k) = psi^bitrev(k), output in bit-reversed
 * order. (solve.cu:204-244, stages 0..logN-1) */
void pmo_ntt_fwd(const pmo_plan* P, uint32_t* R);
/* In-place inverse, Gentleman-Sande with psi^-bitrev(k), times N^-1. */
void pmo_ntt_inv(const pmo_plan* P, uint32_t* R);
/* C = A * B mod p_j elementwise, [K][N]. (solve.cu:374) */
void pmo_pointwise(const pmo_plan* P, const uint32_t* A, const uint32_t* B, uint32_t* C);
/* From X[j][n] = x_n mod p_j (x the integer negacyclic product) to
 * c_n = x_n mod q, [N][L]. (solve.cu:380-501) Returns the number of
 * coefficients whose CRT sum was not within 1/8 of an integer: nonzero only
 * for operands outside [0, q), where the result is not defined. */
int pmo_crt(const pmo_plan* P, const uint32_t* X, uint32_t* c);
/* The whole multiply; returns what pmo_crt returns. */
int pmo_multiply(const pmo_plan* P, const uint32_t* a, const uint32_t* b, uint32_t* c);

#ifdef __cplusplus
}
#endif

#endif
