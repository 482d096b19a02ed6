/* pm_oracle.c -- CPU restatement of the reference's negacyclic multiply.
 *
 * TEST INFRASTRUCTURE ONLY (the checker). Nothing in the product links or
 * calls this file; see pm_oracle.h.
 *
 * Pinned by tests/test_oracle.py against: schoolbook negacyclic products
 * (N <= 256), an independent Kronecker-substitution bigint product, the
 * all-(q-1) closed form c_k = (2k + 2 - N) mod q, sparse/monomial brute force,
 * textbook O(N^2) negacyclic DFTs per prime, bigint residues for the RNS stage,
 * a bigint CRT for the CRT stage, and the reference binary itself (oracle/_ref)
 * on point directories.
 */
#include "pm_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

typedef uint32_t u32;
typedef uint64_t u64;
typedef unsigned __int128 u128;

/* PMO_MUTATION (0 by default: none) plants one deliberate mistake, so that
 * tests/test_oracle.py can show that the independent checks pinning this file
 * catch such slips:
 *   1 one prime fewer than the K rule
 *   2 the inverse butterfly's difference taken with the wrong sign
 *   3 the T[v] = (-v M) mod q term dropped from the CRT
 *   4 the N^-1 scaling dropped from the inverse NTT
 *   5 the final conditional subtraction of q dropped
 *   6 a's transform squared instead of multiplied by b's
 *   7 the fold adding z >> W without the factor delta */
#ifndef PMO_MUTATION
#define PMO_MUTATION 0
#endif

struct pmo_plan {
    u32 N, W, L;
    int logN, K;
    u64 delta;
    u32* q;      /* [L] */
    u32* p;      /* [K] */
    u32* psi;    /* [K] */
    u32* tw;     /* [K][N] psi^bitrev(k) mod p */
    u32* itw;    /* [K][N] psi^-bitrev(k) mod p */
    u32* ninv;   /* [K] N^-1 mod p */
    u32* mpinv;  /* [K] (M/p_j)^-1 mod p_j */
    u32* Q;      /* [K][L] (M/p_j) mod q */
    u32* T;      /* [K+1][L] (-v M) mod q */
    u32* c32;    /* [K][L] 2^(32 l) mod p_j */
};

/* ---- word arithmetic (solve.cu:509-537) ---- */

static u32 powmod(u64 b, u64 e, u32 m) {
    u64 r = 1 % m, x = b % m;
    while (e) {
        if (e & 1) r = r * x % m;
        x = x * x % m;
        e >>= 1;
    }
    return (u32)r;
}

/* Trial division by the primes to 37, then Miller-Rabin to bases 2, 7, 61,
 * which is deterministic below 4,759,123,141 (solve.cu:519-537). */
static int is_prime_u32(u32 n) {
    static const u32 small[] = {2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37};
    static const u32 bases[] = {2, 7, 61};
    if (n < 2) return 0;
    for (size_t i = 0; i < sizeof small / sizeof small[0]; ++i)
        if (n % small[i] == 0) return n == small[i];
    u32 d = n - 1;
    int s = 0;
    while (!(d & 1)) { d >>= 1; ++s; }
    for (size_t i = 0; i < 3; ++i) {
        u64 x = powmod(bases[i], d, n);
        if (x == 1 || x == n - 1) continue;
        int composite = 1;
        for (int r = 1; r < s; ++r) {
            x = x * x % n;
            if (x == n - 1) { composite = 0; break; }
        }
        if (composite) return 0;
    }
    return 1;
}

static u32 bitrev(u32 k, int bits) {
    u32 r = 0;
    for (int b = 0; b < bits; ++b) r |= ((k >> b) & 1u) << (bits - 1 - b);
    return r;
}

/* ---- multi-limb arithmetic, little-endian u32 limbs ---- */

static int big_bitlen(const u32* a, int n) {
    for (int i = n - 1; i >= 0; --i)
        if (a[i]) return 32 * i + 32 - __builtin_clz(a[i]);
    return 0;
}

static int big_cmp(const u32* a, const u32* b, int n) {
    for (int i = n - 1; i >= 0; --i)
        if (a[i] != b[i]) return a[i] > b[i] ? 1 : -1;
    return 0;
}

/* a -= b over n limbs (a >= b). */
static void big_sub(u32* a, const u32* b, int n) {
    u64 borrow = 0;
    for (int i = 0; i < n; ++i) {
        const u64 d = (u64)a[i] - b[i] - borrow;
        a[i] = (u32)d;
        borrow = (d >> 63) & 1u;
    }
}

/* r = x mod q by binary long division; x has nx limbs, q has L limbs and
 * r has L limbs (solve.cu:559-587 restated). */
static void big_mod(const u32* x, int nx, const u32* q, int L, u32* r) {
    u32* acc = (u32*)calloc((size_t)L + 1, 4);
    u32* qq = (u32*)calloc((size_t)L + 1, 4);
    memcpy(qq, q, (size_t)L * 4);
    for (int i = big_bitlen(x, nx) - 1; i >= 0; --i) {
        u32 carry = (x[i / 32] >> (i % 32)) & 1u;
        for (int k = 0; k <= L; ++k) {
            const u32 v = acc[k];
            acc[k] = (v << 1) | carry;
            carry = v >> 31;
        }
        if (big_cmp(acc, qq, L + 1) >= 0) big_sub(acc, qq, L + 1);
    }
    memcpy(r, acc, (size_t)L * 4);
    free(acc);
    free(qq);
}

/* r = (a * m) mod q, a an L-limb residue, m a word. */
static void big_mulword_mod(const u32* a, u32 m, const u32* q, int L, u32* r) {
    u32* t = (u32*)calloc((size_t)L + 1, 4);
    u64 carry = 0;
    for (int i = 0; i < L; ++i) {
        const u64 s = (u64)a[i] * m + carry;
        t[i] = (u32)s;
        carry = s >> 32;
    }
    t[L] = (u32)carry;
    big_mod(t, L + 1, q, L, r);
    free(t);
}

static void set_err(char* err, size_t n, const char* msg) {
    if (err && n) snprintf(err, n, "%s", msg);
}

void pmo_plan_free(pmo_plan* P) {
    if (!P) return;
    free(P->q); free(P->p); free(P->psi); free(P->tw); free(P->itw);
    free(P->ninv); free(P->mpinv); free(P->Q); free(P->T); free(P->c32);
    free(P);
}

/* Point validation, delta, prime selection and tables: solve.cu:869-1020. */
int pmo_plan_create(u32 N, u32 W, u32 L, const u32* q, pmo_plan** out, char* err, size_t errlen) {
    *out = NULL;
    if (N == 0 || (N & (N - 1)) != 0) { set_err(err, errlen, "N must be a power of two"); return -1; }
    if (W < 2 || L != (W + 31) / 32) { set_err(err, errlen, "inconsistent W / L"); return -1; }
    if (!q) { set_err(err, errlen, "q must have L limbs"); return -1; }
    const int qb = big_bitlen(q, (int)L);
    if (qb > (int)W || qb < 2 || !(q[0] & 1u)) { set_err(err, errlen, "q out of range"); return -1; }
    if (qb <= 32 * ((int)L - 1)) { set_err(err, errlen, "q top limb is zero"); return -1; }
    int logN = 0;
    while ((1u << logN) < N) ++logN;

    /* delta = 2^W - q, at most 63 bits and below 2^(W-1) (solve.cu:883-897). */
    u32* dl = (u32*)calloc((size_t)L + 1, 4);
    dl[W / 32] |= 1u << (W % 32);
    {
        u32* qx = (u32*)calloc((size_t)L + 1, 4);
        memcpy(qx, q, (size_t)L * 4);
        big_sub(dl, qx, (int)L + 1);
        free(qx);
    }
    const int dbits = big_bitlen(dl, (int)L + 1);
    if (dbits > 63 || dbits > (int)W - 1) {
        free(dl);
        set_err(err, errlen, "q too far below 2^W for this reduction");
        return -1;
    }
    const u64 delta = (u64)dl[0] | ((L + 1 > 1) ? ((u64)dl[1] << 32) : 0);
    free(dl);

    /* Primes p = 1 + k 2N < 2^31 for descending k until sum log2 p >= 2W + logN + 3
     * (solve.cu:899-913): prod p > 2^(2W+logN+3) > 8 N q^2. */
    const double need = 2.0 * W + logN + 3;
    const u64 step = 2ull * N;
    u32* primes = (u32*)malloc(sizeof(u32) * 1024);
    int K = 0;
    double bits = 0;
    for (u64 k = ((1ull << 31) - 2) / step; k >= 1 && bits < need + 1e-6 && K < 1024; --k) {
        const u64 pp = 1 + k * step;
        if (pp < (1ull << 31) && is_prime_u32((u32)pp)) {
            primes[K++] = (u32)pp;
            bits += log2((double)pp);
        }
    }
    if (bits < need + 1e-6) {
        free(primes);
        set_err(err, errlen, K >= 1024 ? "too many RNS primes" : "not enough NTT primes for this point");
        return -1;
    }

#if PMO_MUTATION == 1
    if (K > 1) --K;
#endif
    pmo_plan* P = (pmo_plan*)calloc(1, sizeof(pmo_plan));
    P->N = N; P->W = W; P->L = L; P->logN = logN; P->K = K; P->delta = delta;
    P->q = (u32*)malloc((size_t)L * 4);
    memcpy(P->q, q, (size_t)L * 4);
    P->p = primes;
    P->psi = (u32*)malloc((size_t)K * 4);
    P->tw = (u32*)malloc((size_t)K * N * 4);
    P->itw = (u32*)malloc((size_t)K * N * 4);
    P->ninv = (u32*)malloc((size_t)K * 4);
    P->mpinv = (u32*)malloc((size_t)K * 4);
    P->Q = (u32*)calloc((size_t)K * L, 4);
    P->T = (u32*)calloc((size_t)(K + 1) * L, 4);
    P->c32 = (u32*)malloc((size_t)K * L * 4);

    u32* pw = (u32*)malloc((size_t)N * 4);
    u32* ipw = (u32*)malloc((size_t)N * 4);
    for (int j = 0; j < K; ++j) {
        const u32 p = primes[j];
        /* psi: the first g = 2, 3, ... whose g^((p-1)/2N) has order 2N
         * (solve.cu:942-946). */
        u32 psi = 0;
        for (u32 g = 2;; ++g) {
            const u32 x = powmod(g, (p - 1) / step, p);
            if (powmod(x, N, p) == p - 1) { psi = x; break; }
        }
        P->psi[j] = psi;
        const u32 ipsi = powmod(psi, p - 2, p);
        pw[0] = 1; ipw[0] = 1;
        for (u32 e = 1; e < N; ++e) {
            pw[e] = (u32)((u64)pw[e - 1] * psi % p);
            ipw[e] = (u32)((u64)ipw[e - 1] * ipsi % p);
        }
        for (u32 k = 0; k < N; ++k) {
            const u32 r = bitrev(k, logN);
            P->tw[(size_t)j * N + k] = pw[r];
            P->itw[(size_t)j * N + k] = ipw[r];
        }
        P->ninv[j] = powmod(N % p, p - 2, p);
        u64 mp = 1;
        for (int i = 0; i < K; ++i)
            if (i != j) mp = mp * (primes[i] % p) % p;
        P->mpinv[j] = powmod(mp, p - 2, p);
        for (u32 l = 0; l < L; ++l) P->c32[(size_t)j * L + l] = powmod(2, 32ull * l, p);
    }
    free(pw);
    free(ipw);

    /* Q_j = (M/p_j) mod q by prefix and suffix products, T_v = (-v M) mod q
     * (solve.cu:969-1007). */
    u32* pre = (u32*)calloc((size_t)(K + 1) * L, 4);
    u32* suf = (u32*)calloc((size_t)(K + 1) * L, 4);
    pre[0] = 1;
    suf[(size_t)K * L] = 1;
    for (int j = 0; j < K; ++j) big_mulword_mod(pre + (size_t)j * L, primes[j], q, (int)L, pre + (size_t)(j + 1) * L);
    for (int j = K - 1; j >= 0; --j) big_mulword_mod(suf + (size_t)(j + 1) * L, primes[j], q, (int)L, suf + (size_t)j * L);
    u32* prod = (u32*)calloc((size_t)2 * L, 4);
    for (int j = 0; j < K; ++j) {
        const u32* a = pre + (size_t)j * L;
        const u32* b = suf + (size_t)(j + 1) * L;
        memset(prod, 0, (size_t)2 * L * 4);
        for (u32 i = 0; i < L; ++i) {
            u64 carry = 0;
            for (u32 k = 0; k < L; ++k) {
                const u64 t = (u64)a[i] * b[k] + prod[i + k] + carry;
                prod[i + k] = (u32)t;
                carry = t >> 32;
            }
            prod[i + L] = (u32)carry;
        }
        big_mod(prod, 2 * (int)L, q, (int)L, P->Q + (size_t)j * L);
    }
    free(prod);
    const u32* Mq = pre + (size_t)K * L;
    u32* acc = (u32*)calloc((size_t)L + 1, 4);
    u32* qx = (u32*)calloc((size_t)L + 1, 4);
    memcpy(qx, q, (size_t)L * 4);
    for (int v = 1; v <= K; ++v) {
        u64 carry = 0;
        for (u32 l = 0; l <= L; ++l) {
            const u64 s = (u64)acc[l] + (l < L ? Mq[l] : 0u) + carry;
            acc[l] = (u32)s;
            carry = s >> 32;
        }
        if (big_cmp(acc, qx, (int)L + 1) >= 0) big_sub(acc, qx, (int)L + 1);
        int zero = 1;
        for (u32 l = 0; l < L; ++l) zero = zero && acc[l] == 0;
        if (!zero) {
            memcpy(P->T + (size_t)v * L, q, (size_t)L * 4);
            big_sub(P->T + (size_t)v * L, acc, (int)L);
        }
    }
    free(acc);
    free(qx);
    free(pre);
    free(suf);
    *out = P;
    return 0;
}

int pmo_plan_info(const pmo_plan* P, u32* K, u64* delta, u32* primes, u32* psi) {
    if (K) *K = (u32)P->K;
    if (delta) *delta = P->delta;
    if (primes) memcpy(primes, P->p, (size_t)P->K * 4);
    if (psi) memcpy(psi, P->psi, (size_t)P->K * 4);
    return 0;
}

/* a_n mod p_j = sum_l limb_l (2^(32 l) mod p_j) mod p_j; the 128-bit sum
 * cannot overflow for L < 2^33. (solve.cu:147-202 computes the same residue.) */
void pmo_rns(const pmo_plan* P, const u32* a, u32* R) {
    const u32 N = P->N, L = P->L;
    for (int j = 0; j < P->K; ++j) {
        const u32 p = P->p[j];
        const u32* c = P->c32 + (size_t)j * L;
        u32* out = R + (size_t)j * N;
        for (u32 n = 0; n < N; ++n) {
            const u32* x = a + (size_t)n * L;
            u128 s = 0;
            for (u32 l = 0; l < L; ++l) s += (u64)x[l] * c[l];
            out[n] = (u32)(s % p);
        }
    }
}

/* Cooley-Tukey over stages s = 0..logN-1: blocks of 2t with t = N >> (s+1),
 * block i multiplied by tw(2^s + i). (solve.cu:204-244 and 254-378 split the
 * same loop into column and row passes.) */
void pmo_ntt_fwd(const pmo_plan* P, u32* R) {
    const u32 N = P->N;
    for (int j = 0; j < P->K; ++j) {
        const u32 p = P->p[j];
        const u32* tw = P->tw + (size_t)j * N;
        u32* x = R + (size_t)j * N;
        for (u32 m = 1, t = N; m < N; m <<= 1) {
            t >>= 1;
            for (u32 i = 0; i < m; ++i) {
                const u64 w = tw[m + i];
                const u32 j1 = 2 * i * t;
                for (u32 k = j1; k < j1 + t; ++k) {
                    const u32 U = x[k];
                    const u32 V = (u32)(x[k + t] * w % p);
                    x[k] = U + V >= p ? U + V - p : U + V;
                    x[k + t] = U >= V ? U - V : U + p - V;
                }
            }
        }
    }
}

/* Gentleman-Sande in reverse stage order with psi^-bitrev, then N^-1.
 * (The reference uses -tw of the mirrored index, psi^-bitrev(2^s + i) =
 * -psi^bitrev(2^s + 2^s - 1 - i), and folds N^-1 into its CRT constant.) */
void pmo_ntt_inv(const pmo_plan* P, u32* R) {
    const u32 N = P->N;
    for (int j = 0; j < P->K; ++j) {
        const u32 p = P->p[j];
        const u32* itw = P->itw + (size_t)j * N;
        u32* x = R + (size_t)j * N;
        for (u32 m = N, t = 1; m > 1; m >>= 1) {
            const u32 h = m >> 1;
            for (u32 i = 0; i < h; ++i) {
                const u64 w = itw[h + i];
                const u32 j1 = 2 * i * t;
                for (u32 k = j1; k < j1 + t; ++k) {
                    const u32 U = x[k], V = x[k + t];
                    x[k] = U + V >= p ? U + V - p : U + V;
#if PMO_MUTATION == 2
                    x[k + t] = (u32)((u64)(V >= U ? V - U : V + p - U) * w % p);
#else
                    x[k + t] = (u32)((u64)(U >= V ? U - V : U + p - V) * w % p);
#endif
                }
            }
            t <<= 1;
        }
        const u64 ni = P->ninv[j];
        if (PMO_MUTATION != 4)
            for (u32 k = 0; k < N; ++k) x[k] = (u32)(x[k] * ni % p);
    }
}

void pmo_pointwise(const pmo_plan* P, const u32* A, const u32* B, u32* C) {
    const u32 N = P->N;
    for (int j = 0; j < P->K; ++j) {
        const u64 p = P->p[j];
        for (u32 n = 0; n < N; ++n) {
            const size_t i = (size_t)j * N + n;
            C[i] = (u32)((u64)A[i] * (PMO_MUTATION == 6 ? A[i] : B[i]) % p);
        }
    }
}

/* CRT straight to mod q (solve.cu:380-501):
 *   y_j = x_j (M/p_j)^-1 mod p_j,  v = round(sum_j y_j / p_j),
 *   z = sum_j y_j [(M/p_j) mod q] + [(-v M) mod q],
 *   fold z = (z mod 2^W) + (z >> W) delta until z < 2^W, then one -q.
 * v is computed exactly here with 64 fractional bits per term (error below
 * K 2^-64), where the reference uses fp32; both are exact because the sum is
 * within 1/8 of an integer when |x| < M/8. */
int pmo_crt(const pmo_plan* P, const u32* X, u32* c) {
    const u32 N = P->N, L = P->L, W = P->W;
    const int K = P->K;
    const int ZL = (int)L + 4;
    u32* z = (u32*)malloc((size_t)ZL * 4);
    u32* y = (u32*)malloc((size_t)K * 4);
    int bad = 0;
    for (u32 n = 0; n < N; ++n) {
        u128 F = 0;
        for (int j = 0; j < K; ++j) {
            const u32 p = P->p[j];
            y[j] = (u32)((u64)X[(size_t)j * N + n] * P->mpinv[j] % p);
            F += ((u128)y[j] << 64) / p;
        }
        const u64 v = (u64)((F + ((u128)1 << 63)) >> 64);
        /* margin: the fraction must lie within 1/8 of the integer v */
        {
            const u128 V = (u128)v << 64;
            const u128 dist = F > V ? F - V : V - F;
            if (dist > ((u128)1 << 61) + (u128)K) ++bad;
        }
        memset(z, 0, (size_t)ZL * 4);
        for (int j = 0; j < K; ++j) {
            const u32* Qj = P->Q + (size_t)j * L;
            u64 carry = 0;
            for (int l = 0; l < ZL; ++l) {
                const u64 s = (u64)z[l] + (l < (int)L ? (u64)y[j] * Qj[l] : 0) + carry;
                z[l] = (u32)s;
                carry = s >> 32;
            }
        }
        if (PMO_MUTATION != 3) {
            const u32* Tv = P->T + (size_t)v * L;
            u64 carry = 0;
            for (int l = 0; l < ZL; ++l) {
                const u64 s = (u64)z[l] + (l < (int)L ? Tv[l] : 0u) + carry;
                z[l] = (u32)s;
                carry = s >> 32;
            }
        }
        const int wl = (int)(W >> 5), wb = (int)(W & 31);
        for (;;) {
            /* hi = z >> W (fits 64 bits: z < 2^(W+64) throughout) */
            u128 hi = 0;
            for (int l = ZL - 1; l >= wl; --l) hi = (hi << 32) | z[l];
            hi >>= wb;
            if (hi == 0) break;
            z[wl] = wb ? (z[wl] & ((1u << wb) - 1u)) : 0u;
            for (int l = wl + 1; l < ZL; ++l) z[l] = 0;
            const u128 add = hi * (u128)(PMO_MUTATION == 7 ? 1u : P->delta);
            u128 carry = 0;
            for (int l = 0; l < ZL; ++l) {
                const u128 s = (u128)z[l] + (l < 4 ? (u32)(add >> (32 * l)) : 0u) + carry;
                z[l] = (u32)s;
                carry = s >> 32;
            }
        }
        if (PMO_MUTATION != 5 && big_cmp(z, P->q, (int)L) >= 0) big_sub(z, P->q, (int)L);
        memcpy(c + (size_t)n * L, z, (size_t)L * 4);
    }
    free(z);
    free(y);
    return bad;
}

int pmo_multiply(const pmo_plan* P, const u32* a, const u32* b, u32* c) {
    const size_t kn = (size_t)P->K * P->N;
    u32* A = (u32*)malloc(kn * 4);
    u32* B = (u32*)malloc(kn * 4);
    pmo_rns(P, a, A);
    pmo_rns(P, b, B);
    pmo_ntt_fwd(P, A);
    pmo_ntt_fwd(P, B);
    pmo_pointwise(P, A, B, A);
    pmo_ntt_inv(P, A);
    const int bad = pmo_crt(P, A, c);
    free(A);
    free(B);
    return bad;
}
