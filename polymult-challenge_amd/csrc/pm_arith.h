// Word-level modular arithmetic shared by the kernels. All values are
// canonical residues in [0, p) with p < 2^31 unless stated; the "min" forms
// rely on 2p < 2^32, so a wrapped difference is always larger than p.
#pragma once

#include <cstdint>

#if defined(__HIPCC__)
#define PM_HD __host__ __device__ __forceinline__
#else
#define PM_HD inline
#endif

namespace pm {

PM_HD uint32_t umin32(uint32_t a, uint32_t b) { return a < b ? a : b; }

PM_HD uint32_t mulhi32(uint32_t a, uint32_t b) { return (uint32_t)(((uint64_t)a * b) >> 32); }

// (a + b) mod p
PM_HD uint32_t add_mod(uint32_t a, uint32_t b, uint32_t p) {
    const uint32_t s = a + b;
    return umin32(s, s - p);
}

// (a - b) mod p
PM_HD uint32_t sub_mod(uint32_t a, uint32_t b, uint32_t p) {
    const uint32_t d = a - b;
    return umin32(d, d + p);
}

// x * w mod p for any 32-bit x, w < p, wq = floor(w 2^32 / p) (Shoup).
// x w - floor(x wq / 2^32) p lies in [0, 2p).
PM_HD uint32_t shoup_mul(uint32_t x, uint32_t w, uint32_t wq, uint32_t p) {
    const uint32_t qt = mulhi32(x, wq);
    const uint32_t r = x * w - qt * p;
    return umin32(r, r - p);
}

// v 2^-32 mod p for v < p 2^32 (Montgomery), pinv = -p^-1 mod 2^32.
PM_HD uint32_t redc64(uint64_t v, uint32_t p, uint32_t pinv) {
    const uint32_t m = (uint32_t)v * pinv;
    const uint32_t r = (uint32_t)((v + (uint64_t)m * p) >> 32);
    return umin32(r, r - p);
}

// a b 2^-32 mod p.
PM_HD uint32_t mont_mul(uint32_t a, uint32_t b, uint32_t p, uint32_t pinv) {
    return redc64((uint64_t)a * b, p, pinv);
}

// Cooley-Tukey: (x0, x1) -> (x0 + w x1, x0 - w x1).
PM_HD void ct_butterfly(uint32_t& x0, uint32_t& x1, uint32_t w, uint32_t wq, uint32_t p) {
    const uint32_t v = shoup_mul(x1, w, wq, p);
    const uint32_t u = x0;
    x0 = add_mod(u, v, p);
    x1 = sub_mod(u, v, p);
}

// The same butterfly for outputs that the next stage only multiplies by a
// twiddle: Shoup's product takes any 32-bit input, so x0 = u + v in [0, 2p)
// and x1 = u - v + p in [1, 2p) stay unreduced (x0 must come in below p).
PM_HD void ct_butterfly_lazy(uint32_t& x0, uint32_t& x1, uint32_t w, uint32_t wq, uint32_t p) {
    const uint32_t v = shoup_mul(x1, w, wq, p);
    const uint32_t u = x0;
    x0 = u + v;
    x1 = u + (p - v);
}

// Gentleman-Sande with the mirrored twiddle m = -w^-1:
// (x0, x1) -> (x0 + x1, (x0 - x1) w^-1) = (x0 + x1, (x1 - x0) m).
PM_HD void gs_butterfly(uint32_t& x0, uint32_t& x1, uint32_t m, uint32_t mq, uint32_t p) {
    const uint32_t u = x0, v = x1;
    x0 = add_mod(u, v, p);
    // v - u + p lies in [1, 2p), below 2^32: Shoup's product takes it as it
    // is, so the difference needs no reduction of its own.
    x1 = shoup_mul(v + (p - u), m, mq, p);
}

}  // namespace pm
