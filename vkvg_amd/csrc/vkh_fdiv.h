// Correctly rounded float division by a divisor whose reciprocal was computed once (per draw or
// per gradient stop) instead of once per pixel. Shared by the HIP kernels and the CPU test that
// compares it bit for bit with IEEE division (tests/test_fdiv.py).
//
// With r = RN(1/b), q0 = RN(a r) lies within 1.5 ulp of a/b; one correction q1 = RN(q0 + (a - b q0) r)
// is within 1/2 ulp + 2^-24 ulp, a faithful quotient; a second correction with the correctly
// rounded reciprocal is RN(a/b) by Markstein's theorem (P. Markstein, "Computation of elementary
// functions on the IBM RISC System/6000 processor", IBM J. R&D 34(1), 1990; also Muller et al.,
// Handbook of Floating-Point Arithmetic, section on Newton-Raphson based division with an FMA).
// The residuals a - b q are exact in an FMA when neither they nor the products leave the normal
// range, which the bounds below guarantee: 2^-50 <= |b| <= 2^50 (otherwise the reciprocal is
// stored as 0) and 2^-50 <= |quotient| <= 2^100. Every other case (r = 0, zero, tiny, huge,
// infinite or NaN quotients) falls back to the IEEE division.
#ifndef VKH_FDIV_H
#define VKH_FDIV_H

#include <math.h>

#ifdef __HIPCC__
#define VKH_FD __host__ __device__ __forceinline__
#else
#define VKH_FD static inline
#endif

// 1/b rounded to nearest when b is inside the fast path's range, 0 otherwise
VKH_FD float vkh_rcp_for_div(float b) {
    const float m = fabsf(b);
    return (m >= 0x1p-50f && m <= 0x1p50f) ? 1.0f / b : 0.0f;
}

// a / b rounded to nearest even, given r = vkh_rcp_for_div(b)
VKH_FD float vkh_div_rcp(float a, float b, float r) {
    float q = a * r;
    q = fmaf(fmaf(-b, q, a), r, q);
    q = fmaf(fmaf(-b, q, a), r, q);
    const float m = fabsf(q);
    if (!(m >= 0x1p-50f && m <= 0x1p100f)) {
        __asm__ volatile(""); // keeps the rare IEEE division a branch (not computed for every call and selected)
        q = a / b;
    }
    return q;
}

#endif
