// Float arctangent with a proven error bound, for decisions that only need to know on which side
// of a threshold an angle lies (the flatness test of the AGG subdivision). Shared by the HIP
// kernels and the CPU test that measures its error (tests/test_fatan.py).
//
// atan(t), t in [0, 1], is t P(t^2) with P a degree-7 polynomial fitted by iterated weighted least
// squares; evaluated in float with FMA its error is measured against double atan over every float
// t in [0, 1]. atan2 reduces to t = min(|x|, |y|) / max(|x|, |y|), computed on the device with the
// hardware reciprocal (relative error far below 2^-21, which moves atan by less than 2^-22), and
// reconstructs the angle with pi/2 - a, pi - a and the sign of y. VKH_ATAN2_ERR bounds the total
// difference to the true atan2 for the arguments it accepts: finite, with
// 2^-60 <= max(|x|, |y|) <= 2^60 (*ok is cleared otherwise).
#ifndef VKH_FATAN_H
#define VKH_FATAN_H

#include <math.h>

#ifdef __HIPCC__
#define VKH_FA __host__ __device__ __forceinline__
#else
#define VKH_FA static inline
#endif

#define VKH_ATAN2_ERR 1.0e-6f

VKH_FA float vkh_atan_poly(float t) {
    const float s = t * t;
    float p = -0.004054570570588112f;
    p = fmaf(p, s, 0.021862968802452087f);
    p = fmaf(p, s, -0.05591234192252159f);
    p = fmaf(p, s, 0.0964219868183136f);
    p = fmaf(p, s, -0.1390863060951233f);
    p = fmaf(p, s, 0.19946566224098206f);
    p = fmaf(p, s, -0.33329859375953674f);
    p = fmaf(p, s, 0.9999993443489075f);
    return p * t;
}

VKH_FA float vkh_rcp_approx(float x) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_rcpf(x);
#else
    return 1.0f / x;
#endif
}

// reduced-argument core with the reciprocal supplied (the CPU test perturbs it)
VKH_FA float vkh_atan2_fast_r(float y, float x, float (*rcp)(float), int *ok) {
    const float ax = fabsf(x), ay = fabsf(y);
    const float mx = fmaxf(ax, ay), mn = fminf(ax, ay);
    if (!(ax <= 0x1p60f && ay <= 0x1p60f && mx >= 0x1p-60f)) // also NaN
        *ok = 0;
    const float t = mn * rcp(mx);
    float a = vkh_atan_poly(t);
    if (ay > ax)
        a = 1.57079637f - a; // float pi/2
    if (x < 0.f)
        a = 3.14159274f - a; // float pi
    return copysignf(a, y);
}

VKH_FA float vkh_atan2_fast(float y, float x, int *ok) {
    const float ax = fabsf(x), ay = fabsf(y);
    const float mx = fmaxf(ax, ay), mn = fminf(ax, ay);
    if (!(ax <= 0x1p60f && ay <= 0x1p60f && mx >= 0x1p-60f)) // also NaN
        *ok = 0;
    const float t = mn * vkh_rcp_approx(mx);
    float a = vkh_atan_poly(t);
    if (ay > ax)
        a = 1.57079637f - a;
    if (x < 0.f)
        a = 3.14159274f - a;
    return copysignf(a, y);
}

#endif
