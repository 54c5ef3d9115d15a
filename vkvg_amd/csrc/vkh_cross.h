// Column crossings of one 24.8 fixed-point edge over the sample rows of a tile, in exact double
// arithmetic. Edge coordinates are clamped to +-2^23 (+-32768 px), so every product below is an
// integer under 2^50 and every sum under 2^51: doubles hold them exactly, and double FMA (full rate
// on CDNA) replaces the 64-bit integer multiplies of the straightforward form. Shared by the fine
// kernel and the CPU test that compares it with 128-bit integer arithmetic (tests/test_cross.py).
// Samples sit at (256 c + 128, 256 r + 128) in fixed point.
#ifndef VKH_CROSS_H
#define VKH_CROSS_H

#include <math.h>
#include <stdint.h>

#ifdef __HIPCC__
#define VKH_CR __host__ __device__ __forceinline__
#else
#define VKH_CR static inline
#endif

// n / d for the quotients below, each of which is followed by an exact correction of one step in
// either direction, so it only has to be within one of the exact quotient: a single-precision
// reciprocal (relative error below 2^-21, with d rounded to float) refined by one Newton step
// (below 2^-40), times n. The quotients stay below 2^25 in magnitude, so the error stays below
// 2^-15. Cheaper than the correctly rounded division, which the corrections make unnecessary.
VKH_CR double cross_quot(double n, double d) {
#if defined(__HIP_DEVICE_COMPILE__)
    double r = (double)__builtin_amdgcn_rcpf((float)d);
#else
    double r = (double)(1.0f / (float)d);
#endif
    r = fma(fma(-d, r, 1.0), r, r);
    return n * r;
}
// Vertical crossing. For an edge crossing the sample column xs (lx <= xs < rx, (lx, ly) its left
// end and (rx, ry) its right end), the sample (xs, y) lies above the edge when
// E(y) = (rx - lx)(y - ly) - (ry - ly)(xs - lx) < 0, or E(y) = 0 and ry > ly (the tie-break that
// perturbs samples consistently). E increases with y, so the samples of rows y = y0 + 256 r are
// above the edge exactly for r < r*. Returns r*, clamped to [lo, hi].
VKH_CR int32_t cross_first_row(int32_t lx, int32_t ly, int32_t rx, int32_t ry, int32_t xs, int32_t y0, int32_t lo,
                               int32_t hi) {
    const double A = (double)(rx - lx);                       // in (0, 2^24]
    const double C = (double)(ry - ly) * (double)(xs - lx);   // |C| < 2^48
    double f = floor(cross_quot(C, A)); // |C / A| < |ry - ly| <= 2^24; within one, corrected
    double rem = fma(-f, A, C);
    if (rem < 0.0) {
        f -= 1.0;
        rem += A;
    } else if (rem >= A) {
        f += 1.0;
        rem -= A;
    }
    // least u = y - ly that is not above: A u > C, or A u == C when ry <= ly
    const double t = (ry <= ly && rem == 0.0) ? f : f + 1.0;
    double r = ceil((t + (double)(ly - y0)) * (1.0 / 256.0));
    r = fmin(fmax(r, (double)lo), (double)hi);
    return (int32_t)r;
}

// Horizontal crossings. For an edge going down from (xt, yt) to (xb, yb), yt < yb, the first
// column whose sample at row y lies right of the edge is q = ceil(N / D) with
// N = dy (xt - 128) + (xb - xt)(y - yt), D = 256 dy, dy = yb - yt. One row down adds
// dN = 256 (xb - xt) to N, so with dq = floor(dN / D), dr = dN - dq D in [0, D) and rem = q D - N in
// [0, D): q' = q + dq + (dr > rem), rem' = rem - dr, plus D when it carried.
typedef struct {
    int32_t q, dq;      // column at the current row, whole-column step per row
    double rem, dr, D;  // exact integers below 2^33
} vkh_dda;

// q = ceil(N / D) at row y, with rem = q D - N
VKH_CR double cross_column_rem(int32_t xt, int32_t yt, int32_t xb, int32_t yb, int32_t y, double *rem_out) {
    const double dy = (double)(yb - yt);
    const double D = 256.0 * dy;
    const double N = fma(dy, (double)(xt - 128), (double)(xb - xt) * (double)(y - yt)); // exact, < 2^50
    double q = ceil(cross_quot(N, D)); // off by at most one
    double rem = fma(q, D, -N);
    if (rem < 0.0) {
        q += 1.0;
        rem += D;
    } else if (rem >= D) {
        q -= 1.0;
        rem -= D;
    }
    *rem_out = rem;
    return q;
}

// the first column whose sample at row y lies right of the edge (yt <= y < yb)
VKH_CR int32_t cross_column(int32_t xt, int32_t yt, int32_t xb, int32_t yb, int32_t y) {
    double rem;
    return (int32_t)cross_column_rem(xt, yt, xb, yb, y, &rem);
}

VKH_CR vkh_dda cross_dda_start(int32_t xt, int32_t yt, int32_t xb, int32_t yb, int32_t y) {
    vkh_dda s;
    const double D = 256.0 * (double)(yb - yt);
    double rem;
    const double q = cross_column_rem(xt, yt, xb, yb, y, &rem);
    const double dN = 256.0 * (double)(xb - xt);
    double dq = floor(cross_quot(dN, D));
    double dr = fma(-dq, D, dN);
    if (dr < 0.0) {
        dq -= 1.0;
        dr += D;
    } else if (dr >= D) {
        dq += 1.0;
        dr -= D;
    }
    s.q = (int32_t)q;
    s.dq = (int32_t)dq;
    s.rem = rem;
    s.dr = dr;
    s.D = D;
    return s;
}

VKH_CR void cross_dda_step(vkh_dda *s) {
    const int carry = s->dr > s->rem;
    s->q += s->dq + (carry ? 1 : 0);
    s->rem = carry ? s->rem + (s->D - s->dr) : s->rem - s->dr;
}

#endif
