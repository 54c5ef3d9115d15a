// Device-side building blocks shared by the vkvg HIP kernels (gfx950).
//
// Every float expression keeps the operand order of the reference so that, compiled with
// -ffp-contract=off and IEEE division / square root, the kernels produce the same bits as
// the reference's host code:
//   vec2 helpers        src/vectors.h:92-165
//   AGG flattening      src/vkvg_context_internal.c:1313-1461
//   joins / caps        src/vkvg_context_internal.c:924-1239
// Transcendental functions are evaluated in double and rounded once to float.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string.h>

#include "vkvg_hip.h"
#include "vkh_fatan.h"

#define VKH_PIF   3.14159265358979323846f
#define VKH_PIF_2 1.57079632679489661923f
#define VKH_2_PIF 0.63661977236758134308f /* 2/pi where 2*pi was meant (reproduced) */
#define VKH_FLT_EPSILON 1.19209290e-7f

// the vector helpers and the flattening below also compile for the host (vkvg_hip_item_extents)
#define VKH_HD __host__ __device__ __forceinline__

struct v2 {
    float x, y;
};

VKH_HD v2 V(float x, float y) { return v2{x, y}; }
VKH_HD v2 vadd(v2 a, v2 b) { return V(a.x + b.x, a.y + b.y); }
VKH_HD v2 vsub(v2 a, v2 b) { return V(a.x - b.x, a.y - b.y); }
VKH_HD v2 vmul(v2 a, float m) { return V(a.x * m, a.y * m); }
VKH_HD v2 vdiv(v2 a, float m) { return V(a.x / m, a.y / m); }
VKH_HD v2 vperp(v2 a) { return V(a.y, -a.x); }
VKH_HD float vlen(v2 v) { return sqrtf(v.x * v.x + v.y * v.y); }
VKH_HD v2 vnorm(v2 a) {
    float m = sqrtf(a.x * a.x + a.y * a.y);
    return V(a.x / m, a.y / m);
}
VKH_HD float vdot(v2 a, v2 b) { return (a.x * b.x) + (a.y * b.y); }
VKH_HD bool equf(float a, float b) { return fabsf(a - b) <= VKH_FLT_EPSILON; }

VKH_HD float d_atan2f(float y, float x) { return (float)atan2((double)y, (double)x); }
VKH_HD float d_cosf(float a) { return (float)cos((double)a); }
VKH_HD float d_sinf(float a) { return (float)sin((double)a); }
VKH_HD float d_acosf(float a) { return (float)acos((double)a); }

// ---------------------------------------------------------------------------------------
// Arc items: the interior points of vkvg_arc / vkvg_arc_negative (src/vkvg_context.c:431-444,
// :485-498) and of elliptic arcs (src/vkvg_context_internal.c:1555-1574), one point per step of
// the reference's sequential angle recurrence. The float (circle) or double (ellipse) angle is
// advanced by one addition per point, as the reference's loop does, so every point repeats the
// reference's operations exactly; the host counted the points with the same additions.
VKH_HD uint32_t item_points(const vkh_item &it) {
    return (it.flags & (VKH_ITEM_ARC | VKH_ITEM_EARC)) ? it.flags >> VKH_ITEM_COUNT_SHIFT : 1u;
}
template <class F> VKH_HD void arc_points(const vkh_item &it, F emit) {
    const uint32_t n = it.flags >> VKH_ITEM_COUNT_SHIFT;
    if (it.flags & VKH_ITEM_ARC) {
        const float xc = it.p[0], yc = it.p[1], r = it.p[2], s = it.p[4];
        float a = it.p[3];
        for (uint32_t k = 0; k < n; k++) {
            emit(V(d_cosf(a) * r + xc, d_sinf(a) * r + yc));
            a += s;
        }
    } else {
        double th;
        memcpy(&th, it.p, sizeof(th));
        const float s = it.p[2], rx = it.p[3], ry = it.p[4], cph = it.p[5], sph = it.p[6], cx = it.p[7], cy = it.tol;
        for (uint32_t k = 0; k < n; k++) {
            const float ex = rx * d_cosf((float)th), ey = ry * d_sinf((float)th);
            emit(V(((cph * ex) + (-sph * ey)) + cx, ((sph * ex) + (cph * ey)) + cy));
            th += s;
        }
    }
}

// ---------------------------------------------------------------------------------------
// AGG adaptive subdivision. The reference recurses depth-first (left child first) and emits
// at the leaves. Here the traversal is stackless: the current node is identified by its path
// of left/right choices from the root (up to 101 levels), and its control points are
// recomputed from the root, which repeats exactly the reference's float operations.
struct Cubic {
    float x1, y1, x2, y2, x3, y3, x4, y4;
};

VKH_HD void agg_split(const Cubic &c, bool right, Cubic &o) {
    float x12 = (c.x1 + c.x2) / 2, y12 = (c.y1 + c.y2) / 2;
    float x23 = (c.x2 + c.x3) / 2, y23 = (c.y2 + c.y3) / 2;
    float x34 = (c.x3 + c.x4) / 2, y34 = (c.y3 + c.y4) / 2;
    float x123 = (x12 + x23) / 2, y123 = (y12 + y23) / 2;
    float x234 = (x23 + x34) / 2, y234 = (y23 + y34) / 2;
    float x1234 = (x123 + x234) / 2, y1234 = (y123 + y234) / 2;
    if (!right)
        o = Cubic{c.x1, c.y1, x12, y12, x123, y123, x1234, y1234};
    else
        o = Cubic{x1234, y1234, x234, y234, x34, y34, c.x4, c.y4};
}

// Leaf test of one node at `level` (>= 1). Returns the number of points emitted (0 means
// "subdivide further"), writing them to pts. The comparisons against the reference's double
// constants (1.7, 0.01) promote the float operand exactly as C does.
VKH_HD int agg_leaf(const Cubic &c, float tol, v2 *pts) {
    float x1 = c.x1, y1 = c.y1, x2 = c.x2, y2 = c.y2, x3 = c.x3, y3 = c.y3, x4 = c.x4, y4 = c.y4;
    float dx = x4 - x1, dy = y4 - y1;
    float d2 = fabsf(((x2 - x4) * dy - (y2 - y4) * dx));
    float d3 = fabsf(((x3 - x4) * dy - (y3 - y4) * dx));
    float da1, da2;
    if ((double)d2 > 1.7 && (double)d3 > 1.7) {
        if ((d2 + d3) * (d2 + d3) <= (dx * dx + dy * dy) * tol) {
            float a23 = d_atan2f(y3 - y2, x3 - x2);
            da1 = fabsf(a23 - d_atan2f(y2 - y1, x2 - x1));
            da2 = fabsf(d_atan2f(y4 - y3, x4 - x3) - a23);
            if (da1 >= VKH_PIF)
                da1 = VKH_2_PIF - da1;
            if (da2 >= VKH_PIF)
                da2 = VKH_2_PIF - da2;
            if (da1 + da2 < (float)0.01) {
                float x12 = (x1 + x2) / 2, y12 = (y1 + y2) / 2, x23 = (x2 + x3) / 2, y23 = (y2 + y3) / 2;
                float x34 = (x3 + x4) / 2, y34 = (y3 + y4) / 2;
                float x123 = (x12 + x23) / 2, y123 = (y12 + y23) / 2, x234 = (x23 + x34) / 2, y234 = (y23 + y34) / 2;
                pts[0] = V((x123 + x234) / 2, (y123 + y234) / 2);
                return 1;
            }
            if ((double)da1 > 0.01) {
                pts[0] = V(x2, y2);
                return 1;
            }
            if ((double)da2 > 0.01) {
                pts[0] = V(x3, y3);
                return 1;
            }
        }
    } else {
        if ((double)d2 > 1.7) {
            if (d2 * d2 <= tol * (dx * dx + dy * dy)) {
                da1 = fabsf(d_atan2f(y3 - y2, x3 - x2) - d_atan2f(y2 - y1, x2 - x1));
                if (da1 >= VKH_PIF)
                    da1 = VKH_2_PIF - da1;
                if ((double)da1 < 0.01) {
                    pts[0] = V(x2, y2);
                    pts[1] = V(x3, y3);
                    return 2;
                }
                if ((double)da1 > 0.01) {
                    pts[0] = V(x2, y2);
                    return 1;
                }
            }
        } else if ((double)d3 > 1.7) {
            if (d3 * d3 <= tol * (dx * dx + dy * dy)) {
                da1 = fabsf(d_atan2f(y4 - y3, x4 - x3) - d_atan2f(y3 - y2, x3 - x2));
                if (da1 >= VKH_PIF)
                    da1 = VKH_2_PIF - da1;
                if ((double)da1 < 0.01) {
                    pts[0] = V(x2, y2);
                    pts[1] = V(x3, y3);
                    return 2;
                }
                if ((double)da1 > 0.01) {
                    pts[0] = V(x3, y3);
                    return 1;
                }
            }
        } else {
            float x12 = (x1 + x2) / 2, y12 = (y1 + y2) / 2, x23 = (x2 + x3) / 2, y23 = (y2 + y3) / 2;
            float x34 = (x3 + x4) / 2, y34 = (y3 + y4) / 2;
            float x123 = (x12 + x23) / 2, y123 = (y12 + y23) / 2, x234 = (x23 + x34) / 2, y234 = (y23 + y34) / 2;
            float x1234 = (x123 + x234) / 2, y1234 = (y123 + y234) / 2;
            float ex = x1234 - (x1 + x4) / 2;
            float ey = y1234 - (y1 + y4) / 2;
            if (ex * ex + ey * ey <= tol) {
                pts[0] = V(x1234, y1234);
                return 1;
            }
        }
    }
    return 0;
}

// The same leaf test with the float arctangent of vkh_fatan.h. Each of its angles differs from the
// double-rounded one by at most E = VKH_ATAN2_ERR + 2^-22; a difference of two angles (below 2 pi)
// by at most 2E + 2^-21 after both roundings, the wrapped value (2/pi - d, below 8 in magnitude)
// by 2^-21 more, and a sum of two wrapped values (below 16) by twice that plus 2^-20. A comparison
// is taken only when its operand lies further than that from the threshold; otherwise, or when an
// argument is outside the fast arctangent's range, the result is -1 and the caller runs agg_leaf.
// Any other result is the one agg_leaf gives.
#define VKH_ANG_E (VKH_ATAN2_ERR + 0x1p-22f)
#define VKH_ANG_M1 (2.f * VKH_ANG_E + 0x1p-21f)
#define VKH_ANG_M1W (VKH_ANG_M1 + 0x1p-21f)
#define VKH_ANG_M2 (2.f * VKH_ANG_M1W + 0x1p-20f)
VKH_HD int agg_leaf_fast(const Cubic &c, float tol, v2 *pts) {
    float x1 = c.x1, y1 = c.y1, x2 = c.x2, y2 = c.y2, x3 = c.x3, y3 = c.y3, x4 = c.x4, y4 = c.y4;
    float dx = x4 - x1, dy = y4 - y1;
    float d2 = fabsf(((x2 - x4) * dy - (y2 - y4) * dx));
    float d3 = fabsf(((x3 - x4) * dy - (y3 - y4) * dx));
    const bool b2 = (double)d2 > 1.7, b3 = (double)d3 > 1.7;
    const float len2 = dx * dx + dy * dy;
    // the reference's cases: both control points off the chord (f1), only the first (f2), only the
    // second (f3), each when flat enough; the three leg angles are computed once for all of them
    const bool f1 = b2 && b3 && (d2 + d3) * (d2 + d3) <= len2 * tol;
    const bool f2 = b2 && !b3 && d2 * d2 <= tol * len2;
    const bool f3 = !b2 && b3 && d3 * d3 <= tol * len2;
    if (f1 || f2 || f3) {
        int ok12 = 1, ok23 = 1, ok34 = 1;
        const float a12 = vkh_atan2_fast(y2 - y1, x2 - x1, &ok12);
        const float a23 = vkh_atan2_fast(y3 - y2, x3 - x2, &ok23);
        const float a34 = vkh_atan2_fast(y4 - y3, x4 - x3, &ok34);
        if (f1) {
            float da1 = fabsf(a23 - a12);
            float da2 = fabsf(a34 - a23);
            if (!(ok12 && ok23 && ok34) || fabsf(da1 - VKH_PIF) <= VKH_ANG_M1 || fabsf(da2 - VKH_PIF) <= VKH_ANG_M1)
                return -1;
            if (da1 >= VKH_PIF)
                da1 = VKH_2_PIF - da1;
            if (da2 >= VKH_PIF)
                da2 = VKH_2_PIF - da2;
            const float sum = da1 + da2;
            if (fabsf(sum - (float)0.01) <= VKH_ANG_M2)
                return -1;
            if (sum < (float)0.01) {
                float x12 = (x1 + x2) / 2, y12 = (y1 + y2) / 2, x23 = (x2 + x3) / 2, y23 = (y2 + y3) / 2;
                float x34 = (x3 + x4) / 2, y34 = (y3 + y4) / 2;
                float x123 = (x12 + x23) / 2, y123 = (y12 + y23) / 2, x234 = (x23 + x34) / 2, y234 = (y23 + y34) / 2;
                pts[0] = V((x123 + x234) / 2, (y123 + y234) / 2);
                return 1;
            }
            if (fabs((double)da1 - 0.01) <= (double)VKH_ANG_M1W)
                return -1;
            if ((double)da1 > 0.01) {
                pts[0] = V(x2, y2);
                return 1;
            }
            if (fabs((double)da2 - 0.01) <= (double)VKH_ANG_M1W)
                return -1;
            if ((double)da2 > 0.01) {
                pts[0] = V(x3, y3);
                return 1;
            }
            return 0;
        }
        // f2: angle(2-3) - angle(1-2); f3: angle(3-4) - angle(2-3)
        float da1 = f2 ? fabsf(a23 - a12) : fabsf(a34 - a23);
        if (!(ok23 && (f2 ? ok12 : ok34)) || fabsf(da1 - VKH_PIF) <= VKH_ANG_M1)
            return -1;
        if (da1 >= VKH_PIF)
            da1 = VKH_2_PIF - da1;
        if (fabs((double)da1 - 0.01) <= (double)VKH_ANG_M1W)
            return -1;
        if ((double)da1 < 0.01) {
            pts[0] = V(x2, y2);
            pts[1] = V(x3, y3);
            return 2;
        }
        if ((double)da1 > 0.01) {
            pts[0] = V(f2 ? x2 : x3, f2 ? y2 : y3);
            return 1;
        }
        return 0;
    }
    if (!b2 && !b3) {
        float x12 = (x1 + x2) / 2, y12 = (y1 + y2) / 2, x23 = (x2 + x3) / 2, y23 = (y2 + y3) / 2;
        float x34 = (x3 + x4) / 2, y34 = (y3 + y4) / 2;
        float x123 = (x12 + x23) / 2, y123 = (y12 + y23) / 2, x234 = (x23 + x34) / 2, y234 = (y23 + y34) / 2;
        float x1234 = (x123 + x234) / 2, y1234 = (y123 + y234) / 2;
        float ex = x1234 - (x1 + x4) / 2;
        float ey = y1234 - (y1 + y4) / 2;
        if (ex * ex + ey * ey <= tol) {
            pts[0] = V(x1234, y1234);
            return 1;
        }
    }
    return 0;
}

// Depth-first traversal; calls emit(v2) for every emitted point in the reference's order.
// Path bits: bit (level-1) of `path` is 1 when the node at `level` is a right child. The parent
// and grandparent of the current node are kept, so moving to the sibling or to the parent's
// sibling takes one split; deeper returns recompute the node from the root. Every node is thus
// produced by the same chain of splits as in the reference's recursion.
template <typename Emit>
__host__ __device__ void agg_flatten(const Cubic &root, float tol, Emit emit) {
    uint64_t path_lo = 0, path_hi = 0; // 128 bits of choices, levels 1..101
    int level = 1;                      // the root (level 0) always subdivides
    Cubic node, par = root, gpar = root;
    bool gpar_ok = false; // the parent is always known; the grandparent not after a two-level move
    agg_split(root, false, node);
    while (true) {
        bool leaf;
        if (level > 100) {
            leaf = true; // recursion limit: returns without emitting
        } else {
            v2 pts[2];
            int n = agg_leaf(node, tol, pts);
            leaf = n > 0;
            for (int i = 0; i < n; i++)
                emit(pts[i]);
        }
        if (!leaf) {
            // descend to the left child (bit level stays 0)
            gpar = par;
            gpar_ok = true;
            par = node;
            Cubic ch;
            agg_split(node, false, ch);
            node = ch;
            level++;
            continue;
        }
        // backtrack: pop right children until a left child is found, then go to its sibling
        int popped = 0;
        while (level >= 1) {
            int b = level - 1;
            bool is_right = b < 64 ? ((path_lo >> b) & 1ull) : ((path_hi >> (b - 64)) & 1ull);
            if (!is_right)
                break;
            if (b < 64)
                path_lo &= ~(1ull << b);
            else
                path_hi &= ~(1ull << (b - 64));
            level--;
            popped++;
        }
        if (level < 1)
            return;
        {
            int b = level - 1;
            if (b < 64)
                path_lo |= (1ull << b);
            else
                path_hi |= (1ull << (b - 64));
        }
        Cubic nx;
        if (popped == 0) { // sibling of a left child: same parent
            agg_split(par, true, nx);
            node = nx;
        } else if (popped == 1 && gpar_ok) { // sibling of the parent: the grandparent's right child
            agg_split(gpar, true, nx);
            node = nx;
            par = gpar;
            gpar_ok = false;
        } else { // recompute the node at `level` from the root, keeping its two ancestors
            Cubic cur = root, p1 = root, p2 = root;
            for (int l = 1; l <= level; l++) {
                int b = l - 1;
                bool r = b < 64 ? ((path_lo >> b) & 1ull) : ((path_hi >> (b - 64)) & 1ull);
                p2 = p1;
                p1 = cur;
                agg_split(cur, r, nx);
                cur = nx;
            }
            node = cur;
            par = p1;
            gpar = p2;
            gpar_ok = level >= 2;
        }
    }
}

// The same traversal with the ancestors of the current node on a stack (levels 0 .. D-1 kept by the
// caller, e.g. in LDS): every move costs one split. The parent of a right sibling at `level` is
// read from the stack, or, beyond level D-1, recomputed from the level D-1 ancestor along the path
// bits; all nodes come from the same chain of splits as in the recursion. Every iteration tests one
// node and makes exactly one split (the left child of an inner node, or the right sibling of the
// first left child above a leaf), so lanes of a wavefront that take different moves still share it;
// the right children above a leaf are popped in one count of the path's leading ones.
template <int D, typename Emit, typename Store, typename Load>
__host__ __device__ void agg_flatten_stack(const Cubic &root, float tol, Emit emit, Store store, Load load) {
    uint64_t path_lo = 0, path_hi = 0; // bit (l-1): the node at level l is a right child
    int level = 1;
    store(0, root);
    Cubic node;
    agg_split(root, false, node);
    while (true) {
        bool leaf;
        if (level > 100) {
            leaf = true; // recursion limit: returns without emitting
        } else {
            v2 pts[2];
            int n = agg_leaf_fast(node, tol, pts);
            if (n < 0)
                n = agg_leaf(node, tol, pts);
            leaf = n > 0;
            if (n > 0)
                emit(pts[0]);
            if (n > 1)
                emit(pts[1]);
        }
        Cubic src;
        bool right;
        if (!leaf) { // descend to the left child
            if (level < D)
                store(level, node);
            src = node;
            right = false;
            level++;
        } else {
            // pop the right children above the leaf: the run of ones from bit level-1 downwards
            if (level <= 64) {
                const uint64_t up = path_lo << (64 - level); // bit level-1 at the top
                const int ones = up == ~0ull ? level : __builtin_clzll(~up);
                level -= ones;
                path_lo &= level >= 64 ? ~0ull : ((1ull << level) - 1ull);
            } else {
                while (level >= 1) {
                    const int b = level - 1;
                    const bool is_right = b < 64 ? ((path_lo >> b) & 1ull) : ((path_hi >> (b - 64)) & 1ull);
                    if (!is_right)
                        break;
                    if (b < 64)
                        path_lo &= ~(1ull << b);
                    else
                        path_hi &= ~(1ull << (b - 64));
                    level--;
                }
            }
            if (level < 1)
                return;
            // the first left child becomes its right sibling
            {
                const int b = level - 1;
                if (b < 64)
                    path_lo |= (1ull << b);
                else
                    path_hi |= (1ull << (b - 64));
            }
            if (level - 1 < D) {
                src = load(level - 1);
            } else {
                src = load(D - 1);
                for (int l = D; l <= level - 1; l++) {
                    const int b = l - 1;
                    const bool r = b < 64 ? ((path_lo >> b) & 1ull) : ((path_hi >> (b - 64)) & 1ull);
                    Cubic nx;
                    agg_split(src, r, nx);
                    src = nx;
                }
            }
            right = true;
        }
        agg_split(src, right, node);
    }
}

// ---------------------------------------------------------------------------------------
// The dash loop of _draw_dashed_segment (src/vkvg_context_internal.c:1240-1264) without its caps:
// while the offset is below the segment length, flip dashOn, add the current dash to the offset in
// float and move to the next dash. Each addition rounds, so the offsets follow from the additions
// in order; but while every partial sum is a multiple of 2^g (g the exponent of the lowest set bit
// of the offset and of the dashes) below 2^(g+24), every one of them is a float and the additions
// are exact. A run of whole periods (a period, from any dash, holds every dash once) then adds
// exactly their number times the period's sum, flips dashOn once per dash and leaves the dash index
// where it was, so such a run is taken in one step with the same result as the loop.

// exponent of the lowest set bit of a finite v > 0
VKH_HD int f32_lowbit(float v) {
    uint32_t b;
    memcpy(&b, &v, 4);
    int e = (int)((b >> 23) & 0xFFu);
    uint32_t m = b & 0x7FFFFFu;
    if (e) {
        m |= 0x800000u;
        e -= 150;
    } else {
        e = -149;
    }
    return e + __builtin_ctz(m);
}

struct DashPeriod {
    double sum; // exact sum of the dashes
    int lowbit; // lowest set bit over the non-zero dashes
    bool ok;    // dashes finite, >= 0, with a non-zero exact sum
};

VKH_HD DashPeriod dash_period(const float *dashes, uint32_t n) {
    DashPeriod p{0.0, 1 << 20, n > 0};
    for (uint32_t i = 0; i < n; i++) {
        const float v = dashes[i];
        if (!(v >= 0.f) || !(v <= 3.0e38f)) {
            p.ok = false;
        } else if (v > 0.f) {
            p.sum += (double)v;
            const int lb = f32_lowbit(v);
            p.lowbit = lb < p.lowbit ? lb : p.lowbit;
        }
    }
    // every partial sum is a multiple of 2^lowbit; below 2^(lowbit + 52) the double additions were exact
    if (!(p.sum > 0.0) || !(p.sum < ldexp(1.0, p.lowbit + 52)))
        p.ok = false;
    return p;
}

// Whole periods the loop takes from offset off < segLen with every float addition exact and every
// offset before an addition below segLen: the largest k with off + k * sum < min(2^(g+24), segLen).
// (A k of at least 1 needs sum < 2^(g+24), so off + k * sum is a multiple of 2^g below 2^(g+25) and
// exact in double; the quotient's rounding is corrected by the exact test.)
VKH_HD uint32_t dash_whole_periods(const DashPeriod &p, float off, float segLen) {
    if (!p.ok || !(off >= 0.f))
        return 0;
    int g = p.lowbit;
    if (off > 0.f) {
        const int lo = f32_lowbit(off);
        g = lo < g ? lo : g;
    }
    const double lim = fmin(ldexp(1.0, g + 24), (double)segLen);
    const double room = lim - (double)off;
    if (!(room > p.sum))
        return 0;
    const double kd = floor(room / p.sum);
    uint32_t k = kd < 16777216.0 ? (uint32_t)kd : 16777216u;
    if ((double)off + (double)k * p.sum >= lim)
        k--;
    return k;
}

// The loop over one segment: advances (on, cur, off) and returns the number of caps it makes. On a
// segment some 2^24 dash lengths long or longer the offset's float additions stop advancing it
// (each dash rounds away), and the reference's loop never ends; here the walk stops once it has
// made VKH_DASH_MAX_CAPS caps and sets `over`, and the frame is refused.
#define VKH_DASH_MAX_CAPS (1u << 24)
VKH_HD uint32_t dash_walk_segment(const float *dashes, uint32_t n, const DashPeriod &p, float segLen, bool &on,
                                  uint32_t &cur, float &off, bool &over) {
    uint32_t caps = 0;
    while (off < segLen) {
        if (caps >= VKH_DASH_MAX_CAPS) {
            over = true;
            break;
        }
        uint32_t w = dash_whole_periods(p, off, segLen);
        if (w > VKH_DASH_MAX_CAPS / n) // fewer periods are as exact; keeps caps within 32 bits
            w = VKH_DASH_MAX_CAPS / n;
        if (w) {
            off = (float)((double)off + (double)w * p.sum);
            caps += w * n;
            on = on != (((w * n) & 1u) != 0);
        }
        on = !on;
        off += dashes[cur];
        if (++cur == n)
            cur = 0;
        caps++;
    }
    return caps;
}

// ---------------------------------------------------------------------------------------
// Vertex transform and 1/256 snap (shaders/vkvg_main.vert:74-79 + viewport transform).
#define VKH_INVALID_FIX ((int32_t)0x80000000)

__device__ __forceinline__ void snap_vertex(float x, float y, const float *m, float sw, float sh, int32_t &fx,
                                            int32_t &fy) {
    float px = (m[0] * x + m[2] * y) + m[4];
    float py = (m[1] * x + m[3] * y) + m[5];
    float nx = (px * 2.0f) / sw - 1.0f;
    float ny = (py * 2.0f) / sh - 1.0f;
    float hw = sw * 0.5f, hh = sh * 0.5f;
    float X = nx * hw + hw;
    float Y = ny * hh + hh;
    if (isnan(X) || isnan(Y)) {
        fx = fy = VKH_INVALID_FIX;
        return;
    }
    X = fminf(fmaxf(X, -32768.0f), 32768.0f);
    Y = fminf(fmaxf(Y, -32768.0f), 32768.0f);
    fx = (int32_t)rintf(X * 256.0f);
    fy = (int32_t)rintf(Y * 256.0f);
}

// ---------------------------------------------------------------------------------------
// Exact winding predicates on 24.8 fixed-point edges. Samples sit at (256c+128, 256r+128)
// perturbed by (+eps, +eps^2), which reproduces the top-left tie rule: no sample is ever on
// an edge. Winding uses a ray towards -x; an edge counts +1 when it runs towards +y.
struct Edge {
    int32_t x0, y0, x1, y1;
};

// ---------------------------------------------------------------------------------------
// float colour channel -> UNORM8 (clamped, rounded to nearest)
__device__ __forceinline__ uint8_t f2u8(float f) {
    // NaN and f <= 0 give 0 (fmaxf returns the non-NaN operand), f >= 1 gives 255
    return (uint8_t)rintf(fminf(fmaxf(f, 0.f), 1.f) * 255.0f);
}
// Four colour channels -> packed premultiplied RGBA8, each byte f2u8 of its channel, with the
// hardware conversion: v_cvt_pk_u8_f32(x * 255) rounds to nearest even and saturates to [0, 255]
// (NaN -> 0), which equals f2u8(x) for every float x (checked over all 2^32 bit patterns by
// vkh_check_u8_conversion, run by tests/test_gpu_convert.py).
__device__ __forceinline__ uint32_t f2u8x4(float r, float g, float b, float a) {
    uint32_t s = __builtin_amdgcn_cvt_pk_u8_f32(r * 255.0f, 0, 0u);
    s = __builtin_amdgcn_cvt_pk_u8_f32(g * 255.0f, 1, s);
    s = __builtin_amdgcn_cvt_pk_u8_f32(b * 255.0f, 2, s);
    return __builtin_amdgcn_cvt_pk_u8_f32(a * 255.0f, 3, s);
}
__device__ __forceinline__ uint32_t draw_key(uint32_t n) {
    uint32_t x = n + 0x9E3779B9u;
    x ^= x >> 16;
    x *= 0x85EBCA6Bu;
    x ^= x >> 13;
    x *= 0xC2B2AE35u;
    x ^= x >> 16;
    return x | 1u;
}
