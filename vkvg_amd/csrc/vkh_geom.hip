// Geometry stage of the vkvg HIP backend: flattening, stroke expansion, edge generation.
//
//   flatten      one thread per path item in work-sorted order; cubics run the AGG subdivision
//                once into per-cubic scratch slots, a prefix sum places them, a copy pass moves
//                them (src/vkvg_context.c:541-566, src/vkvg_context_internal.c:1313-1461)
//   stroke       one thread per step: the join or cap at one point of a solid sub-path, one
//                segment piece of a dashed sub-path; counted, prefix-summed, then emitted into
//                vertex / triangle buffers exactly as the reference's caches fill
//                (src/vkvg_context.c:822-948, src/vkvg_context_internal.c:924-1282)
//   edges        fills: closed polygon edges of every subpath with more than 2 points;
//                strokes: 3 edges per triangle, oriented so that coverage count = winding
#include "vkh_internal.h"

// ---------------------------------------------------------------------------------------- flatten
// Flattening in one subdivision pass. Cubics are processed in descending order of an estimate
// of their subdivision work (64 buckets of log2(1 + |second differences| / tolerance), counting
// sort), so the lanes of a wavefront carry similar work. Every cubic item takes a scratch slot of
// VKH_FLAT_SLOT points (one atomic per wavefront), runs the AGG traversal once, keeps its first
// VKH_FLAT_SLOT points there and records its point count; the rare cubics with more points are
// listed. After the scan of the counts, k_flatten_copy moves the slots to their final positions,
// k_flatten_overflow runs the listed cubics again into theirs, and k_flatten_lines writes the
// explicit points.
__device__ __forceinline__ uint32_t flat_key(const vkh_item &it) {
    if (!(it.flags & VKH_ITEM_CUBIC))
        return 63;
    const float ax = fabsf(it.p[0] - 2.f * it.p[2] + it.p[4]), ay = fabsf(it.p[1] - 2.f * it.p[3] + it.p[5]);
    const float bx = fabsf(it.p[2] - 2.f * it.p[4] + it.p[6]), by = fabsf(it.p[3] - 2.f * it.p[5] + it.p[7]);
    const float d = fmaxf(fmaxf(ax, ay), fmaxf(bx, by)) / fmaxf(it.tol, 1e-12f);
    const float b = 4.f * log2f(1.f + fminf(d, 1e18f)); // NaN -> fminf gives 1e18
    return 62u - (uint32_t)fminf(fmaxf(b, 0.f), 62.f);  // heavy cubics first
}

// rank of an item within its workgroup's bucket: items of the line bucket (every item of a frame
// without curves) take their rank from one atomic per wavefront; cubics one atomic each
__device__ __forceinline__ uint32_t flat_rank(uint32_t *h, bool valid, uint32_t key) {
    const bool line = valid && key == 63;
    const uint64_t m = __ballot(line);
    uint32_t rank = 0;
    if (m) {
        const uint32_t lane = threadIdx.x & 63, leader = (uint32_t)__ffsll((unsigned long long)m) - 1;
        uint32_t base = 0;
        if (lane == leader)
            base = atomicAdd(&h[63], (uint32_t)__popcll(m));
        base = __shfl(base, (int)leader);
        rank = base + (uint32_t)__popcll(m & ((1ull << lane) - 1ull));
    }
    if (valid && !line)
        rank = atomicAdd(&h[key], 1u);
    return rank;
}

// Per workgroup of 256 items, the number of items in each bucket, stored bucket-major
// (whist[key * n_wg + workgroup]): one exclusive scan of that array gives every workgroup its
// first position in each bucket, without global atomics.
__global__ __launch_bounds__(256) void k_flat_hist(const vkh_item *__restrict__ items, uint32_t n,
                                                   uint32_t *__restrict__ whist, uint32_t n_wg) {
    __shared__ uint32_t h[64];
    if (threadIdx.x < 64)
        h[threadIdx.x] = 0;
    __syncthreads();
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    (void)flat_rank(h, i < n, i < n ? flat_key(items[i]) : 0u);
    __syncthreads();
    if (threadIdx.x < 64)
        whist[(size_t)threadIdx.x * n_wg + blockIdx.x] = h[threadIdx.x];
}

// every item to its position: the workgroup's first position in its bucket + its rank there
__global__ __launch_bounds__(256) void k_flat_scatter(const vkh_item *__restrict__ items, uint32_t n,
                                                      const uint32_t *__restrict__ wcur, uint32_t n_wg,
                                                      uint32_t *__restrict__ perm) {
    __shared__ uint32_t h[64];
    if (threadIdx.x < 64)
        h[threadIdx.x] = 0;
    __syncthreads();
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t key = i < n ? flat_key(items[i]) : 0u;
    const uint32_t rank = flat_rank(h, i < n, key);
    if (i < n)
        perm[wcur[(size_t)key * n_wg + blockIdx.x] + rank] = i;
}

#define VKH_FLAT_STACK 6   // ancestor levels kept in LDS per thread by the subdivision pass
#define VKH_FLAT_BLOCK 64 // threads per workgroup of the subdivision pass: one wavefront, 12 KB of stacks

__global__ __launch_bounds__(VKH_FLAT_BLOCK) void k_flatten_stage(const vkh_item *__restrict__ items, uint32_t n, const uint32_t *__restrict__ perm,
                                uint32_t *__restrict__ counts, uint32_t *__restrict__ slot_alloc,
                                uint32_t *__restrict__ slot_item, uint32_t *__restrict__ item_slot,
                                v2 *__restrict__ scratch, uint32_t *__restrict__ ovf_count,
                                uint32_t *__restrict__ ovf_list) {
    __shared__ float4 s_stk[VKH_FLAT_STACK][2][VKH_FLAT_BLOCK]; // ancestors, two float4 per cubic
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t i = t < n ? perm[t] : 0u;
    vkh_item it;
    bool cubic = false;
    if (t < n) {
        it = items[i];
        cubic = (it.flags & VKH_ITEM_CUBIC) != 0;
    }
    // one slot allocation per wavefront
    const uint64_t m = __ballot(cubic);
    const uint32_t lane = threadIdx.x & 63;
    uint32_t base = 0;
    if (m) {
        const uint32_t leader = (uint32_t)__ffsll((unsigned long long)m) - 1;
        if (lane == leader)
            base = atomicAdd(slot_alloc, (uint32_t)__popcll(m));
        base = __shfl(base, (int)leader);
    }
    if (t >= n)
        return;
    if (!cubic) {
        counts[i] = item_points(it);
        return;
    }
    const uint32_t slot = base + (uint32_t)__popcll(m & ((1ull << lane) - 1ull));
    slot_item[slot] = i;
    item_slot[i] = slot;
    v2 *out = scratch + (size_t)slot * VKH_FLAT_SLOT;
    Cubic cb{it.p[0], it.p[1], it.p[2], it.p[3], it.p[4], it.p[5], it.p[6], it.p[7]};
    uint32_t k = 0;
    agg_flatten_stack<VKH_FLAT_STACK>(
        cb, it.tol,
        [&](v2 p) {
            if (k < VKH_FLAT_SLOT)
                out[k] = p;
            k++;
        },
        [&](int l, const Cubic &c) {
            s_stk[l][0][threadIdx.x] = make_float4(c.x1, c.y1, c.x2, c.y2);
            s_stk[l][1][threadIdx.x] = make_float4(c.x3, c.y3, c.x4, c.y4);
        },
        [&](int l) {
            const float4 a = s_stk[l][0][threadIdx.x], b = s_stk[l][1][threadIdx.x];
            return Cubic{a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w};
        });
    if (!(it.flags & VKH_ITEM_NO_END)) {
        if (k < VKH_FLAT_SLOT)
            out[k] = V(it.p[6], it.p[7]);
        k++;
    }
    counts[i] = k;
    if (k > VKH_FLAT_SLOT)
        ovf_list[atomicAdd(ovf_count, 1u)] = i;
}

// VKH_FLAT_COPY_LANES lanes per slot (a cubic has about 10 points): its points to their final positions
#define VKH_FLAT_COPY_LANES 16
__global__ void k_flatten_copy(const vkh_item *__restrict__ items, uint32_t n_slots,
                               const uint32_t *__restrict__ slot_item, const uint32_t *__restrict__ counts,
                               const uint32_t *__restrict__ offs, const v2 *__restrict__ scratch,
                               v2 *__restrict__ pts, uint8_t *__restrict__ pflags) {
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t slot = (uint32_t)(t / VKH_FLAT_COPY_LANES), lane = (uint32_t)(t % VKH_FLAT_COPY_LANES);
    if (slot >= n_slots)
        return;
    const uint32_t i = slot_item[slot];
    const uint32_t n = counts[i];
    if (n > VKH_FLAT_SLOT)
        return; // overflowing cubics are written by k_flatten_overflow
    const uint32_t o = offs[i];
    const uint8_t f = (items[i].flags & VKH_ITEM_CURVED) ? 1 : 0;
    const v2 *src = scratch + (size_t)slot * VKH_FLAT_SLOT;
    for (uint32_t k = lane; k < n; k += VKH_FLAT_COPY_LANES) {
        pts[o + k] = src[k];
        pflags[o + k] = f;
    }
}

__global__ void k_flatten_lines(const vkh_item *__restrict__ items, uint32_t n, const uint32_t *__restrict__ offs,
                                v2 *__restrict__ pts, uint8_t *__restrict__ pflags) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    const vkh_item it = items[i];
    if (it.flags & VKH_ITEM_CUBIC)
        return;
    const uint32_t o = offs[i];
    const uint8_t f = (it.flags & VKH_ITEM_CURVED) ? 1 : 0;
    if (it.flags & (VKH_ITEM_ARC | VKH_ITEM_EARC)) {
        uint32_t k = o;
        arc_points(it, [&](v2 p) {
            pts[k] = p;
            pflags[k] = f;
            k++;
        });
        return;
    }
    pts[o] = V(it.p[0], it.p[1]);
    pflags[o] = f;
}

// the listed cubics with more points than a slot holds, traversed again into their positions
__global__ void k_flatten_overflow(const vkh_item *__restrict__ items, const uint32_t *__restrict__ list,
                                   uint32_t n_list, const uint32_t *__restrict__ offs, v2 *__restrict__ pts,
                                   uint8_t *__restrict__ pflags) {
    uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n_list)
        return;
    const uint32_t i = list[t];
    const vkh_item it = items[i];
    const uint8_t f = (it.flags & VKH_ITEM_CURVED) ? 1 : 0;
    Cubic cb{it.p[0], it.p[1], it.p[2], it.p[3], it.p[4], it.p[5], it.p[6], it.p[7]};
    uint32_t k = offs[i];
    agg_flatten(cb, it.tol, [&](v2 p) {
        pts[k] = p;
        pflags[k] = f;
        k++;
    });
    if (!(it.flags & VKH_ITEM_NO_END)) {
        pts[k] = V(it.p[6], it.p[7]);
        pflags[k] = f;
    }
}

// Diagnostic, not part of the public headers: the subdivision points of one cubic item (without its
// end point), computed on the host by the flattening kernel's stack traversal (stack != 0, the
// ancestors in a local array as the kernel keeps them in LDS) or by the stackless traversal of the
// overflow kernel; returns the number of points, of which the first cap are written to xy.
extern "C" uint32_t vkvg_hip_flatten_points(const vkh_item *it, int stack, float *xy, uint32_t cap) {
    const Cubic cb{it->p[0], it->p[1], it->p[2], it->p[3], it->p[4], it->p[5], it->p[6], it->p[7]};
    uint32_t k = 0;
    auto emit = [&](v2 p) {
        if (k < cap) {
            xy[2 * k] = p.x;
            xy[2 * k + 1] = p.y;
        }
        k++;
    };
    if (stack) {
        Cubic stk[VKH_FLAT_STACK];
        agg_flatten_stack<VKH_FLAT_STACK>(
            cb, it->tol, emit, [&](int l, const Cubic &c) { stk[l] = c; }, [&](int l) { return stk[l]; });
    } else {
        agg_flatten(cb, it->tol, emit);
    }
    return k;
}

// The dash walk of one segment as the dash-walk kernel runs it (whole periods taken at once where the
// float additions are exact), computed on the host: state (on, cur, off) in and out; returns the caps.
extern "C" uint32_t vkvg_hip_dash_walk(const float *dashes, uint32_t n, float seg_len, int *on, uint32_t *cur,
                                       float *off, int *over) {
    *over = 0;
    if (!n)
        return 0;
    const DashPeriod p = dash_period(dashes, n);
    bool b = *on != 0, o = false;
    const uint32_t caps = dash_walk_segment(dashes, n, p, seg_len, b, *cur, *off, o);
    *on = b ? 1 : 0;
    *over = o ? 1 : 0;
    return caps;
}

extern "C" void vkvg_hip_item_extents(const vkh_item *it, float ext[4]) {
    auto widen = [&](v2 p) {
        if (p.x < ext[0])
            ext[0] = p.x;
        if (p.x > ext[2])
            ext[2] = p.x;
        if (p.y < ext[1])
            ext[1] = p.y;
        if (p.y > ext[3])
            ext[3] = p.y;
    };
    if (it->flags & (VKH_ITEM_ARC | VKH_ITEM_EARC)) {
        arc_points(*it, widen);
        return;
    }
    if (!(it->flags & VKH_ITEM_CUBIC)) {
        widen(V(it->p[0], it->p[1]));
        return;
    }
    const Cubic cb{it->p[0], it->p[1], it->p[2], it->p[3], it->p[4], it->p[5], it->p[6], it->p[7]};
    agg_flatten(cb, it->tol, widen);
    if (!(it->flags & VKH_ITEM_NO_END))
        widen(V(it->p[6], it->p[7]));
}

__global__ void k_subpath_ranges(const vkh_subpath *__restrict__ sp, uint32_t n, const uint32_t *__restrict__ item_off,
                                 uint32_t *__restrict__ sp_first, uint32_t *__restrict__ sp_count) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    uint32_t a = item_off[sp[i].first_item];
    uint32_t b = item_off[sp[i].first_item + sp[i].n_items];
    sp_first[i] = a;
    sp_count[i] = b - a;
}

// ---------------------------------------------------------------------------------------- stroke
struct CountSink {
    uint32_t nv = 0, ni = 0;
    __device__ void addv(v2) { nv++; }
    __device__ void addt(uint32_t, uint32_t, uint32_t) { ni += 3; }
    __device__ void patch(uint32_t, bool) {}
};
struct EmitSink {
    v2 *verts;
    uint32_t *idx;
    uint32_t nv = 0, ni = 0;
    __device__ void addv(v2 p) { verts[nv++] = p; }
    __device__ void addt(uint32_t a, uint32_t b, uint32_t c) {
        idx[ni] = a;
        idx[ni + 1] = b;
        idx[ni + 2] = c;
        ni += 3;
    }
    // closing patch of src/vkvg_context.c:917-932 on the last 6 indices of this subpath
    __device__ void patch(uint32_t ii, bool inverse) {
        if (ni < 6)
            return;
        uint32_t *ind = idx + ni - 6;
        if (inverse) {
            ind[1] = ii + 1;
            ind[4] = ii + 1;
            ind[5] = ii;
        } else {
            ind[1] = ii;
            ind[4] = ii;
            ind[5] = ii + 1;
        }
    }
};

struct StrokeCtx {
    const v2 *P;
    const uint8_t *F;
    uint32_t n;
    float hw, lhMax, arcStep;
    int join, cap;
    const float *dashes;
    uint32_t dashCount;
};

// _build_vb_step, src/vkvg_context_internal.c:924-1163
template <typename S>
__device__ bool vb_step(const StrokeCtx &c, S &s, uint32_t iL, uint32_t cp, uint32_t iR, bool isCurve) {
    v2 p0 = c.P[cp];
    v2 v0 = vsub(p0, c.P[iL]);
    v2 v1 = vsub(c.P[iR], p0);
    float lv0 = vlen(v0), lv1 = vlen(v1);
    if (lv0 < VKH_FLT_EPSILON || lv1 < VKH_FLT_EPSILON)
        return false;
    v2 v0n = vdiv(v0, lv0), v1n = vdiv(v1, lv1);
    float dot = vdot(v0n, v1n);
    float det = v0n.x * v1n.y - v0n.y * v1n.x;
    if (equf(dot, 1.0f))
        return false;
    uint32_t idx = s.nv;
    if (equf(dot, -1.0f)) {
        v2 vPerp = vmul(vperp(v0n), c.hw);
        s.addv(vadd(p0, vPerp));
        s.addv(vsub(p0, vPerp));
        s.addt(idx, idx + 1, idx + 2);
        s.addt(idx, idx + 2, idx + 3);
        return true;
    }
    v2 bisec_n = vnorm(vadd(v0n, v1n));
    float alpha = d_acosf(dot);
    if (det < 0)
        alpha = -alpha;
    float halfAlpha = alpha / 2.f;
    float cosHalfAlpha = d_cosf(halfAlpha);
    float lh = c.hw / cosHalfAlpha;
    v2 bisec_n_perp = vperp(bisec_n);
    float rlh = lh;
    if (dot < 0.f)
        rlh = fminf(rlh, fminf(lv0, lv1));
    v2 bisec = vmul(bisec_n_perp, rlh);
    v2 in_pos, out_pos;
    if (rlh < lh) {
        v2 vnPerp = (lv0 < lv1) ? vperp(v1n) : vperp(v0n);
        v2 vHwPerp = vmul(vnPerp, c.hw);
        float lbc = cosHalfAlpha * rlh;
        if (det < 0.f) {
            in_pos = vadd(vadd(vmul(vnPerp, -lbc), vadd(p0, bisec)), vHwPerp);
            out_pos = vsub(p0, vmul(bisec_n_perp, lh));
        } else {
            in_pos = vsub(vadd(vmul(vnPerp, lbc), vsub(p0, bisec)), vHwPerp);
            out_pos = vadd(p0, vmul(bisec_n_perp, lh));
        }
    } else {
        if (det < 0.0f) {
            in_pos = vadd(p0, bisec);
            out_pos = vsub(p0, bisec);
        } else {
            in_pos = vsub(p0, bisec);
            out_pos = vadd(p0, bisec);
        }
    }
    int join = c.join;
    if (isCurve)
        join = (dot < 0.8f) ? 1 : 0;
    if (join == 0) {
        if (lh > c.lhMax) {
            float x = (lh - c.lhMax) * cosHalfAlpha;
            v2 bisecPerp = vmul(bisec_n, x);
            bisec = vmul(bisec_n_perp, c.lhMax);
            if (det < 0) {
                s.addv(in_pos);
                v2 p = vsub(p0, bisec);
                s.addv(vsub(p, bisecPerp));
                s.addv(vadd(p, bisecPerp));
                s.addt(idx, idx + 2, idx + 1);
                s.addt(idx + 2, idx + 4, idx);
                s.addt(idx, idx + 3, idx + 4);
                return true;
            } else {
                v2 p = vadd(p0, bisec);
                s.addv(vsub(p, bisecPerp));
                s.addv(in_pos);
                s.addv(vadd(p, bisecPerp));
                s.addt(idx, idx + 2, idx + 1);
                s.addt(idx + 2, idx + 3, idx + 1);
                s.addt(idx + 1, idx + 3, idx + 4);
                return false;
            }
        } else {
            if (det < 0) {
                s.addv(in_pos);
                s.addv(out_pos);
            } else {
                s.addv(out_pos);
                s.addv(in_pos);
            }
            s.addt(idx, idx + 2, idx + 1);
            s.addt(idx + 1, idx + 2, idx + 3);
            return false;
        }
    }
    v2 vp = vperp(v0n);
    if (det < 0) {
        s.addv((dot < 0 && rlh < lh) ? in_pos : vadd(p0, bisec));
        s.addv(vsub(p0, vmul(vp, c.hw)));
    } else {
        s.addv(vadd(p0, vmul(vp, c.hw)));
        s.addv((dot < 0 && rlh < lh) ? in_pos : vsub(p0, bisec));
    }
    if (join == 2) {
        if (det < 0) {
            s.addt(idx, idx + 2, idx + 1);
            s.addt(idx + 2, idx + 4, idx + 0);
            s.addt(idx, idx + 3, idx + 4);
        } else {
            s.addt(idx, idx + 2, idx + 1);
            s.addt(idx + 2, idx + 3, idx + 1);
            s.addt(idx + 1, idx + 3, idx + 4);
        }
    } else if (join == 1) {
        float a = d_acosf(vp.x);
        if (vp.y < 0)
            a = -a;
        if (det < 0) {
            a += VKH_PIF;
            float a1 = a + alpha;
            a -= c.arcStep;
            while (a > a1) {
                s.addv(V(d_cosf(a) * c.hw + p0.x, d_sinf(a) * c.hw + p0.y));
                a -= c.arcStep;
            }
        } else {
            float a1 = a + alpha;
            a += c.arcStep;
            while (a < a1) {
                s.addv(V(d_cosf(a) * c.hw + p0.x, d_sinf(a) * c.hw + p0.y));
                a += c.arcStep;
            }
        }
        uint32_t p0Idx = s.nv;
        s.addt(idx, idx + 2, idx + 1);
        if (det < 0) {
            for (uint32_t p = idx + 2; p < p0Idx; p++)
                s.addt(p, p + 1, idx);
            s.addt(p0Idx, p0Idx + 2, idx);
            s.addt(idx, p0Idx + 1, p0Idx + 2);
        } else {
            for (uint32_t p = idx + 2; p < p0Idx; p++)
                s.addt(p, p + 1, idx + 1);
            s.addt(p0Idx, p0Idx + 1, idx + 1);
            s.addt(idx + 1, p0Idx + 1, p0Idx + 2);
        }
    }
    vp = vmul(vperp(v1n), c.hw);
    s.addv(det < 0 ? vsub(p0, vp) : vadd(p0, vp));
    return (det < 0);
}

// _draw_stoke_cap, src/vkvg_context_internal.c:1165-1239
template <typename S>
__device__ void stroke_cap(const StrokeCtx &c, S &s, v2 p0, v2 n, bool isStart) {
    uint32_t firstIdx = s.nv;
    v2 vhw = vmul(n, c.hw);
    if (isStart) {
        if (c.cap == 2)
            p0 = vsub(p0, vhw);
        vhw = vperp(vhw);
        if (c.cap == 1) {
            float a = d_acosf(n.x) + VKH_PIF_2;
            if (n.y < 0)
                a = VKH_PIF - a;
            float a1 = a + VKH_PIF;
            a += c.arcStep;
            while (a < a1) {
                s.addv(V(d_cosf(a) * c.hw + p0.x, d_sinf(a) * c.hw + p0.y));
                a += c.arcStep;
            }
            uint32_t p0Idx = s.nv;
            for (uint32_t p = firstIdx; p < p0Idx; p++)
                s.addt(p0Idx + 1, p, p + 1);
            firstIdx = p0Idx;
        }
        s.addv(vadd(p0, vhw));
        s.addv(vsub(p0, vhw));
        s.addt(firstIdx, firstIdx + 2, firstIdx + 1);
        s.addt(firstIdx + 1, firstIdx + 2, firstIdx + 3);
    } else {
        if (c.cap == 2)
            p0 = vadd(p0, vhw);
        vhw = vperp(vhw);
        s.addv(vadd(p0, vhw));
        s.addv(vsub(p0, vhw));
        firstIdx = s.nv;
        if (c.cap == 1) {
            float a = d_acosf(n.x) + VKH_PIF_2;
            if (n.y < 0)
                a = VKH_PIF - a;
            float a1 = a - VKH_PIF;
            a -= c.arcStep;
            while (a > a1) {
                s.addv(V(d_cosf(a) * c.hw + p0.x, d_sinf(a) * c.hw + p0.y));
                a -= c.arcStep;
            }
            uint32_t p0Idx = s.nv - 1;
            for (uint32_t p = firstIdx - 1; p < p0Idx; p++)
                s.addt(p + 1, p, firstIdx - 2);
        }
    }
}

struct DashState {
    bool dashOn;
    uint32_t curDash;
    float curDashOffset, totDashLength;
    v2 normal;
};

// _draw_dashed_segment, src/vkvg_context_internal.c:1240-1264
template <typename S>
__device__ void dashed_segment(const StrokeCtx &c, S &s, DashState &dc, uint32_t iL, uint32_t cp, uint32_t iR,
                               bool isCurve) {
    v2 p = c.P[cp];
    v2 pR = c.P[iR];
    if (!dc.dashOn)
        vb_step(c, s, iL, cp, iR, isCurve);
    v2 d = vsub(pR, p);
    dc.normal = vnorm(d);
    float segLen = vlen(d);
    while (dc.curDashOffset < segLen) {
        v2 p0 = vadd(p, vmul(dc.normal, dc.curDashOffset));
        stroke_cap(c, s, p0, dc.normal, dc.dashOn);
        dc.dashOn = !dc.dashOn;
        dc.curDashOffset += c.dashes[dc.curDash];
        if (++dc.curDash == c.dashCount)
            dc.curDash = 0;
    }
    dc.curDashOffset -= segLen;
    dc.curDashOffset = fmodf(dc.curDashOffset, dc.totDashLength);
}

// A dashed sub-path of _stroke_preserve (src/vkvg_context.c:836-945) in pieces: piece k < last is
// the dashed segment from point k (iL is the last point for k = 0, as there), a closed path adds
// the closing segment as piece `last`, and the last piece also emits the final cap. Run one after
// another over all pieces, dash_start / dash_piece perform the reference's operations in its order;
// k_dash_walk records the dash state at each piece's start so that the pieces can be counted (one
// thread each, dash_piece_count) and emitted independently: by one thread each (k_stroke_steps),
// or, for a piece of many caps, by a wavefront (k_dash_pieces).
__device__ __forceinline__ uint32_t dash_pieces(uint32_t n, bool closed) { return n < 2 ? 0u : (closed ? n : n - 1); }

__device__ __forceinline__ DashState dash_start(const StrokeCtx &c, float dashOffset) {
    DashState dc;
    dc.dashOn = true;
    dc.curDash = 0;
    dc.totDashLength = 0;
    for (uint32_t i = 0; i < c.dashCount; i++)
        dc.totDashLength += c.dashes[i];
    dc.curDashOffset = fmodf(dashOffset, dc.totDashLength);
    dc.normal = V(0, 0);
    return dc;
}

// the segment of piece k: its previous, start and end points and whether it comes from a curve
__device__ __forceinline__ void dash_piece_points(const StrokeCtx &c, uint32_t k, uint32_t &iL, uint32_t &cp,
                                                  uint32_t &iR, bool &curved) {
    const uint32_t last = c.n - 1;
    if (k < last) {
        cp = k;
        iR = k + 1;
        iL = k == 0 ? last : k - 1;
        curved = c.F[cp] != 0;
    } else { // closing segment of a closed path
        cp = last;
        iR = 0;
        iL = last - 1;
        curved = false;
    }
}

// the final cap of a dashed sub-path ending inside a gap (src/vkvg_context.c:936-944), after its
// last piece with the state that piece leaves
template <typename S>
__device__ void dash_final_cap(const StrokeCtx &c, S &s, DashState &dc, uint32_t iR) {
    int32_t prevDash = (int32_t)dc.curDash - 1;
    if (prevDash < 0)
        dc.curDash = c.dashCount - 1;
    float dprev = c.dashes[prevDash < 0 ? (int32_t)c.dashCount - 1 : prevDash];
    float m = fminf(dprev - dc.curDashOffset, c.dashes[dc.curDash]);
    v2 p = vsub(c.P[iR], vmul(dc.normal, m));
    stroke_cap(c, s, p, dc.normal, false);
}

// piece k in one thread, in the reference's order of operations
template <typename S>
__device__ void dash_piece(const StrokeCtx &c, bool closed, uint32_t k, DashState &dc, S &s) {
    uint32_t iL, cp, iR;
    bool curved;
    dash_piece_points(c, k, iL, cp, iR, curved);
    dashed_segment(c, s, dc, iL, cp, iR, curved);
    if (k + 1 == dash_pieces(c.n, closed) && !dc.dashOn)
        dash_final_cap(c, s, dc, iR);
}

// The dash state after piece k without emitting anything: dashed_segment's state updates only
// (the join and caps it emits do not change the state), with the same float operations.
// Returns the number of caps the piece emits (the loop's iterations).
__device__ __forceinline__ uint32_t dash_walk_piece(const StrokeCtx &c, const DashPeriod &per, uint32_t k,
                                                   DashState &dc, bool &over) {
    const uint32_t last = c.n - 1;
    const uint32_t cp = k < last ? k : last, iR = k < last ? k + 1 : 0;
    v2 d = vsub(c.P[iR], c.P[cp]);
    float segLen = vlen(d);
    const uint32_t caps =
        dash_walk_segment(c.dashes, c.dashCount, per, segLen, dc.dashOn, dc.curDash, dc.curDashOffset, over);
    dc.curDashOffset -= segLen;
    dc.curDashOffset = fmodf(dc.curDashOffset, dc.totDashLength);
    return caps;
}

// The vertex and index counts of piece k (dash_piece's) from its start state, without running its
// caps one by one: a cap's counts depend only on whether it starts or ends a dash and on the
// piece's direction (a round cap's arc steps run from the direction's angle; the cap's position
// does not enter), the caps alternate between the two from the start state, and the dash walk
// gives their number and the state the piece ends in.
__device__ void dash_piece_count(const StrokeCtx &c, const DashPeriod &per, bool closed, uint32_t k, DashState dc,
                                 CountSink &s) {
    uint32_t iL, cp, iR;
    bool curved;
    dash_piece_points(c, k, iL, cp, iR, curved);
    if (!dc.dashOn)
        vb_step(c, s, iL, cp, iR, curved);
    const v2 d = vsub(c.P[iR], c.P[cp]);
    const v2 n = vnorm(d);
    const float segLen = vlen(d);
    const bool on0 = dc.dashOn;
    bool over = false; // the dash walk of the same piece has flagged the frame
    const uint32_t caps =
        dash_walk_segment(c.dashes, c.dashCount, per, segLen, dc.dashOn, dc.curDash, dc.curDashOffset, over);
    CountSink cs, ce;
    stroke_cap(c, cs, V(0.f, 0.f), n, true);
    stroke_cap(c, ce, V(0.f, 0.f), n, false);
    const uint32_t starts = on0 ? (caps + 1) / 2 : caps / 2, ends = caps - starts;
    s.nv += starts * cs.nv + ends * ce.nv;
    s.ni += starts * cs.ni + ends * ce.ni;
    if (k + 1 == dash_pieces(c.n, closed) && !dc.dashOn) { // the final cap ends a dash
        s.nv += ce.nv;
        s.ni += ce.ni;
    }
}

// A piece's start state: dashOn in bit 0, the dash index above it, and in bit 31 whether the piece
// has enough caps to be emitted by a whole wavefront (k_dash_pieces) rather than by one thread.
#define VKH_DASH_WAVE_BIT 0x80000000u
#define VKH_DASH_WAVE_MIN 64 // caps from which a piece is given a wavefront
__device__ __forceinline__ uint2 dash_pack(const DashState &dc, bool wave) {
    return make_uint2((dc.dashOn ? 1u : 0u) | (dc.curDash << 1) | (wave ? VKH_DASH_WAVE_BIT : 0u),
                      __float_as_uint(dc.curDashOffset));
}
__device__ __forceinline__ DashState dash_unpack(const StrokeCtx &c, float dashOffset, uint2 ds) {
    DashState dc = dash_start(c, dashOffset);
    dc.dashOn = (ds.x & 1u) != 0;
    dc.curDash = (ds.x & ~VKH_DASH_WAVE_BIT) >> 1;
    dc.curDashOffset = __uint_as_float(ds.y);
    return dc;
}

__device__ __forceinline__ StrokeCtx make_stroke_ctx(const vkh_stroke_style &st, const float *dashes, const v2 *pts,
                                                     const uint8_t *pf, uint32_t first, uint32_t n) {
    StrokeCtx c;
    c.P = pts + first;
    c.F = pf + first;
    c.n = n;
    c.hw = st.hw;
    c.lhMax = st.lhMax;
    c.arcStep = st.arcStep;
    c.join = (int)st.join;
    c.cap = (int)st.cap;
    c.dashes = dashes + st.dash_first;
    c.dashCount = st.dash_count;
    return c;
}

// The dash array of a thread's stroke copied to LDS: the dash walks read one entry per dash
// boundary and the next boundary depends on it, so a global load there is a full memory latency
// per dash. Each thread owns a row of VKH_DASH_LDS + 1 words (an odd stride puts the lanes of a
// wavefront in distinct banks); longer arrays stay in global memory.
#define VKH_DASH_LDS 16
#define VKH_DASH_ROW (VKH_DASH_LDS + 1)
__device__ __forceinline__ void stage_dashes(StrokeCtx &c, float *s_dash) {
    if (c.dashCount == 0 || c.dashCount > VKH_DASH_LDS)
        return;
    float *mine = s_dash + threadIdx.x * VKH_DASH_ROW;
    for (uint32_t i = 0; i < c.dashCount; i++)
        mine[i] = c.dashes[i];
    c.dashes = mine;
}

// One step of a non-dashed sub-path: step 0 and step n-1 are the caps of an open path, the
// others a join at point i; a closed path has a join at every point, the last one closing
// (src/vkvg_context.c:869-935). A step's output depends only on its three points, so steps
// run in parallel; the vertex indices a step emits are global (base = the step's vertex
// offset), so the references a rect makes to the next step's first two vertices resolve to
// the same vertices as in the reference's sequential emission.
template <typename S>
__device__ bool stroke_step(const StrokeCtx &c, bool closed, uint32_t i, S &s) {
    const uint32_t last = c.n - 1;
    if (!closed) {
        if (i == 0 || i == last) {
            v2 a = i == 0 ? c.P[0] : c.P[last - 1], b = i == 0 ? c.P[1] : c.P[last];
            v2 d = V(b.x - a.x, b.y - a.y);
            float md = sqrtf(d.x * d.x + d.y * d.y);
            d.x /= md;
            d.y /= md;
            stroke_cap(c, s, i == 0 ? a : b, d, i == 0);
            return false;
        }
        return vb_step(c, s, i - 1, i, i + 1, c.F[i] != 0);
    }
    if (i == last)
        return vb_step(c, s, last - 1, last, 0, false);
    return vb_step(c, s, i == 0 ? last : i - 1, i, i + 1, c.F[i] != 0);
}

// per unit: fill -> number of polygon edges; stroke -> number of work steps (one per point for
// a solid stroke, one per piece of a dashed sub-path)
__global__ void k_unit_count(const Unit *__restrict__ units, uint32_t n_units, const vkh_draw *__restrict__ draws,
                             const vkh_subpath *__restrict__ sps, const uint32_t *__restrict__ sp_count,
                             const vkh_stroke_style *__restrict__ styles,
                             uint32_t *__restrict__ fill_edges, uint32_t *__restrict__ steps) {
    uint32_t u = blockIdx.x * blockDim.x + threadIdx.x;
    if (u >= n_units)
        return;
    Unit un = units[u];
    const vkh_draw &d = draws[un.draw];
    uint32_t n = sp_count[un.subpath];
    if (d.kind == VKH_DRAW_STROKE) {
        fill_edges[u] = 0;
        const bool closed = (sps[un.subpath].flags & VKH_SP_CLOSED) != 0;
        steps[u] = n < 2 ? 0u : (styles[d.style].dash_count ? dash_pieces(n, closed) : n);
    } else {
        fill_edges[u] = n > 2 ? n : 0;
        steps[u] = 0;
    }
}

__device__ __forceinline__ uint32_t unit_of(const uint32_t *off, uint32_t n, uint32_t v) {
    uint32_t lo = 0, hi = n; // first index with off[i] > v
    while (lo < hi) {
        uint32_t mid = (lo + hi) >> 1;
        if (off[mid] <= v)
            lo = mid + 1;
        else
            hi = mid;
    }
    return lo - 1;
}

// per dashed sub-path (its first piece's step): the dash state at the start of every piece
__global__ void k_dash_walk(const Unit *__restrict__ units, uint32_t n_units, const uint32_t *__restrict__ step_off,
                            uint32_t n_steps, const vkh_draw *__restrict__ draws, const vkh_subpath *__restrict__ sps,
                            const uint32_t *__restrict__ sp_first, const uint32_t *__restrict__ sp_count,
                            const v2 *__restrict__ pts, const uint8_t *__restrict__ pf,
                            const vkh_stroke_style *__restrict__ styles, const float *__restrict__ dashes,
                            uint2 *__restrict__ dash_st, uint4 *__restrict__ dlist, uint32_t *__restrict__ dcount) {
    __shared__ float s_dash[256 * VKH_DASH_ROW];
    const uint32_t u = blockIdx.x * blockDim.x + threadIdx.x;
    if (u >= n_units)
        return;
    const uint32_t s0 = step_off[u], s1 = step_off[u + 1];
    if (s1 <= s0)
        return;
    Unit un = units[u];
    const vkh_draw &d = draws[un.draw];
    if (d.kind != VKH_DRAW_STROKE)
        return;
    const vkh_stroke_style &st = styles[d.style];
    if (!st.dash_count)
        return;
    StrokeCtx c = make_stroke_ctx(st, dashes, pts, pf, sp_first[un.subpath], sp_count[un.subpath]);
    stage_dashes(c, s_dash);
    const DashPeriod per = dash_period(c.dashes, c.dashCount);
    // pieces of many caps join the list that k_dash_pieces takes a wavefront per entry from
    DashState dc = dash_start(c, st.dash_offset);
    // a segment too long for its dash pattern stops the walk and refuses the frame (dcount[1]); the
    // later pieces keep the last state (a valid dash index), so their counting ends at once
    bool over = false;
    for (uint32_t k = 0; k < s1 - s0; k++) {
        const DashState d0 = dc;
        const uint32_t caps = over ? 0u : dash_walk_piece(c, per, k, dc, over);
        const bool wave = !over && caps >= VKH_DASH_WAVE_MIN;
        dash_st[s0 + k] = dash_pack(d0, wave);
        if (wave)
            dlist[atomicAdd(dcount, 1u)] = make_uint4(s0 + k, caps, u, 0u);
    }
    if (over)
        atomicOr(dcount + 1, 1u);
}

// The dash updates of dashed_segment's loop without the cap: the state at the next boundary
__device__ __forceinline__ void dash_advance(const StrokeCtx &c, float dreg, DashState &dc) {
    dc.dashOn = !dc.dashOn;
    dc.curDashOffset += c.dashCount <= 64 ? __shfl(dreg, (int)dc.curDash) : c.dashes[dc.curDash];
    if (++dc.curDash == c.dashCount)
        dc.curDash = 0;
}

// wave-wide exclusive prefix sum of v (all 64 lanes active); total in *tot
__device__ __forceinline__ uint32_t wave_exclusive(uint32_t v, uint32_t lane, uint32_t *tot) {
    uint32_t x = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t y = __shfl_up(x, (unsigned)o);
        if ((int)lane >= o)
            x += y;
    }
    *tot = __shfl(x, 63);
    return x - v;
}

// One wavefront per dashed piece (entries of the list written by k_dash_walk): the piece's
// operations of dashed_segment (and the final cap after the sub-path's last piece), with its caps
// spread over the lanes. Lane l takes the caps l, l + 64,
// ...; it reaches cap j's dash state by applying the loop's updates one by one from the piece's
// start state (the same float additions in the same order as the sequential loop), and the cap
// exists while the offset stays below the segment length. A cap's vertex and index counts come
// from running it against a counting sink; their wave-wide prefix places it, so the caps land where
// the sequential emission puts them and their relative vertex references (a start cap refers to
// the next cap's first two vertices) resolve to the same vertices. The join before the first cap
// and the final cap are emitted by lane 0. (The piece's totals come from dash_piece_count.)
__global__ __launch_bounds__(256) void k_dash_pieces(const Unit *__restrict__ units, uint32_t n_units,
                                                     const uint32_t *__restrict__ step_off,
                                                     const vkh_draw *__restrict__ draws,
                                                     const vkh_subpath *__restrict__ sps,
                                                     const uint32_t *__restrict__ sp_first,
                                                     const uint32_t *__restrict__ sp_count, const v2 *__restrict__ pts,
                                                     const uint8_t *__restrict__ pf,
                                                     const vkh_stroke_style *__restrict__ styles,
                                                     const float *__restrict__ dashes, uint32_t *__restrict__ step_nv,
                                                     uint32_t *__restrict__ step_nt, const uint32_t *__restrict__ sv_off,
                                                     const uint32_t *__restrict__ st_off, v2 *__restrict__ verts,
                                                     uint32_t *__restrict__ tris, const uint2 *__restrict__ dash_st,
                                                     const uint4 *__restrict__ dlist,
                                                     const uint32_t *__restrict__ dcount) {
    const uint32_t lane = threadIdx.x & 63;
    const uint32_t n_list = *dcount;
    const uint32_t stride = gridDim.x * (blockDim.x >> 6);
    for (uint32_t li = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6); li < n_list; li += stride) {
        const uint4 ent = dlist[li]; // step, caps, unit
        const uint32_t sidx = ent.x, u = ent.z;
        const Unit un = units[u];
        const vkh_draw &d = draws[un.draw];
        const vkh_stroke_style &st = styles[d.style];
        const StrokeCtx c = make_stroke_ctx(st, dashes, pts, pf, sp_first[un.subpath], sp_count[un.subpath]);
        const bool closed = (sps[un.subpath].flags & VKH_SP_CLOSED) != 0;
        const uint32_t k = sidx - step_off[u];
        const float dreg = lane < c.dashCount ? c.dashes[lane] : 0.f; // lane i holds dash i
        uint32_t iL, cp, iR;
        bool curved;
        dash_piece_points(c, k, iL, cp, iR, curved);
        DashState dc0 = dash_unpack(c, st.dash_offset, dash_st[sidx]);
        uint32_t nv = sv_off[sidx], ni = 3u * st_off[sidx];
        const v2 p = c.P[cp], pR = c.P[iR];
        if (!dc0.dashOn) { // the join at the piece's start point
            CountSink cs;
            vb_step(c, cs, iL, cp, iR, curved);
            if (lane == 0) {
                EmitSink es;
                es.verts = verts;
                es.nv = nv;
                es.idx = tris + ni;
                es.ni = 0;
                vb_step(c, es, iL, cp, iR, curved);
            }
            nv += cs.nv;
            ni += cs.ni;
        }
        const v2 dv = vsub(pR, p);
        dc0.normal = vnorm(dv);
        const float segLen = vlen(dv);
        // lane l: the state at cap l
        DashState my = dc0;
        const uint32_t setup = min(ent.y, 63u); // lane min(caps, 63) is the last one that matters
        for (uint32_t t = 0; t < setup; t++) { // every lane takes part in the shuffles of the update
            DashState nx = my;
            dash_advance(c, dreg, nx);
            if (t < lane)
                my = nx;
        }
        DashState fin;
        while (true) {
            const bool act = my.curDashOffset < segLen;
            const uint64_t am = __ballot(act);
            CountSink cs;
            const v2 p0 = vadd(p, vmul(my.normal, my.curDashOffset));
            if (act)
                stroke_cap(c, cs, p0, my.normal, my.dashOn);
            uint32_t tv, ti;
            const uint32_t ov = wave_exclusive(cs.nv, lane, &tv), oi = wave_exclusive(cs.ni, lane, &ti);
            if (act) {
                EmitSink es;
                es.verts = verts;
                es.nv = nv + ov;
                es.idx = tris + ni + oi;
                es.ni = 0;
                stroke_cap(c, es, p0, my.normal, my.dashOn);
            }
            nv += tv;
            ni += ti;
            if (am != ~0ull) {
                // the first lane without a cap holds the state after the last cap's update
                const int l0 = __ffsll((unsigned long long)~am) - 1;
                fin.dashOn = __shfl((int)my.dashOn, l0) != 0;
                fin.curDash = (uint32_t)__shfl((int)my.curDash, l0);
                fin.curDashOffset = __shfl(my.curDashOffset, l0);
                fin.totDashLength = my.totDashLength;
                fin.normal = my.normal;
                break;
            }
            for (uint32_t t = 0; t < 64; t++)
                dash_advance(c, dreg, my);
        }
        fin.curDashOffset -= segLen;
        fin.curDashOffset = fmodf(fin.curDashOffset, fin.totDashLength);
        if (k + 1 == dash_pieces(c.n, closed) && !fin.dashOn) {
            CountSink cs;
            DashState f1 = fin;
            dash_final_cap(c, cs, f1, iR);
            if (lane == 0) {
                EmitSink es;
                es.verts = verts;
                es.nv = nv;
                es.idx = tris + ni;
                es.ni = 0;
                DashState f2 = fin;
                dash_final_cap(c, es, f2, iR);
            }
            nv += cs.nv;
            ni += cs.ni;
        }
    }
}

template <bool EMIT>
__global__ void k_stroke_steps(const Unit *__restrict__ units, uint32_t n_units, const uint32_t *__restrict__ step_off,
                               uint32_t n_steps, const vkh_draw *__restrict__ draws,
                               const vkh_subpath *__restrict__ sps, const uint32_t *__restrict__ sp_first,
                               const uint32_t *__restrict__ sp_count, const v2 *__restrict__ pts,
                               const uint8_t *__restrict__ pf, const vkh_stroke_style *__restrict__ styles,
                               const float *__restrict__ dashes, uint32_t *__restrict__ step_nv,
                               uint32_t *__restrict__ step_nt, const uint32_t *__restrict__ sv_off,
                               const uint32_t *__restrict__ st_off, v2 *__restrict__ verts,
                               uint32_t *__restrict__ tris, uint2 *__restrict__ dash_st) {
    __shared__ float s_dash[256 * VKH_DASH_ROW];
    uint32_t sidx = blockIdx.x * blockDim.x + threadIdx.x;
    if (sidx >= n_steps)
        return;
    uint32_t u = unit_of(step_off, n_units, sidx);
    Unit un = units[u];
    const vkh_draw &d = draws[un.draw];
    const vkh_stroke_style &st = styles[d.style];
    uint2 ds = make_uint2(0u, 0u);
    if (st.dash_count) {
        ds = dash_st[sidx];
        if (EMIT && (ds.x & VKH_DASH_WAVE_BIT))
            return; // a piece of many caps: emitted by k_dash_pieces
    }
    StrokeCtx c = make_stroke_ctx(st, dashes, pts, pf, sp_first[un.subpath], sp_count[un.subpath]);
    stage_dashes(c, s_dash);
    bool closed = (sps[un.subpath].flags & VKH_SP_CLOSED) != 0;
    uint32_t i = sidx - step_off[u];
    if (!EMIT) {
        CountSink k;
        if (st.dash_count)
            dash_piece_count(c, dash_period(c.dashes, c.dashCount), closed, i,
                             dash_unpack(c, st.dash_offset, ds), k);
        else
            stroke_step(c, closed, i, k);
        step_nv[sidx] = k.nv;
        step_nt[sidx] = k.ni / 3;
    } else {
        EmitSink k;
        k.verts = verts;
        k.nv = sv_off[sidx];
        k.idx = tris + 3 * (size_t)st_off[sidx];
        k.ni = 0;
        if (st.dash_count) {
            DashState dc = dash_unpack(c, st.dash_offset, ds);
            dash_piece(c, closed, i, dc, k);
        } else {
            stroke_step(c, closed, i, k);
        }
    }
}

// closing patch of src/vkvg_context.c:917-932 for closed solid strokes, applied on the last two
// triangles of the sub-path once every step has written its triangles
__global__ void k_stroke_patch(const Unit *__restrict__ units, uint32_t n_units, const uint32_t *__restrict__ step_off,
                               const vkh_draw *__restrict__ draws, const vkh_subpath *__restrict__ sps,
                               const uint32_t *__restrict__ sp_first, const uint32_t *__restrict__ sp_count,
                               const v2 *__restrict__ pts, const uint8_t *__restrict__ pf,
                               const vkh_stroke_style *__restrict__ styles, const float *__restrict__ dashes,
                               const uint32_t *__restrict__ sv_off, const uint32_t *__restrict__ st_off,
                               uint32_t *__restrict__ tris) {
    uint32_t u = blockIdx.x * blockDim.x + threadIdx.x;
    if (u >= n_units)
        return;
    Unit un = units[u];
    const vkh_draw &d = draws[un.draw];
    if (d.kind != VKH_DRAW_STROKE || !(sps[un.subpath].flags & VKH_SP_CLOSED))
        return;
    const vkh_stroke_style &st = styles[d.style];
    if (st.dash_count)
        return;
    uint32_t s0 = step_off[u], s1 = step_off[u + 1];
    if (s1 <= s0)
        return;
    uint32_t t0 = st_off[s0], t1 = st_off[s1];
    if (t1 - t0 < 2)
        return;
    StrokeCtx c = make_stroke_ctx(st, dashes, pts, pf, sp_first[un.subpath], sp_count[un.subpath]);
    CountSink k;
    bool inverse = stroke_step(c, true, s1 - s0 - 1, k);
    uint32_t ii = sv_off[s0];
    uint32_t *ind = tris + 3 * (size_t)t1 - 6;
    if (inverse) {
        ind[1] = ii + 1;
        ind[4] = ii + 1;
        ind[5] = ii;
    } else {
        ind[1] = ii;
        ind[4] = ii;
        ind[5] = ii + 1;
    }
}

// per unit: first triangle and vertex range of its stroke steps
// Per unit: its first triangle and its vertex range. The first unit of each draw also opens the
// draw's edge range at its first edge slot (its first fill edge, or its first triangle's first
// edge after the ebase fill edges): the edge kernels append the draw's kept edges there
// (erange[d].y is the append cursor), so the range holds no dropped edges.
__global__ void k_unit_ranges(const Unit *__restrict__ units, uint32_t n_units, const vkh_draw *__restrict__ draws,
                              const uint32_t *__restrict__ fill_off, const uint32_t *__restrict__ step_off,
                              const uint32_t *__restrict__ sv_off, const uint32_t *__restrict__ st_off,
                              uint32_t *__restrict__ u_toff, uint32_t *__restrict__ u_vbeg,
                              uint32_t *__restrict__ u_vend, uint32_t ebase, uint2 *__restrict__ erange) {
    uint32_t u = blockIdx.x * blockDim.x + threadIdx.x;
    if (u > n_units)
        return;
    uint32_t s0 = step_off[u];
    const uint32_t t0 = st_off[s0];
    u_toff[u] = t0;
    if (u < n_units) {
        u_vbeg[u] = sv_off[s0];
        u_vend[u] = sv_off[step_off[u + 1]];
        const uint32_t d = units[u].draw;
        if (u == 0 || units[u - 1].draw != d) {
            const uint32_t b = draws[d].kind == VKH_DRAW_STROKE ? ebase + 3u * t0 : fill_off[u];
            erange[d] = make_uint2(b, b);
        }
    }
}

// ---------------------------------------------------------------------------------------- edges
__device__ __forceinline__ uint32_t upper_bound_u32(const uint32_t *a, uint32_t n, uint32_t v) {
    // first index i in [0,n) with a[i] > v (a non-decreasing)
    uint32_t lo = 0, hi = n;
    while (lo < hi) {
        uint32_t mid = (lo + hi) >> 1;
        if (a[mid] <= v)
            lo = mid + 1;
        else
            hi = mid;
    }
    return lo;
}

// upper_bound_u32(a, n, t) - 1 for the wavefront's consecutive indices t = t0 + lane, with the
// dependent global loads of one search shared by the wavefront: a 64-way search (each lane probes
// one position, a ballot narrows the range 64 times) finds lane 0's entry in about log64(n)
// round trips; each lane then finds its own among the next 64 entries by a search over lanes
// (shuffles). A wavefront spanning more than 64 entries finishes those lanes with the plain search.
// Every lane of the wavefront must call it.
__device__ __forceinline__ uint32_t wave_entry_of(const uint32_t *__restrict__ a, uint32_t n, uint32_t t) {
    const uint32_t lane = threadIdx.x & 63;
    const uint32_t t0 = (uint32_t)__shfl((int)t, 0);
    uint32_t lo = 0, hi = n; // upper_bound(t0) in [lo, hi]
    while (hi - lo > 64) {
        const uint32_t step = (hi - lo + 63) / 64;
        const uint32_t p = lo + (lane + 1) * step;
        const bool le = p < hi && a[p] <= t0;
        const uint32_t c = (uint32_t)__popcll((unsigned long long)__ballot(le)); // probes <= t0 form a prefix
        const uint32_t nlo = c ? lo + c * step + 1 : lo;
        const uint32_t pn = lo + (c + 1) * step;
        hi = (c < 64 && pn < hi) ? pn : hi;
        lo = nlo;
    }
    {
        const uint32_t p = lo + lane;
        const bool le = p < hi && a[p] <= t0;
        lo += (uint32_t)__popcll((unsigned long long)__ballot(le));
    }
    const uint32_t u0 = lo - 1; // entry of t0
    const uint32_t q = u0 + 1 + lane;
    const uint32_t v = q < n ? a[q] : 0xFFFFFFFFu; // starts of the next 64 entries, ascending over lanes
    uint32_t k = 0;
#pragma unroll
    for (int sh = 32; sh > 0; sh >>= 1) {
        const uint32_t vv = (uint32_t)__shfl((int)v, (int)(k + (uint32_t)sh - 1));
        if (vv <= t)
            k += (uint32_t)sh;
    }
    if (k == 63 && (uint32_t)__shfl((int)v, 63) <= t) // beyond the 64 entries
        return upper_bound_u32(a, n, t) - 1;
    return u0 + k;
}

// Appends each lane's n edges of draw d to the draw's edge range (erange[d].y is its cursor) and
// reduces the draw's bbox: one set of atomics per distinct draw among the lanes, the lanes of a
// draw placed by a wave-wide prefix of their counts. Returns the lane's first slot. Lanes without
// edges pass VKH_NO_DRAW; every lane of the wavefront must call it.
__device__ __forceinline__ uint32_t wave_draw_append(uint32_t d, uint32_t n, int32_t x0, int32_t y0, int32_t x1,
                                                     int32_t y1, int4 *__restrict__ bbox, uint2 *__restrict__ erange) {
    const uint32_t lane = threadIdx.x & 63;
    uint32_t at = 0;
    uint64_t pending = __ballot(d != VKH_NO_DRAW);
    while (pending) {
        const uint32_t leader = (uint32_t)__ffsll((unsigned long long)pending) - 1;
        const uint32_t dl = __shfl(d, (int)leader);
        const bool mine = d == dl;
        int32_t a = mine ? x0 : INT32_MAX, b = mine ? y0 : INT32_MAX;
        int32_t c = mine ? x1 : INT32_MIN, e = mine ? y1 : INT32_MIN;
        uint32_t x = mine ? n : 0u; // inclusive prefix of the counts of this draw's lanes
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) {
            a = min(a, __shfl_xor(a, o));
            b = min(b, __shfl_xor(b, o));
            c = max(c, __shfl_xor(c, o));
            e = max(e, __shfl_xor(e, o));
        }
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const uint32_t y = __shfl_up(x, o);
            if ((int)lane >= o)
                x += y;
        }
        const uint32_t total = __shfl(x, 63);
        uint32_t b0 = 0;
        if (lane == leader) {
            atomicMin(&bbox[dl].x, a);
            atomicMin(&bbox[dl].y, b);
            atomicMax(&bbox[dl].z, c);
            atomicMax(&bbox[dl].w, e);
            b0 = atomicAdd(&erange[dl].y, total);
        }
        b0 = (uint32_t)__shfl((int)b0, (int)leader);
        if (mine)
            at = b0 + x - n;
        pending &= ~__ballot(mine);
    }
    return at;
}

// Fill edges, one per point of each sub-path (the point to its successor, closing); a degenerate
// edge or one with a vertex that does not snap is dropped. The kept edges are appended to their
// draw's range (opened by k_unit_ranges at the draw's first edge slot).
__global__ void k_fill_edges(const Unit *__restrict__ units, uint32_t n_units, const uint32_t *__restrict__ fill_off,
                             uint32_t total, const vkh_draw *__restrict__ draws, const TargetInfo *__restrict__ tgts,
                             const uint32_t *__restrict__ sp_first, const uint32_t *__restrict__ sp_count,
                             const v2 *__restrict__ pts, Edge *__restrict__ edges, int4 *__restrict__ bbox,
                             uint2 *__restrict__ erange) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t dr = VKH_NO_DRAW;
    Edge e{0, 0, 0, 0};
    const uint32_t u = wave_entry_of(fill_off, n_units, i);
    if (i < total) {
        Unit un = units[u];
        const vkh_draw &d = draws[un.draw];
        uint32_t j = i - fill_off[u];
        uint32_t n = sp_count[un.subpath];
        uint32_t f = sp_first[un.subpath];
        v2 a = pts[f + j], b = pts[f + (j + 1 == n ? 0 : j + 1)];
        float sw = (float)tgts[d.target].width, sh = (float)tgts[d.target].height;
        snap_vertex(a.x, a.y, d.mat, sw, sh, e.x0, e.y0);
        snap_vertex(b.x, b.y, d.mat, sw, sh, e.x1, e.y1);
        bool bad = e.x0 == VKH_INVALID_FIX || e.x1 == VKH_INVALID_FIX || (e.x0 == e.x1 && e.y0 == e.y1);
        dr = bad ? VKH_NO_DRAW : un.draw;
    }
    const uint32_t at = wave_draw_append(dr, dr != VKH_NO_DRAW ? 1u : 0u, min(e.x0, e.x1), min(e.y0, e.y1),
                                         max(e.x0, e.x1), max(e.y0, e.y1), bbox, erange);
    if (dr != VKH_NO_DRAW)
        edges[at] = e;
}

__device__ __forceinline__ bool tri_has_edge(const int32_t (&N)[3][4], int32_t x0, int32_t y0, int32_t x1,
                                              int32_t y1) {
    bool h = false;
#pragma unroll
    for (int k = 0; k < 3; k++)
        h |= N[k][0] == x0 && N[k][1] == y0 && N[k][2] == x1 && N[k][3] == y1;
    return h;
}

// Stroke triangles -> 3 edges each, oriented so every triangle winds +1 (ray to -x, +y edges
// count +1; area > 0 is traversed 0 -> 2 -> 1); indices outside the sub-path's own vertices
// (never drawn by a valid stream) or a degenerate triangle produce no edge. A stroke's coverage count is the sum of its triangles'
// windings, and an edge a->b and its reverse b->a contribute exact negatives to every winding
// term, so such pairs between neighbouring triangles of the same draw (the diagonal of each
// segment quad, the spokes of round joins and caps) are dropped. A pair (t, t+1) is dropped only
// when t-1 does not hold b->a and t+2 does not hold a->b, and only when t-1 .. t+2 lie in the
// same wavefront, so the two triangles always take the same decision. The kept edges are appended
// to their draw's range (k_unit_ranges opened it at the draw's first triangle slot), so the range
// is dense; their order inside it depends on which wavefront appends first, and nothing downstream
// depends on it (every per-tile quantity is an integer sum over the draw's edges).
__global__ __launch_bounds__(256) void k_tri_edges(const Unit *__restrict__ units, uint32_t n_units,
                                                   const uint32_t *__restrict__ u_toff,
                                                   const uint32_t *__restrict__ u_vbeg,
                                                   const uint32_t *__restrict__ u_vend, uint32_t total_tris,
                                                   const vkh_draw *__restrict__ draws,
                                                   const TargetInfo *__restrict__ tgts, const v2 *__restrict__ verts,
                                                   const uint32_t *__restrict__ tris, Edge *__restrict__ edges,
                                                   int4 *__restrict__ bbox, uint2 *__restrict__ erange) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    const bool live = t < total_tris;
    uint32_t draw = VKH_NO_DRAW;
    int32_t X[3] = {0, 0, 0}, Y[3] = {0, 0, 0};
    const uint32_t u = wave_entry_of(u_toff, n_units, t);
    if (live) {
        Unit un = units[u];
        const vkh_draw &d = draws[un.draw];
        uint32_t vb = u_vbeg[u], ve = u_vend[u];
        uint32_t i0 = tris[3 * (size_t)t], i1 = tris[3 * (size_t)t + 1], i2 = tris[3 * (size_t)t + 2];
        bool bad = i0 < vb || i1 < vb || i2 < vb || i0 >= ve || i1 >= ve || i2 >= ve;
        if (!bad) {
            float sw = (float)tgts[d.target].width, sh = (float)tgts[d.target].height;
            snap_vertex(verts[i0].x, verts[i0].y, d.mat, sw, sh, X[0], Y[0]);
            snap_vertex(verts[i1].x, verts[i1].y, d.mat, sw, sh, X[1], Y[1]);
            snap_vertex(verts[i2].x, verts[i2].y, d.mat, sw, sh, X[2], Y[2]);
            bad = X[0] == VKH_INVALID_FIX || X[1] == VKH_INVALID_FIX || X[2] == VKH_INVALID_FIX;
        }
        if (!bad) {
            int64_t a2 = (int64_t)(X[1] - X[0]) * (Y[2] - Y[0]) - (int64_t)(Y[1] - Y[0]) * (X[2] - X[0]);
            bad = (a2 == 0);
            if (a2 > 0) { // positive area winds -1: traverse 0 -> 2 -> 1
                int32_t tx = X[1], ty = Y[1];
                X[1] = X[2];
                Y[1] = Y[2];
                X[2] = tx;
                Y[2] = ty;
            }
        }
        if (!bad)
            draw = un.draw;
    }
    int32_t E[3][4];
#pragma unroll
    for (int k = 0; k < 3; k++) {
        const int k1 = k == 2 ? 0 : k + 1;
        E[k][0] = X[k];
        E[k][1] = Y[k];
        E[k][2] = X[k1];
        E[k][3] = Y[k1];
    }
    // neighbours t-2 .. t+2 through lane shuffles (all 64 lanes take part)
    const int lane = (int)(threadIdx.x & 63);
    bool rev_m1[3], rev_p1[3], same_m2[3], same_p2[3];
#pragma unroll
    for (int o = -2; o <= 2; o++) {
        if (o == 0)
            continue;
        const int src = min(max(lane + o, 0), 63);
        const uint32_t nd = __shfl(draw, src);
        int32_t N[3][4];
#pragma unroll
        for (int k = 0; k < 3; k++)
#pragma unroll
            for (int c = 0; c < 4; c++)
                N[k][c] = __shfl(E[k][c], src);
        const bool ok = draw != VKH_NO_DRAW && nd == draw && lane + o >= 0 && lane + o <= 63;
#pragma unroll
        for (int k = 0; k < 3; k++) {
            const bool rev = ok && tri_has_edge(N, E[k][2], E[k][3], E[k][0], E[k][1]);
            const bool same = ok && tri_has_edge(N, E[k][0], E[k][1], E[k][2], E[k][3]);
            if (o == -1)
                rev_m1[k] = rev;
            if (o == 1)
                rev_p1[k] = rev;
            if (o == -2)
                same_m2[k] = same;
            if (o == 2)
                same_p2[k] = same;
        }
    }
    const bool fwd_visible = lane >= 1 && lane <= 61; // t-1 .. t+2 in this wavefront
    const bool bwd_visible = lane >= 2 && lane <= 62; // t-2 .. t+1 in this wavefront
    int32_t bx0 = INT32_MAX, by0 = INT32_MAX, bx1 = INT32_MIN, by1 = INT32_MIN;
    uint32_t kmask = 0;
#pragma unroll
    for (int k = 0; k < 3; k++) {
        const bool fwd = fwd_visible && rev_p1[k] && !rev_m1[k] && !same_p2[k];
        const bool bwd = bwd_visible && rev_m1[k] && !same_m2[k] && !rev_p1[k];
        if (draw != VKH_NO_DRAW && !fwd && !bwd) {
            kmask |= 1u << k;
            bx0 = min(bx0, min(E[k][0], E[k][2]));
            bx1 = max(bx1, max(E[k][0], E[k][2]));
            by0 = min(by0, min(E[k][1], E[k][3]));
            by1 = max(by1, max(E[k][1], E[k][3]));
        }
    }
    const uint32_t nk = (uint32_t)__builtin_popcount(kmask);
    uint32_t at = wave_draw_append(nk ? draw : VKH_NO_DRAW, nk, bx0, by0, bx1, by1, bbox, erange);
#pragma unroll
    for (int k = 0; k < 3; k++)
        if (kmask & (1u << k)) {
            edges[at] = Edge{E[k][0], E[k][1], E[k][2], E[k][3]};
            at++;
        }
}

// ---------------------------------------------------------------------------------------- launchers
namespace vkh {
uint32_t flat_order_groups(uint32_t n) { return (n + 255) / 256; }
void launch_flat_hist(hipStream_t s, const vkh_item *items, uint32_t n, uint32_t *whist) {
    if (n)
        hipLaunchKernelGGL(k_flat_hist, dim3(flat_order_groups(n)), dim3(256), 0, s, items, n, whist,
                           flat_order_groups(n));
}
void launch_flat_scatter(hipStream_t s, const vkh_item *items, uint32_t n, const uint32_t *wcur, uint32_t *perm) {
    if (n)
        hipLaunchKernelGGL(k_flat_scatter, dim3(flat_order_groups(n)), dim3(256), 0, s, items, n, wcur,
                           flat_order_groups(n), perm);
}
void launch_flatten_stage(hipStream_t s, const vkh_item *items, uint32_t n, const uint32_t *perm, uint32_t *counts,
                          uint32_t *slot_alloc, uint32_t *slot_item, uint32_t *item_slot, v2 *scratch,
                          uint32_t *ovf_count, uint32_t *ovf_list) {
    if (n)
        hipLaunchKernelGGL(k_flatten_stage, dim3((n + VKH_FLAT_BLOCK - 1) / VKH_FLAT_BLOCK), dim3(VKH_FLAT_BLOCK), 0, s,
                           items, n, perm, counts, slot_alloc,
                           slot_item, item_slot, scratch, ovf_count, ovf_list);
}
void launch_flatten_copy(hipStream_t s, const vkh_item *items, uint32_t n_slots, const uint32_t *slot_item,
                         const uint32_t *counts, const uint32_t *offs, const v2 *scratch, v2 *pts, uint8_t *pf) {
    const size_t threads = (size_t)n_slots * VKH_FLAT_COPY_LANES;
    if (threads)
        hipLaunchKernelGGL(k_flatten_copy, dim3((unsigned)((threads + 255) / 256)), dim3(256), 0, s, items, n_slots,
                           slot_item, counts, offs, scratch, pts, pf);
}
void launch_flatten_lines(hipStream_t s, const vkh_item *items, uint32_t n, const uint32_t *offs, v2 *pts,
                          uint8_t *pf) {
    if (n)
        hipLaunchKernelGGL(k_flatten_lines, dim3((n + 255) / 256), dim3(256), 0, s, items, n, offs, pts, pf);
}
void launch_flatten_overflow(hipStream_t s, const vkh_item *items, const uint32_t *list, uint32_t n_list,
                             const uint32_t *offs, v2 *pts, uint8_t *pf) {
    if (n_list)
        hipLaunchKernelGGL(k_flatten_overflow, dim3((n_list + 255) / 256), dim3(256), 0, s, items, list, n_list, offs,
                           pts, pf);
}
void launch_subpath_ranges(hipStream_t s, const vkh_subpath *sp, uint32_t n, const uint32_t *item_off,
                           uint32_t *first, uint32_t *count) {
    if (n)
        hipLaunchKernelGGL(k_subpath_ranges, dim3((n + 255) / 256), dim3(256), 0, s, sp, n, item_off, first, count);
}
void launch_unit_count(hipStream_t s, const Unit *units, uint32_t n_units, const vkh_draw *draws,
                       const vkh_subpath *sps, const uint32_t *spc, const vkh_stroke_style *styles, uint32_t *fe,
                       uint32_t *steps) {
    if (n_units)
        hipLaunchKernelGGL(k_unit_count, dim3((n_units + 255) / 256), dim3(256), 0, s, units, n_units, draws, sps, spc,
                           styles, fe, steps);
}
void launch_stroke_steps(hipStream_t s, bool emit, const Unit *units, uint32_t n_units, const uint32_t *step_off,
                         uint32_t n_steps, const vkh_draw *draws, const vkh_subpath *sps, const uint32_t *spf,
                         const uint32_t *spc, const v2 *pts, const uint8_t *pf, const vkh_stroke_style *styles,
                         const float *dashes, uint32_t *step_nv, uint32_t *step_nt, const uint32_t *sv_off,
                         const uint32_t *st_off, v2 *verts, uint32_t *tris, uint2 *dash_st, uint4 *dlist,
                         uint32_t *dcount) {
    if (!n_steps)
        return;
    if (!emit)
        hipLaunchKernelGGL(k_dash_walk, dim3((n_units + 255) / 256), dim3(256), 0, s, units, n_units, step_off, n_steps,
                           draws, sps, spf, spc, pts, pf, styles, dashes, dash_st, dlist, dcount);
    // a wavefront per dashed piece, at most 4096 wavefronts walking the list
    const uint32_t dblocks = (uint32_t)std::min<uint64_t>(((uint64_t)n_steps + 3) / 4, 1024);
    if (emit) {
        hipLaunchKernelGGL(k_stroke_steps<true>, dim3((n_steps + 255) / 256), dim3(256), 0, s, units, n_units, step_off,
                           n_steps, draws, sps, spf, spc, pts, pf, styles, dashes, step_nv, step_nt, sv_off, st_off,
                           verts, tris, dash_st);
        hipLaunchKernelGGL(k_dash_pieces, dim3(dblocks), dim3(256), 0, s, units, n_units, step_off, draws, sps, spf,
                           spc, pts, pf, styles, dashes, step_nv, step_nt, sv_off, st_off, verts, tris, dash_st, dlist,
                           dcount);
    } else {
        hipLaunchKernelGGL(k_stroke_steps<false>, dim3((n_steps + 255) / 256), dim3(256), 0, s, units, n_units,
                           step_off, n_steps, draws, sps, spf, spc, pts, pf, styles, dashes, step_nv, step_nt, sv_off,
                           st_off, verts, tris, dash_st);
    }
}
void launch_stroke_patch(hipStream_t s, const Unit *units, uint32_t n_units, const uint32_t *step_off,
                         const vkh_draw *draws, const vkh_subpath *sps, const uint32_t *spf, const uint32_t *spc,
                         const v2 *pts, const uint8_t *pf, const vkh_stroke_style *styles, const float *dashes,
                         const uint32_t *sv_off, const uint32_t *st_off, uint32_t *tris) {
    if (n_units)
        hipLaunchKernelGGL(k_stroke_patch, dim3((n_units + 255) / 256), dim3(256), 0, s, units, n_units, step_off,
                           draws, sps, spf, spc, pts, pf, styles, dashes, sv_off, st_off, tris);
}
void launch_unit_ranges(hipStream_t s, const Unit *units, uint32_t n_units, const vkh_draw *draws,
                        const uint32_t *fill_off, const uint32_t *step_off, const uint32_t *sv_off,
                        const uint32_t *st_off, uint32_t *u_toff, uint32_t *u_vbeg, uint32_t *u_vend, uint32_t ebase,
                        uint2 *erange) {
    hipLaunchKernelGGL(k_unit_ranges, dim3((n_units + 1 + 255) / 256), dim3(256), 0, s, units, n_units, draws,
                       fill_off, step_off, sv_off, st_off, u_toff, u_vbeg, u_vend, ebase, erange);
}
void launch_fill_edges(hipStream_t s, const Unit *units, uint32_t n_units, const uint32_t *fill_off, uint32_t total,
                       const vkh_draw *draws, const TargetInfo *tg, const uint32_t *spf, const uint32_t *spc,
                       const v2 *pts, Edge *edges, int4 *bbox, uint2 *erange) {
    if (total)
        hipLaunchKernelGGL(k_fill_edges, dim3((total + 255) / 256), dim3(256), 0, s, units, n_units, fill_off, total,
                           draws, tg, spf, spc, pts, edges, bbox, erange);
}
void launch_tri_edges(hipStream_t s, const Unit *units, uint32_t n_units, const uint32_t *u_toff,
                      const uint32_t *u_vbeg, const uint32_t *u_vend, uint32_t total_tris, const vkh_draw *draws,
                      const TargetInfo *tg, const v2 *verts, const uint32_t *tris, Edge *edges, int4 *bbox,
                      uint2 *erange) {
    if (total_tris)
        hipLaunchKernelGGL(k_tri_edges, dim3((total_tris + 255) / 256), dim3(256), 0, s, units, n_units, u_toff,
                           u_vbeg, u_vend, total_tris, draws, tg, verts, tris, edges, bbox, erange);
}
} // namespace vkh
