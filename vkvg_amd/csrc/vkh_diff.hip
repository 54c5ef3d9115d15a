// Differentiable renderer of the vkvg HIP backend: analytic area coverage into FP32 RGBA and its
// backward pass (C ABI of include/vkvg_diff.h). Semantics are stated in that header.
//
// Pipeline per forward call:
//   k_dflatten      one thread per polyline point: Bernstein sample of its cubic
//   k_dnormals      one thread per point: unit left normal of the segment leaving it (strokes)
//   k_dedges        one thread per point: the fill edge, or the stroke's rectangle + 2 bevel
//                   triangles (10 edges), each endpoint tagged (point, normal, sign) for backward
//   k_dbbox         one workgroup per path: bbox of its edges (deterministic reduction)
//   draw tiles, scan, k_dbin (per path, LDS counts / scan / cursors; every tile from the edge's
//   left end to the path's right bbox end, so a tile's list holds all edges left of it), coarse
//   k_dcmd_tile     the tile of every command
//   k_dunit_count, scan, k_dunit_cmd: every command's segment list cut into work units of at
//                   most 64 segments
//   k_dcov          16 units per workgroup pass, whatever their commands and tiles: one thread per
//                   (unit, pixel row) accumulates the row's area deltas in a fixed order; a
//                   one-unit command is prefix-summed into the coverage buffer (A per command and
//                   pixel) at once, the units of longer ones are added in unit order by k_dcov_reduce
//   k_work_order    the tiles heaviest first (commands per tile), for k_dcomp and k_dback
//   k_dcomp         one workgroup per 16x16 tile, one thread per pixel: composites the tile's
//                   commands in draw order from the coverage buffer. With checkpoints on, the pixel
//                   state before every batch of 8 commands is kept, and so is the coverage buffer.
// Backward:
//   k_dback         per tile, batches in reverse: reads the batch's coverage, replays the
//                   batch's compositing from its checkpoint, walks it backwards to get dL/dA per
//                   (command, pixel) and the colour gradients, then turns dL/dA into edge-endpoint
//                   gradients with the boundary integral over every edge piece inside a pixel (one
//                   thread per edge touching the tile)
//   k_dg_normal_sums, k_dg_points, k_dg_ctrl, k_dg_paths: chain rule to normals, half-widths,
//                   points, control points, colours and widths, all as gathers
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "vkh_internal.h"
#include "vkvg_diff.h"

#define VKD_NSEG 16 // polyline segments per cubic
#define VKD_T 16    // tile size in pixels
#define VKD_B 8     // commands per batch of the backward pass (and per compositing checkpoint)
#define VKD_FB 16   // work units per coverage pass of the forward: 16 units x 16 rows = 256 threads
#define VKD_U 64      // segments per work unit of the coverage pass
#define VKD_UC 16     // segments of each unit staged in LDS at a time
#define VKD_BIN_MARGIN 0.01f // px added around an edge when binning (rounding safety)

namespace {

struct DSrc { // endpoint tags of an edge: point index, normal index | (sign + 1) << 30
    uint32_t ia, ka, ib, kb;
};

__device__ __forceinline__ float2 f2(float x, float y) { return make_float2(x, y); }

// geometry in R = float (product path) or double (the FP64 evaluation used by finite differences)
template <typename R> struct alignas(2 * sizeof(R)) RV2 {
    R x, y;
};
template <typename R> struct alignas(4 * sizeof(R)) RV4 {
    R x, y, z, w;
};
template <typename R> __device__ __forceinline__ RV2<R> r2(R x, R y) { return RV2<R>{x, y}; }

// --------------------------------------------------------------------------------- geometry
template <typename R>
__global__ void k_dflatten(uint32_t n_pts, const uint32_t *__restrict__ pt_path,
                           const uint32_t *__restrict__ path_first_pt, const uint32_t *__restrict__ path_first_ctrl,
                           const uint32_t *__restrict__ path_ncubic, const float2 *__restrict__ ctrl,
                           RV2<R> *__restrict__ pts) {
    uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n_pts)
        return;
    uint32_t path = pt_path[p];
    uint32_t j = p - path_first_pt[path];
    uint32_t q = j / VKD_NSEG, i = j % VKD_NSEG;
    uint32_t n3 = 3 * path_ncubic[path];
    const float2 *c = ctrl + path_first_ctrl[path];
    const float2 c0 = c[3 * q], c1 = c[3 * q + 1], c2 = c[3 * q + 2], c3 = c[(3 * q + 3) % n3];
    const R t = (R)i / (R)VKD_NSEG, u = (R)1 - t;
    const R b0 = u * u * u, b1 = (R)3 * u * u * t, b2 = (R)3 * u * t * t, b3 = t * t * t;
    pts[p] = r2<R>(b0 * (R)c0.x + b1 * (R)c1.x + b2 * (R)c2.x + b3 * (R)c3.x,
                   b0 * (R)c0.y + b1 * (R)c1.y + b2 * (R)c2.y + b3 * (R)c3.y);
}

__device__ __forceinline__ uint32_t next_pt(uint32_t p, uint32_t first, uint32_t end) {
    return p + 1 == end ? first : p + 1;
}
__device__ __forceinline__ uint32_t prev_pt(uint32_t p, uint32_t first, uint32_t end) {
    return p == first ? end - 1 : p - 1;
}

template <typename R>
__global__ void k_dnormals(uint32_t n_pts, const uint32_t *__restrict__ pt_path,
                           const uint32_t *__restrict__ path_first_pt, const RV2<R> *__restrict__ pts,
                           RV2<R> *__restrict__ nrm) {
    uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n_pts)
        return;
    uint32_t path = pt_path[p];
    uint32_t nx = next_pt(p, path_first_pt[path], path_first_pt[path + 1]);
    R dx = pts[nx].x - pts[p].x, dy = pts[nx].y - pts[p].y;
    R L = sqrt(dx * dx + dy * dy);
    nrm[p] = L > (R)0 ? r2<R>(-dy / L, dx / L) : r2<R>((R)0, (R)0);
}

__device__ __forceinline__ uint32_t tag(uint32_t k, int s) { return k | ((uint32_t)(s + 1) << 30); }
template <typename R>
__device__ __forceinline__ RV2<R> vpos(const RV2<R> *pts, const RV2<R> *nrm, R h, uint32_t i, uint32_t k, int s) {
    if (!s)
        return pts[i];
    R sh = (R)s * h;
    return r2<R>(pts[i].x + sh * nrm[k].x, pts[i].y + sh * nrm[k].y);
}

template <typename R>
__global__ void k_dedges(uint32_t n_pts, const uint32_t *__restrict__ pt_path,
                         const uint32_t *__restrict__ path_first_pt, const uint32_t *__restrict__ path_first_edge,
                         const uint32_t *__restrict__ path_stroke, const float *__restrict__ width,
                         const RV2<R> *__restrict__ pts, const RV2<R> *__restrict__ nrm, RV4<R> *__restrict__ edges,
                         DSrc *__restrict__ src) {
    uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n_pts)
        return;
    uint32_t path = pt_path[p];
    uint32_t first = path_first_pt[path], end = path_first_pt[path + 1];
    uint32_t j = p - first;
    uint32_t nx = next_pt(p, first, end);
    if (!path_stroke[path]) {
        uint32_t e = path_first_edge[path] + j;
        edges[e] = RV4<R>{pts[p].x, pts[p].y, pts[nx].x, pts[nx].y};
        if (src)
            src[e] = DSrc{p, tag(0, 0), nx, tag(0, 0)};
        return;
    }
    const R h = (R)0.5 * (R)width[path];
    uint32_t e = path_first_edge[path] + 10 * j;
    uint32_t vi[4] = {p, nx, nx, p}, vk[4] = {p, p, p, p};
    int vs[4] = {1, 1, -1, -1};
    for (int a = 0; a < 4; a++) {
        int b = (a + 1) & 3;
        RV2<R> A = vpos(pts, nrm, h, vi[a], vk[a], vs[a]), Bv = vpos(pts, nrm, h, vi[b], vk[b], vs[b]);
        edges[e] = RV4<R>{A.x, A.y, Bv.x, Bv.y};
        if (src)
            src[e] = DSrc{vi[a], tag(vk[a], vs[a]), vi[b], tag(vk[b], vs[b])};
        e++;
    }
    uint32_t pp = prev_pt(p, first, end);
    RV2<R> n0 = nrm[pp], n1 = nrm[p];
    bool swap = n0.x * n1.y - n0.y * n1.x > (R)0;
    for (int s = 1; s >= -1; s -= 2) {
        uint32_t ti[3] = {p, p, p}, tk[3] = {0, pp, p};
        int ts[3] = {0, s, s};
        if (swap) {
            tk[1] = p;
            tk[2] = pp;
        }
        for (int a = 0; a < 3; a++) {
            int b = a == 2 ? 0 : a + 1;
            RV2<R> A = vpos(pts, nrm, h, ti[a], tk[a], ts[a]), Bv = vpos(pts, nrm, h, ti[b], tk[b], ts[b]);
            edges[e] = RV4<R>{A.x, A.y, Bv.x, Bv.y};
            if (src)
                src[e] = DSrc{ti[a], tag(tk[a], ts[a]), ti[b], tag(tk[b], ts[b])};
            e++;
        }
    }
}

// bbox of a path's edges in 24.8 fixed point (floor / ceil), one workgroup per path
template <typename R>
__global__ __launch_bounds__(256) void k_dbbox(const uint32_t *__restrict__ path_first_edge,
                                               const RV4<R> *__restrict__ edges, int4 *__restrict__ bbox) {
    __shared__ int32_t red[4][4];
    uint32_t d = blockIdx.x, tid = threadIdx.x;
    float x0 = 3.0e38f, y0 = 3.0e38f, x1 = -3.0e38f, y1 = -3.0e38f;
    for (uint32_t e = path_first_edge[d] + tid; e < path_first_edge[d + 1]; e += 256) {
        const RV4<R> v = edges[e];
        if (v.y == v.w || isnan(v.x) || isnan(v.y) || isnan(v.z) || isnan(v.w))
            continue; // horizontal edges contribute nothing
        // widened by the binning margin so that rounding to float never loses a tile
        x0 = fminf(x0, (float)fmin(v.x, v.z) - VKD_BIN_MARGIN);
        x1 = fmaxf(x1, (float)fmax(v.x, v.z) + VKD_BIN_MARGIN);
        y0 = fminf(y0, (float)fmin(v.y, v.w) - VKD_BIN_MARGIN);
        y1 = fmaxf(y1, (float)fmax(v.y, v.w) + VKD_BIN_MARGIN);
    }
    for (int o = 32; o > 0; o >>= 1) {
        x0 = fminf(x0, __shfl_xor(x0, o));
        y0 = fminf(y0, __shfl_xor(y0, o));
        x1 = fmaxf(x1, __shfl_xor(x1, o));
        y1 = fmaxf(y1, __shfl_xor(y1, o));
    }
    auto fx = [](float v, bool up) -> int32_t {
        float s = v * 256.0f;
        s = up ? ceilf(s) : floorf(s);
        return (int32_t)fminf(fmaxf(s, -8.0e6f), 8.0e6f);
    };
    if ((tid & 63) == 0) {
        int w = tid >> 6;
        red[w][0] = x0 > x1 ? INT32_MAX : fx(x0, false);
        red[w][1] = x0 > x1 ? INT32_MAX : fx(y0, false);
        red[w][2] = x0 > x1 ? INT32_MIN : fx(x1, true);
        red[w][3] = x0 > x1 ? INT32_MIN : fx(y1, true);
    }
    __syncthreads();
    if (tid == 0) {
        int4 b = make_int4(INT32_MAX, INT32_MAX, INT32_MIN, INT32_MIN);
        for (int w = 0; w < 4; w++) {
            b.x = min(b.x, red[w][0]);
            b.y = min(b.y, red[w][1]);
            b.z = max(b.z, red[w][2]);
            b.w = max(b.w, red[w][3]);
        }
        bbox[d] = b;
    }
}

// ---------------------------------------------------------------------------------- binning
// One workgroup per path, in bands of whole tile rows fitting the LDS table: an edge goes into
// every tile of the rows it spans from its leftmost tile to the bbox's right end.
#define VKD_BIN_LDS 4096
__device__ __forceinline__ int32_t tile_of(float v) { // clamped before the conversion
    return (int32_t)fminf(fmaxf(floorf(v / VKD_T), -1.0e6f), 1.0e6f);
}
// tile rows and first tile column of an edge, widened by VKD_BIN_MARGIN px; empty (r0 > r1) for
// edges with a NaN coordinate
template <typename R>
__device__ __forceinline__ void dbin_range(const RV4<R> &v, const DrawTiles &dt, int32_t ty_lo, int32_t ty_hi,
                                           int32_t &r0, int32_t &r1, int32_t &c0) {
    if (isnan(v.x) || isnan(v.y) || isnan(v.z) || isnan(v.w)) {
        r0 = 1;
        r1 = 0;
        c0 = 0;
        return;
    }
    const float ymin = (float)fmin(v.y, v.w) - VKD_BIN_MARGIN, ymax = (float)fmax(v.y, v.w) + VKD_BIN_MARGIN;
    const float xmin = (float)fmin(v.x, v.z) - VKD_BIN_MARGIN;
    r0 = max(tile_of(ymin), ty_lo);
    r1 = min(tile_of(ymax), ty_hi);
    c0 = max(tile_of(xmin), dt.bx0);
}

__device__ __forceinline__ uint32_t block_excl_scan(uint32_t *a, uint32_t n, uint32_t *s_wave) {
    const uint32_t tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const uint32_t beg = min(tid * 16u, n), end = min(beg + 16u, n);
    uint32_t sum = 0;
    for (uint32_t k = beg; k < end; k++)
        sum += a[k];
    uint32_t x = sum;
    for (int o = 1; o < 64; o <<= 1) {
        uint32_t y = __shfl_up(x, o);
        if ((int)lane >= o)
            x += y;
    }
    if (lane == 63)
        s_wave[w] = x;
    __syncthreads();
    uint32_t run = x - sum, total = 0;
    for (uint32_t k = 0; k < 4; k++) {
        run += k < w ? s_wave[k] : 0u;
        total += s_wave[k];
    }
    for (uint32_t k = beg; k < end; k++) {
        uint32_t v = a[k];
        a[k] = run;
        run += v;
    }
    __syncthreads();
    return total;
}

// Per tile the list holds the edges touching the tile's columns first (PathTile.reserved of
// them), then the edges entirely left of it, which only carry their rows' dy.
template <typename R>
__device__ __forceinline__ int32_t last_touch_col(const RV4<R> &v) {
    return tile_of((float)fmax(v.x, v.z) + VKD_BIN_MARGIN);
}

template <typename R>
__global__ __launch_bounds__(256) void k_dbin(const uint32_t *__restrict__ path_first_edge,
                                              const RV4<R> *__restrict__ edges, const DrawTiles *__restrict__ dts,
                                              PathTile *__restrict__ ptl, uint32_t *__restrict__ segs, uint32_t seg_cap,
                                              uint32_t *__restrict__ seg_alloc) {
    __shared__ uint32_t s_cnt[VKD_BIN_LDS];   // counts, then offsets, then cursors of touching edges
    __shared__ uint32_t s_touch[VKD_BIN_LDS]; // touching counts, then cursors of left-only edges
    __shared__ uint32_t s_wave[4];
    __shared__ uint32_t s_base;
    const uint32_t d = blockIdx.x, tid = threadIdx.x;
    const DrawTiles dt = dts[d];
    if (dt.bw == 0 || dt.bx0 > dt.bx1 || dt.by0 > dt.by1)
        return;
    const uint32_t e0 = path_first_edge[d], e1 = path_first_edge[d + 1];
    const uint32_t bw = dt.bw, bh = (uint32_t)(dt.by1 - dt.by0 + 1);
    const uint32_t band_rows = max(1u, VKD_BIN_LDS / bw);
    for (uint32_t rb = 0; rb < bh; rb += band_rows) {
        const uint32_t nr = min(band_rows, bh - rb), nt = nr * bw;
        const int32_t ty_lo = dt.by0 + (int32_t)rb, ty_hi = ty_lo + (int32_t)nr - 1;
        for (uint32_t k = tid; k < nt; k += 256) {
            s_cnt[k] = 0;
            s_touch[k] = 0;
        }
        __syncthreads();
        for (uint32_t e = e0 + tid; e < e1; e += 256) {
            const RV4<R> v = edges[e];
            if (v.y == v.w)
                continue;
            int32_t r0, r1, c0;
            dbin_range(v, dt, ty_lo, ty_hi, r0, r1, c0);
            const int32_t ct = last_touch_col(v);
            for (int32_t r = r0; r <= r1; r++)
                for (int32_t c = c0; c <= dt.bx1; c++) {
                    const uint32_t k = (uint32_t)(r - ty_lo) * bw + (uint32_t)(c - dt.bx0);
                    atomicAdd(&s_cnt[k], 1u);
                    if (c <= ct)
                        atomicAdd(&s_touch[k], 1u);
                }
        }
        __syncthreads();
        const uint32_t total = block_excl_scan(s_cnt, nt, s_wave);
        if (tid == 0)
            s_base = atomicAdd(seg_alloc, total);
        __syncthreads();
        const uint32_t base = s_base;
        const bool fits = (uint64_t)base + total <= (uint64_t)seg_cap;
        PathTile *out = ptl + dt.pt_base + rb * bw;
        for (uint32_t k = tid; k < nt; k += 256) {
            uint32_t off = s_cnt[k], next = k + 1 < nt ? s_cnt[k + 1] : total;
            PathTile pt;
            pt.backdrop = 0;
            pt.nseg = next - off;
            pt.seg_off = base + off;
            pt.reserved = s_touch[k]; // edges touching the tile, listed first
            out[k] = pt;
        }
        __syncthreads();
        for (uint32_t k = tid; k < nt; k += 256)
            s_touch[k] += s_cnt[k]; // cursor of the left-only part
        __syncthreads();
        if (fits) {
            for (uint32_t e = e0 + tid; e < e1; e += 256) {
                const RV4<R> v = edges[e];
                if (v.y == v.w)
                    continue;
                int32_t r0, r1, c0;
                dbin_range(v, dt, ty_lo, ty_hi, r0, r1, c0);
                const int32_t ct = last_touch_col(v);
                for (int32_t r = r0; r <= r1; r++)
                    for (int32_t c = c0; c <= dt.bx1; c++) {
                        const uint32_t k = (uint32_t)(r - ty_lo) * bw + (uint32_t)(c - dt.bx0);
                        const uint32_t pos = c <= ct ? atomicAdd(&s_cnt[k], 1u) : atomicAdd(&s_touch[k], 1u);
                        segs[base + pos] = e;
                    }
            }
        }
        __syncthreads();
    }
}

// --------------------------------------------------------------------------- edge pieces
// Calls f(lc, dy, fm, s0, s1) for the pieces of edge v inside pixel row [yr, yr + 1): lc = -1 for
// the part left of the tile column X0 (merged), 0..T-1 for tile cells; dy = (s1 - s0)(by - ay)
// signed; fm the piece's mean x inside its cell; s0 < s1 its parameter range on the edge.
template <typename R, typename F>
__device__ __forceinline__ void row_pieces(const RV4<R> &v, R yr, R X0, F f) {
    const R ax = v.x, ay = v.y, bx = v.z, by = v.w;
    if (ay == by)
        return;
    const R ylo = fmax(fmin(ay, by), yr), yhi = fmin(fmax(ay, by), yr + (R)1);
    if (!(ylo < yhi))
        return;
    const R dyt = by - ay, dxt = bx - ax;
    const R sa = (ylo - ay) / dyt, sb = (yhi - ay) / dyt;
    const R s0 = fmin(sa, sb), s1 = fmax(sa, sb);
    const R x0 = ax + s0 * dxt, x1 = ax + s1 * dxt;
    const R xl = fmin(x0, x1), xr = fmax(x0, x1);
    const bool vert = bx == ax;
    // parameters at the row segment's ends come from the y clip; x-derived split points are used
    // only strictly inside (clamped), which keeps nearly vertical edges well conditioned
    const R sL = x0 <= x1 ? s0 : s1, sR = x0 <= x1 ? s1 : s0;
    // a piece spanning its whole row segment takes dy from the y clip itself, so the carries of a
    // closed outline cancel exactly to the right of it
    const R dy_row = dyt > (R)0 ? yhi - ylo : ylo - yhi;
    auto emit = [&](int lc, R xa, R xb) {
        R p0 = s0, p1 = s1;
        bool whole = true;
        if (!vert) {
            const bool at_l = xa <= xl, at_r = xb >= xr;
            R u = at_l ? sL : fmin(fmax((xa - ax) / dxt, s0), s1);
            R w = at_r ? sR : fmin(fmax((xb - ax) / dxt, s0), s1);
            p0 = fmin(u, w);
            p1 = fmax(u, w);
            whole = at_l && at_r;
        }
        R dy = whole ? dy_row : (p1 - p0) * dyt;
        R fm = lc >= 0 ? (R)0.5 * (xa + xb) - (X0 + (R)lc) : (R)0;
        f(lc, dy, fm, p0, p1);
    };
    if (xl < X0)
        emit(-1, xl, fmin(xr, X0));
    const int32_t cl = (int32_t)floor(xl) - (int32_t)X0;
    const int32_t cr = (xr > xl ? (int32_t)floor(xr) : (int32_t)floor(xl)) - (int32_t)X0;
    for (int32_t c = max(cl, 0); c <= min(cr, VKD_T - 1); c++) {
        const R cx = X0 + (R)c;
        const R xa = fmax(xl, cx), xb = fmin(xr, cx + (R)1);
        if (xb < xa || (xb == xa && xr > xl))
            continue;
        emit(c, xa, xb);
    }
}

// ---------------------------------------------------------------------------------- forward
__device__ __forceinline__ float4 premul(float4 c) { return make_float4(c.x * c.w, c.y * c.w, c.z * c.w, c.w); }

// stages a batch's command records and the prefix of their segment counts
__device__ __forceinline__ void stage_batch(uint32_t base, uint32_t nb, const Cmd *__restrict__ cmds,
                                            const PathTile *__restrict__ ptl, uint32_t *c_draw, uint32_t *c_nseg,
                                            uint32_t *c_ntouch, uint32_t *c_segoff, uint32_t *c_item) {
    const uint32_t tid = threadIdx.x;
    if (tid < nb) {
        const Cmd cm = cmds[base + tid];
        const PathTile p = ptl[cm.pt];
        c_draw[tid] = cm.draw;
        c_nseg[tid] = p.nseg;
        c_ntouch[tid] = p.reserved;
        c_segoff[tid] = p.seg_off;
    }
    __syncthreads();
    if (tid == 0) {
        uint32_t acc_n = 0;
        for (uint32_t j = 0; j < nb; j++) {
            c_item[j] = acc_n;
            acc_n += c_nseg[j];
        }
        c_item[nb] = acc_n;
    }
    __syncthreads();
}

// the tile of every command (commands are stored tile by tile)
__global__ void k_dcmd_tile(uint32_t n_tiles, uint32_t nch, const uint32_t *__restrict__ tile_off,
                            uint32_t *__restrict__ cmd_tile) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n_tiles)
        return;
    for (uint32_t c = tile_off[(size_t)t * nch]; c < tile_off[(size_t)(t + 1) * nch]; c++)
        cmd_tile[c] = t;
}

// Work units of the coverage pass: every command's segment list cut into pieces of at most VKD_U
// segments (at least one unit per command), so a command with thousands of segments is spread
// over many row groups instead of holding one. unit_cnt[c] = that count.
__global__ void k_dunit_count(uint32_t n_cmds, const Cmd *__restrict__ cmds, const PathTile *__restrict__ ptl,
                              uint32_t *__restrict__ unit_cnt) {
    const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n_cmds)
        return;
    const uint32_t n = ptl[cmds[c].pt].nseg;
    unit_cnt[c] = n ? (n + VKD_U - 1) / VKD_U : 1u;
}

// the command of every work unit
__global__ void k_dunit_cmd(uint32_t n_cmds, uint32_t max_units, const uint32_t *__restrict__ unit_off,
                            uint32_t *__restrict__ unit_cmd) {
    const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n_cmds)
        return;
    for (uint32_t u = unit_off[c]; u < min(unit_off[c + 1], max_units); u++)
        unit_cmd[u] = c;
}

// Coverage of every command, independent of compositing order: 16 work units per workgroup pass
// (whatever their commands and tiles), one thread per (unit, pixel row). The units' edges are
// staged in LDS VKD_UC per unit at a time, so all 16 row groups advance together; each row thread
// adds the pieces of its unit's segments in list order. A command held by one unit gets its rows
// prefix-summed into cov[c * 256 + pixel] here; the rows of a command cut into several units go to
// part[u * 256 + row * 16 + column] and k_dcov_reduce adds them in unit order, so the result does
// not depend on scheduling. Grid-stride over units: the unit count is read on the device.
template <typename R>
__global__ __launch_bounds__(256) void k_dcov(uint32_t n_cmds, uint32_t max_units, uint32_t tilesW,
                                              const uint32_t *__restrict__ unit_off,
                                              const uint32_t *__restrict__ unit_cmd,
                                              const uint32_t *__restrict__ cmd_tile, const Cmd *__restrict__ cmds,
                                              const PathTile *__restrict__ ptl, const uint32_t *__restrict__ segs,
                                              const RV4<R> *__restrict__ edges, R *__restrict__ cov,
                                              R *__restrict__ part) {
    __shared__ R acc[VKD_FB][VKD_T][VKD_T + 1];
    __shared__ RV4<R> s_edge[VKD_FB][VKD_UC];
    __shared__ uint32_t u_cmd[VKD_FB], u_s0[VKD_FB], u_n[VKD_FB], u_nt[VKD_FB], u_off[VKD_FB], u_tile[VKD_FB];
    __shared__ uint32_t u_single[VKD_FB];
    const uint32_t tid = threadIdx.x, j = tid / VKD_T, r = tid % VKD_T;
    const uint32_t total = min(unit_off[n_cmds], max_units); // never below unit_off[n_cmds]: the host's bound
    R *a = acc[j][r];
    for (uint32_t base = blockIdx.x * VKD_FB; base < total; base += gridDim.x * VKD_FB) {
        const uint32_t nb = min((uint32_t)VKD_FB, total - base);
        if (tid < nb) {
            const uint32_t u = base + tid, c = unit_cmd[u];
            const PathTile p = ptl[cmds[c].pt];
            const uint32_t s0 = (u - unit_off[c]) * VKD_U;
            u_cmd[tid] = c;
            u_s0[tid] = s0;
            u_n[tid] = min(p.nseg - min(p.nseg, s0), (uint32_t)VKD_U);
            u_nt[tid] = p.reserved;
            u_off[tid] = p.seg_off;
            u_tile[tid] = cmd_tile[c];
            u_single[tid] = unit_off[c + 1] - unit_off[c] == 1u;
        }
        for (int c = 0; c <= VKD_T; c++)
            a[c] = (R)0;
        __syncthreads();
        const bool active = j < nb;
        const uint32_t n = active ? u_n[j] : 0u, s0 = active ? u_s0[j] : 0u, nt = active ? u_nt[j] : 0u;
        const uint32_t t = active ? u_tile[j] : 0u;
        const R X0 = (R)((t % tilesW) * VKD_T), yr = (R)((t / tilesW) * VKD_T + r);
        uint32_t nmax = 0;
        for (uint32_t q = 0; q < nb; q++)
            nmax = max(nmax, u_n[q]);
        for (uint32_t k0 = 0; k0 < nmax; k0 += VKD_UC) {
            if (k0 + r < n) // thread (j, r) stages item k0 + r of unit j
                s_edge[j][r] = edges[segs[u_off[j] + s0 + k0 + r]];
            __syncthreads();
            const uint32_t m = n > k0 ? min((uint32_t)VKD_UC, n - k0) : 0u;
            for (uint32_t q = 0; q < m; q++) {
                const RV4<R> v = s_edge[j][q];
                if (s0 + k0 + q < nt) {
                    row_pieces<R>(v, yr, X0, [&](int lc, R dy, R fm, R, R) {
                        if (lc < 0) {
                            a[0] += dy;
                        } else {
                            a[lc] += dy * ((R)1 - fm);
                            a[lc + 1] += dy * fm;
                        }
                    });
                } else { // edges left of the tile: the row's whole y-extent, as row_pieces gives for them
                    const R ylo = fmax(fmin(v.y, v.w), yr), yhi = fmin(fmax(v.y, v.w), yr + (R)1);
                    if (ylo < yhi)
                        a[0] += v.w > v.y ? yhi - ylo : ylo - yhi;
                }
            }
            __syncthreads();
        }
        if (active) {
            if (u_single[j]) {
                R *out = cov + ((size_t)u_cmd[j] * VKD_T + r) * VKD_T;
                R run = (R)0;
#pragma unroll
                for (int c = 0; c < VKD_T; c++) {
                    run += a[c];
                    out[c] = run;
                }
            } else {
                R *out = part + ((size_t)(base + j) * VKD_T + r) * VKD_T;
#pragma unroll
                for (int c = 0; c < VKD_T; c++)
                    out[c] = a[c];
            }
        }
        __syncthreads();
    }
}

// rows of the commands cut into several units: partial rows added in unit order, then prefix-summed
template <typename R>
__global__ void k_dcov_reduce(uint32_t n_cmds, const uint32_t *__restrict__ unit_off, const R *__restrict__ part,
                              R *__restrict__ cov) {
    const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t c = g / VKD_T, r = g % VKD_T;
    if (c >= n_cmds)
        return;
    const uint32_t u0 = unit_off[c], u1 = unit_off[c + 1];
    if (u1 - u0 < 2)
        return;
    R sum[VKD_T];
    const R *p = part + ((size_t)u0 * VKD_T + r) * VKD_T;
#pragma unroll
    for (int k = 0; k < VKD_T; k++)
        sum[k] = p[k];
    for (uint32_t u = u0 + 1; u < u1; u++) {
        p = part + ((size_t)u * VKD_T + r) * VKD_T;
#pragma unroll
        for (int k = 0; k < VKD_T; k++)
            sum[k] += p[k];
    }
    R *out = cov + ((size_t)c * VKD_T + r) * VKD_T;
    R run = (R)0;
#pragma unroll
    for (int k = 0; k < VKD_T; k++) {
        run += sum[k];
        out[k] = run;
    }
}

// Compositing in draw order: one workgroup per 16x16 tile, one thread per pixel; coverage comes
// from k_dcov. With checkpoints on, the pixel state before every batch of VKD_B commands is kept
// for the backward pass.
template <typename R>
__global__ __launch_bounds__(256) void k_dcomp(uint32_t W, uint32_t H, uint32_t tilesW, uint32_t nch,
                                               const uint32_t *__restrict__ tile_off, const Cmd *__restrict__ cmds,
                                               const R *__restrict__ cov, const float4 *__restrict__ rgba,
                                               RV4<R> *__restrict__ image, float4 *__restrict__ ckpt,
                                               const uint32_t *__restrict__ batch_off,
                                               const uint32_t *__restrict__ tile_order) {
    const uint32_t t = tile_order[blockIdx.x], tid = threadIdx.x;
    const uint32_t c0 = tile_off[(size_t)t * nch], c1 = tile_off[(size_t)(t + 1) * nch];
    const int32_t px = (int32_t)((t % tilesW) * VKD_T + tid % VKD_T), py = (int32_t)((t / tilesW) * VKD_T + tid / VKD_T);
    RV4<R> st{(R)1, (R)1, (R)1, (R)1};
    for (uint32_t c = c0; c < c1; c++) {
        if (ckpt && ((c - c0) % VKD_B) == 0)
            ckpt[(size_t)(batch_off[t] + (c - c0) / VKD_B) * 256 + tid] =
                make_float4((float)st.x, (float)st.y, (float)st.z, (float)st.w);
        const R A = cov[(size_t)c * 256 + tid];
        const R k = fmin((R)1, fabs(A));
        const float4 cf = rgba[cmds[c].draw];
        const R ca = (R)cf.w;
        const R cx = (R)cf.x * ca, cy = (R)cf.y * ca, cz = (R)cf.z * ca;
        const R ia = (R)1 - ca * k;
        st = RV4<R>{cx * k + st.x * ia, cy * k + st.y * ia, cz * k + st.z * ia, ca * k + st.w * ia};
    }
    if (px < (int32_t)W && py < (int32_t)H)
        image[(size_t)py * W + px] = st;
}

// --------------------------------------------------------------------------------- backward
// sums each of v[0..31] over the 64 lanes of the wavefront with 31 + 1 shuffles; lane l returns
// the total of value index (l >> 1) & 31, written to *idx
__device__ __forceinline__ float wave_multi_sum32(float (&v)[32], uint32_t lane, uint32_t *idx) {
    uint32_t base = 0;
#pragma unroll
    for (int step = 0; step < 5; step++) {
        const int o = 32 >> step;  // 32, 16, 8, 4, 2
        const int half = 16 >> step; // values kept after this step
        const bool hi = (lane & (uint32_t)o) != 0;
#pragma unroll
        for (int i = 0; i < half; i++) {
            float send = hi ? v[i] : v[i + half];
            float keep = hi ? v[i + half] : v[i];
            v[i] = keep + __shfl_xor(send, o);
        }
        if (hi)
            base += (uint32_t)half;
    }
    *idx = base;
    return v[0] + __shfl_xor(v[0], 1);
}

__global__ __launch_bounds__(256) void k_dback(uint32_t W, uint32_t H, uint32_t tilesW, uint32_t nch,
                                               const uint32_t *__restrict__ tile_off, const Cmd *__restrict__ cmds,
                                               const PathTile *__restrict__ ptl, const uint32_t *__restrict__ segs,
                                               const RV4<float> *__restrict__ edges, const float4 *__restrict__ rgba,
                                               const float4 *__restrict__ grad_img, const float4 *__restrict__ ckpt,
                                               const uint32_t *__restrict__ batch_off, const float *__restrict__ cov,
                                               float4 *__restrict__ g_edge, float *__restrict__ g_col,
                                               const uint32_t *__restrict__ tile_order) {
    __shared__ float G[VKD_B][VKD_T][VKD_T];
    __shared__ float s_red[4][32];
    __shared__ uint32_t c_item[VKD_B + 1], c_sitem[VKD_B + 1];
    __shared__ uint32_t c_nseg[VKD_B], c_ntouch[VKD_B], c_segoff[VKD_B], c_draw[VKD_B];
    const uint32_t t = tile_order[blockIdx.x], tid = threadIdx.x, lane = tid & 63;
    const uint32_t c0 = tile_off[(size_t)t * nch], c1 = tile_off[(size_t)(t + 1) * nch];
    if (c0 == c1)
        return;
    const int32_t X0 = (int32_t)(t % tilesW) * VKD_T, Y0 = (int32_t)(t / tilesW) * VKD_T;
    const int32_t px = X0 + (int32_t)(tid % VKD_T), py = Y0 + (int32_t)(tid / VKD_T);
    const bool valid = px < (int32_t)W && py < (int32_t)H;
    float4 g = valid ? grad_img[(size_t)py * W + px] : make_float4(0.f, 0.f, 0.f, 0.f);
    const uint32_t nbatch = (c1 - c0 + VKD_B - 1) / VKD_B;
    for (uint32_t bb = nbatch; bb-- > 0;) {
        const uint32_t base = c0 + bb * VKD_B;
        const uint32_t nb = min((uint32_t)VKD_B, c1 - base);
        stage_batch(base, nb, cmds, ptl, c_draw, c_nseg, c_ntouch, c_segoff, c_sitem);
        // replay the batch's compositing from its checkpoint, then walk it backwards
        float4 st[VKD_B + 1];
        st[0] = ckpt[(size_t)(batch_off[t] + bb) * 256 + tid];
        float Av[VKD_B];
#pragma unroll
        for (uint32_t j = 0; j < VKD_B; j++) {
            Av[j] = j < nb ? cov[(size_t)(base + j) * 256 + tid] : 0.f;
            if (j < nb) {
                const float k = fminf(1.0f, fabsf(Av[j]));
                const float4 c = premul(rgba[c_draw[j]]);
                const float ia = 1.0f - c.w * k;
                st[j + 1] = make_float4(c.x * k + st[j].x * ia, c.y * k + st[j].y * ia, c.z * k + st[j].z * ia,
                                        c.w * k + st[j].w * ia);
            } else {
                st[j + 1] = st[j];
            }
        }
        float gc[32];
#pragma unroll
        for (int q = 0; q < 32; q++)
            gc[q] = 0.f;
#pragma unroll
        for (int jj = VKD_B - 1; jj >= 0; jj--) {
            const uint32_t j = (uint32_t)jj;
            if (j >= nb)
                continue;
            const float A = Av[j];
            const float k = fminf(1.0f, fabsf(A));
            const float fp = fabsf(A) < 1.0f ? (A > 0.f ? 1.f : (A < 0.f ? -1.f : 0.f)) : 0.f;
            const float4 c = premul(rgba[c_draw[j]]);
            const float4 s = st[j];
            float dk = 0.f, gs = 0.f;
            if (valid) {
                dk = g.x * (c.x - c.w * s.x) + g.y * (c.y - c.w * s.y) + g.z * (c.z - c.w * s.z) + g.w * (c.w - c.w * s.w);
                gs = g.x * s.x + g.y * s.y + g.z * s.z + g.w * s.w;
                gc[4 * j + 0] = g.x * k;
                gc[4 * j + 1] = g.y * k;
                gc[4 * j + 2] = g.z * k;
                gc[4 * j + 3] = g.w * k - gs * k;
            }
            G[j][tid / VKD_T][tid % VKD_T] = valid ? dk * fp : 0.f;
            const float ia = 1.0f - c.w * k;
            g = make_float4(g.x * ia, g.y * ia, g.z * ia, g.w * ia);
        }
        // colour gradients: reduce the 8 commands x 4 channels over the tile
        uint32_t idx;
        const float tot = wave_multi_sum32(gc, lane, &idx);
        if ((lane & 1) == 0)
            s_red[tid >> 6][idx] = tot;
        __syncthreads();
        if (tid < 4 * nb) {
            const float v = s_red[0][tid] + s_red[1][tid] + s_red[2][tid] + s_red[3][tid];
            if (v != 0.f)
                atomicAdd(&g_col[4 * c_draw[tid >> 2] + (tid & 3)], v);
        }
        // edge-endpoint gradients: one thread per touching segment of the batch integrates dL/dA
        // along its pieces in every row of the tile it spans (one set of atomics per edge and tile)
        if (tid == 0) {
            uint32_t acc_n = 0;
            for (uint32_t j = 0; j < nb; j++) {
                c_item[j] = acc_n;
                acc_n += c_ntouch[j];
            }
            c_item[nb] = acc_n;
        }
        __syncthreads();
        for (uint32_t it = tid; it < c_item[nb]; it += 256) {
            uint32_t lo = 0, hi = nb; // last j with c_item[j] <= it
            while (hi - lo > 1) {
                const uint32_t mid = (lo + hi) >> 1;
                if (c_item[mid] <= it)
                    lo = mid;
                else
                    hi = mid;
            }
            const uint32_t j = lo;
            const uint32_t ei = segs[c_segoff[j] + (it - c_item[j])];
            const RV4<float> v = edges[ei];
            const int32_t r0 = max((int32_t)floorf(fminf(v.y, v.w)) - Y0, 0);
            const int32_t r1 = min((int32_t)floorf(fmaxf(v.y, v.w)) - Y0, VKD_T - 1);
            float ga = 0.f, gb = 0.f; // scalar factors of (-dy, dx)
            for (int32_t r = r0; r <= r1; r++) {
                row_pieces<float>(v, (float)(Y0 + r), (float)X0, [&](int lc, float, float, float p0, float p1) {
                    if (lc < 0)
                        return;
                    const float Gv = G[j][r][lc];
                    const float I1 = 0.5f * (p1 * p1 - p0 * p0);
                    const float I0 = (p1 - p0) - I1;
                    ga += Gv * I0;
                    gb += Gv * I1;
                });
            }
            if (ga != 0.f || gb != 0.f) {
                const float dx = v.z - v.x, dy = v.w - v.y;
                float *ge = (float *)(g_edge + ei);
                atomicAdd(ge + 0, -dy * ga);
                atomicAdd(ge + 1, dx * ga);
                atomicAdd(ge + 2, -dy * gb);
                atomicAdd(ge + 3, dx * gb);
            }
        }
        __syncthreads();
    }
}

// --------------------------------------------------------------------------- chain rule
// Chain rule from edge endpoints to points, normals and half-widths, as gathers (no atomics).
// Edge layout per path: a fill has edge j = p_j -> p_j+1; a stroke has, per point j, 10 edges
// [10j, 10j+10): the rectangle of segment j (points j, j+1, normal j), then the two bevel triangles
// at vertex j (point j, normals j-1 and j). So the endpoints that use point p or normal p lie in the
// edges of points p-1 and p (point p) or p and p+1 (normal p).
__device__ __forceinline__ void endpoint(const DSrc &s, int which, uint32_t &i, uint32_t &k, int &sg) {
    i = which ? s.ib : s.ia;
    const uint32_t t = which ? s.kb : s.ka;
    k = t & 0x3FFFFFFFu;
    sg = (int)(t >> 30) - 1;
}

// per stroke point p: g_n[p] and the half-width partial sum of the endpoints using normal p
__global__ void k_dg_normal_sums(uint32_t n_pts, const uint32_t *__restrict__ pt_path,
                                 const uint32_t *__restrict__ path_first_pt,
                                 const uint32_t *__restrict__ path_first_edge, const uint32_t *__restrict__ path_stroke,
                                 const float *__restrict__ width, const DSrc *__restrict__ src,
                                 const float4 *__restrict__ g_edge, const float2 *__restrict__ nrm,
                                 float2 *__restrict__ g_n, float *__restrict__ g_hp) {
    uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n_pts)
        return;
    const uint32_t path = pt_path[p];
    if (!path_stroke[path]) {
        g_n[p] = f2(0.f, 0.f);
        g_hp[p] = 0.f;
        return;
    }
    const uint32_t first = path_first_pt[path], end = path_first_pt[path + 1];
    const float h = 0.5f * width[path];
    float gx = 0.f, gy = 0.f, gh = 0.f;
    const uint32_t owners[2] = {p, next_pt(p, first, end)};
    for (int o = 0; o < 2; o++) {
        const uint32_t e0 = path_first_edge[path] + 10 * (owners[o] - first);
        for (uint32_t e = e0; e < e0 + 10; e++) {
            const DSrc sr = src[e];
            const float4 ge = g_edge[e];
            for (int w = 0; w < 2; w++) {
                uint32_t i, k;
                int sg;
                endpoint(sr, w, i, k, sg);
                if (sg == 0 || k != p)
                    continue;
                const float vx = w ? ge.z : ge.x, vy = w ? ge.w : ge.y;
                gx += (float)sg * h * vx;
                gy += (float)sg * h * vy;
                gh += (float)sg * (nrm[p].x * vx + nrm[p].y * vy);
            }
        }
    }
    g_n[p] = f2(gx, gy);
    g_hp[p] = gh;
}

// per point p: endpoint gradients using point p, plus, for strokes, the normals' dependence on
// the segments p-1 -> p and p -> p+1: n = (-d.y, d.x)/|d|, g_d = ((g.y, -g.x) - d (n.g)/|d|)/|d|
__device__ __forceinline__ float2 seg_grad(const float2 *pts, const float2 *nrm, const float2 *g_n, uint32_t a,
                                           uint32_t b) {
    const float2 gn = g_n[a];
    const float dx = pts[b].x - pts[a].x, dy = pts[b].y - pts[a].y;
    const float L = sqrtf(dx * dx + dy * dy);
    if (!(L > 0.f) || (gn.x == 0.f && gn.y == 0.f))
        return f2(0.f, 0.f);
    const float ng = nrm[a].x * gn.x + nrm[a].y * gn.y;
    return f2((gn.y - dx * ng / L) / L, (-gn.x - dy * ng / L) / L);
}

__global__ void k_dg_points(uint32_t n_pts, const uint32_t *__restrict__ pt_path,
                            const uint32_t *__restrict__ path_first_pt, const uint32_t *__restrict__ path_first_edge,
                            const uint32_t *__restrict__ path_stroke, const DSrc *__restrict__ src,
                            const float4 *__restrict__ g_edge, const float2 *__restrict__ pts,
                            const float2 *__restrict__ nrm, const float2 *__restrict__ g_n,
                            float2 *__restrict__ g_pt) {
    uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n_pts)
        return;
    const uint32_t path = pt_path[p];
    const uint32_t first = path_first_pt[path], end = path_first_pt[path + 1];
    const uint32_t pp = prev_pt(p, first, end), nx = next_pt(p, first, end);
    const bool stroke = path_stroke[path] != 0;
    const uint32_t per = stroke ? 10u : 1u;
    float gx = 0.f, gy = 0.f;
    const uint32_t owners[2] = {pp, p};
    for (int o = 0; o < 2; o++) {
        const uint32_t e0 = path_first_edge[path] + per * (owners[o] - first);
        for (uint32_t e = e0; e < e0 + per; e++) {
            const DSrc sr = src[e];
            const float4 ge = g_edge[e];
            if (sr.ia == p) {
                gx += ge.x;
                gy += ge.y;
            }
            if (sr.ib == p) {
                gx += ge.z;
                gy += ge.w;
            }
        }
    }
    if (stroke) {
        const float2 din = seg_grad(pts, nrm, g_n, pp, p); // segment pp -> p: +g_d at its end p
        const float2 dout = seg_grad(pts, nrm, g_n, p, nx); // segment p -> nx: -g_d at its start p
        gx += din.x - dout.x;
        gy += din.y - dout.y;
    }
    g_pt[p] = f2(gx, gy);
}

// one thread per control point: gathers the Bernstein-weighted point gradients of its cubic(s)
__global__ void k_dg_ctrl(uint32_t n_ctrl, const uint32_t *__restrict__ ctrl_path,
                          const uint32_t *__restrict__ path_first_ctrl, const uint32_t *__restrict__ path_first_pt,
                          const uint32_t *__restrict__ path_ncubic, const float2 *__restrict__ g_pt,
                          float2 *__restrict__ g_ctrl) {
    uint32_t ci = blockIdx.x * blockDim.x + threadIdx.x;
    if (ci >= n_ctrl)
        return;
    const uint32_t path = ctrl_path[ci];
    const uint32_t local = ci - path_first_ctrl[path], n = path_ncubic[path];
    const uint32_t q = local / 3, m = local % 3;
    const float2 *gp = g_pt + path_first_pt[path];
    float gx = 0.f, gy = 0.f;
    for (int i = 0; i < VKD_NSEG; i++) {
        const float t = (float)i / (float)VKD_NSEG, u = 1.0f - t;
        const float B[4] = {u * u * u, 3.0f * u * u * t, 3.0f * u * t * t, t * t * t};
        const float2 a = gp[q * VKD_NSEG + i];
        gx += B[m] * a.x;
        gy += B[m] * a.y;
        if (m == 0) { // also the end point of the previous cubic
            const uint32_t qp = (q + n - 1) % n;
            const float2 bq = gp[qp * VKD_NSEG + i];
            gx += B[3] * bq.x;
            gy += B[3] * bq.y;
        }
    }
    g_ctrl[ci] = f2(gx, gy);
}

// per path: colour chain c = (r a, g a, b a, a) and width = 2 h
__global__ void k_dg_paths(uint32_t n_paths, const float4 *__restrict__ rgba, const float *__restrict__ g_col,
                           const float *__restrict__ g_hp, const uint32_t *__restrict__ path_first_pt,
                           const uint32_t *__restrict__ path_stroke, float4 *__restrict__ g_rgba,
                           float *__restrict__ g_width) {
    uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n_paths)
        return;
    const float4 c = rgba[p];
    const float g0 = g_col[4 * p], g1 = g_col[4 * p + 1], g2 = g_col[4 * p + 2], g3 = g_col[4 * p + 3];
    g_rgba[p] = make_float4(c.w * g0, c.w * g1, c.w * g2, c.x * g0 + c.y * g1 + c.z * g2 + g3);
    float gh = 0.f;
    if (path_stroke[p])
        for (uint32_t k = path_first_pt[p]; k < path_first_pt[p + 1]; k++)
            gh += g_hp[k];
    g_width[p] = 0.5f * gh;
}

__global__ void k_dbatches(uint32_t n_tiles, uint32_t nch, const uint32_t *__restrict__ tile_off,
                           uint32_t *__restrict__ nb) {
    uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n_tiles)
        nb[t] = (tile_off[(size_t)(t + 1) * nch] - tile_off[(size_t)t * nch] + VKD_B - 1) / VKD_B;
}

__global__ void k_dzero(uint32_t *p, size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n)
        p[i] = 0;
}

struct Buf {
    void *p = nullptr;
    size_t cap = 0;
};

} // namespace

struct vkvg_diff_scene_t {
    uint32_t W = 0, H = 0, n_paths = 0, n_pts = 0, n_ctrl = 0, n_edges = 0, tilesW = 0, tilesH = 0, n_tiles = 0;
    uint32_t bin_total = 0, nch = 1, chunk = VKH_COARSE_CHUNK;
    hipStream_t stream = nullptr;
    hipEvent_t ev[4] = {};
    std::string err;
    float ms_fwd = 0, ms_bwd = 0;
    uint32_t n_segs = 0, n_cmds = 0, n_batches = 0;
    bool have_ckpt = false;
    // static topology
    Buf pt_path, ctrl_path, first_pt, first_ctrl, first_edge, ncubic, stroke, draws, tg;
    // per call
    Buf pts, nrm, edges, pts64, nrm64, edges64, src, bbox, dtiles, pt_count, pt_off, ptiles, segs, seg_alloc, tile_cnt, tile_off, cmds, nbat,
        batch_off, ckpt, cmd_tile, unit_cnt, unit_off, unit_cmd, cov, part, tile_order, scan_tmp, host_tot;
    Buf g_edge, g_pt, g_n, g_h, g_col;
};

namespace {

bool ck(vkvg_diff_scene_t *s, hipError_t e, const char *what) {
    if (e == hipSuccess)
        return true;
    s->err = std::string(what) + ": " + hipGetErrorString(e);
    return false;
}
bool ensure(vkvg_diff_scene_t *s, Buf &b, size_t bytes) {
    if (b.p && bytes <= b.cap)
        return true;
    if (b.p)
        (void)hipFree(b.p);
    b.p = nullptr;
    b.cap = 0;
    size_t cap = bytes + bytes / 4 + 256;
    if (!ck(s, hipMalloc(&b.p, cap), "hipMalloc"))
        return false;
    b.cap = cap;
    return true;
}
template <typename T> T *P(Buf &b) { return (T *)b.p; }
bool upload(vkvg_diff_scene_t *s, Buf &b, const void *h, size_t bytes) {
    return ensure(s, b, bytes) && ck(s, hipMemcpy(b.p, h, bytes, hipMemcpyHostToDevice), "upload");
}
bool read_u32(vkvg_diff_scene_t *s, const uint32_t *dptr, uint32_t *out) {
    if (!ck(s, hipMemcpyAsync(out, dptr, 4, hipMemcpyDeviceToHost, s->stream), "readback"))
        return false;
    return ck(s, hipStreamSynchronize(s->stream), "sync");
}
bool scan(vkvg_diff_scene_t *s, const uint32_t *in, uint32_t n, uint32_t *out) {
    if (!ensure(s, s->scan_tmp, vkh::scan_temp_words(n) * 4))
        return false;
    vkh::scan_u32(s->stream, in, n, out, P<uint32_t>(s->scan_tmp));
    return ck(s, hipGetLastError(), "scan");
}
void zero(vkvg_diff_scene_t *s, Buf &b, size_t words) {
    if (words)
        hipLaunchKernelGGL(k_dzero, dim3((unsigned)((words + 255) / 256)), dim3(256), 0, s->stream, P<uint32_t>(b),
                           words);
}
unsigned blocks(size_t n) { return (unsigned)((n + 255) / 256); }

} // namespace

extern "C" {

VkvgDiffScene vkvg_diff_create(uint32_t width, uint32_t height, uint32_t n_paths, const uint32_t *ncubic,
                               const uint32_t *is_stroke) {
    if (!width || !height || width > 16384 || height > 16384 || !n_paths || !ncubic || !is_stroke)
        return nullptr;
    auto *s = new vkvg_diff_scene_t();
    s->W = width;
    s->H = height;
    s->n_paths = n_paths;
    s->tilesW = (width + VKD_T - 1) / VKD_T;
    s->tilesH = (height + VKD_T - 1) / VKD_T;
    s->n_tiles = s->tilesW * s->tilesH;
    std::vector<uint32_t> first_pt(n_paths + 1), first_ctrl(n_paths + 1), first_edge(n_paths + 1), pt_path, ctrl_path;
    std::vector<uint32_t> stroke(n_paths), nc(n_paths);
    uint64_t np = 0, nctl = 0, ne = 0;
    for (uint32_t i = 0; i < n_paths; i++) {
        if (ncubic[i] == 0 || ncubic[i] > (1u << 20)) {
            delete s;
            return nullptr;
        }
        first_pt[i] = (uint32_t)np;
        first_ctrl[i] = (uint32_t)nctl;
        first_edge[i] = (uint32_t)ne;
        nc[i] = ncubic[i];
        stroke[i] = is_stroke[i] ? 1u : 0u;
        uint64_t m = (uint64_t)ncubic[i] * VKD_NSEG;
        np += m;
        nctl += 3ull * ncubic[i];
        ne += stroke[i] ? 10 * m : m;
        if (np > (1u << 28) || ne > (1u << 30)) {
            delete s;
            return nullptr;
        }
        pt_path.insert(pt_path.end(), (size_t)m, i);
        ctrl_path.insert(ctrl_path.end(), 3 * (size_t)ncubic[i], i);
    }
    first_pt[n_paths] = (uint32_t)np;
    first_ctrl[n_paths] = (uint32_t)nctl;
    first_edge[n_paths] = (uint32_t)ne;
    s->n_pts = (uint32_t)np;
    s->n_ctrl = (uint32_t)nctl;
    s->n_edges = (uint32_t)ne;
    // draw records for the shared tile / coarse kernels: one target, kind per path
    std::vector<vkh_draw> dr(n_paths);
    memset(dr.data(), 0, sizeof(vkh_draw) * n_paths);
    for (uint32_t i = 0; i < n_paths; i++) {
        dr[i].kind = stroke[i] ? VKH_DRAW_STROKE : VKH_DRAW_FILL_NZ;
        dr[i].target = 0;
        dr[i].grad = -1;
        dr[i].source = -1;
    }
    TargetInfo ti;
    memset(&ti, 0, sizeof(ti));
    ti.width = width;
    ti.height = height;
    ti.tilesW = s->tilesW;
    ti.tilesH = s->tilesH;
    ti.tile_base = 0;
    ti.binsW = (s->tilesW + VKH_BIN_TILES - 1) / VKH_BIN_TILES;
    ti.binsH = (s->tilesH + VKH_BIN_TILES - 1) / VKH_BIN_TILES;
    ti.bin_base = 0;
    ti.draw_begin = 0;
    ti.draw_end = n_paths;
    s->bin_total = ti.binsW * ti.binsH;
    s->chunk = vkh::coarse_chunk(n_paths, s->bin_total);
    s->nch = (n_paths + s->chunk - 1) / s->chunk;
    bool ok = ck(s, hipStreamCreate(&s->stream), "stream");
    for (int k = 0; ok && k < 4; k++)
        ok = ck(s, hipEventCreate(&s->ev[k]), "event");
    ok = ok && upload(s, s->pt_path, pt_path.data(), 4ull * np) && upload(s, s->ctrl_path, ctrl_path.data(), 4ull * nctl) &&
         upload(s, s->first_pt, first_pt.data(), 4ull * (n_paths + 1)) &&
         upload(s, s->first_ctrl, first_ctrl.data(), 4ull * (n_paths + 1)) &&
         upload(s, s->first_edge, first_edge.data(), 4ull * (n_paths + 1)) &&
         upload(s, s->ncubic, nc.data(), 4ull * n_paths) && upload(s, s->stroke, stroke.data(), 4ull * n_paths) &&
         upload(s, s->draws, dr.data(), sizeof(vkh_draw) * n_paths) && upload(s, s->tg, &ti, sizeof(ti)) &&
         ensure(s, s->pts, 8ull * np) && ensure(s, s->nrm, 8ull * np) && ensure(s, s->edges, 16ull * ne) &&
         ensure(s, s->src, sizeof(DSrc) * ne) && ensure(s, s->bbox, 16ull * n_paths) &&
         ensure(s, s->dtiles, sizeof(DrawTiles) * n_paths) && ensure(s, s->pt_count, 4ull * (n_paths + 1)) &&
         ensure(s, s->pt_off, 4ull * (n_paths + 1)) && ensure(s, s->seg_alloc, 256) && ensure(s, s->segs, 4096) &&
         ensure(s, s->tile_cnt, 4ull * ((size_t)s->n_tiles * s->nch + 1)) &&
         ensure(s, s->tile_off, 4ull * ((size_t)s->n_tiles * s->nch + 1)) &&
         ensure(s, s->nbat, 4ull * (s->n_tiles + 1)) && ensure(s, s->batch_off, 4ull * (s->n_tiles + 1)) &&
         ensure(s, s->g_edge, 16ull * ne) && ensure(s, s->g_pt, 8ull * np) && ensure(s, s->g_n, 8ull * np) &&
         ensure(s, s->g_h, 4ull * np) && ensure(s, s->g_col, 16ull * n_paths);
    if (!ok) {
        fprintf(stderr, "vkvg_diff_create: %s\n", s->err.c_str());
        vkvg_diff_destroy(s);
        return nullptr;
    }
    return s;
}

void vkvg_diff_destroy(VkvgDiffScene s) {
    if (!s)
        return;
    if (s->stream)
        (void)hipStreamSynchronize(s->stream);
    Buf *bufs[] = {&s->pt_path, &s->ctrl_path, &s->first_pt, &s->first_ctrl, &s->first_edge, &s->ncubic, &s->stroke,
                   &s->draws, &s->tg, &s->pts, &s->nrm, &s->edges, &s->pts64, &s->nrm64, &s->edges64, &s->src, &s->bbox, &s->dtiles, &s->pt_count,
                   &s->pt_off, &s->ptiles, &s->segs, &s->seg_alloc, &s->tile_cnt, &s->tile_off, &s->cmds, &s->nbat,
                   &s->batch_off, &s->ckpt, &s->cmd_tile, &s->unit_cnt, &s->unit_off, &s->unit_cmd, &s->cov, &s->part, &s->tile_order, &s->scan_tmp, &s->host_tot, &s->g_edge, &s->g_pt, &s->g_n, &s->g_h,
                   &s->g_col};
    for (Buf *b : bufs)
        if (b->p)
            (void)hipFree(b->p);
    for (int k = 0; k < 4; k++)
        if (s->ev[k])
            (void)hipEventDestroy(s->ev[k]);
    if (s->stream)
        (void)hipStreamDestroy(s->stream);
    delete s;
}

} // extern "C"

template <typename R>
static int forward_impl(VkvgDiffScene s, const float *ctrl, const float *rgba, const float *width, RV4<R> *image,
                        int keep, Buf &pts, Buf &nrm, Buf &edges) {
    if (!ctrl || !rgba || !width || !image) {
        s->err = "null pointer";
        return -1;
    }
    hipStream_t st = s->stream;
    s->have_ckpt = false;
    if (!ensure(s, pts, 2 * sizeof(R) * (size_t)s->n_pts) || !ensure(s, nrm, 2 * sizeof(R) * (size_t)s->n_pts) ||
        !ensure(s, edges, 4 * sizeof(R) * (size_t)s->n_edges))
        return -1;
    if (!ck(s, hipEventRecord(s->ev[0], st), "event"))
        return -1;
    const uint32_t np = s->n_pts, nd = s->n_paths;
    hipLaunchKernelGGL(k_dflatten<R>, dim3(blocks(np)), dim3(256), 0, st, np, P<uint32_t>(s->pt_path),
                       P<uint32_t>(s->first_pt), P<uint32_t>(s->first_ctrl), P<uint32_t>(s->ncubic),
                       (const float2 *)ctrl, P<RV2<R>>(pts));
    hipLaunchKernelGGL(k_dnormals<R>, dim3(blocks(np)), dim3(256), 0, st, np, P<uint32_t>(s->pt_path),
                       P<uint32_t>(s->first_pt), P<RV2<R>>(pts), P<RV2<R>>(nrm));
    hipLaunchKernelGGL(k_dedges<R>, dim3(blocks(np)), dim3(256), 0, st, np, P<uint32_t>(s->pt_path),
                       P<uint32_t>(s->first_pt), P<uint32_t>(s->first_edge), P<uint32_t>(s->stroke), width,
                       P<RV2<R>>(pts), P<RV2<R>>(nrm), P<RV4<R>>(edges),
                       sizeof(R) == sizeof(float) ? P<DSrc>(s->src) : (DSrc *)nullptr);
    hipLaunchKernelGGL(k_dbbox<R>, dim3(nd), dim3(256), 0, st, P<uint32_t>(s->first_edge), P<RV4<R>>(edges),
                       P<int4>(s->bbox));
    vkh::launch_draw_tiles(st, P<vkh_draw>(s->draws), nd, P<TargetInfo>(s->tg), P<int4>(s->bbox), VKD_T,
                           P<DrawTiles>(s->dtiles), P<uint32_t>(s->pt_count));
    if (!scan(s, P<uint32_t>(s->pt_count), nd, P<uint32_t>(s->pt_off)))
        return -1;
    uint32_t n_ptiles = 0;
    if (!read_u32(s, P<uint32_t>(s->pt_off) + nd, &n_ptiles))
        return -1;
    vkh::launch_set_pt_base(st, P<DrawTiles>(s->dtiles), P<uint32_t>(s->pt_off), nd);
    if (!ensure(s, s->ptiles, sizeof(PathTile) * ((size_t)n_ptiles + 1)))
        return -1;
    for (int pass = 0; pass < 2; pass++) {
        const uint32_t cap = (uint32_t)std::min<size_t>(s->segs.cap / 4, 0xFFFFFFFFull);
        zero(s, s->seg_alloc, 1);
        hipLaunchKernelGGL(k_dbin<R>, dim3(nd), dim3(256), 0, st, P<uint32_t>(s->first_edge), P<RV4<R>>(edges),
                           P<DrawTiles>(s->dtiles), P<PathTile>(s->ptiles), P<uint32_t>(s->segs), cap,
                           P<uint32_t>(s->seg_alloc));
        if (!read_u32(s, P<uint32_t>(s->seg_alloc), &s->n_segs))
            return -1;
        if (s->n_segs <= cap)
            break;
        if (pass == 1 || !ensure(s, s->segs, 4ull * s->n_segs + 4)) {
            if (s->err.empty())
                s->err = "segment buffer overflow";
            return -1;
        }
    }
    const uint32_t n_tc = s->n_tiles * s->nch;
    zero(s, s->tile_cnt, n_tc + 1);
    vkh::launch_coarse(st, false, s->bin_total, s->nch, s->chunk, P<TargetInfo>(s->tg), 1, P<vkh_draw>(s->draws),
                       P<DrawTiles>(s->dtiles), P<PathTile>(s->ptiles), P<uint32_t>(s->tile_cnt), nullptr, nullptr);
    if (!scan(s, P<uint32_t>(s->tile_cnt), n_tc, P<uint32_t>(s->tile_off)))
        return -1;
    if (!read_u32(s, P<uint32_t>(s->tile_off) + n_tc, &s->n_cmds))
        return -1;
    if (!ensure(s, s->cmds, sizeof(Cmd) * ((size_t)s->n_cmds + 1)))
        return -1;
    vkh::launch_coarse(st, true, s->bin_total, s->nch, s->chunk, P<TargetInfo>(s->tg), 1, P<vkh_draw>(s->draws),
                       P<DrawTiles>(s->dtiles), P<PathTile>(s->ptiles), nullptr, P<uint32_t>(s->tile_off),
                       P<Cmd>(s->cmds));
    float4 *ck_ptr = nullptr;
    if (keep && sizeof(R) == sizeof(float)) {
        hipLaunchKernelGGL(k_dbatches, dim3(blocks(s->n_tiles)), dim3(256), 0, st, s->n_tiles, s->nch,
                           P<uint32_t>(s->tile_off), P<uint32_t>(s->nbat));
        if (!scan(s, P<uint32_t>(s->nbat), s->n_tiles, P<uint32_t>(s->batch_off)))
            return -1;
        if (!read_u32(s, P<uint32_t>(s->batch_off) + s->n_tiles, &s->n_batches))
            return -1;
        if (!ensure(s, s->ckpt, 16ull * 256 * ((size_t)s->n_batches + 1)))
            return -1;
        ck_ptr = P<float4>(s->ckpt);
    }
    // work units: at most n_cmds + n_segs / VKD_U of them (one per command plus one per full VKD_U)
    const size_t max_units = (size_t)s->n_cmds + s->n_segs / VKD_U + 1;
    if (!ensure(s, s->cmd_tile, 4ull * ((size_t)s->n_cmds + 1)) || !ensure(s, s->unit_cnt, 4ull * ((size_t)s->n_cmds + 1)) ||
        !ensure(s, s->unit_off, 4ull * ((size_t)s->n_cmds + 1)) || !ensure(s, s->unit_cmd, 4ull * max_units) ||
        !ensure(s, s->cov, sizeof(R) * 256 * ((size_t)s->n_cmds + 1)) || !ensure(s, s->part, sizeof(R) * 256 * max_units))
        return -1;
    hipLaunchKernelGGL(k_dcmd_tile, dim3(blocks(s->n_tiles)), dim3(256), 0, st, s->n_tiles, s->nch,
                       P<uint32_t>(s->tile_off), P<uint32_t>(s->cmd_tile));
    if (s->n_cmds) {
        hipLaunchKernelGGL(k_dunit_count, dim3(blocks(s->n_cmds)), dim3(256), 0, st, s->n_cmds, P<Cmd>(s->cmds),
                           P<PathTile>(s->ptiles), P<uint32_t>(s->unit_cnt));
        if (!scan(s, P<uint32_t>(s->unit_cnt), s->n_cmds, P<uint32_t>(s->unit_off)))
            return -1;
        hipLaunchKernelGGL(k_dunit_cmd, dim3(blocks(s->n_cmds)), dim3(256), 0, st, s->n_cmds, (uint32_t)max_units,
                           P<uint32_t>(s->unit_off), P<uint32_t>(s->unit_cmd));
        const unsigned grid = (unsigned)std::min<size_t>((max_units + VKD_FB - 1) / VKD_FB, 8192);
        hipLaunchKernelGGL(k_dcov<R>, dim3(grid), dim3(256), 0, st, s->n_cmds, (uint32_t)max_units, s->tilesW, P<uint32_t>(s->unit_off),
                           P<uint32_t>(s->unit_cmd), P<uint32_t>(s->cmd_tile), P<Cmd>(s->cmds), P<PathTile>(s->ptiles),
                           P<uint32_t>(s->segs), P<RV4<R>>(edges), P<R>(s->cov), P<R>(s->part));
        hipLaunchKernelGGL(k_dcov_reduce<R>, dim3(blocks((size_t)s->n_cmds * VKD_T)), dim3(256), 0, st, s->n_cmds,
                           P<uint32_t>(s->unit_off), P<R>(s->part), P<R>(s->cov));
    }
    // tiles heaviest first (commands per tile vary several-fold), for compositing and the backward
    if (!ensure(s, s->tile_order, 4ull * vkh::order_words(s->n_tiles)))
        return -1;
    vkh::launch_tile_order(st, s->n_tiles, s->nch, P<uint32_t>(s->tile_off), P<uint32_t>(s->tile_order));
    hipLaunchKernelGGL(k_dcomp<R>, dim3(s->n_tiles), dim3(256), 0, st, s->W, s->H, s->tilesW, s->nch,
                       P<uint32_t>(s->tile_off), P<Cmd>(s->cmds), P<R>(s->cov), (const float4 *)rgba, image, ck_ptr,
                       P<uint32_t>(s->batch_off), P<uint32_t>(s->tile_order));
    if (!ck(s, hipGetLastError(), "forward launch") || !ck(s, hipEventRecord(s->ev[1], st), "event") ||
        !ck(s, hipStreamSynchronize(st), "forward"))
        return -1;
    (void)hipEventElapsedTime(&s->ms_fwd, s->ev[0], s->ev[1]);
    s->have_ckpt = ck_ptr != nullptr;
    return 0;
}

extern "C" {

int vkvg_diff_forward(VkvgDiffScene s, const float *ctrl, const float *rgba, const float *width, float *image,
                      int keep) {
    if (!s)
        return -1;
    return forward_impl<float>(s, ctrl, rgba, width, (RV4<float> *)image, keep, s->pts, s->nrm, s->edges);
}

int vkvg_diff_forward_f64(VkvgDiffScene s, const float *ctrl, const float *rgba, const float *width,
                          double *image) {
    if (!s)
        return -1;
    return forward_impl<double>(s, ctrl, rgba, width, (RV4<double> *)image, 0, s->pts64, s->nrm64, s->edges64);
}

int vkvg_diff_backward(VkvgDiffScene s, const float *ctrl, const float *rgba, const float *width,
                       const float *grad_image, float *grad_ctrl, float *grad_rgba, float *grad_width) {
    if (!s)
        return -1;
    if (!s->have_ckpt) {
        s->err = "backward needs a preceding forward with keep != 0";
        return -1;
    }
    if (!ctrl || !rgba || !width || !grad_image || !grad_ctrl || !grad_rgba || !grad_width) {
        s->err = "null pointer";
        return -1;
    }
    (void)ctrl;
    hipStream_t st = s->stream;
    if (!ck(s, hipEventRecord(s->ev[2], st), "event"))
        return -1;
    zero(s, s->g_edge, 4ull * s->n_edges);
    zero(s, s->g_col, 4ull * s->n_paths);
    hipLaunchKernelGGL(k_dback, dim3(s->n_tiles), dim3(256), 0, st, s->W, s->H, s->tilesW, s->nch,
                       P<uint32_t>(s->tile_off),
                       P<Cmd>(s->cmds), P<PathTile>(s->ptiles), P<uint32_t>(s->segs), P<RV4<float>>(s->edges),
                       (const float4 *)rgba, (const float4 *)grad_image, P<float4>(s->ckpt),
                       P<uint32_t>(s->batch_off), P<float>(s->cov), P<float4>(s->g_edge), P<float>(s->g_col),
                       P<uint32_t>(s->tile_order));
    hipLaunchKernelGGL(k_dg_normal_sums, dim3(blocks(s->n_pts)), dim3(256), 0, st, s->n_pts,
                       P<uint32_t>(s->pt_path), P<uint32_t>(s->first_pt), P<uint32_t>(s->first_edge),
                       P<uint32_t>(s->stroke), width, P<DSrc>(s->src), P<float4>(s->g_edge), P<float2>(s->nrm),
                       P<float2>(s->g_n), P<float>(s->g_h));
    hipLaunchKernelGGL(k_dg_points, dim3(blocks(s->n_pts)), dim3(256), 0, st, s->n_pts, P<uint32_t>(s->pt_path),
                       P<uint32_t>(s->first_pt), P<uint32_t>(s->first_edge), P<uint32_t>(s->stroke),
                       P<DSrc>(s->src), P<float4>(s->g_edge), P<float2>(s->pts), P<float2>(s->nrm),
                       P<float2>(s->g_n), P<float2>(s->g_pt));
    hipLaunchKernelGGL(k_dg_ctrl, dim3(blocks(s->n_ctrl)), dim3(256), 0, st, s->n_ctrl, P<uint32_t>(s->ctrl_path),
                       P<uint32_t>(s->first_ctrl), P<uint32_t>(s->first_pt), P<uint32_t>(s->ncubic),
                       P<float2>(s->g_pt), (float2 *)grad_ctrl);
    hipLaunchKernelGGL(k_dg_paths, dim3(blocks(s->n_paths)), dim3(256), 0, st, s->n_paths, (const float4 *)rgba,
                       P<float>(s->g_col), P<float>(s->g_h), P<uint32_t>(s->first_pt), P<uint32_t>(s->stroke),
                       (float4 *)grad_rgba, grad_width);
    if (!ck(s, hipGetLastError(), "backward launch") || !ck(s, hipEventRecord(s->ev[3], st), "event") ||
        !ck(s, hipStreamSynchronize(st), "backward"))
        return -1;
    (void)hipEventElapsedTime(&s->ms_bwd, s->ev[2], s->ev[3]);
    return 0;
}

float vkvg_diff_last_ms(VkvgDiffScene s, int backward) { return s ? (backward ? s->ms_bwd : s->ms_fwd) : 0.f; }

void vkvg_diff_stats(VkvgDiffScene s, uint32_t out[4]) {
    if (!s || !out)
        return;
    out[0] = s->n_edges;
    out[1] = s->n_segs;
    out[2] = s->n_cmds;
    out[3] = s->n_batches;
}

const char *vkvg_diff_last_error(VkvgDiffScene s) { return s ? s->err.c_str() : "null scene"; }

} // extern "C"
