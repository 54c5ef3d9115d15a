// Raster stage of the vkvg HIP backend: tile binning, coarse command lists, fine raster.
//
// Coverage semantics (replaces Vulkan raster + stencil + blend of src/vkvg_device_internal.c:182-401):
// for every draw the winding w of its oriented edges is evaluated exactly, in 24.8 fixed
// point, at each pixel centre perturbed by (+eps, +eps^2) (= top-left tie rule), and the
// draw's source is composited k times with k = (w != 0) for non-zero fills, (w & 1) for
// even-odd fills, w for strokes (one blend per covering triangle) and 1 for paint.
//
// Per tile (T x T pixels) and draw, with (X0, Y0) the tile's first pixel:
//   w(c, r) = W(tile) + V(r) + H(c, r)
//   W  winding at (X0, Y0): row-wise prefix over tiles of the crossings of sample row Y0
//   V  signed crossings of the column X0 between rows Y0 and r (edges inside the tile)
//   H  signed crossings of row r between columns X0 and c (edges inside the tile)
// so only the edges touching a tile are read when rasterizing it.
#include "vkh_internal.h"
#include "vkh_blend.h"
#include "vkh_fdiv.h"
#include "vkh_cross.h"

#include <hip/hip_runtime.h>

#include <algorithm>

// Tiles are T = 1 << lt pixels, so a tile spans 1 << (lt + 8) in 24.8 fixed point and floor
// division by the span is an arithmetic shift.

// ----------------------------------------------------------------------------------- draw bbox
__global__ void k_bbox_init(int4 *bbox, uint2 *erange, uint32_t n) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        bbox[i] = make_int4(INT32_MAX, INT32_MAX, INT32_MIN, INT32_MIN);
        erange[i] = make_uint2(0xFFFFFFFFu, 0u);
    }
}

__global__ void k_draw_tiles(const vkh_draw *__restrict__ draws, uint32_t n, const TargetInfo *__restrict__ tg,
                             const int4 *__restrict__ bbox, uint32_t T, DrawTiles *__restrict__ dt,
                             uint32_t *__restrict__ pt_count) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    const vkh_draw &d = draws[i];
    const TargetInfo &t = tg[d.target];
    DrawTiles o;
    o.pt_base = 0;
    if (!kind_has_tiles(d.kind) || kind_is_clip(d.kind)) {
        // paint, clear and the stencil draws cover the target without path tiles; a clip changes
        // every pixel (outside its path CLIP is set), so it takes a path tile everywhere
        o.bx0 = 0;
        o.by0 = 0;
        o.bx1 = (int32_t)t.tilesW - 1;
        o.by1 = (int32_t)t.tilesH - 1;
        o.bw = kind_is_clip(d.kind) ? t.tilesW : 0;
        dt[i] = o;
        pt_count[i] = o.bw * t.tilesH;
        return;
    }
    int4 b = bbox[i];
    const int32_t ssh = 8 + (int32_t)__builtin_ctz(T); // log2 of the tile span
    if (b.x > b.z) {
        o.bx0 = o.by0 = 0;
        o.bx1 = o.by1 = -1;
    } else {
        o.bx0 = max(b.x >> ssh, 0);
        o.by0 = max(b.y >> ssh, 0);
        o.bx1 = min(b.z >> ssh, (int32_t)t.tilesW - 1);
        o.by1 = min(b.w >> ssh, (int32_t)t.tilesH - 1);
    }
    if (o.bx0 > o.bx1 || o.by0 > o.by1) {
        o.bx0 = o.by0 = 0;
        o.bx1 = o.by1 = -1;
        o.bw = 0;
        pt_count[i] = 0;
    } else {
        o.bw = (uint32_t)(o.bx1 - o.bx0 + 1);
        pt_count[i] = o.bw * (uint32_t)(o.by1 - o.by0 + 1);
    }
    dt[i] = o;
}

__global__ void k_set_pt_base(DrawTiles *dt, const uint32_t *pt_off, uint32_t n) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n)
        dt[i].pt_base = pt_off[i];
}

// ----------------------------------------------------------------------------------- edge binning
// One workgroup per draw bins the draw's edges into its bbox tile grid, in bands of whole tile
// rows that fit the LDS tables (VKH_BIN_LDS_TILES tiles). An edge is enumerated into the tiles
// it may touch: per tile row, the edge's x-extent inside the row's slab widened by one pixel.
// Count pass (LDS atomics): segments per tile, and the edge's crossing of each tile row's first
// sample row added to the backdrop delta of the first tile whose first column lies right of the
// crossing. The counts are scanned in LDS, the band's segment range is taken from a global
// allocator with one atomic, the backdrop deltas are prefix-summed along each row and the path
// tiles are written; the scatter pass then writes the edges themselves (16 bytes, so the fine
// kernel reads a tile's segments without a dependent load) through LDS cursors. A band
// whose range does not fit the segment buffer writes no segments; the host sees the allocator
// beyond the capacity, grows the buffer and bins again.
#define VKH_BIN_LDS_TILES 4096
#define VKH_BIN_U 4 // edges per thread per load step

// blends a draw of this kind makes at a pixel of winding w (1 for paint and clear)
__device__ __forceinline__ int coverage_count(uint32_t kind, int32_t w) {
    switch (kind) {
    case VKH_DRAW_FILL_NZ:
        return w != 0 ? 1 : 0;
    case VKH_DRAW_FILL_EO:
        return w & 1;
    case VKH_DRAW_STROKE:
        return w > 0 ? w : 0;
    default:
        return 1;
    }
}

template <bool SCATTER>
__device__ __forceinline__ void bin_edge(const Edge &e, const DrawTiles &dt, int32_t lt, int32_t ty_lo,
                                         int32_t ty_hi, uint32_t *s_cnt, int32_t *s_back, Edge *segs,
                                         uint32_t base) {
    const int32_t ssh = lt + 8, span = 1 << ssh;
    const int32_t ymin = min(e.y0, e.y1), ymax = max(e.y0, e.y1);
    const int32_t ty0 = max(ymin >> ssh, ty_lo);
    const int32_t ty1 = min(ymax >> ssh, ty_hi);
    if (ty0 > ty1) // outside the band of tile rows
        return;
    const bool horiz = e.y0 == e.y1;
    const double invdy = horiz ? 0.0 : 1.0 / ((double)e.y1 - (double)e.y0);
    const int dir = e.y0 < e.y1 ? 1 : -1;
    const int32_t xt = e.y0 < e.y1 ? e.x0 : e.x1, yt = ymin;
    const int32_t xb = e.y0 < e.y1 ? e.x1 : e.x0, yb = ymax;
    for (int32_t ty = ty0; ty <= ty1; ty++) {
        int32_t xlo, xhi;
        if (horiz) {
            xlo = min(e.x0, e.x1);
            xhi = max(e.x0, e.x1);
        } else {
            double ya = fmax((double)ty * span, (double)ymin);
            double yb2 = fmin((double)(ty + 1) * span, (double)ymax);
            double xa = (double)e.x0 + ((double)e.x1 - e.x0) * ((ya - e.y0) * invdy);
            double xc = (double)e.x0 + ((double)e.x1 - e.x0) * ((yb2 - e.y0) * invdy);
            xlo = (int32_t)floor(fmin(xa, xc)) - 256;
            xhi = (int32_t)ceil(fmax(xa, xc)) + 256;
        }
        const int32_t tx0 = max(xlo >> ssh, dt.bx0);
        const int32_t tx1 = min(xhi >> ssh, dt.bx1);
        const uint32_t rowbase = (uint32_t)(ty - ty_lo) * dt.bw;
        for (int32_t tx = tx0; tx <= tx1; tx++) {
            const uint32_t k = rowbase + (uint32_t)(tx - dt.bx0);
            if (!SCATTER)
                atomicAdd(&s_cnt[k], 1u);
            else
                segs[base + atomicAdd(&s_cnt[k], 1u)] = e;
        }
        if (!SCATTER && !horiz) {
            const int32_t sy = ty * span + 128; // first sample row of the tile row
            if (yt <= sy && sy < yb) {
                const int32_t c = cross_column(xt, yt, xb, yb, sy);
                int32_t t0 = -((-c) >> lt); // first tile tx with tx*T >= c: ceil(c / T)
                if (t0 < dt.bx0)
                    t0 = dt.bx0;
                if (t0 <= dt.bx1)
                    atomicAdd(&s_back[rowbase + (uint32_t)(t0 - dt.bx0)], dir);
            }
        }
    }
}

// exclusive scan in place of a[0 .. n), n <= VKH_BIN_LDS_TILES, by a 256-thread workgroup (each
// thread sums a run of VKH_BIN_LDS_TILES / 256 entries); returns the total
static_assert(VKH_BIN_LDS_TILES % 256 == 0, "binning table: whole runs per thread");
__device__ __forceinline__ uint32_t block_exclusive_scan(uint32_t *a, uint32_t n, uint32_t *s_wave) {
    constexpr uint32_t RUN = VKH_BIN_LDS_TILES / 256;
    const uint32_t tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const uint32_t beg = min(tid * RUN, n), end = min(beg + RUN, n);
    uint32_t sum = 0;
    for (uint32_t k = beg; k < end; k++)
        sum += a[k];
    uint32_t x = sum;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t y = __shfl_up(x, o);
        if ((int)lane >= o)
            x += y;
    }
    if (lane == 63)
        s_wave[w] = x;
    __syncthreads();
    uint32_t run = x - sum, total = 0;
#pragma unroll
    for (uint32_t k = 0; k < 4; k++) {
        run += k < w ? s_wave[k] : 0u;
        total += s_wave[k];
    }
    for (uint32_t k = beg; k < end; k++) {
        const uint32_t v = a[k];
        a[k] = run;
        run += v;
    }
    __syncthreads();
    return total;
}

// Besides the path tiles and segments, every path tile that will hold a command (segments, a
// non-zero blend count from its backdrop, or a clip) adds one to its tile's command count in the
// draw's chunk (tile_cnt[tile * nch + chunk], the coarse pass's layout), so the coarse pass needs
// no counting pass of its own.
__global__ __launch_bounds__(256) void k_bin_draw(const Edge *__restrict__ edges, const uint2 *__restrict__ erange,
                                                  const DrawTiles *__restrict__ dts, uint32_t T,
                                                  PathTile *__restrict__ pts, Edge *__restrict__ segs,
                                                  uint32_t seg_cap, uint32_t *__restrict__ seg_alloc,
                                                  const uint32_t *__restrict__ draw_order,
                                                  const vkh_draw *__restrict__ draws, const TargetInfo *__restrict__ tg,
                                                  uint32_t nch, uint32_t chunk_draws, uint32_t *__restrict__ tile_cnt) {
    __shared__ uint32_t s_cnt[VKH_BIN_LDS_TILES];
    __shared__ int32_t s_back[VKH_BIN_LDS_TILES];
    __shared__ uint32_t s_wave[4];
    __shared__ uint32_t s_base;
    const uint32_t d = draw_order ? draw_order[blockIdx.x] : blockIdx.x, tid = threadIdx.x;
    const DrawTiles dt = dts[d];
    if (dt.bw == 0 || dt.bx0 > dt.bx1 || dt.by0 > dt.by1)
        return; // paint / clear, or no edges
    const uint2 er = erange[d];
    const int32_t lt = (int32_t)__builtin_ctz(T);
    const vkh_draw &dr = draws[d];
    const uint32_t kind = dr.kind;
    const bool clip = kind_is_clip(kind);
    const TargetInfo &TI = tg[dr.target];
    const uint32_t chunk = (d - TI.draw_begin) / chunk_draws;
    const uint32_t bw = dt.bw, bh = (uint32_t)(dt.by1 - dt.by0 + 1);
    // The draw's edges (its range holds only its own, kept edges), VKH_BIN_U per thread per step
    // with every load issued before any is used. A draw of at most 256 * VKH_BIN_U edges (most of
    // them) keeps its edges in registers for both passes of every band; a larger one streams them
    // each pass.
    const bool cached = er.y - er.x <= 256u * VKH_BIN_U;
    Edge ce[VKH_BIN_U];
    auto load = [&](uint32_t b0) {
#pragma unroll
        for (int u = 0; u < VKH_BIN_U; u++) {
            const uint32_t i = b0 + (uint32_t)u * 256u + tid;
            ce[u] = i < er.y ? edges[i] : Edge{0, 0, 0, 0};
        }
    };
    auto for_edges = [&](auto &&f) {
        if (cached) {
#pragma unroll
            for (int u = 0; u < VKH_BIN_U; u++)
                if (er.x + (uint32_t)u * 256u + tid < er.y)
                    f(ce[u]);
            return;
        }
        for (uint32_t b0 = er.x; b0 < er.y; b0 += 256u * VKH_BIN_U) {
            load(b0);
#pragma unroll
            for (int u = 0; u < VKH_BIN_U; u++)
                if (b0 + (uint32_t)u * 256u + tid < er.y)
                    f(ce[u]);
        }
    };
    if (cached)
        load(er.x);
    const uint32_t band_rows = max(1u, VKH_BIN_LDS_TILES / bw); // bw <= 4096 (width <= 32768, T >= 8)
    for (uint32_t r0 = 0; r0 < bh; r0 += band_rows) {
        const uint32_t nr = min(band_rows, bh - r0);
        const uint32_t nt = nr * bw;
        const int32_t ty_lo = dt.by0 + (int32_t)r0, ty_hi = ty_lo + (int32_t)nr - 1;
        for (uint32_t k = tid; k < nt; k += 256) {
            s_cnt[k] = 0;
            s_back[k] = 0;
        }
        __syncthreads();
        for_edges([&](const Edge &e) { bin_edge<false>(e, dt, lt, ty_lo, ty_hi, s_cnt, s_back, nullptr, 0); });
        __syncthreads();
        const uint32_t total = block_exclusive_scan(s_cnt, nt, s_wave);
        if (tid == 0)
            s_base = atomicAdd(seg_alloc, total);
        for (uint32_t r = tid; r < nr; r += 256) {
            int32_t acc = 0;
            for (uint32_t k = r * bw; k < (r + 1) * bw; k++) {
                acc += s_back[k];
                s_back[k] = acc;
            }
        }
        __syncthreads();
        const uint32_t base = s_base;
        const bool fits = (uint64_t)base + total <= (uint64_t)seg_cap;
        PathTile *out = pts + dt.pt_base + r0 * bw;
        for (uint32_t k = tid; k < nt; k += 256) {
            const uint32_t off = s_cnt[k];
            const uint32_t next = k + 1 < nt ? s_cnt[k + 1] : total;
            PathTile p;
            p.backdrop = s_back[k];
            p.nseg = next - off;
            p.seg_off = base + off;
            p.reserved = 0;
            out[k] = p;
            if (p.nseg || clip || coverage_count(kind, p.backdrop) != 0) {
                const uint32_t r = k / bw, c = k - r * bw;
                const uint32_t gtile = TI.tile_base + (uint32_t)(ty_lo + (int32_t)r) * TI.tilesW + (uint32_t)dt.bx0 + c;
                atomicAdd(&tile_cnt[(size_t)gtile * nch + chunk], 1u);
            }
        }
        __syncthreads();
        if (fits)
            for_edges([&](const Edge &e) { bin_edge<true>(e, dt, lt, ty_lo, ty_hi, s_cnt, s_back, segs, base); });
        __syncthreads();
    }
}

__device__ __forceinline__ uint32_t find_target_by(const TargetInfo *tg, uint32_t n, uint32_t v, bool bins) {
    uint32_t lo = 0, hi = n;
    while (lo < hi) {
        uint32_t mid = (lo + hi) >> 1;
        uint32_t base = bins ? tg[mid].bin_base : tg[mid].tile_base;
        if (base <= v)
            lo = mid + 1;
        else
            hi = mid;
    }
    return lo - 1;
}

// The coarse pass's per (tile, chunk) command counts before binning adds the path tiles': the draws
// without path tiles (paint, clear, stencil draws) hold a command in every tile of their target,
// table[target * nch + chunk] of them in the chunk
__global__ void k_tile_cnt_init(const TargetInfo *__restrict__ tg, uint32_t n_targets, uint32_t nch,
                                const uint32_t *__restrict__ table, uint32_t *__restrict__ tile_cnt, uint32_t n_tc) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i > n_tc)
        return;
    if (i == n_tc) {
        tile_cnt[i] = 0;
        return;
    }
    const uint32_t gtile = i / nch, chunk = i - gtile * nch;
    const uint32_t t = find_target_by(tg, n_targets, gtile, false);
    tile_cnt[i] = table[(size_t)t * nch + chunk];
}

__global__ void k_zero_u32(uint32_t *p, uint32_t n) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n)
        p[i] = 0;
}

// ----------------------------------------------------------------------------------- coarse
// One workgroup per (bin of 16 x 16 tiles, chunk of `chunk` draws), one thread per
// tile; walks the chunk's draws in order. Counts are laid out tile-major (tile * nch + chunk), so
// one scan gives every tile a contiguous command list ordered by draw, without any sort. The
// rendering pipeline takes the counts from binning (k_tile_cnt_init, k_bin_draw) and runs only the
// writing pass; the differentiable renderer runs both.
// Per 256 draws: the draws whose tile bbox meets the bin are compacted into LDS (in draw order,
// with their bbox and kind), then every tile walks that list, loading the path tiles of four
// draws at a time so their latencies overlap.
__global__ __launch_bounds__(256) __attribute__((amdgpu_num_sgpr(80))) void k_coarse(bool write, const TargetInfo *__restrict__ tg, uint32_t n_targets,
                                                const vkh_draw *__restrict__ draws, const DrawTiles *__restrict__ dts,
                                                const PathTile *__restrict__ pts, uint32_t nch, uint32_t chunk_draws,
                                                uint32_t *__restrict__ tile_count, const uint32_t *__restrict__ tile_off,
                                                Cmd *__restrict__ cmds) {
    __shared__ uint64_t masks[4];
    __shared__ DrawTiles s_dt[256];
    __shared__ uint32_t s_draw[256], s_kind[256];
    uint32_t bin = blockIdx.x, chunk = blockIdx.y;
    uint32_t t = find_target_by(tg, n_targets, bin, true);
    const TargetInfo &T = tg[t];
    uint32_t lb = bin - T.bin_base;
    int32_t bxi = (int32_t)(lb % T.binsW), byi = (int32_t)(lb / T.binsW);
    uint32_t tid = threadIdx.x;
    const uint32_t lane = tid & 63, wv = tid >> 6;
    int32_t tx = bxi * VKH_BIN_TILES + (int32_t)(tid & 15), ty = byi * VKH_BIN_TILES + (int32_t)(tid >> 4);
    bool valid = tx < (int32_t)T.tilesW && ty < (int32_t)T.tilesH;
    uint32_t gtile = T.tile_base + (uint32_t)ty * T.tilesW + (uint32_t)tx;
    int32_t bx0 = bxi * VKH_BIN_TILES, by0 = byi * VKH_BIN_TILES;
    int32_t bx1 = bx0 + VKH_BIN_TILES - 1, by1 = by0 + VKH_BIN_TILES - 1;
    const uint32_t d_begin = T.draw_begin + chunk * chunk_draws;
    if (d_begin >= T.draw_end)
        return; // uniform over the workgroup
    const uint32_t d_end = min(T.draw_end, d_begin + chunk_draws);
    uint32_t count = 0;
    uint32_t out = (write && valid) ? tile_off[(size_t)gtile * nch + chunk] : 0;
    // the commands counted for this (tile, chunk): a write never leaves that range
    const uint32_t room = (write && valid) ? tile_off[(size_t)gtile * nch + chunk + 1] - out : 0u;
    for (uint32_t c0 = d_begin; c0 < d_end; c0 += 256) {
        const uint32_t d = c0 + tid;
        bool hit = false;
        DrawTiles b;
        if (d < d_end) {
            b = dts[d];
            hit = b.bx0 <= b.bx1 && b.bx0 <= bx1 && b.bx1 >= bx0 && b.by0 <= by1 && b.by1 >= by0;
        }
        const uint64_t m = __ballot(hit);
        if (lane == 0)
            masks[wv] = m;
        __syncthreads();
        uint32_t before = 0, nh = 0;
        for (uint32_t w = 0; w < 4; w++) {
            const uint32_t c = (uint32_t)__popcll((unsigned long long)masks[w]);
            before += w < wv ? c : 0u;
            nh += c;
        }
        if (hit) {
            const uint32_t r = before + (uint32_t)__popcll((unsigned long long)(m & ((1ull << lane) - 1ull)));
            s_dt[r] = b;
            s_draw[r] = d;
            s_kind[r] = draws[d].kind;
        }
        __syncthreads();
        for (uint32_t h0 = 0; h0 < nh; h0 += 4) {
            uint32_t ptid[4];
            PathTile p[4];
            bool in[4];
#pragma unroll
            for (int u = 0; u < 4; u++) {
                in[u] = false;
                ptid[u] = VKH_NO_DRAW;
                p[u] = PathTile{0, 0, 0, 0};
                const uint32_t h = h0 + (uint32_t)u;
                if (h < nh && valid) {
                    const DrawTiles hb = s_dt[h];
                    in[u] = tx >= hb.bx0 && tx <= hb.bx1 && ty >= hb.by0 && ty <= hb.by1;
                    const uint32_t kind = s_kind[h];
                    if (in[u] && kind_has_tiles(kind)) {
                        ptid[u] = hb.pt_base + (uint32_t)(ty - hb.by0) * hb.bw + (uint32_t)(tx - hb.bx0);
                        p[u] = pts[ptid[u]];
                    }
                }
            }
#pragma unroll
            for (int u = 0; u < 4; u++) {
                if (!in[u])
                    continue;
                const uint32_t h = h0 + (uint32_t)u;
                const uint32_t kind = s_kind[h];
                const bool emit = ptid[u] == VKH_NO_DRAW || p[u].nseg > 0 || kind_is_clip(kind) ||
                                  coverage_count(kind, p[u].backdrop) != 0;
                if (emit) {
                    if (write && count < room)
                        cmds[out + count] = Cmd{s_draw[h], ptid[u]};
                    count++;
                }
            }
        }
        __syncthreads();
    }
    if (!write && valid)
        tile_count[(size_t)gtile * nch + chunk] = count;
}

// ----------------------------------------------------------------------------------- work order
// Kernels that give one workgroup to each item of very uneven cost (the fine kernel's tiles, the
// binning kernel's draws) take the items heaviest first (longest-processing-time order): an item
// with many times the mean cost that started late would hold its CU long after the others
// finished. Weights are bucketed exactly below 128 and in 16 steps per octave above (to 2^15),
// buckets reversed so that bucket 0 is the heaviest; items in one bucket keep no particular order
// (each item writes its own outputs, so no result depends on it).
#define VKH_ORDER_BUCKETS 256
__device__ __forceinline__ uint32_t order_bucket(uint32_t w) {
    uint32_t b = w;
    if (w >= 128u) {
        const uint32_t e = 31u - (uint32_t)__clz((int)w); // >= 7
        b = min(128u + (e - 7u) * 16u + ((w >> (e - 4u)) & 15u), 255u);
    }
    return (VKH_ORDER_BUCKETS - 1) - b;
}
// commands of a tile
struct TileWeight {
    uint32_t nch;
    const uint32_t *tile_off;
    __device__ uint32_t operator()(uint32_t t) const {
        return tile_off[(size_t)(t + 1) * nch] - tile_off[(size_t)t * nch];
    }
};
// edges of a draw
struct DrawWeight {
    const uint2 *erange;
    __device__ uint32_t operator()(uint32_t d) const {
        const uint2 e = erange[d];
        return e.y > e.x ? e.y - e.x : 0u;
    }
};

// pass 0: bucket sizes (LDS histogram per workgroup, one global add per non-empty bucket);
// pass 1: each workgroup reserves its range in every bucket and writes its items there
template <bool SCATTER, class W>
__global__ __launch_bounds__(256) void k_work_order(uint32_t n, W weight, uint32_t *__restrict__ bucket,
                                                    uint32_t *__restrict__ order) {
    __shared__ uint32_t h[VKH_ORDER_BUCKETS];
    const uint32_t tid = threadIdx.x, i = blockIdx.x * 256u + tid;
    h[tid] = 0u;
    __syncthreads();
    const uint32_t b = i < n ? order_bucket(weight(i)) : 0u;
    const uint32_t rank = i < n ? atomicAdd(&h[b], 1u) : 0u;
    __syncthreads();
    if (!SCATTER) {
        if (h[tid])
            atomicAdd(&bucket[tid], h[tid]);
        return;
    }
    if (h[tid])
        h[tid] = atomicAdd(&bucket[tid], h[tid]); // the workgroup's first slot in bucket tid
    __syncthreads();
    if (i < n)
        order[h[b] + rank] = i;
}

// exclusive scan of the bucket sizes into bucket starts (one workgroup)
__global__ __launch_bounds__(256) void k_work_order_scan(uint32_t *__restrict__ bucket) {
    __shared__ uint32_t s_wave[4];
    const uint32_t tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const uint32_t v = bucket[tid];
    uint32_t x = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t y = __shfl_up(x, o);
        if ((int)lane >= o)
            x += y;
    }
    if (lane == 63)
        s_wave[w] = x;
    __syncthreads();
    uint32_t before = 0;
    for (uint32_t k = 0; k < w; k++)
        before += s_wave[k];
    bucket[tid] = before + x - v;
}

// ----------------------------------------------------------------------------------- fine
__device__ __forceinline__ float clamp01(float x) { return fminf(fmaxf(x, 0.f), 1.f); }
// smoothstep(e0, e1, x) with the division (x - e0) / (e1 - e0) taken through r, the prepared
// reciprocal of e1 - e0 (same rounding as the IEEE division)
__device__ __forceinline__ float smoothstep_r(float e0, float e1, float r, float x) {
    float t = clamp01(vkh_div_rcp(x - e0, e1 - e0, r));
    return t * t * (3.f - 2.f * t);
}
__device__ __forceinline__ void mix4(float *c, const float *d, float a) {
#pragma unroll
    for (int i = 0; i < 4; i++)
        c[i] = c[i] * (1.f - a) + d[i] * a;
}

// Per draw: premultiplied solid colour, signature key, normalised kind/op, and the gradient's
// pixel-independent terms of shaders/vkvg_main.frag:83-144 (p = gl_FragCoord.xy / size), each
// computed with the same float operations the per-pixel evaluation used to repeat.
__global__ void k_draw_prep(const vkh_draw *__restrict__ draws, uint32_t n, const TargetInfo *__restrict__ tg,
                            const vkh_gradient *__restrict__ grads, const vkh_surface_source *__restrict__ srcs,
                            DrawPrep *__restrict__ out, GradRec *__restrict__ rec) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    const vkh_draw &d = draws[i];
    DrawPrep p;
    const uint32_t kind = d.kind;
    const uint32_t op = kind == VKH_DRAW_CLEAR ? 0u : d.op;
    const bool surf = d.source >= 0 && kind <= VKH_DRAW_PAINT; // fills, strokes, paint
    p.meta = kind | (op << 4) | (surf ? VKH_DP_SURF : (d.grad >= 0 ? VKH_DP_GRAD : 0u));
    p.key = draw_key(d.key_index);
    p.grad = d.grad;
    p.opacity = d.opacity;
    const uint32_t c = d.color;
    const float o = d.opacity;
    p.solid = (uint32_t)f2u8(((float)(c & 0xFF) / 255.0f) * o) |
              ((uint32_t)f2u8(((float)((c >> 8) & 0xFF) / 255.0f) * o) << 8) |
              ((uint32_t)f2u8(((float)((c >> 16) & 0xFF) / 255.0f) * o) << 16) |
              ((uint32_t)f2u8(((float)((c >> 24) & 0xFF) / 255.0f) * o) << 24);
    if (kind >= VKH_DRAW_CLIP_SAVE)
        p.solid = (uint32_t)d.style; // the saved bit or copy index of a stencil draw
    p.gflags = 0;
    for (int k = 0; k < 6; k++)
        p.g[k] = 0.f;
    p.rg[0] = p.rg[1] = 0.f;
    p.pad[0] = p.pad[1] = 0;
    if (d.grad >= 0) {
        const vkh_gradient &g = grads[d.grad];
        const float W = (float)tg[d.target].width, H = (float)tg[d.target].height;
        if (g.type == 2) {
            float p0x = g.cp[0][0] / W, p0y = g.cp[0][1] / H;
            float p1x = g.cp[0][2] / W, p1y = g.cp[0][3] / H;
            float dx = p1x - p0x, dy = p1y - p0y;
            float l = sqrtf(dx * dx + dy * dy);
            float ux = dx / l, uy = dy / l;
            p.g[0] = p0x;
            p.g[1] = l;
            if (uy == 0) {
                p.gflags = VKH_GF_UY0 | (ux < 0 ? VKH_GF_UXNEG : 0u);
            } else {
                float m = -ux / uy;
                p.g[2] = m;
                p.g[3] = p0y - m * p0x;
                p.g[4] = sqrtf(1 + m * m);
                p.gflags = uy < 0 ? VKH_GF_UYNEG : 0u;
                p.rg[1] = vkh_rcp_for_div(p.g[4]);
            }
            p.rg[0] = vkh_rcp_for_div(l);
        } else {
            float c0x = g.cp[0][0] / W, c0y = g.cp[0][1] / H;
            float c1x = g.cp[1][0] / W, c1y = g.cp[1][1] / H;
            float r0 = g.cp[0][2] / W, r1 = g.cp[1][2] / W;
            float diffx = c0x - c1x, diffy = c0y - c1y;
            p.g[0] = c0x;
            p.g[1] = c0y;
            p.g[2] = r0;
            p.g[3] = diffx;
            p.g[4] = diffy;
            p.g[5] = (diffx * diffx + diffy * diffy) - r1 * r1;
        }
    }
    out[i] = p;
    if (d.grad >= 0) {
        const vkh_gradient &g = grads[d.grad];
        GradRec r;
        r.type = g.type;
        r.count = g.count;
        r.gflags = p.gflags;
        r.opacity = p.opacity;
        for (int k = 0; k < 6; k++)
            r.g[k] = p.g[k];
        r.rg[0] = p.rg[0];
        r.rg[1] = p.rg[1];
        for (int k = 0; k < VKH_GR_INLINE; k++) {
            r.stops[k] = g.stops[k];
            r.rspan[k] = k ? vkh_rcp_for_div(g.stops[k] - g.stops[k - 1]) : 0.f;
            for (int c = 0; c < 4; c++)
                r.colors[k][c] = g.colors[k][c];
        }
        r.grad = d.grad;
        r.pad[0] = r.pad[1] = r.pad[2] = 0;
        rec[i] = r;
    }
    if (surf) {
        const vkh_surface_source &src = srcs[d.source];
        SurfRec r;
        r.img = (const uint32_t *)src.rgba;
        r.width = src.width;
        r.height = src.height;
        r.extend = src.extend;
        r.linear = (src.filter == 2u || src.filter == 4u) ? 1u : 0u; // VKVG_FILTER_BEST, _BILINEAR
        for (int k = 0; k < 4; k++)
            r.rect[k] = src.rect[k];
        for (int k = 0; k < 6; k++)
            r.m[k] = src.mat_inv[k];
        r.opacity = d.opacity;
        for (int k = 0; k < 23; k++)
            r.pad[k] = 0;
        *reinterpret_cast<SurfRec *>(rec + i) = r;
    }
}

__global__ void k_grad_prep(const vkh_gradient *__restrict__ grads, uint32_t n, GradPrep *__restrict__ out) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    GradPrep p;
    p.rspan[0] = 0.f;
    for (int k = 1; k < 16; k++)
        p.rspan[k] = vkh_rcp_for_div(grads[i].stops[k] - grads[i].stops[k - 1]);
    out[i] = p;
}

// shaders/vkvg_main.frag:83-144 at p = (px, py), from the draw's staged record R (invariant terms,
// first stops with their spans' reciprocals and colours); later stops come from the gradient table
// g and its reciprocals gp. Both gradient types apply colour step 1 and then steps 2 .. count-1
// (the linear loop of the shader indexes the same steps from 0).
__device__ void gradient_color(const GradRec &P, const vkh_gradient *__restrict__ grads,
                               const GradPrep *__restrict__ gprep, float px, float py, float *col) {
    float x;
    if (P.type == 2) {
        if (P.gflags & VKH_GF_UY0) {
            x = vkh_div_rcp(px - P.g[0], P.g[1], P.rg[0]);
            if (P.gflags & VKH_GF_UXNEG)
                x = -x; // -(a) / b and -(a / b) round alike
        } else {
            x = vkh_div_rcp(vkh_div_rcp(py - P.g[2] * px - P.g[3], P.g[4], P.rg[1]), P.g[1], P.rg[0]);
            if (P.gflags & VKH_GF_UYNEG)
                x = -x;
        }
    } else {
        const float c0x = P.g[0], c0y = P.g[1], r0 = P.g[2], diffx = P.g[3], diffy = P.g[4], cc = P.g[5];
        float gradLength = 1.0f;
        float rx = px - c0x, ry = py - c0y;
        float rl = sqrtf(rx * rx + ry * ry);
        float rdx = rx / rl, rdy = ry / rl;
        float a = rdx * rdx + rdy * rdy;
        float b = 2.0f * (rdx * diffx + rdy * diffy);
        float disc = b * b - 4.0f * a * cc;
        if (disc >= 0.0f) {
            float t = (-b + sqrtf(fabsf(disc))) / (2.0f * a);
            float prx = c0x + rdx * t, pry = c0y + rdy * t;
            float ddx = prx - c0x, ddy = pry - c0y;
            gradLength = sqrtf(ddx * ddx + ddy * ddy) - r0;
        }
        float qx = px - c0x, qy = py - c0y;
        x = (sqrtf(qx * qx + qy * qy) - r0) / gradLength;
    }
    for (int i = 0; i < 4; i++)
        col[i] = P.colors[0][i];
    const int steps = max((int)P.count, 2);
    const int inl = min(steps, VKH_GR_INLINE);
#pragma unroll
    for (int i = 1; i < VKH_GR_INLINE; i++) {
        if (i >= inl)
            break;
        const float e0 = P.stops[i - 1], e1 = P.stops[i];
        mix4(col, P.colors[i], smoothstep_r(e0, e1, P.rspan[i], x));
    }
    if (steps > VKH_GR_INLINE) {
        const vkh_gradient &g = grads[P.grad];
        const GradPrep &gp = gprep[P.grad];
        for (int i = VKH_GR_INLINE; i < steps; i++)
            mix4(col, g.colors[i], smoothstep_r(g.stops[i - 1], g.stops[i], gp.rspan[i], x));
    }
}

__device__ __forceinline__ uint32_t uniform_u32(uint32_t v) { return (uint32_t)__builtin_amdgcn_readfirstlane((int)v); }

// Surface source sampling: texture(source, uv) of shaders/vkvg_main.frag:72-82 under the sampler of
// src/vkvg_context_internal.c:730-755, with the Vulkan specification's unnormalized-coordinate
// formulas (u' = u * width; NEAREST: floor(u'); LINEAR: floor(u' - 0.5) and its fraction, texels
// summed ((w00 t00 + w10 t10) + w01 t01) + w11 t11). Address modes: extend NONE = CLAMP_TO_BORDER
// (transparent black), REPEAT, REFLECT = MIRRORED_REPEAT, PAD = CLAMP_TO_EDGE.
__device__ __forceinline__ int32_t surf_floor_index(float x) {
    return (int32_t)fminf(fmaxf(floorf(x), -1073741824.f), 1073741824.f);
}
__device__ __forceinline__ int32_t surf_wrap(int32_t i, int32_t n, uint32_t extend) {
    switch (extend) {
    case 1: { // REPEAT
        const int32_t r = i % n;
        return r < 0 ? r + n : r;
    }
    case 2: { // MIRRORED_REPEAT
        int32_t t = i % (2 * n);
        t = t < 0 ? t + 2 * n : t;
        return t < n ? t : 2 * n - 1 - t;
    }
    case 3: // CLAMP_TO_EDGE
        return min(max(i, 0), n - 1);
    default: // CLAMP_TO_BORDER
        return (i < 0 || i >= n) ? -1 : i;
    }
}
__device__ __forceinline__ void surf_texel(const SurfRec &R, int32_t i, int32_t j, float *t) {
    const int32_t wi = surf_wrap(i, (int32_t)R.width, R.extend), wj = surf_wrap(j, (int32_t)R.height, R.extend);
    uint32_t v = 0u;
    if (wi >= 0 && wj >= 0)
        v = R.img[(size_t)wj * R.width + (size_t)wi];
#pragma unroll
    for (int k = 0; k < 4; k++)
        t[k] = (float)((v >> (8 * k)) & 0xFFu) / 255.0f;
}
__device__ __forceinline__ void surface_color(const SurfRec &R, float fx, float fy, float *col) {
    const float px = fx - R.rect[0], py = fy - R.rect[1];
    float u = R.m[0] * px + R.m[2] * py + R.m[4];
    float v = R.m[1] * px + R.m[3] * py + R.m[5];
    u = u / R.rect[2];
    v = v / R.rect[3];
    const float uu = u * (float)R.width, vv = v * (float)R.height;
    if (R.linear) {
        const float a = uu - 0.5f, b = vv - 0.5f;
        const float al = a - floorf(a), be = b - floorf(b);
        const int32_t i0 = surf_floor_index(a), j0 = surf_floor_index(b);
        float t00[4], t10[4], t01[4], t11[4];
        surf_texel(R, i0, j0, t00);
        surf_texel(R, i0 + 1, j0, t10);
        surf_texel(R, i0, j0 + 1, t01);
        surf_texel(R, i0 + 1, j0 + 1, t11);
        const float w00 = (1.f - al) * (1.f - be), w10 = al * (1.f - be), w01 = (1.f - al) * be, w11 = al * be;
#pragma unroll
        for (int k = 0; k < 4; k++)
            col[k] = ((w00 * t00[k] + w10 * t10[k]) + w01 * t01[k]) + w11 * t11[k];
    } else {
        surf_texel(R, surf_floor_index(uu), surf_floor_index(vv), col);
    }
}

// One pixel's stencil word under a stencil draw without geometry: byte 0 holds FILL (bit 0, always
// 0 between draws), CLIP (bit 1) and the CLIP bits of six saved levels (bits 2..7); bytes 1..3 are
// the copies the reference keeps in extra stencil images beyond six levels.
__device__ __forceinline__ uint32_t stencil_apply(uint32_t kind, uint32_t prm, uint32_t st) {
    switch (kind) {
    case VKH_DRAW_CLIP_SAVE: { // pass REPLACE / fail ZERO on the saved bit, where CLIP is set
        const uint32_t b = 4u << prm;
        return (st & ~b) | ((st & 2u) ? b : 0u);
    }
    case VKH_DRAW_CLIP_RESTORE: { // the same with the roles of the two bits exchanged
        const uint32_t b = 4u << prm;
        return (st & ~2u) | ((st & b) ? 2u : 0u);
    }
    case VKH_DRAW_CLIP_RESET:
        return st & ~0xFFu;
    case VKH_DRAW_STENCIL_SPILL: {
        const uint32_t sh = 8u * (prm + 1u);
        return (st & ~(0xFFu << sh)) | ((st & 0xFFu) << sh);
    }
    case VKH_DRAW_STENCIL_UNSPILL:
        return (st & ~0xFFu) | ((st >> (8u * (prm + 1u))) & 0xFFu);
    default:
        return st;
    }
}


// Blends one source into a covered pixel: once, and for strokes once more per further covering
// triangle (w = the winding count, >= 1 here); the signature adds the draw's key per blend.
__device__ __forceinline__ void blend_count(const vkh_src16 &S, bool stroke, int32_t w, uint32_t key, uint32_t &v,
                                            uint32_t &sg) {
    v = blend_px16(v, S);
    sg += key;
    if (stroke)
        for (int32_t n = 1; n < w; n++) {
            v = blend_px16(v, S);
            sg += key;
        }
}

// One workgroup per T x T tile. Commands are processed in batches of B: the batch's command
// records (draw record + path tile) are staged in LDS with one parallel load; every segment of the
// batch (one thread each) adds its edge's crossings of the tile's sample rows into per-command
// LDS row deltas, walking only the rows the edge spans; each row is prefix-summed once; then every
// pixel applies the batch's commands in order, reading each command's record as wave-uniform
// values. The dependent global loads (command -> draw / path tile -> segment -> edge) are thus
// paid once per batch and once per segment.
// WORDS: row-delta words of LDS per workgroup (B = WORDS / pixels commands per batch)
// EXT: the frame has surface sources or targets with a stencil plane (the plain variant compiles
// neither). With a stencil plane each pixel also carries its stencil word
// (loaded from the plane or zero, stored back at the end): draws blend only where CLIP is 0, clear
// zeroes the stencil byte, and the clip and stencil draws rewrite it (src/vkvg_context.c:698-795,
// :1251-1512; pipeline states of src/vkvg_device_internal.c:226-246).
template <int T, int WORDS, bool EXT>
__global__ __launch_bounds__(256) __attribute__((amdgpu_num_sgpr(80))) void k_fine(const TargetInfo *__restrict__ tg, uint32_t n_targets,
                                              const DrawPrep *__restrict__ dprep,
                                              const GradRec *__restrict__ grec,
                                              const vkh_gradient *__restrict__ grads,
                                              const GradPrep *__restrict__ gprep,
                                              const PathTile *__restrict__ pts, const Edge *__restrict__ segs,
                                              uint32_t nch,
                                              const uint32_t *__restrict__ tile_off, const Cmd *__restrict__ cmds,
                                              const uint32_t *__restrict__ tile_order) {
    constexpr int NPIX = T * T;
    constexpr int NT = NPIX >= 256 ? 256 : NPIX;  // threads
    constexpr int PPT = NPIX / NT;                // pixels per thread
    constexpr int B = WORDS / NPIX;               // commands per batch
    __shared__ int32_t sd[B][T][T];               // row deltas, then per-row inclusive prefix
    __shared__ int32_t srow[B][T];                // V term: +-1 at the first row below a column crossing
    __shared__ uint32_t c_item[B + 1];            // prefix of the commands' segment counts
    __shared__ uint32_t c_nseg[B], c_segoff[B], c_draw[B];
    __shared__ int32_t c_back[B];                 // path tile backdrop (the row prefix adds it)
    __shared__ uint4 c_rec[B];  // meta, key, winding added in compositing (the backdrop without segments),
                                // covering winding bits (the parameter of a stencil draw)
    __shared__ uint4 c_src[B];  // blend operands of a solid source (vkh_src16)

    const uint32_t gtile = tile_order ? tile_order[blockIdx.x] : blockIdx.x;
    uint32_t c0 = tile_off[(size_t)gtile * nch], c1 = tile_off[(size_t)(gtile + 1) * nch];
    if (!EXT && c0 == c1)
        return;
    const uint32_t tid = threadIdx.x;
    uint32_t t = find_target_by(tg, n_targets, gtile, false);
    const TargetInfo &TI = tg[t];
    uint32_t lt = gtile - TI.tile_base;
    int32_t tx = (int32_t)(lt % TI.tilesW), ty = (int32_t)(lt / TI.tilesW);
    int32_t X0 = tx * T, Y0 = ty * T;
    if (EXT && c0 == c1) {
        // no command: the plane keeps its words, or is zeroed when the frame starts from zero
        if (TI.stencil && !TI.stencil_load)
            for (int q = 0; q < NPIX / NT; q++) {
                const int p = (int)tid + q * NT, x = X0 + p % T, y = Y0 + p / T;
                if (x < (int32_t)TI.width && y < (int32_t)TI.height)
                    TI.stencil[(size_t)y * TI.width + (size_t)x] = 0u;
            }
        return;
    }
    const float Wf = (float)TI.width, Hf = (float)TI.height;
    const int32_t xs0 = 256 * X0 + 128, s0y = 256 * Y0 + 128;

    int lx[PPT], ly[PPT];
    bool valid[PPT];
    uint32_t px[PPT], sig[PPT];
    float gx[PPT], gy[PPT]; // gradient coordinates gl_FragCoord.xy / size
#pragma unroll
    for (int q = 0; q < PPT; q++) {
        int p = (int)tid + q * NT;
        lx[q] = p % T;
        ly[q] = p / T;
        valid[q] = (X0 + lx[q]) < (int32_t)TI.width && (Y0 + ly[q]) < (int32_t)TI.height;
        sig[q] = 0;
        gx[q] = ((float)(X0 + lx[q]) + 0.5f) / Wf;
        gy[q] = ((float)(Y0 + ly[q]) + 0.5f) / Hf;
    }
    // the destination need not be read when the first command overwrites every pixel (under a
    // stencil only a clear does: any other draw may be clipped)
    bool need_read = true;
    {
        const Cmd f = cmds[c0];
        if (f.pt == VKH_NO_DRAW) {
            const DrawPrep &d0 = dprep[f.draw];
            const uint32_t kind0 = d0.meta & 15, op0 = (d0.meta >> 4) & 15;
            if (kind0 == VKH_DRAW_CLEAR)
                need_read = false;
            else if (!EXT && kind0 == VKH_DRAW_PAINT && op0 == 0)
                need_read = false;
            else if (!EXT && kind0 == VKH_DRAW_PAINT && !(d0.meta & VKH_DP_GRAD) && (op0 == 1 || op0 == 2) &&
                     (d0.solid >> 24) == 255u)
                need_read = false;
        }
    }
    uint32_t st[PPT]; // stencil words
#pragma unroll
    for (int q = 0; q < PPT; q++) {
        st[q] = 0u;
        if (EXT && TI.stencil && TI.stencil_load && valid[q])
            st[q] = TI.stencil[(size_t)(Y0 + ly[q]) * TI.width + (size_t)(X0 + lx[q])];
    }
    uint32_t *img = (uint32_t *)TI.rgba;
#pragma unroll
    for (int q = 0; q < PPT; q++) {
        size_t o = (size_t)(Y0 + ly[q]) * TI.width + (size_t)(X0 + lx[q]);
        px[q] = (need_read && valid[q]) ? img[o] : 0u;
    }

    // the row deltas start at zero; afterwards every batch leaves them at zero (the row prefix
    // clears srow as it reads it, the compositing clears each pixel's word of sd after its last read)
    for (uint32_t k = tid; k < (uint32_t)(B * NPIX); k += NT)
        (&sd[0][0][0])[k] = 0;
    for (uint32_t k = tid; k < (uint32_t)(B * T); k += NT)
        (&srow[0][0])[k] = 0;
    for (uint32_t base = c0; base < c1; base += B) {
        const uint32_t nb = min((uint32_t)B, c1 - base);
        // ---- stage the batch's records
        for (uint32_t j = tid; j < nb; j += NT) {
            const Cmd cm = cmds[base + j];
            const DrawPrep &d = dprep[cm.draw];
            uint32_t meta = d.meta;
            const uint32_t kind = meta & 15;
            if (cm.pt == VKH_NO_DRAW) {
                if (kind <= VKH_DRAW_CLEAR)
                    meta = (meta & ~15u) | (kind == VKH_DRAW_CLEAR ? VKH_DRAW_CLEAR : VKH_DRAW_PAINT);
                c_back[j] = 0;
                c_rec[j].z = kind <= VKH_DRAW_CLEAR ? 1u : 0u; // paint and clear: winding 1 everywhere
                c_nseg[j] = 0;
                c_segoff[j] = 0;
            } else {
                const PathTile p = pts[cm.pt];
                c_back[j] = p.backdrop;
                c_rec[j].z = p.nseg ? 0u : (uint32_t)p.backdrop;
                c_nseg[j] = p.nseg;
                c_segoff[j] = p.seg_off;
                if (p.nseg)
                    meta |= VKH_DP_HASSEG;
            }
            c_rec[j].x = meta;
            c_rec[j].y = d.key;
            // the winding bits that make a sample covered: odd for even-odd fills, any for non-zero
            // fills, strokes (whose winding, a count of covering triangles, is never negative), paint
            // and clear; the stencil draws read their parameter instead
            c_rec[j].w = kind >= VKH_DRAW_CLIP_NZ ? d.solid : kind == VKH_DRAW_FILL_EO ? 1u : 0xFFFFFFFFu;
            const vkh_src16 S = vkh_src16_of((meta >> 4) & 15, d.solid);
            c_src[j] = make_uint4(vkh_u32_of(S.ia), vkh_u32_of(S.b_lo), vkh_u32_of(S.b_hi), S.x);
            c_draw[j] = cm.draw;
        }
        __syncthreads();
        if (tid == 0) {
            uint32_t acc = 0;
            for (uint32_t j = 0; j < nb; j++) {
                c_item[j] = acc;
                acc += c_nseg[j];
            }
            c_item[nb] = acc;
        }
        __syncthreads();
        // ---- crossings: one thread per segment of the batch, walking only the sample rows its
        // edge spans inside the tile
        const uint32_t total = c_item[nb];
        for (uint32_t it = tid; it < total; it += NT) {
            uint32_t lo = 0, hi = nb; // last j with c_item[j] <= it
            while (hi - lo > 1) {
                uint32_t mid = (lo + hi) >> 1;
                if (c_item[mid] <= it)
                    lo = mid;
                else
                    hi = mid;
            }
            const uint32_t j = lo;
            const Edge e = segs[c_segoff[j] + (it - c_item[j])];
            // V term: the edge separates (xs0, s0y) from (xs0, sy(r)) for every row r >= r*
            if (min(e.x0, e.x1) <= xs0 && xs0 < max(e.x0, e.x1)) {
                const bool lr = e.x0 < e.x1;
                const int32_t r = cross_first_row(lr ? e.x0 : e.x1, lr ? e.y0 : e.y1, lr ? e.x1 : e.x0,
                                                  lr ? e.y1 : e.y0, xs0, s0y, 0, T);
                if (r >= 1 && r < T) // r* = 0: (xs0, s0y) is not above the edge
                    atomicAdd(&srow[j][r], lr ? -1 : 1);
            }
            // H term: rows with ytop <= sy(r) < ybot
            const bool down = e.y0 < e.y1;
            const int32_t xt = down ? e.x0 : e.x1, yt = down ? e.y0 : e.y1;
            const int32_t xb = down ? e.x1 : e.x0, yb = down ? e.y1 : e.y0;
            const int32_t r_lo = max(((yt - 128 + 255) >> 8) - Y0, 0);
            const int32_t r_hi = min(((yb - 128 + 255) >> 8) - 1 - Y0, T - 1);
            if (r_lo > r_hi)
                continue;
            const int32_t sgn = down ? 1 : -1;
            vkh_dda dda = cross_dda_start(xt, yt, xb, yb, s0y + 256 * r_lo);
            for (int32_t r = r_lo;; r++) {
                const int32_t lc = dda.q - X0;
                if (lc >= 1 && lc < T)
                    atomicAdd(&sd[j][r][lc], sgn);
                if (r == r_hi)
                    break;
                cross_dda_step(&dda);
            }
        }
        __syncthreads();
        // ---- per (command, row): inclusive prefix with the row's V term. The T rows of a command
        // are T consecutive lanes (nb * T and the loop stride are multiples of T, so a group is
        // wholly active or inactive): a segmented shuffle scan sums srow[j][1 .. r] (srow[j][0] is 0).
        for (uint32_t q = tid; q < nb * (uint32_t)T; q += NT) {
            const uint32_t j = q / T, r = q % T;
            int32_t acc = srow[j][r];
            srow[j][r] = 0;
#pragma unroll
            for (int o = 1; o < T; o <<= 1) {
                const int32_t y = __shfl_up(acc, (unsigned)o, T);
                if ((int)r >= o)
                    acc += y;
            }
            if (!c_nseg[j])
                continue;
            acc += c_back[j];
#pragma unroll
            for (int c = 0; c < T; c++) {
                acc += sd[j][r][c];
                sd[j][r][c] = acc;
            }
        }
        __syncthreads();
        // ---- ordered compositing. A command's coverage count k is 0 or 1 for fills (no loop) and
        // the winding number for strokes (a pixel covered by k triangles is blended k times); the
        // blend operands of a solid source were computed once per command when the batch was staged.
        for (uint32_t j = 0; j < nb; j++) {
            const uint4 rec = c_rec[j];
            const uint32_t meta = uniform_u32(rec.x);
            const uint32_t kind = meta & 15, op = (meta >> 4) & 15;
            const int32_t bk = (int32_t)rec.z; // 0 when the winding comes from the row deltas
            const uint32_t key = rec.y;
            if (EXT && kind >= VKH_DRAW_CLIP_NZ) { // stencil draws (uniform branch)
                if (kind_is_clip(kind)) {
#pragma unroll
                    for (int q = 0; q < PPT; q++) {
                        // the row deltas of a command without segments are zero
                        const int32_t w = sd[j][ly[q]][lx[q]] + bk;
                        sd[j][ly[q]][lx[q]] = 0;
                        const bool cov = kind == VKH_DRAW_CLIP_NZ ? w != 0 : (w & 1) != 0;
                        if (cov && !(st[q] & 2u)) {
                            if (valid[q])
                                sig[q] += key; // the clip's 1-sample mask
                        } else {
                            st[q] |= 2u;
                        }
                    }
                } else {
                    const uint32_t prm = uniform_u32(rec.w);
#pragma unroll
                    for (int q = 0; q < PPT; q++)
                        st[q] = stencil_apply(kind, prm, st[q]);
                }
                continue;
            }
            const bool stroke = kind == VKH_DRAW_STROKE;
            const uint32_t rule = rec.w; // covering winding bits
            int32_t wq[PPT];
            bool on[PPT];
#pragma unroll
            for (int q = 0; q < PPT; q++) {
                // the row deltas of a command without segments are zero
                const int32_t w = sd[j][ly[q]][lx[q]] + bk;
                sd[j][ly[q]][lx[q]] = 0;
                wq[q] = w;
                on[q] = ((uint32_t)w & rule) != 0u && valid[q];
                if (EXT && on[q]) {
                    if (kind == VKH_DRAW_CLEAR)
                        st[q] &= ~0xFFu; // vkCmdClearAttachments ignores the stencil test
                    else if (st[q] & 2u)
                        on[q] = false; // clipped
                }
            }
            // blend operands: once per command for a solid source, per pixel for a gradient or surface
            vkh_src16 S[PPT];
            if (!(meta & (EXT ? (VKH_DP_GRAD | VKH_DP_SURF) : VKH_DP_GRAD))) {
                const uint4 sv = c_src[j];
                const vkh_src16 S0 = {vkh_u16x2_of(sv.x), vkh_u16x2_of(sv.y), vkh_u16x2_of(sv.z), sv.w};
#pragma unroll
                for (int q = 0; q < PPT; q++)
                    S[q] = S0;
            } else {
#pragma unroll
                for (int q = 0; q < PPT; q++) {
                    if (!on[q])
                        continue;
                    float col[4], op_;
                    if (!EXT || (meta & VKH_DP_GRAD)) {
                        const GradRec &P = grec[uniform_u32(c_draw[j])]; // scalar loads
                        op_ = P.opacity;
                        gradient_color(P, grads, gprep, gx[q], gy[q], col);
                    } else {
                        const SurfRec &R = *reinterpret_cast<const SurfRec *>(&grec[uniform_u32(c_draw[j])]);
                        op_ = R.opacity;
                        surface_color(R, (float)(X0 + lx[q]) + 0.5f, (float)(Y0 + ly[q]) + 0.5f, col);
                    }
                    S[q] = vkh_src16_of(op, f2u8x4(col[0] * op_, col[1] * op_, col[2] * op_, col[3] * op_));
                }
            }
#pragma unroll
            for (int q = 0; q < PPT; q++)
                if (on[q])
                    blend_count(S[q], stroke, wq[q], key, px[q], sig[q]);
        }
        __syncthreads();
    }
#pragma unroll
    for (int q = 0; q < PPT; q++) {
        if (!valid[q])
            continue;
        size_t o = (size_t)(Y0 + ly[q]) * TI.width + (size_t)(X0 + lx[q]);
        img[o] = px[q];
        if (TI.sig)
            TI.sig[o] += sig[q];
        if (EXT && TI.stencil)
            TI.stencil[o] = st[q];
    }
}

// Diagnostic: f2u8x4's byte conversions against f2u8 over all 2^32 float bit patterns, in every
// byte position (2^24 threads x 256 values); counts mismatches.
__global__ void k_check_u8_conversion(unsigned long long *__restrict__ bad) {
    const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    unsigned long long nb = 0;
    for (uint32_t j = 0; j < 256; j++) {
        const float x = __uint_as_float((uint32_t)(t * 256u + j));
        const uint32_t want = f2u8(x);
        const uint32_t got = f2u8x4(x, x, x, x);
        nb += (got != want * 0x01010101u) ? 1u : 0u;
    }
    if (nb)
        atomicAdd(bad, nb);
}

// Diagnostic: the device form of blend_px16 (byte permutes) against the per-channel blend over every
// operator x source alpha x destination channel x source channel, with the other three channels of
// both pixels derived from the varied ones (2^18 threads x 256 source channels); counts mismatches.
__global__ void k_check_blend(unsigned long long *__restrict__ bad) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; // op << 16 | sa << 8 | dc
    const uint32_t op = t >> 16, sa = (t >> 8) & 255u, dc = t & 255u;
    unsigned long long nb = 0;
    const uint32_t d = dc | (((dc * 7u + 3u) & 255u) << 8) | ((255u - dc) << 16) | (((dc * 13u) & 255u) << 24);
    for (uint32_t sc = 0; sc < 256; sc++) {
        const uint32_t s = sc | (((sc * 5u + 1u) & 255u) << 8) | (((sc ^ 0x5Au) & 255u) << 16) | (sa << 24);
        nb += blend_px16(d, vkh_src16_of(op, s)) != blend_px_channels(op, d, s) ? 1u : 0u;
        nb += blend_px16(s, vkh_src16_of(op, d)) != blend_px_channels(op, s, d) ? 1u : 0u;
    }
    if (nb)
        atomicAdd(bad, nb);
}

// readback with the reference's truncating unpremultiply (src/vkvg_surface.c:441-452)
__global__ void k_unpremultiply(const uint8_t *__restrict__ src, uint8_t *__restrict__ dst, size_t npix) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= npix)
        return;
    uchar4 p = ((const uchar4 *)src)[i];
    uchar4 q;
    if (p.w == 0) {
        q = make_uchar4(0, 0, 0, 0);
    } else {
        double alpha = (double)p.w / (double)255.f;
        q.x = (uint8_t)(int32_t)((double)p.x / alpha);
        q.y = (uint8_t)(int32_t)((double)p.y / alpha);
        q.z = (uint8_t)(int32_t)((double)p.z / alpha);
        q.w = p.w;
    }
    ((uchar4 *)dst)[i] = q;
}

// clears the stencil byte (byte 0) of each pixel's stencil word; bytes 1..3 hold the saved copies a
// context spilled, which live in images of their own in the reference and survive a surface clear
__global__ void k_stencil_clear(uint32_t *__restrict__ words, size_t npix) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < npix)
        words[i] &= 0xFFFFFF00u;
}

namespace vkh {
void launch_draw_bbox_init(hipStream_t s, int4 *bbox, uint2 *erange, uint32_t n) {
    if (n)
        hipLaunchKernelGGL(k_bbox_init, dim3((n + 255) / 256), dim3(256), 0, s, bbox, erange, n);
}
void launch_draw_tiles(hipStream_t s, const vkh_draw *draws, uint32_t n, const TargetInfo *tg, const int4 *bbox,
                       uint32_t tile, DrawTiles *dt, uint32_t *pt_count) {
    if (n)
        hipLaunchKernelGGL(k_draw_tiles, dim3((n + 255) / 256), dim3(256), 0, s, draws, n, tg, bbox, tile, dt,
                           pt_count);
}
void launch_set_pt_base(hipStream_t s, DrawTiles *dt, const uint32_t *pt_off, uint32_t n) {
    if (n)
        hipLaunchKernelGGL(k_set_pt_base, dim3((n + 255) / 256), dim3(256), 0, s, dt, pt_off, n);
}
void launch_zero_u32(hipStream_t s, uint32_t *p, uint32_t n) {
    if (n)
        hipLaunchKernelGGL(k_zero_u32, dim3((n + 255) / 256), dim3(256), 0, s, p, n);
}
void launch_bin_draws(hipStream_t s, const Edge *edges, const uint2 *erange,
                      const DrawTiles *dt, uint32_t n_draws, uint32_t tile, PathTile *pts, Edge *segs,
                      uint32_t seg_cap, uint32_t *seg_alloc, const uint32_t *order, const vkh_draw *draws,
                      const TargetInfo *tg, uint32_t nch, uint32_t chunk_draws, uint32_t *tile_cnt) {
    if (n_draws)
        hipLaunchKernelGGL(k_bin_draw, dim3(n_draws), dim3(256), 0, s, edges, erange, dt, tile, pts, segs,
                           seg_cap, seg_alloc, order, draws, tg, nch, chunk_draws, tile_cnt);
}
void launch_tile_cnt_init(hipStream_t s, const TargetInfo *tg, uint32_t n_targets, uint32_t nch,
                          const uint32_t *table, uint32_t *tile_cnt, uint32_t n_tc) {
    hipLaunchKernelGGL(k_tile_cnt_init, dim3((n_tc + 256) / 256), dim3(256), 0, s, tg, n_targets, nch, table, tile_cnt,
                       n_tc);
}
uint32_t coarse_chunk(uint32_t most_draws, uint32_t n_bins) {
    const uint32_t want = (1024 + std::max(n_bins, 1u) - 1) / std::max(n_bins, 1u); // chunks per bin
    uint32_t c = (most_draws + want - 1) / want;
    c = (c + 63) / 64 * 64;
    return std::min<uint32_t>(std::max<uint32_t>(c, 64), VKH_COARSE_CHUNK);
}
void launch_coarse(hipStream_t s, bool write, uint32_t n_bins, uint32_t nch, uint32_t chunk, const TargetInfo *tg,
                   uint32_t n_targets, const vkh_draw *draws, const DrawTiles *dt, const PathTile *pts,
                   uint32_t *tile_count, const uint32_t *tile_off, Cmd *cmds) {
    if (n_bins && nch)
        hipLaunchKernelGGL(k_coarse, dim3(n_bins, nch), dim3(256), 0, s, write, tg, n_targets, draws, dt, pts, nch,
                           chunk, tile_count, tile_off, cmds);
}
void launch_draw_prep(hipStream_t s, const vkh_draw *draws, uint32_t n, const TargetInfo *tg,
                      const vkh_gradient *grads, uint32_t n_grads, const vkh_surface_source *srcs, DrawPrep *out,
                      GradRec *rec, GradPrep *gout) {
    if (n)
        hipLaunchKernelGGL(k_draw_prep, dim3((n + 255) / 256), dim3(256), 0, s, draws, n, tg, grads, srcs, out, rec);
    if (n_grads)
        hipLaunchKernelGGL(k_grad_prep, dim3((n_grads + 255) / 256), dim3(256), 0, s, grads, n_grads, gout);
}
template <int T, int W>
static void fine_launch(hipStream_t s, bool ext, uint32_t n_tiles, const TargetInfo *tg, uint32_t n_targets,
                        const DrawPrep *dprep, const GradRec *grec, const vkh_gradient *grads, const GradPrep *gprep,
                        const PathTile *pts, const Edge *segs, uint32_t nch, const uint32_t *tile_off,
                        const Cmd *cmds, const uint32_t *order) {
    constexpr int NT = T * T >= 256 ? 256 : T * T;
    if (ext)
        hipLaunchKernelGGL((k_fine<T, W, true>), dim3(n_tiles), dim3(NT), 0, s, tg, n_targets, dprep, grec, grads, gprep,
                           pts, segs, nch, tile_off, cmds, order);
    else
        hipLaunchKernelGGL((k_fine<T, W, false>), dim3(n_tiles), dim3(NT), 0, s, tg, n_targets, dprep, grec, grads,
                           gprep, pts, segs, nch, tile_off, cmds, order);
}
template <int T>
static void fine_words(hipStream_t s, bool ext, uint32_t words, uint32_t n_tiles, const TargetInfo *tg, uint32_t n_targets,
                       const DrawPrep *dprep, const GradRec *grec, const vkh_gradient *grads, const GradPrep *gprep,
                       const PathTile *pts, const Edge *segs, uint32_t nch, const uint32_t *tile_off,
                       const Cmd *cmds, const uint32_t *order) {
    switch (words) {
    case 2048:
        fine_launch<T, 2048>(s, ext, n_tiles, tg, n_targets, dprep, grec, grads, gprep, pts, segs, nch, tile_off, cmds, order);
        break;
    case 8192:
        fine_launch<T, 8192>(s, ext, n_tiles, tg, n_targets, dprep, grec, grads, gprep, pts, segs, nch, tile_off, cmds, order);
        break;
    default:
        fine_launch<T, 4096>(s, ext, n_tiles, tg, n_targets, dprep, grec, grads, gprep, pts, segs, nch, tile_off, cmds, order);
        break;
    }
}
void launch_fine(hipStream_t s, bool ext, uint32_t tile, uint32_t words, uint32_t n_tiles, const TargetInfo *tg,
                 uint32_t n_targets, const DrawPrep *dprep, const GradRec *grec, const vkh_gradient *grads,
                 const GradPrep *gprep, const PathTile *pts, const Edge *segs, uint32_t nch,
                 const uint32_t *tile_off, const Cmd *cmds, const uint32_t *order) {
    if (!n_tiles)
        return;
    switch (tile) {
    case 8:
        fine_words<8>(s, ext, words, n_tiles, tg, n_targets, dprep, grec, grads, gprep, pts, segs, nch, tile_off, cmds, order);
        break;
    case 32:
        fine_words<32>(s, ext, words, n_tiles, tg, n_targets, dprep, grec, grads, gprep, pts, segs, nch, tile_off, cmds, order);
        break;
    default:
        fine_words<16>(s, ext, words, n_tiles, tg, n_targets, dprep, grec, grads, gprep, pts, segs, nch, tile_off, cmds, order);
        break;
    }
}
size_t order_words(uint32_t n) { return (size_t)n + VKH_ORDER_BUCKETS; }
template <class W>
static void launch_order(hipStream_t s, uint32_t n, W weight, uint32_t *work) {
    if (!n)
        return;
    uint32_t *bucket = work + n, *order = work;
    const unsigned g = (n + 255u) / 256u;
    launch_zero_u32(s, bucket, VKH_ORDER_BUCKETS);
    hipLaunchKernelGGL((k_work_order<false, W>), dim3(g), dim3(256), 0, s, n, weight, bucket, order);
    hipLaunchKernelGGL(k_work_order_scan, dim3(1), dim3(256), 0, s, bucket);
    hipLaunchKernelGGL((k_work_order<true, W>), dim3(g), dim3(256), 0, s, n, weight, bucket, order);
}
void launch_tile_order(hipStream_t s, uint32_t n_tiles, uint32_t nch, const uint32_t *tile_off, uint32_t *work) {
    launch_order(s, n_tiles, TileWeight{nch, tile_off}, work);
}
void launch_draw_order(hipStream_t s, uint32_t n_draws, const uint2 *erange, uint32_t *work) {
    launch_order(s, n_draws, DrawWeight{erange}, work);
}
void launch_check_u8_conversion(hipStream_t s, unsigned long long *bad) {
    hipLaunchKernelGGL(k_check_u8_conversion, dim3(1u << 16), dim3(256), 0, s, bad);
}
void launch_check_blend(hipStream_t s, unsigned long long *bad) {
    hipLaunchKernelGGL(k_check_blend, dim3(1u << 10), dim3(256), 0, s, bad); // 2^18 = 4 ops x 256 x 256
}
void launch_stencil_clear(hipStream_t s, uint32_t *words, size_t npix) {
    if (npix)
        hipLaunchKernelGGL(k_stencil_clear, dim3((unsigned)((npix + 255) / 256)), dim3(256), 0, s, words, npix);
}
void launch_unpremultiply(hipStream_t s, const uint8_t *src, uint8_t *dst, size_t npix) {
    if (npix)
        hipLaunchKernelGGL(k_unpremultiply, dim3((unsigned)((npix + 255) / 256)), dim3(256), 0, s, src, dst, npix);
}
} // namespace vkh
