// Engine of the vkvg HIP backend: device buffers, the frame pipeline and the C ABI of
// include/vkvg_hip.h. One engine per device; a frame may hold draws for many targets.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "vkh_internal.h"

// Staging array of vkvg_hip_image_read_unpremultiplied_noalloc: one RGBA8 image of up to 4096 x 4096
// pixels, held by every GPU the library runs on until the process ends. Host code uses the address
// hipGetSymbolAddress returns per device (vkh_engine::noalloc_scratch), never the name.
#define VKH_NOALLOC_SCRATCH_BYTES (64u << 20)
static __device__ __align__(256) uint8_t scratch[VKH_NOALLOC_SCRATCH_BYTES];

// ----------------------------------------------------------------------------------------- scan
// Exclusive prefix sum over n uint32 values into out[0..n] (out[n] = total): block scans of
// 1024 values (256 threads x 4), then the block totals added back (summed per block when there
// are at most 1024 of them, scanned recursively otherwise).
__global__ __launch_bounds__(256) void k_scan_block(const uint32_t *__restrict__ in, uint32_t n,
                                                    uint32_t *__restrict__ out, uint32_t *__restrict__ sums) {
    __shared__ uint32_t wsum[4];
    uint32_t base = blockIdx.x * 1024u + threadIdx.x * 4u;
    uint32_t v[4];
#pragma unroll
    for (int k = 0; k < 4; k++)
        v[k] = (base + k < n) ? in[base + k] : 0u;
    uint32_t t = v[0] + v[1] + v[2] + v[3];
    uint32_t lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    uint32_t incl = t;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        uint32_t y = __shfl_up(incl, o);
        if (lane >= (uint32_t)o)
            incl += y;
    }
    if (lane == 63)
        wsum[wid] = incl;
    __syncthreads();
    uint32_t woff = 0;
    for (uint32_t w = 0; w < wid; w++)
        woff += wsum[w];
    uint32_t run = woff + incl - t;
#pragma unroll
    for (int k = 0; k < 4; k++) {
        if (base + k <= n)
            out[base + k] = run;
        run += v[k];
    }
    if (threadIdx.x == 255 && sums)
        sums[blockIdx.x] = woff + incl;
}
// block b adds the total of blocks 0 .. b-1 (at most 1024 block sums, reduced by the workgroup)
__global__ __launch_bounds__(256) void k_scan_add_sums(uint32_t *__restrict__ out, uint32_t n,
                                                       const uint32_t *__restrict__ sums) {
    __shared__ uint32_t wsum[4];
    const uint32_t b = blockIdx.x, tid = threadIdx.x;
    uint32_t x = 0;
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const uint32_t j = tid * 4u + (uint32_t)k;
        x += j < b ? sums[j] : 0u;
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1)
        x += __shfl_xor(x, o);
    if ((tid & 63) == 0)
        wsum[tid >> 6] = x;
    __syncthreads();
    const uint32_t o = wsum[0] + wsum[1] + wsum[2] + wsum[3];
    const uint32_t i = b * 1024u + tid;
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const uint32_t j = i + 256u * k;
        if (j <= n)
            out[j] += o;
    }
}
__global__ void k_scan_add(uint32_t *__restrict__ out, uint32_t n, const uint32_t *__restrict__ offs) {
    uint32_t i = blockIdx.x * 1024u + threadIdx.x;
    uint32_t o = offs[blockIdx.x];
#pragma unroll
    for (int k = 0; k < 4; k++) {
        uint32_t j = i + 256u * k;
        if (j <= n)
            out[j] += o;
    }
}

namespace vkh {
// temp words needed by scan_u32 for n values
size_t scan_temp_words(uint32_t n) {
    const uint32_t nb = (n + 1 + 1023) / 1024;
    if (nb <= 1)
        return 1;
    if (nb <= 1024)
        return nb + 1;
    return 2ull * nb + 1 + scan_temp_words(nb);
}
// exclusive scan of n values; out must hold n+1 entries (out[n] = total); tmp holds
// scan_temp_words(n) words. Up to 1024 blocks of 1024 values: block scans, then each block adds
// the sum of the blocks before it (two launches); more: the block sums are scanned recursively.
void scan_u32(hipStream_t s, const uint32_t *in, uint32_t n, uint32_t *out, uint32_t *tmp) {
    const uint32_t nb = (n + 1 + 1023) / 1024;
    if (nb <= 1) {
        hipLaunchKernelGGL(k_scan_block, dim3(1), dim3(256), 0, s, in, n, out, nullptr);
        return;
    }
    uint32_t *sums = tmp;
    hipLaunchKernelGGL(k_scan_block, dim3(nb), dim3(256), 0, s, in, n, out, sums);
    if (nb <= 1024) {
        hipLaunchKernelGGL(k_scan_add_sums, dim3(nb), dim3(256), 0, s, out, n, sums);
        return;
    }
    uint32_t *sscan = tmp + nb; // nb + 1 entries
    scan_u32(s, sums, nb, sscan, tmp + 2ull * nb + 1);
    hipLaunchKernelGGL(k_scan_add, dim3(nb), dim3(256), 0, s, out, n, sscan);
}
} // namespace vkh

namespace {

struct DevBuf {
    void *p = nullptr;
    size_t cap = 0;
};

struct HostBuf {
    void *p = nullptr;
    size_t cap = 0;
};

} // namespace

struct vkh_engine {
    int device = 0;
    uint8_t *noalloc_scratch = nullptr; // this device's address of `scratch` (allocation-free readback)
    hipStream_t stream = nullptr;
    hipStream_t own = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    uint32_t tile = 16;
    std::string err;
    vkh_stats stats{};

    DevBuf blob, units, tgts, item_cnt, item_off, pts, pflags, sp_first, sp_count;
    DevBuf fe, nv, nt, fe_off, nv_off, nt_off, verts, tris, edges;
    DevBuf bbox, erange, dtiles, pt_count, pt_off, ptiles, counters, segs, dprep; // counters: [0] segment allocator, [1] flatten slots, [2] overflow cubics, [3] dashed pieces
    DevBuf tile_cnt, tile_off, tile_order, draw_order, cmds, scan_tmp, readback, item_slot, slot_item, flat_scratch, flat_perm, flat_ovf,
        flat_wh, dash_list, tl_table;
    DevBuf step_nv, step_nt, sv_off, st_off, u_toff, u_vbeg, u_vend, dash_st; // dash_st: per dashed piece, its starting dash state
    HostBuf hblob, htot;
    hipEvent_t ev_begin = nullptr, ev_end = nullptr;
    hipEvent_t ev_stage[6] = {};

    // graph replay of the resident frame (sizes cached by the last sizing run)
    bool graph_on = false;
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    uint64_t gkey = 0, gfail_key = 0;
    uint64_t frame_gen = 0, alloc_gen = 0;
    uint32_t sz[10] = {}; // points, overflow cubics, fill edges, stroke steps, stroke vertices, triangles, dash
                          // walk stopped, path tiles, segments, commands
    uint32_t fine_words = 4096;
    std::vector<TargetInfo> h_tg;
    // draws without path tiles (target, index within the target's draws), and per (target, coarse
    // chunk) how many of them there are: every tile of the target holds their commands
    std::vector<uint2> h_tileless;
    std::vector<uint32_t> h_tl_table;

    // the uploaded (resident) frame
    bool resident = false;
    size_t o_items, o_sp, o_draws, o_styles, o_dashes, o_grads, o_tg, o_units, o_srcs;
    uint32_t n_items, n_sp, n_draws, n_styles, n_dashes, n_grads, n_targets, n_units;
    uint32_t tile_total, bin_total, frame_tile, coarse_nch = 1, coarse_chunk = VKH_COARSE_CHUNK, n_cubic_items = 0;
    uint32_t most_draws = 1; // most draws on one target
    bool frame_ext = false; // the frame has surface sources or a target with a stencil plane
};

static bool check(vkh_engine *e, hipError_t r, const char *what) {
    if (r == hipSuccess)
        return true;
    e->err = std::string(what) + ": " + hipGetErrorString(r);
    return false;
}
#define CK(x)                                                                                                          \
    do {                                                                                                               \
        if (!check(eng, (x), #x))                                                                                      \
            return -1;                                                                                                 \
    } while (0)

static bool ensure(vkh_engine *eng, DevBuf &b, size_t bytes) {
    if (bytes <= b.cap && b.p)
        return true;
    eng->alloc_gen++; // invalidates a captured graph that points at the old buffer
    if (b.p)
        (void)hipFree(b.p);
    b.p = nullptr;
    size_t cap = bytes + bytes / 2 + 256;
    if (!check(eng, hipMalloc(&b.p, cap), "hipMalloc")) {
        b.cap = 0;
        return false;
    }
    b.cap = cap;
    return true;
}
static bool ensure_host(vkh_engine *eng, HostBuf &b, size_t bytes) {
    if (bytes <= b.cap && b.p)
        return true;
    if (b.p)
        (void)hipHostFree(b.p);
    b.p = nullptr;
    size_t cap = bytes + bytes / 2 + 256;
    if (!check(eng, hipHostMalloc(&b.p, cap, 0), "hipHostMalloc")) {
        b.cap = 0;
        return false;
    }
    b.cap = cap;
    return true;
}
template <typename T> static T *P(DevBuf &b) { return (T *)b.p; }

// in capture mode a buffer must already be large enough: nothing is allocated
static bool need(vkh_engine *eng, bool cap, DevBuf &b, size_t bytes) {
    if (!cap)
        return ensure(eng, b, bytes);
    if (b.p && bytes <= b.cap)
        return true;
    eng->err = "graph capture would need an allocation";
    return false;
}

static int scan_u32(vkh_engine *eng, const uint32_t *in, uint32_t n, uint32_t *out, bool cap = false) {
    if (!need(eng, cap, eng->scan_tmp, vkh::scan_temp_words(n) * sizeof(uint32_t)))
        return -1;
    vkh::scan_u32(eng->stream, in, n, out, P<uint32_t>(eng->scan_tmp));
    return check(eng, hipGetLastError(), "scan") ? 0 : -1;
}

static int read_totals(vkh_engine *eng, const std::vector<const uint32_t *> &srcs, uint32_t *dst) {
    if (!ensure_host(eng, eng->htot, 64 * sizeof(uint32_t)))
        return -1;
    uint32_t *h = (uint32_t *)eng->htot.p;
    for (size_t i = 0; i < srcs.size(); i++)
        CK(hipMemcpyAsync(h + i, srcs[i], 4, hipMemcpyDeviceToHost, eng->stream));
    CK(hipStreamSynchronize(eng->stream));
    for (size_t i = 0; i < srcs.size(); i++)
        dst[i] = h[i];
    return 0;
}

static size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

extern "C" {

vkh_engine *vkvg_hip_engine_create(int device_ordinal) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= device_ordinal || device_ordinal < 0)
        return nullptr;
    if (hipSetDevice(device_ordinal) != hipSuccess)
        return nullptr;
    vkh_engine *e = new vkh_engine();
    e->device = device_ordinal;
    if (hipStreamCreateWithFlags(&e->own, hipStreamNonBlocking) != hipSuccess) {
        delete e;
        return nullptr;
    }
    e->stream = e->own;
    // the allocation-free readback's staging array: its address on this device (each GPU holds its
    // own copy of the array; the name itself is a host-side stand-in, not a device address)
    if (hipGetSymbolAddress((void **)&e->noalloc_scratch, HIP_SYMBOL(scratch)) != hipSuccess) {
        (void)hipGetLastError();
        e->noalloc_scratch = nullptr;
    }
    (void)hipEventCreate(&e->ev0);
    (void)hipEventCreate(&e->ev1);
    (void)hipEventCreate(&e->ev_begin);
    (void)hipEventCreate(&e->ev_end);
    for (hipEvent_t &ev : e->ev_stage)
        (void)hipEventCreate(&ev);
    return e;
}

void vkvg_hip_engine_destroy(vkh_engine *e) {
    if (!e)
        return;
    (void)hipSetDevice(e->device);
    if (e->stream)
        (void)hipStreamSynchronize(e->stream);
    if (e->gexec)
        (void)hipGraphExecDestroy(e->gexec);
    if (e->graph)
        (void)hipGraphDestroy(e->graph);
    DevBuf *bufs[] = {&e->blob, &e->units, &e->tgts, &e->item_cnt, &e->item_off, &e->pts, &e->pflags,
                      &e->sp_first, &e->sp_count, &e->fe, &e->nv, &e->nt, &e->fe_off, &e->nv_off, &e->nt_off,
                      &e->verts, &e->tris, &e->edges, &e->bbox, &e->dtiles, &e->pt_count, &e->pt_off,
                      &e->erange, &e->ptiles, &e->counters, &e->segs, &e->dprep, &e->tile_cnt,
                      &e->tile_off, &e->tile_order, &e->draw_order, &e->cmds, &e->scan_tmp, &e->readback, &e->step_nv, &e->step_nt,
                      &e->sv_off, &e->st_off, &e->dash_st, &e->u_toff, &e->u_vbeg, &e->u_vend, &e->item_slot,
                      &e->slot_item, &e->flat_scratch, &e->flat_perm, &e->flat_ovf, &e->flat_wh, &e->dash_list, &e->tl_table};
    for (DevBuf *b : bufs)
        if (b->p)
            (void)hipFree(b->p);
    if (e->hblob.p)
        (void)hipHostFree(e->hblob.p);
    if (e->htot.p)
        (void)hipHostFree(e->htot.p);
    if (e->ev0)
        (void)hipEventDestroy(e->ev0);
    if (e->ev1)
        (void)hipEventDestroy(e->ev1);
    if (e->ev_begin)
        (void)hipEventDestroy(e->ev_begin);
    if (e->ev_end)
        (void)hipEventDestroy(e->ev_end);
    for (hipEvent_t ev : e->ev_stage)
        if (ev)
            (void)hipEventDestroy(ev);
    if (e->own)
        (void)hipStreamDestroy(e->own);
    delete e;
}

int vkvg_hip_set_tile_size(vkh_engine *e, uint32_t tile) {
    if (!e)
        return -1;
    if (tile == 0)
        tile = 16;
    if (tile != 8 && tile != 16 && tile != 32)
        return -1;
    e->tile = tile;
    return 0;
}

void vkvg_hip_set_stream(vkh_engine *e, void *stream) {
    if (e)
        e->stream = stream ? (hipStream_t)stream : e->own;
}

int vkvg_hip_set_fine_words(vkh_engine *e, uint32_t words) {
    if (!e)
        return -1;
    if (words == 0)
        words = 4096;
    if (words != 2048 && words != 4096 && words != 8192)
        return -1;
    e->fine_words = words;
    return 0;
}

void vkvg_hip_set_graph(vkh_engine *e, int on) {
    if (e)
        e->graph_on = on != 0;
}

int vkvg_hip_image_alloc(vkh_engine *eng, uint32_t w, uint32_t h, void **dev_ptr) {
    if (!eng || !dev_ptr)
        return -1;
    (void)hipSetDevice(eng->device);
    size_t bytes = (size_t)w * h * 4;
    CK(hipMalloc(dev_ptr, bytes ? bytes : 4));
    CK(hipMemsetAsync(*dev_ptr, 0, bytes ? bytes : 4, eng->stream));
    CK(hipStreamSynchronize(eng->stream));
    return 0;
}
void vkvg_hip_image_free(vkh_engine *eng, void *p) {
    if (eng && p) {
        (void)hipSetDevice(eng->device);
        (void)hipStreamSynchronize(eng->stream);
        (void)hipFree(p);
    }
}
int vkvg_hip_image_clear(vkh_engine *eng, void *p, size_t bytes) {
    if (!eng || !p)
        return -1;
    CK(hipMemsetAsync(p, 0, bytes, eng->stream));
    CK(hipStreamSynchronize(eng->stream));
    return 0;
}
int vkvg_hip_stencil_clear(vkh_engine *eng, void *p, size_t npix) {
    if (!eng || !p)
        return -1;
    vkh::launch_stencil_clear(eng->stream, (uint32_t *)p, npix);
    CK(hipGetLastError());
    CK(hipStreamSynchronize(eng->stream));
    return 0;
}
int vkvg_hip_image_download(vkh_engine *eng, const void *p, void *host, size_t bytes) {
    if (!eng || !p || !host)
        return -1;
    CK(hipMemcpyAsync(host, p, bytes, hipMemcpyDeviceToHost, eng->stream));
    CK(hipStreamSynchronize(eng->stream));
    return 0;
}
int vkvg_hip_image_upload(vkh_engine *eng, void *p, const void *host, size_t bytes) {
    if (!eng || !p || !host)
        return -1;
    CK(hipMemcpyAsync(p, host, bytes, hipMemcpyHostToDevice, eng->stream));
    CK(hipStreamSynchronize(eng->stream));
    return 0;
}
int vkvg_hip_image_read_unpremultiplied_noalloc(vkh_engine *eng, const void *p, uint32_t w, uint32_t h, void *host) {
    if (!eng || !p || !host)
        return -1;
    const size_t bytes = (size_t)w * h * 4;
    if (bytes > VKH_NOALLOC_SCRATCH_BYTES) {
        eng->err = "vkvg_hip_image_read_unpremultiplied_noalloc: " + std::to_string(bytes) +
                   " bytes exceed the " + std::to_string(VKH_NOALLOC_SCRATCH_BYTES) + "-byte staging array";
        return -1;
    }
    if (!eng->noalloc_scratch) {
        eng->err = "vkvg_hip_image_read_unpremultiplied_noalloc: the staging array has no address on this device";
        return -1;
    }
    (void)hipSetDevice(eng->device);
    vkh::launch_unpremultiply(eng->stream, (const uint8_t *)p, eng->noalloc_scratch, (size_t)w * h);
    CK(hipGetLastError());
    CK(hipMemcpyAsync(host, eng->noalloc_scratch, bytes, hipMemcpyDeviceToHost, eng->stream));
    CK(hipStreamSynchronize(eng->stream));
    return 0;
}
int vkvg_hip_image_read_unpremultiplied(vkh_engine *eng, const void *p, uint32_t w, uint32_t h, void *host) {
    if (!eng || !p || !host)
        return -1;
    size_t npix = (size_t)w * h;
    if (!ensure(eng, eng->readback, npix * 4))
        return -1;
    vkh::launch_unpremultiply(eng->stream, (const uint8_t *)p, P<uint8_t>(eng->readback), npix);
    CK(hipGetLastError());
    CK(hipMemcpyAsync(host, eng->readback.p, npix * 4, hipMemcpyDeviceToHost, eng->stream));
    CK(hipStreamSynchronize(eng->stream));
    return 0;
}

void vkvg_hip_get_stats(vkh_engine *e, vkh_stats *out) {
    if (e && out)
        *out = e->stats;
}
const char *vkvg_hip_last_error(vkh_engine *e) { return e ? e->err.c_str() : "null engine"; }

// coarse work split for the current tiling: draws per chunk and chunks per bin
static int set_coarse_chunks(vkh_engine *eng) {
    eng->coarse_chunk = vkh::coarse_chunk(eng->most_draws, eng->bin_total);
    eng->coarse_nch = (eng->most_draws + eng->coarse_chunk - 1) / eng->coarse_chunk;
    const size_t n = (size_t)std::max<uint32_t>(eng->n_targets, 1u) * eng->coarse_nch;
    eng->h_tl_table.assign(n, 0u);
    for (const uint2 &tl : eng->h_tileless)
        eng->h_tl_table[(size_t)tl.x * eng->coarse_nch + tl.y / eng->coarse_chunk]++;
    if (!ensure(eng, eng->tl_table, 4 * n))
        return -1;
    CK(hipMemcpyAsync(eng->tl_table.p, eng->h_tl_table.data(), 4 * n, hipMemcpyHostToDevice, eng->stream));
    CK(hipStreamSynchronize(eng->stream));
    return 0;
}

int vkvg_hip_upload(vkh_engine *eng, const vkh_frame *f) {
    if (!eng || !f)
        return -1;
    (void)hipSetDevice(eng->device);
    hipStream_t s = eng->stream;
    const uint32_t T = eng->tile;
    eng->resident = false;
    CK(hipStreamSynchronize(s)); // the staging blob may still feed an earlier copy

    // ---- host-side frame tables: work units and targets
    bool any_stencil = false, any_source = false;
    std::vector<Unit> units;
    units.reserve(f->n_subpaths + 16);
    for (uint32_t d = 0; d < f->n_draws; d++) {
        const vkh_draw &dr = f->draws[d];
        if (dr.target >= f->n_targets) {
            eng->err = "draw target out of range";
            return -2;
        }
        if (d && dr.target < f->draws[d - 1].target) {
            eng->err = "draws must be grouped by target in ascending order";
            return -2;
        }
        if (dr.kind > VKH_DRAW_STENCIL_UNSPILL) {
            eng->err = "unknown draw kind";
            return -2;
        }
        if ((dr.kind >= VKH_DRAW_CLIP_NZ) && !f->targets[dr.target].stencil) {
            eng->err = "clip or stencil draw on a target without a stencil plane";
            return -2;
        }
        if (dr.kind >= VKH_DRAW_CLIP_SAVE) {
            const uint32_t lim = dr.kind >= VKH_DRAW_STENCIL_SPILL ? VKH_STENCIL_SPILLS : 6u;
            if (dr.style < 0 || (uint32_t)dr.style >= lim) {
                eng->err = "stencil draw index out of range";
                return -2;
            }
        }
        // sources are checked for every draw (paints take them too)
        if (dr.grad >= 0 && (uint32_t)dr.grad >= f->n_grads) {
            eng->err = "gradient index out of range";
            return -2;
        }
        if (dr.source >= 0) {
            if ((uint32_t)dr.source >= f->n_sources || !f->sources) {
                eng->err = "surface source index out of range";
                return -2;
            }
            const vkh_surface_source &src = f->sources[dr.source];
            any_source = true;
            if (!src.rgba || src.width == 0 || src.height == 0 || src.width > 32768 || src.height > 32768) {
                eng->err = "invalid surface source";
                return -2;
            }
        }
        if (!kind_has_tiles(dr.kind))
            continue;
        if (dr.kind == VKH_DRAW_STROKE && (dr.style < 0 || (uint32_t)dr.style >= f->n_styles)) {
            eng->err = "stroke draw without a style";
            return -2;
        }
        if ((uint64_t)dr.first_subpath + dr.n_subpaths > f->n_subpaths) {
            eng->err = "subpath range out of bounds";
            return -2;
        }
        for (uint32_t k = 0; k < dr.n_subpaths; k++)
            units.push_back(Unit{d, dr.first_subpath + k});
    }
    for (uint32_t i = 0; i < f->n_subpaths; i++) {
        if ((uint64_t)f->subpaths[i].first_item + f->subpaths[i].n_items > f->n_items || !f->subpaths[i].n_items) {
            eng->err = "item range of a subpath out of bounds";
            return -2;
        }
    }
    for (uint32_t i = 0; i < f->n_styles; i++) {
        if ((uint64_t)f->styles[i].dash_first + f->styles[i].dash_count > f->n_dashes) {
            eng->err = "dash range out of bounds";
            return -2;
        }
        if (!(f->styles[i].arcStep >= VKH_MIN_ARC_STEP)) {
            eng->err = "stroke style arc step below VKH_MIN_ARC_STEP";
            return -2;
        }
    }
    std::vector<TargetInfo> tg(f->n_targets);
    uint32_t tile_total = 0, bin_total = 0;
    {
        uint32_t d = 0;
        for (uint32_t t = 0; t < f->n_targets; t++) {
            TargetInfo &ti = tg[t];
            ti.width = f->targets[t].width;
            ti.height = f->targets[t].height;
            if (ti.width == 0 || ti.height == 0 || ti.width > 32768 || ti.height > 32768 || !f->targets[t].rgba) {
                eng->err = "invalid render target";
                return -2;
            }
            ti.tilesW = (ti.width + T - 1) / T;
            ti.tilesH = (ti.height + T - 1) / T;
            ti.tile_base = tile_total;
            tile_total += ti.tilesW * ti.tilesH;
            ti.binsW = (ti.tilesW + VKH_BIN_TILES - 1) / VKH_BIN_TILES;
            ti.binsH = (ti.tilesH + VKH_BIN_TILES - 1) / VKH_BIN_TILES;
            ti.bin_base = bin_total;
            bin_total += ti.binsW * ti.binsH;
            ti.draw_begin = d;
            while (d < f->n_draws && f->draws[d].target == t)
                d++;
            ti.draw_end = d;
            ti.rgba = (uint8_t *)f->targets[t].rgba;
            ti.sig = f->targets[t].sig;
            ti.stencil = f->targets[t].stencil;
            ti.stencil_load = f->targets[t].stencil_load ? 1u : 0u;
            ti.pad = 0;
            if (ti.stencil)
                any_stencil = true;
        }
    }
    const uint32_t n_units = (uint32_t)units.size();
    const uint32_t n_items = f->n_items, n_sp = f->n_subpaths, n_draws = f->n_draws;

    // ---- upload everything in one copy through a pinned staging blob
    size_t o_items = 0;
    size_t o_sp = align256(o_items + sizeof(vkh_item) * n_items);
    size_t o_draws = align256(o_sp + sizeof(vkh_subpath) * n_sp);
    size_t o_styles = align256(o_draws + sizeof(vkh_draw) * n_draws);
    size_t o_dashes = align256(o_styles + sizeof(vkh_stroke_style) * f->n_styles);
    size_t o_grads = align256(o_dashes + sizeof(float) * f->n_dashes);
    size_t o_tg = align256(o_grads + sizeof(vkh_gradient) * f->n_grads);
    size_t o_units = align256(o_tg + sizeof(TargetInfo) * f->n_targets);
    size_t o_srcs = align256(o_units + sizeof(Unit) * n_units);
    size_t total = align256(o_srcs + sizeof(vkh_surface_source) * f->n_sources) + 256;
    if (!ensure_host(eng, eng->hblob, total) || !ensure(eng, eng->blob, total))
        return -1;
    uint8_t *hb = (uint8_t *)eng->hblob.p;
    if (n_items)
        memcpy(hb + o_items, f->items, sizeof(vkh_item) * n_items);
    if (n_sp)
        memcpy(hb + o_sp, f->subpaths, sizeof(vkh_subpath) * n_sp);
    if (n_draws)
        memcpy(hb + o_draws, f->draws, sizeof(vkh_draw) * n_draws);
    if (f->n_styles)
        memcpy(hb + o_styles, f->styles, sizeof(vkh_stroke_style) * f->n_styles);
    if (f->n_dashes)
        memcpy(hb + o_dashes, f->dashes, sizeof(float) * f->n_dashes);
    if (f->n_grads)
        memcpy(hb + o_grads, f->grads, sizeof(vkh_gradient) * f->n_grads);
    if (f->n_targets)
        memcpy(hb + o_tg, tg.data(), sizeof(TargetInfo) * f->n_targets);
    if (n_units)
        memcpy(hb + o_units, units.data(), sizeof(Unit) * n_units);
    if (f->n_sources)
        memcpy(hb + o_srcs, f->sources, sizeof(vkh_surface_source) * f->n_sources);
    CK(hipMemcpyAsync(eng->blob.p, hb, total, hipMemcpyHostToDevice, s));
    eng->o_items = o_items;
    eng->o_sp = o_sp;
    eng->o_draws = o_draws;
    eng->o_styles = o_styles;
    eng->o_dashes = o_dashes;
    eng->o_grads = o_grads;
    eng->o_tg = o_tg;
    eng->o_units = o_units;
    eng->o_srcs = o_srcs;
    eng->n_items = n_items;
    eng->n_sp = n_sp;
    eng->n_draws = n_draws;
    eng->n_styles = f->n_styles;
    eng->n_dashes = f->n_dashes;
    eng->n_grads = f->n_grads;
    eng->n_targets = f->n_targets;
    eng->n_units = n_units;
    eng->tile_total = tile_total;
    eng->bin_total = bin_total;
    eng->frame_tile = T;
    eng->h_tg = tg;
    eng->n_cubic_items = 0;
    for (uint32_t i = 0; i < n_items; i++)
        eng->n_cubic_items += (f->items[i].flags & VKH_ITEM_CUBIC) ? 1u : 0u;
    eng->most_draws = 1;
    for (const TargetInfo &ti : tg)
        eng->most_draws = std::max(eng->most_draws, ti.draw_end - ti.draw_begin);
    eng->frame_ext = any_stencil || any_source;
    eng->h_tileless.clear();
    for (uint32_t i = 0; i < n_draws; i++)
        if (!kind_has_tiles(f->draws[i].kind))
            eng->h_tileless.push_back(make_uint2(f->draws[i].target, i - tg[f->draws[i].target].draw_begin));
    if (set_coarse_chunks(eng))
        return -1;
    eng->frame_gen++;
    eng->gfail_key = 0;
    eng->resident = true;
    return 0;
}

// totals of the pipeline: read back (sizing run, cached in eng->sz[slot..]) or taken from the
// cache (capture run, which must not synchronise)
static int totals(vkh_engine *eng, bool cap, const std::vector<const uint32_t *> &srcs, uint32_t *dst,
                  uint32_t slot) {
    if (cap) {
        for (size_t i = 0; i < srcs.size(); i++)
            dst[i] = eng->sz[slot + i];
        return 0;
    }
    if (read_totals(eng, srcs, dst))
        return -1;
    for (size_t i = 0; i < srcs.size(); i++)
        eng->sz[slot + i] = dst[i];
    return 0;
}

static int run_pipeline(vkh_engine *eng, bool cap) {
    auto t_start = std::chrono::steady_clock::now();
    (void)hipSetDevice(eng->device);
    hipStream_t s = eng->stream;
    const uint32_t T = eng->frame_tile;
    vkh_stats st{};
    st.tile_size = T;
    st.fine_words = eng->fine_words;
    const uint32_t n_items = eng->n_items, n_sp = eng->n_sp, n_draws = eng->n_draws, n_units = eng->n_units;
    const uint32_t tile_total = eng->tile_total, bin_total = eng->bin_total;
    const uint32_t n_targets = eng->n_targets;
    if (!cap)
        CK(hipEventRecord(eng->ev_begin, s));
    uint8_t *db = (uint8_t *)eng->blob.p;
    const vkh_item *d_items = (const vkh_item *)(db + eng->o_items);
    const vkh_subpath *d_sp = (const vkh_subpath *)(db + eng->o_sp);
    const vkh_draw *d_draws = (const vkh_draw *)(db + eng->o_draws);
    const vkh_stroke_style *d_styles = (const vkh_stroke_style *)(db + eng->o_styles);
    const float *d_dashes = (const float *)(db + eng->o_dashes);
    const vkh_gradient *d_grads = (const vkh_gradient *)(db + eng->o_grads);
    const TargetInfo *d_tg = (const TargetInfo *)(db + eng->o_tg);
    const Unit *d_units = (const Unit *)(db + eng->o_units);

    // ---- flatten
    const uint32_t n_cubics = eng->n_cubic_items;
    if (!need(eng, cap, eng->item_cnt, 4 * (n_items + 1)) || !need(eng, cap, eng->item_off, 4 * (n_items + 1)) ||
        !need(eng, cap, eng->item_slot, 4 * (n_items + 1)) || !need(eng, cap, eng->slot_item, 4ull * (n_cubics + 1)) ||
        !need(eng, cap, eng->flat_scratch, 8ull * VKH_FLAT_SLOT * (n_cubics + 1)) ||
        !need(eng, cap, eng->flat_perm, 4 * (n_items + 1)) || !need(eng, cap, eng->flat_ovf, 4ull * (n_cubics + 1)) ||
        !need(eng, cap, eng->counters, 1024))
        return -1;
    // work order of the flattening: per-workgroup bucket counts, their scan, the scatter
    const uint32_t n_wh = 64u * vkh::flat_order_groups(n_items);
    if (!need(eng, cap, eng->flat_wh, 4ull * (2ull * n_wh + 2)))
        return -1;
    // [0] segments, [1] flatten slots, [2] overflow cubics, [3] long dashed pieces, [4] a dash walk
    // stopped at its cap bound
    uint32_t *ctr = P<uint32_t>(eng->counters);
    vkh::launch_zero_u32(s, ctr + 1, 4);
    if (n_items) {
        uint32_t *whist = P<uint32_t>(eng->flat_wh), *wcur = whist + n_wh + 1;
        vkh::launch_flat_hist(s, d_items, n_items, whist);
        if (scan_u32(eng, whist, n_wh, wcur, cap))
            return -1;
        vkh::launch_flat_scatter(s, d_items, n_items, wcur, P<uint32_t>(eng->flat_perm));
    }
    vkh::launch_flatten_stage(s, d_items, n_items, P<uint32_t>(eng->flat_perm), P<uint32_t>(eng->item_cnt), ctr + 1,
                              P<uint32_t>(eng->slot_item), P<uint32_t>(eng->item_slot), P<v2>(eng->flat_scratch),
                              ctr + 2, P<uint32_t>(eng->flat_ovf));
    if (scan_u32(eng, P<uint32_t>(eng->item_cnt), n_items, P<uint32_t>(eng->item_off), cap))
        return -1;
    uint32_t tot[8];
    if (totals(eng, cap, {P<uint32_t>(eng->item_off) + n_items, ctr + 2}, tot, 0))
        return -1;
    const uint32_t n_pts = tot[0], n_ovf = tot[1];
    st.points = n_pts;
    if (!need(eng, cap, eng->pts, 8ull * n_pts + 8) || !need(eng, cap, eng->pflags, n_pts + 8) ||
        !need(eng, cap, eng->sp_first, 4 * (n_sp + 1)) || !need(eng, cap, eng->sp_count, 4 * (n_sp + 1)))
        return -1;
    vkh::launch_flatten_copy(s, d_items, n_cubics, P<uint32_t>(eng->slot_item), P<uint32_t>(eng->item_cnt),
                             P<uint32_t>(eng->item_off), P<v2>(eng->flat_scratch), P<v2>(eng->pts),
                             P<uint8_t>(eng->pflags));
    vkh::launch_flatten_lines(s, d_items, n_items, P<uint32_t>(eng->item_off), P<v2>(eng->pts),
                              P<uint8_t>(eng->pflags));
    vkh::launch_flatten_overflow(s, d_items, P<uint32_t>(eng->flat_ovf), n_ovf, P<uint32_t>(eng->item_off),
                                 P<v2>(eng->pts), P<uint8_t>(eng->pflags));
    vkh::launch_subpath_ranges(s, d_sp, n_sp, P<uint32_t>(eng->item_off), P<uint32_t>(eng->sp_first),
                               P<uint32_t>(eng->sp_count));
    CK(hipGetLastError());

    if (!cap)
        CK(hipEventRecord(eng->ev_stage[0], s)); // flatten done
    // ---- per-unit geometry: fill edge counts and stroke work steps
    size_t un4 = 4ull * (n_units + 1);
    if (!need(eng, cap, eng->fe, un4) || !need(eng, cap, eng->nv, un4) || !need(eng, cap, eng->fe_off, un4) ||
        !need(eng, cap, eng->nv_off, un4) || !need(eng, cap, eng->u_toff, un4) || !need(eng, cap, eng->u_vbeg, un4) ||
        !need(eng, cap, eng->u_vend, un4))
        return -1;
    uint32_t *d_steps = P<uint32_t>(eng->nv), *d_step_off = P<uint32_t>(eng->nv_off);
    vkh::launch_unit_count(s, d_units, n_units, d_draws, d_sp, P<uint32_t>(eng->sp_count), d_styles,
                           P<uint32_t>(eng->fe), d_steps);
    if (scan_u32(eng, P<uint32_t>(eng->fe), n_units, P<uint32_t>(eng->fe_off), cap) ||
        scan_u32(eng, d_steps, n_units, d_step_off, cap))
        return -1;
    if (totals(eng, cap, {P<uint32_t>(eng->fe_off) + n_units, d_step_off + n_units}, tot, 2))
        return -1;
    const uint32_t n_fill_edges = tot[0], n_steps = tot[1];
    size_t sn4 = 4ull * (n_steps + 1);
    if (!need(eng, cap, eng->step_nv, sn4) || !need(eng, cap, eng->step_nt, sn4) || !need(eng, cap, eng->sv_off, sn4) ||
        !need(eng, cap, eng->st_off, sn4) || !need(eng, cap, eng->dash_st, 2 * sn4) ||
        !need(eng, cap, eng->dash_list, 4 * sn4))
        return -1;
    vkh::launch_stroke_steps(s, false, d_units, n_units, d_step_off, n_steps, d_draws, d_sp, P<uint32_t>(eng->sp_first),
                             P<uint32_t>(eng->sp_count), P<v2>(eng->pts), P<uint8_t>(eng->pflags), d_styles, d_dashes,
                             P<uint32_t>(eng->step_nv), P<uint32_t>(eng->step_nt), nullptr, nullptr, nullptr, nullptr,
                             P<uint2>(eng->dash_st), P<uint4>(eng->dash_list), ctr + 3);
    if (scan_u32(eng, P<uint32_t>(eng->step_nv), n_steps, P<uint32_t>(eng->sv_off), cap) ||
        scan_u32(eng, P<uint32_t>(eng->step_nt), n_steps, P<uint32_t>(eng->st_off), cap))
        return -1;
    if (totals(eng, cap, {P<uint32_t>(eng->sv_off) + n_steps, P<uint32_t>(eng->st_off) + n_steps, ctr + 4}, tot, 4))
        return -1;
    if (tot[2]) {
        eng->err = "dashed segment too long for its dash pattern (the dash loop would not end)";
        return -3;
    }
    const uint32_t n_verts = tot[0], n_tris = tot[1];
    const uint64_t n_edges64 = (uint64_t)n_fill_edges + 3ull * n_tris;
    if (n_edges64 > 0x7FFFFFFFull) {
        eng->err = "frame too large (edges)";
        return -3;
    }
    const uint32_t n_edges = (uint32_t)n_edges64;
    st.stroke_vertices = n_verts;
    st.stroke_triangles = n_tris;
    st.edges = n_edges;
    if (!need(eng, cap, eng->verts, 8ull * n_verts + 8) || !need(eng, cap, eng->tris, 12ull * n_tris + 12) ||
        !need(eng, cap, eng->edges, sizeof(Edge) * (size_t)n_edges + 16))
        return -1;
    vkh::launch_stroke_steps(s, true, d_units, n_units, d_step_off, n_steps, d_draws, d_sp, P<uint32_t>(eng->sp_first),
                             P<uint32_t>(eng->sp_count), P<v2>(eng->pts), P<uint8_t>(eng->pflags), d_styles, d_dashes,
                             nullptr, nullptr, P<uint32_t>(eng->sv_off), P<uint32_t>(eng->st_off), P<v2>(eng->verts),
                             P<uint32_t>(eng->tris), P<uint2>(eng->dash_st), P<uint4>(eng->dash_list), ctr + 3);
    vkh::launch_stroke_patch(s, d_units, n_units, d_step_off, d_draws, d_sp, P<uint32_t>(eng->sp_first),
                             P<uint32_t>(eng->sp_count), P<v2>(eng->pts), P<uint8_t>(eng->pflags), d_styles, d_dashes,
                             P<uint32_t>(eng->sv_off), P<uint32_t>(eng->st_off), P<uint32_t>(eng->tris));
    // the edge kernels also reduce each draw's edge bbox and edge index range (a stroke draw's range
    // is opened by the unit ranges and appended to by the triangle edges)
    size_t dn4 = 4ull * (n_draws + 1);
    if (!need(eng, cap, eng->bbox, sizeof(int4) * (n_draws + 1)) || !need(eng, cap, eng->dtiles, sizeof(DrawTiles) * (n_draws + 1)) ||
        !need(eng, cap, eng->erange, sizeof(uint2) * (n_draws + 1)) || !need(eng, cap, eng->pt_count, dn4) ||
        !need(eng, cap, eng->pt_off, dn4) || !need(eng, cap, eng->counters, 256))
        return -1;
    vkh::launch_draw_bbox_init(s, P<int4>(eng->bbox), P<uint2>(eng->erange), n_draws);
    vkh::launch_unit_ranges(s, d_units, n_units, d_draws, P<uint32_t>(eng->fe_off), d_step_off, P<uint32_t>(eng->sv_off),
                            P<uint32_t>(eng->st_off), P<uint32_t>(eng->u_toff), P<uint32_t>(eng->u_vbeg),
                            P<uint32_t>(eng->u_vend), n_fill_edges, P<uint2>(eng->erange));
    vkh::launch_fill_edges(s, d_units, n_units, P<uint32_t>(eng->fe_off), n_fill_edges, d_draws, d_tg,
                           P<uint32_t>(eng->sp_first), P<uint32_t>(eng->sp_count), P<v2>(eng->pts), P<Edge>(eng->edges),
                           P<int4>(eng->bbox), P<uint2>(eng->erange));
    vkh::launch_tri_edges(s, d_units, n_units, P<uint32_t>(eng->u_toff), P<uint32_t>(eng->u_vbeg),
                          P<uint32_t>(eng->u_vend), n_tris, d_draws, d_tg, P<v2>(eng->verts), P<uint32_t>(eng->tris),
                          P<Edge>(eng->edges), P<int4>(eng->bbox), P<uint2>(eng->erange));
    CK(hipGetLastError());

    if (!cap)
        CK(hipEventRecord(eng->ev_stage[1], s)); // geometry done
    // ---- tiles of every draw
    vkh::launch_draw_tiles(s, d_draws, n_draws, d_tg, P<int4>(eng->bbox), T, P<DrawTiles>(eng->dtiles),
                           P<uint32_t>(eng->pt_count));
    // draw records, gradient records (one per draw), then one GradPrep per gradient
    if (!need(eng, cap, eng->dprep,
              (sizeof(DrawPrep) + sizeof(GradRec)) * (n_draws + 1) + sizeof(GradPrep) * (eng->n_grads + 1)))
        return -1;
    GradRec *d_grec = (GradRec *)(P<DrawPrep>(eng->dprep) + n_draws + 1);
    GradPrep *d_gprep = (GradPrep *)(d_grec + n_draws + 1);
    vkh::launch_draw_prep(s, d_draws, n_draws, d_tg, d_grads, eng->n_grads,
                          (const vkh_surface_source *)(db + eng->o_srcs), P<DrawPrep>(eng->dprep), d_grec,
                          d_gprep);
    if (scan_u32(eng, P<uint32_t>(eng->pt_count), n_draws, P<uint32_t>(eng->pt_off), cap))
        return -1;
    if (totals(eng, cap, {P<uint32_t>(eng->pt_off) + n_draws}, tot, 7))
        return -1;
    const uint32_t n_ptiles = tot[0];
    st.path_tiles = n_ptiles;
    vkh::launch_set_pt_base(s, P<DrawTiles>(eng->dtiles), P<uint32_t>(eng->pt_off), n_draws);
    if (!need(eng, cap, eng->ptiles, sizeof(PathTile) * (size_t)n_ptiles + 16) || !need(eng, cap, eng->segs, 4096))
        return -1;
    if (!cap)
        CK(hipEventRecord(eng->ev_stage[2], s)); // tile setup done
    // ---- binning: per-draw segment lists and backdrops; bins again once if the segment buffer
    // was too small (the allocator then holds the exact total)
    uint32_t n_segs = 0;
    if (!need(eng, cap, eng->draw_order, 4ull * vkh::order_words(n_draws)))
        return -1;
    vkh::launch_draw_order(s, n_draws, P<uint2>(eng->erange), P<uint32_t>(eng->draw_order));
    // the coarse pass's command counts per (tile, chunk): seeded with the draws without path tiles,
    // incremented by binning for every path tile that holds a command
    const uint32_t nch = eng->coarse_nch;
    const uint64_t n_tc64 = (uint64_t)tile_total * nch;
    if (n_tc64 > 0x7FFFFFFFull) {
        eng->err = "frame too large (tiles x draw chunks)";
        return -3;
    }
    const uint32_t n_tc = (uint32_t)n_tc64;
    if (!need(eng, cap, eng->tile_cnt, 4ull * (n_tc + 1)) || !need(eng, cap, eng->tile_off, 4ull * (n_tc + 1)))
        return -1;
    for (int pass = 0; pass < 2; pass++) {
        const uint32_t seg_cap = (uint32_t)std::min<size_t>(eng->segs.cap / sizeof(Edge), 0xFFFFFFFFull);
        vkh::launch_zero_u32(s, P<uint32_t>(eng->counters), 1);
        vkh::launch_tile_cnt_init(s, d_tg, n_targets, nch, P<uint32_t>(eng->tl_table), P<uint32_t>(eng->tile_cnt), n_tc);
        vkh::launch_bin_draws(s, P<Edge>(eng->edges), P<uint2>(eng->erange),
                              P<DrawTiles>(eng->dtiles), n_draws, T, P<PathTile>(eng->ptiles), P<Edge>(eng->segs),
                              seg_cap, P<uint32_t>(eng->counters), P<uint32_t>(eng->draw_order), d_draws, d_tg, nch,
                              eng->coarse_chunk, P<uint32_t>(eng->tile_cnt));
        CK(hipGetLastError());
        if (totals(eng, cap, {P<uint32_t>(eng->counters)}, tot, 8))
            return -1;
        n_segs = tot[0];
        if (n_segs <= seg_cap)
            break;
        if (pass == 1) {
            eng->err = "segment buffer overflow after resize";
            return -1;
        }
        if (!need(eng, cap, eng->segs, sizeof(Edge) * (size_t)n_segs + sizeof(Edge)))
            return -1;
    }
    st.segments = n_segs;

    if (!cap)
        CK(hipEventRecord(eng->ev_stage[3], s)); // binning done
    // ---- coarse: ordered command list per tile (the counts came from binning)
    if (scan_u32(eng, P<uint32_t>(eng->tile_cnt), n_tc, P<uint32_t>(eng->tile_off), cap))
        return -1;
    if (totals(eng, cap, {P<uint32_t>(eng->tile_off) + n_tc}, tot, 9))
        return -1;
    const uint32_t n_cmds = tot[0];
    st.commands = n_cmds;
    st.tiles = tile_total;
    if (!need(eng, cap, eng->cmds, sizeof(Cmd) * (size_t)n_cmds + 16))
        return -1;
    vkh::launch_coarse(s, true, bin_total, nch, eng->coarse_chunk, d_tg, n_targets, d_draws, P<DrawTiles>(eng->dtiles),
                       P<PathTile>(eng->ptiles), nullptr, P<uint32_t>(eng->tile_off), P<Cmd>(eng->cmds));
    if (!need(eng, cap, eng->tile_order, 4ull * vkh::order_words(tile_total)))
        return -1;
    vkh::launch_tile_order(s, tile_total, nch, P<uint32_t>(eng->tile_off), P<uint32_t>(eng->tile_order));
    CK(hipGetLastError());

    // ---- fine: coverage + compositing
    if (!cap)
        CK(hipEventRecord(eng->ev_stage[4], s)); // coarse done
    if (!cap)
        CK(hipEventRecord(eng->ev0, s));
    vkh::launch_fine(s, eng->frame_ext, T, eng->fine_words, tile_total, d_tg, n_targets, P<DrawPrep>(eng->dprep), d_grec, d_grads, d_gprep,
                     P<PathTile>(eng->ptiles),
                     P<Edge>(eng->segs), nch, P<uint32_t>(eng->tile_off),
                     P<Cmd>(eng->cmds), P<uint32_t>(eng->tile_order));
    CK(hipGetLastError());
    if (cap)
        return 0;
    CK(hipEventRecord(eng->ev1, s));
    CK(hipEventRecord(eng->ev_end, s));
    CK(hipStreamSynchronize(s));
    float ms = 0;
    (void)hipEventElapsedTime(&ms, eng->ev0, eng->ev1);
    st.ms_fine = ms;
    ms = 0;
    (void)hipEventElapsedTime(&ms, eng->ev_begin, eng->ev_end);
    st.ms_device = ms;
    hipEvent_t marks[7] = {eng->ev_begin, eng->ev_stage[0], eng->ev_stage[1], eng->ev_stage[2],
                           eng->ev_stage[3], eng->ev_stage[4], eng->ev_end};
    for (int k = 0; k < 6; k++) {
        ms = 0;
        (void)hipEventElapsedTime(&ms, marks[k], marks[k + 1]);
        st.ms_stage[k] = ms;
    }
    st.ms_total = (float)std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_start).count();
    eng->stats = st;
    return 0;
}

static uint64_t graph_key(const vkh_engine *eng) {
    return (eng->frame_gen * 1000003ull + eng->alloc_gen) * 64ull + eng->frame_tile * 2ull + (eng->fine_words >> 11);
}

static void drop_graph(vkh_engine *eng) {
    if (eng->gexec)
        (void)hipGraphExecDestroy(eng->gexec);
    if (eng->graph)
        (void)hipGraphDestroy(eng->graph);
    eng->gexec = nullptr;
    eng->graph = nullptr;
    eng->gkey = 0;
}

// Captures the resident frame's launch sequence (sizes from the sizing run just done) into a
// graph; on any failure the frame keeps running eagerly.
static void capture_graph(vkh_engine *eng) {
    drop_graph(eng);
    hipStream_t s = eng->stream;
    const uint64_t key = graph_key(eng);
    if (hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal) != hipSuccess) {
        eng->gfail_key = key;
        (void)hipGetLastError();
        return;
    }
    std::string keep_err = eng->err;
    int rc = run_pipeline(eng, true);
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(s, &g);
    if (rc == 0 && e == hipSuccess && g && hipGraphInstantiate(&eng->gexec, g, nullptr, nullptr, 0) == hipSuccess) {
        eng->graph = g;
        eng->gkey = graph_key(eng);
    } else {
        if (g)
            (void)hipGraphDestroy(g);
        eng->gexec = nullptr;
        eng->gfail_key = key;
        (void)hipGetLastError();
    }
    eng->err = keep_err;
}

int vkvg_hip_execute(vkh_engine *eng) {
    if (!eng || !eng->resident) {
        if (eng)
            eng->err = "no resident frame";
        return -1;
    }
    (void)hipSetDevice(eng->device);
    hipStream_t s = eng->stream;
    if (eng->graph_on && eng->gexec && eng->gkey == graph_key(eng)) {
        auto t0 = std::chrono::steady_clock::now();
        CK(hipEventRecord(eng->ev_begin, s));
        CK(hipGraphLaunch(eng->gexec, s));
        CK(hipEventRecord(eng->ev_end, s));
        CK(hipStreamSynchronize(s));
        vkh_stats st = eng->stats;
        float ms = 0;
        (void)hipEventElapsedTime(&ms, eng->ev_begin, eng->ev_end);
        st.ms_device = ms;
        st.ms_fine = 0;
        for (float &x : st.ms_stage)
            x = 0;
        st.ms_total = (float)std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        st.graph = 1;
        eng->stats = st;
        return 0;
    }
    int r = run_pipeline(eng, false);
    if (r == 0 && eng->graph_on && eng->gfail_key != graph_key(eng))
        capture_graph(eng);
    return r;
}



int vkvg_hip_retile(vkh_engine *eng, uint32_t T) {
    if (!eng || !eng->resident) {
        if (eng)
            eng->err = "no resident frame";
        return -1;
    }
    if (T != 8 && T != 16 && T != 32)
        return -1;
    if (T == eng->frame_tile)
        return 0;
    (void)hipSetDevice(eng->device);
    hipStream_t s = eng->stream;
    CK(hipStreamSynchronize(s));
    uint32_t tile_total = 0, bin_total = 0;
    for (TargetInfo &ti : eng->h_tg) {
        ti.tilesW = (ti.width + T - 1) / T;
        ti.tilesH = (ti.height + T - 1) / T;
        ti.tile_base = tile_total;
        tile_total += ti.tilesW * ti.tilesH;
        ti.binsW = (ti.tilesW + VKH_BIN_TILES - 1) / VKH_BIN_TILES;
        ti.binsH = (ti.tilesH + VKH_BIN_TILES - 1) / VKH_BIN_TILES;
        ti.bin_base = bin_total;
        bin_total += ti.binsW * ti.binsH;
    }
    uint8_t *hb = (uint8_t *)eng->hblob.p;
    const size_t bytes = sizeof(TargetInfo) * eng->h_tg.size();
    memcpy(hb + eng->o_tg, eng->h_tg.data(), bytes);
    CK(hipMemcpyAsync((uint8_t *)eng->blob.p + eng->o_tg, hb + eng->o_tg, bytes, hipMemcpyHostToDevice, s));
    CK(hipStreamSynchronize(s));
    eng->tile_total = tile_total;
    eng->bin_total = bin_total;
    eng->frame_tile = T;
    if (set_coarse_chunks(eng))
        return -1;
    eng->frame_gen++;
    eng->gfail_key = 0;
    return 0;
}

int vkvg_hip_autotune(vkh_engine *eng, uint32_t reps, vkh_tune *out) {
    if (!eng || !eng->resident) {
        if (eng)
            eng->err = "no resident frame";
        return -1;
    }
    if (reps == 0)
        reps = 3;
    const bool graph = eng->graph_on;
    eng->graph_on = false;
    static const uint32_t tiles[3] = {8, 16, 32}, words[3] = {2048, 4096, 8192};
    vkh_tune res;
    memset(&res, 0, sizeof(res));
    float best = 1e30f;
    for (uint32_t T : tiles) {
        if (vkvg_hip_retile(eng, T))
            return -1;
        for (uint32_t w : words) {
            eng->fine_words = w;
            if (vkvg_hip_execute(eng)) // sizing run, also a warm-up
                return -1;
            float m = 1e30f;
            for (uint32_t r = 0; r < reps; r++) {
                if (vkvg_hip_execute(eng))
                    return -1;
                m = std::min(m, eng->stats.ms_device);
            }
            res.e[res.n] = vkh_tune_entry{T, w, m};
            if (m < best) {
                best = m;
                res.best = res.n;
            }
            res.n++;
        }
    }
    eng->fine_words = res.e[res.best].fine_words;
    if (vkvg_hip_retile(eng, res.e[res.best].tile))
        return -1;
    eng->tile = res.e[res.best].tile;
    eng->graph_on = graph;
    if (out)
        *out = res;
    return 0;
}

int vkvg_hip_render(vkh_engine *eng, const vkh_frame *f) {
    int r = vkvg_hip_upload(eng, f);
    if (r)
        return r;
    return vkvg_hip_execute(eng);
}

// Diagnostics, not part of the public headers (0 expected, -1 on error): mismatches of the hardware
// UNORM8 conversion used by the fine kernel against the reference rounding, over every float; and
// of the device form of the packed blend against the per-channel blend, over every input.
static long long run_check(void (*launch)(hipStream_t, unsigned long long *)) {
    unsigned long long *d = nullptr, h = 0;
    if (hipMalloc(&d, sizeof(*d)) != hipSuccess)
        return -1;
    long long r = -1;
    if (hipMemset(d, 0, sizeof(*d)) == hipSuccess) {
        launch(nullptr, d);
        if (hipGetLastError() == hipSuccess && hipDeviceSynchronize() == hipSuccess &&
            hipMemcpy(&h, d, sizeof(h), hipMemcpyDeviceToHost) == hipSuccess)
            r = (long long)h;
    }
    (void)hipFree(d);
    return r;
}
__attribute__((visibility("default"))) long long vkh_check_u8_conversion(void) {
    return run_check(vkh::launch_check_u8_conversion);
}
__attribute__((visibility("default"))) long long vkh_check_blend(void) {
    return run_check(vkh::launch_check_blend);
}

} // extern "C"
