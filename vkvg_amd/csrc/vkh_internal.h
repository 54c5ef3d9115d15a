// Internal types shared by the vkvg HIP backend translation units.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define VKH_NO_DRAW 0xFFFFFFFFu

#include "vkh_device.h"

#define VKH_BIN_TILES 16        // a coarse bin is 16 x 16 tiles
#define VKH_COARSE_CHUNK 1024   // most draws per coarse workgroup
#define VKH_FLAT_SLOT 128       // flattened points kept per cubic by the subdivision pass

// one (draw, subpath) pair of the frame; fills and strokes are expanded per unit
struct Unit {
    uint32_t draw;
    uint32_t subpath;
};

// per render target, on the device
struct TargetInfo {
    uint32_t width, height;
    uint32_t tilesW, tilesH;
    uint32_t tile_base;  // first global tile index
    uint32_t binsW, binsH;
    uint32_t bin_base;   // first global bin index
    uint32_t draw_begin, draw_end;
    uint8_t *rgba;
    uint32_t *sig;
    uint32_t *stencil;      // stencil words (byte 0: FILL, CLIP, saved CLIP bits; bytes 1..3: copies)
    uint32_t stencil_load;  // 1: start from the plane, 0: from zero
    uint32_t pad;
};

// draw kinds with path tiles (edges binned into tiles); clip draws take every tile of the target
__host__ __device__ __forceinline__ bool kind_has_tiles(uint32_t kind) {
    return kind <= VKH_DRAW_STROKE || kind == VKH_DRAW_CLIP_NZ || kind == VKH_DRAW_CLIP_EO;
}
__host__ __device__ __forceinline__ bool kind_is_clip(uint32_t kind) {
    return kind == VKH_DRAW_CLIP_NZ || kind == VKH_DRAW_CLIP_EO;
}

// per draw, computed on the device
struct DrawTiles {
    int32_t bx0, by0, bx1, by1; // inclusive tile bbox within the target, empty when bx0 > bx1
    uint32_t pt_base;           // first path tile
    uint32_t bw;                // bbox width in tiles
};

// per path tile (draw x tile in its bbox)
struct PathTile {
    int32_t backdrop;  // winding at the tile's first sample (after the row prefix sum)
    uint32_t nseg;     // edges touching the tile
    uint32_t seg_off;  // into the segment index array
    uint32_t reserved; // keeps the record at 16 bytes
};

// per draw, computed once per frame: what the compositing loop needs, with the per-draw parts of
// the source evaluated once (same float operations as per pixel, so the values are identical)
enum : uint32_t {
    VKH_DP_GRAD = 1u << 9,     // meta: gradient source
    VKH_DP_SURF = 1u << 10,    // meta: surface source (the draw's GradRec slot holds a SurfRec)
    VKH_DP_HASSEG = 1u << 8,   // meta (set per command in the fine kernel): the path tile has segments
    VKH_GF_UY0 = 1u,           // linear: direction has no y component
    VKH_GF_UXNEG = 2u,         // linear: ... and points to -x
    VKH_GF_UYNEG = 4u,         // linear: y component < 0
};
struct DrawPrep {
    uint32_t meta;   // kind | op << 4 | VKH_DP_GRAD
    uint32_t solid;  // premultiplied RGBA8 of a solid source, opacity applied
    uint32_t key;    // coverage-signature key of the draw
    int32_t grad;    // gradient index or -1
    float opacity;
    uint32_t gflags; // VKH_GF_*
    float g[6];      // gradient invariants: linear p0x, l, m, bb, sqrt(1+m*m); radial c0x, c0y, r0, dx, dy, cc
    float rg[2];     // linear: vkh_rcp_for_div of g[1] and g[4] (0: divide per pixel)
    uint32_t pad[2];
};

// Per gradient: rspan[i] = vkh_rcp_for_div(stops[i] - stops[i-1]) for the colour step i >= 1
struct GradPrep {
    float rspan[16];
};

// Per gradient draw: everything the per-pixel evaluation reads, in one contiguous 160-byte record
// that the fine kernel reads with independent scalar loads (no chain of dependent loads through
// the draw and gradient tables). Stops beyond the first VKH_GR_INLINE come from the gradient table.
#define VKH_GR_INLINE 4
struct GradRec {
    uint32_t type, count, gflags;
    float opacity;
    float g[6], rg[2];                   // as in DrawPrep
    float stops[VKH_GR_INLINE];
    float rspan[VKH_GR_INLINE];          // as in GradPrep
    float colors[VKH_GR_INLINE][4];
    int32_t grad;
    uint32_t pad[3];
};
static_assert(sizeof(GradRec) == 160, "GradRec layout");

// Per surface-source draw, in the draw's GradRec slot: the source image and the sampler, the
// push constants the fragment shader reads (source rectangle, inverse matrix) and the opacity
struct SurfRec {
    const uint32_t *img;
    uint32_t width, height;
    uint32_t extend;      // vkvg_extend_t
    uint32_t linear;      // 1: LINEAR filtering, 0: NEAREST
    float rect[4];
    float m[6];
    float opacity;
    uint32_t pad[23];
};
static_assert(sizeof(SurfRec) == sizeof(GradRec), "SurfRec shares the GradRec slot");

struct Cmd {
    uint32_t draw;
    uint32_t pt; // path tile, VKH_NO_DRAW for paint / clear
};

namespace vkh {
size_t scan_temp_words(uint32_t n);
void scan_u32(hipStream_t s, const uint32_t *in, uint32_t n, uint32_t *out, uint32_t *tmp);
uint32_t flat_order_groups(uint32_t n); // workgroups of the flattening order's histogram and scatter
void launch_flat_hist(hipStream_t s, const vkh_item *items, uint32_t n, uint32_t *whist);
void launch_flat_scatter(hipStream_t s, const vkh_item *items, uint32_t n, const uint32_t *wcur, uint32_t *perm);
void launch_flatten_stage(hipStream_t s, const vkh_item *items, uint32_t n, const uint32_t *perm, uint32_t *counts,
                          uint32_t *slot_alloc, uint32_t *slot_item, uint32_t *item_slot, v2 *scratch,
                          uint32_t *ovf_count, uint32_t *ovf_list);
void launch_flatten_copy(hipStream_t s, const vkh_item *items, uint32_t n_slots, const uint32_t *slot_item,
                         const uint32_t *counts, const uint32_t *offs, const v2 *scratch, v2 *pts, uint8_t *pf);
void launch_flatten_lines(hipStream_t s, const vkh_item *items, uint32_t n, const uint32_t *offs, v2 *pts,
                          uint8_t *pf);
void launch_flatten_overflow(hipStream_t s, const vkh_item *items, const uint32_t *list, uint32_t n_list,
                             const uint32_t *offs, v2 *pts, uint8_t *pf);
void launch_subpath_ranges(hipStream_t s, const vkh_subpath *sp, uint32_t n, const uint32_t *item_off,
                           uint32_t *first, uint32_t *count);
void launch_unit_count(hipStream_t s, const Unit *units, uint32_t n_units, const vkh_draw *draws,
                       const vkh_subpath *sps, const uint32_t *spc, const vkh_stroke_style *styles, uint32_t *fe,
                       uint32_t *steps);
void launch_stroke_steps(hipStream_t s, bool emit, const Unit *units, uint32_t n_units, const uint32_t *step_off,
                         uint32_t n_steps, const vkh_draw *draws, const vkh_subpath *sps, const uint32_t *spf,
                         const uint32_t *spc, const v2 *pts, const uint8_t *pf, const vkh_stroke_style *styles,
                         const float *dashes, uint32_t *step_nv, uint32_t *step_nt, const uint32_t *sv_off,
                         const uint32_t *st_off, v2 *verts, uint32_t *tris, uint2 *dash_st, uint4 *dlist,
                         uint32_t *dcount);
void launch_stroke_patch(hipStream_t s, const Unit *units, uint32_t n_units, const uint32_t *step_off,
                         const vkh_draw *draws, const vkh_subpath *sps, const uint32_t *spf, const uint32_t *spc,
                         const v2 *pts, const uint8_t *pf, const vkh_stroke_style *styles, const float *dashes,
                         const uint32_t *sv_off, const uint32_t *st_off, uint32_t *tris);
void launch_unit_ranges(hipStream_t s, const Unit *units, uint32_t n_units, const vkh_draw *draws,
                        const uint32_t *fill_off, const uint32_t *step_off, const uint32_t *sv_off,
                        const uint32_t *st_off, uint32_t *u_toff, uint32_t *u_vbeg, uint32_t *u_vend, uint32_t ebase,
                        uint2 *erange);
void launch_fill_edges(hipStream_t s, const Unit *units, uint32_t n_units, const uint32_t *fill_off, uint32_t total,
                       const vkh_draw *draws, const TargetInfo *tg, const uint32_t *spf, const uint32_t *spc,
                       const v2 *pts, Edge *edges, int4 *bbox, uint2 *erange);
void launch_tri_edges(hipStream_t s, const Unit *units, uint32_t n_units, const uint32_t *u_toff,
                      const uint32_t *u_vbeg, const uint32_t *u_vend, uint32_t total_tris, const vkh_draw *draws,
                      const TargetInfo *tg, const v2 *verts, const uint32_t *tris, Edge *edges, int4 *bbox,
                      uint2 *erange);

// raster stage (vkh_raster.hip)
void launch_draw_bbox_init(hipStream_t s, int4 *bbox, uint2 *erange, uint32_t n_draws);
void launch_draw_tiles(hipStream_t s, const vkh_draw *draws, uint32_t n_draws, const TargetInfo *tg,
                       const int4 *bbox, uint32_t tile, DrawTiles *dt, uint32_t *pt_count);
void launch_set_pt_base(hipStream_t s, DrawTiles *dt, const uint32_t *pt_off, uint32_t n_draws);
void launch_zero_u32(hipStream_t s, uint32_t *p, uint32_t n);
void launch_bin_draws(hipStream_t s, const Edge *edges, const uint2 *erange,
                      const DrawTiles *dt, uint32_t n_draws, uint32_t tile, PathTile *pts, Edge *segs,
                      uint32_t seg_cap, uint32_t *seg_alloc, const uint32_t *order, const vkh_draw *draws,
                      const TargetInfo *tg, uint32_t nch, uint32_t chunk_draws, uint32_t *tile_cnt);
void launch_tile_cnt_init(hipStream_t s, const TargetInfo *tg, uint32_t n_targets, uint32_t nch,
                          const uint32_t *table, uint32_t *tile_cnt, uint32_t n_tc);
void launch_draw_prep(hipStream_t s, const vkh_draw *draws, uint32_t n_draws, const TargetInfo *tg,
                      const vkh_gradient *grads, uint32_t n_grads, const vkh_surface_source *srcs, DrawPrep *out,
                      GradRec *rec, GradPrep *gout);
// draws per coarse workgroup: about 1024 workgroups over n_bins bins, 64 .. VKH_COARSE_CHUNK draws
uint32_t coarse_chunk(uint32_t most_draws, uint32_t n_bins);
void launch_coarse(hipStream_t s, bool write, uint32_t n_bins, uint32_t nch, uint32_t chunk, const TargetInfo *tg,
                   uint32_t n_targets, const vkh_draw *draws, const DrawTiles *dt, const PathTile *pts,
                   uint32_t *tile_count, const uint32_t *tile_off, Cmd *cmds);
// ext: the frame has surface sources or a target with a stencil plane (the extended variant runs)
void launch_fine(hipStream_t s, bool ext, uint32_t tile, uint32_t words, uint32_t n_tiles, const TargetInfo *tg,
                 uint32_t n_targets, const DrawPrep *dprep, const GradRec *grec, const vkh_gradient *grads,
                 const GradPrep *gprep, const PathTile *pts, const Edge *segs, uint32_t nch,
                 const uint32_t *tile_off, const Cmd *cmds, const uint32_t *order);
// heaviest first: tiles by command count (fine kernel), draws by edge count (binning kernel); work
// holds order_words(n) words, the order in its first n
size_t order_words(uint32_t n);
void launch_tile_order(hipStream_t s, uint32_t n_tiles, uint32_t nch, const uint32_t *tile_off, uint32_t *work);
void launch_draw_order(hipStream_t s, uint32_t n_draws, const uint2 *erange, uint32_t *work);
void launch_unpremultiply(hipStream_t s, const uint8_t *src, uint8_t *dst, size_t npix);
void launch_stencil_clear(hipStream_t s, uint32_t *words, size_t npix);
void launch_check_u8_conversion(hipStream_t s, unsigned long long *bad);
void launch_check_blend(hipStream_t s, unsigned long long *bad);
} // namespace vkh
