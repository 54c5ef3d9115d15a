/*
 * vkvg_hip.h — the thin C ABI between the C host recorder (libvkvg's vkvg_* entry points)
 * and the gfx950 HIP rendering backend.
 *
 * The host recorder turns vkvg path/paint calls into a *frame*: flat arrays of path items
 * (explicit points and unflattened cubics), subpaths, draws, stroke styles, dash arrays,
 * gradients and render targets. vkvg_hip_render() uploads a frame and runs the kernel
 * pipeline  flatten -> stroke-expand -> edges -> tile binning -> coarse -> fine
 * (integer winding + source-over compositing) into the targets' RGBA8 images in HBM.
 *
 * In the reference the same work is split between host geometry
 * (src/vkvg_context_internal.c:924-1655) and Vulkan draw submission
 * (src/vkvg_context_internal.c:421-603, vkQueueSubmit + fence wait); this header replaces
 * that Vulkan submission boundary. Everything here is plain C: no torch or HIP types.
 */
#ifndef VKVG_HIP_H
#define VKVG_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#ifndef vkvg_public
#define vkvg_public __attribute__((visibility("default")))
#endif

/* ---- path items: one explicit point, one cubic to be flattened on the GPU, or the interior
 * points of an arc, evaluated on the GPU with the reference's sequential angle recurrence */
#define VKH_ITEM_CUBIC  1u /* p[0..7] = current point, c1, c2, end; flattened with AGG */
#define VKH_ITEM_CURVED 2u /* the item's points lie in a curve segment (stroke joins forced) */
#define VKH_ITEM_NO_END 4u /* cubic whose end point was removed by close_path */
/* n = flags >> VKH_ITEM_COUNT_SHIFT points of a circular arc (src/vkvg_context.c:394-503):
 * p[0] xc, p[1] yc, p[2] radius, p[3] first angle a, p[4] signed step s; point k at the angle
 * a + s + ... + s (k float additions): (cos(a)*r + xc, sin(a)*r + yc) */
#define VKH_ITEM_ARC    8u
/* n points of an elliptic arc (src/vkvg_context_internal.c:1539-1580): p[0..1] the first angle
 * theta as a double, p[2] signed step s (float, added to the double theta), p[3] rx, p[4] ry,
 * p[5] cos(phi), p[6] sin(phi), p[7] and tol the centre */
#define VKH_ITEM_EARC   16u
#define VKH_ITEM_COUNT_SHIFT 8
#define VKH_ITEM_MAX_COUNT ((1u << (32 - VKH_ITEM_COUNT_SHIFT)) - 1u)

typedef struct {
    float    p[8];  /* point: p[0], p[1]; cubic: x0 y0 x1 y1 x2 y2 x3 y3 */
    uint32_t flags; /* VKH_ITEM_* */
    float    tol;   /* cubic: squared distance tolerance 0.25/max(sx,sy) (src/vkvg_context.c:556-558) */
} vkh_item;

#define VKH_SP_CLOSED 1u

typedef struct {
    uint32_t first_item, n_items;
    uint32_t flags; /* VKH_SP_CLOSED */
    uint32_t reserved;
} vkh_subpath;

/* draw kinds: how the per-pixel integer winding w of the draw's edges becomes the number k
 * of source-over blends at that pixel */
enum {
    VKH_DRAW_FILL_NZ = 0, /* k = (w != 0)                       src/vkvg_context_internal.c:1720-1793 */
    VKH_DRAW_FILL_EO = 1, /* k = w & 1 (fan INVERT + cover)     src/vkvg_context_internal.c:1583-1655 */
    VKH_DRAW_STROKE = 2,  /* k = w = covering triangles         src/vkvg_context.c:822-948 */
    VKH_DRAW_PAINT = 3,   /* k = 1 on every pixel               src/vkvg_context.c:990-1003 */
    VKH_DRAW_CLEAR = 4,   /* colour := 0, stencil := 0 on every pixel    src/vkvg_context.c:734-753 */
    /* Stencil draws. A target with a stencil plane (vkh_target.stencil) keeps per pixel the CLIP
     * bit (bit 1: draws pass where it is 0), six saved CLIP bits (bits 2..7) and three copies of
     * that byte (bytes 1..3) for save levels beyond six. They follow src/vkvg_context.c:698-795
     * and :1251-1512 with the pipeline states of src/vkvg_device_internal.c:226-246. */
    VKH_DRAW_CLIP_NZ = 5,       /* CLIP |= !(w != 0); key folded where CLIP stays 0   :754-795 */
    VKH_DRAW_CLIP_EO = 6,       /* CLIP |= !(w & 1) */
    VKH_DRAW_CLIP_SAVE = 7,     /* saved bit `style` (0..5) := CLIP                     :1320-1341 */
    VKH_DRAW_CLIP_RESTORE = 8,  /* CLIP := saved bit `style` (0..5)                     :1399-1428 */
    VKH_DRAW_CLIP_RESET = 9,    /* stencil byte := 0                                    :706-718 */
    VKH_DRAW_STENCIL_SPILL = 10,   /* copy `style` (0..2) := stencil byte               :1270-1317 */
    VKH_DRAW_STENCIL_UNSPILL = 11, /* stencil byte := copy `style` (0..2)               :1432-1474 */
};
#define VKH_STENCIL_SPILLS 3 /* byte copies per pixel: 6 * (1 + 3) = 24 nested clip saves */

typedef struct {
    uint32_t kind;           /* VKH_DRAW_* */
    uint32_t op;             /* vkvg_operator_t */
    uint32_t color;          /* premultiplied RGBA8, byte 0 = R (CreateRgbaf) */
    int32_t  grad;           /* gradient index, -1 for the solid colour */
    float    opacity;
    float    mat[6];         /* CTM (xx yx xy yy x0 y0) applied as in shaders/vkvg_main.vert:74-79 */
    uint32_t first_subpath;
    uint32_t n_subpaths;
    int32_t  style;          /* stroke style index (STROKE); bit or copy index (stencil draws) */
    uint32_t target;         /* render target index */
    uint32_t key_index;      /* draw number on its target (coverage-signature key) */
    int32_t  source;         /* surface source index, -1 for none (then grad / color apply) */
} vkh_draw;

/* The smallest arc step a stroke style may carry: round joins and caps step their angle by it in
 * float, so a step below this (a radius beyond some 160 000 px, or a negative width) would take
 * millions of steps per arc, or never end when the steps round away. The recorder raises smaller
 * steps to it; the backend refuses a style below it. */
#define VKH_MIN_ARC_STEP 4.79368996e-5f /* pi / 2^16 */

typedef struct {
    float    hw;          /* line width / 2 */
    float    lhMax;       /* miter limit * line width */
    float    arcStep;     /* _get_arc_step(hw) under the draw's CTM */
    uint32_t join, cap;   /* vkvg_line_join_t, vkvg_line_cap_t */
    uint32_t dash_first;  /* index into the dash array */
    uint32_t dash_count;  /* 0: solid stroke */
    float    dash_offset;
} vkh_stroke_style;

typedef struct {
    uint32_t type;        /* 2 linear, 3 radial (vkvg_pattern_type_t) */
    uint32_t count;       /* colour stops */
    float    colors[16][4]; /* premultiplied at add time (src/vkvg_pattern.c:149-168) */
    float    stops[16];
    float    cp[2][4];    /* control points, transformed by the CTM at set_source time */
} vkh_gradient;

/* a surface used as source (src/vkvg_context_internal.c:705-772, shaders/vkvg_main.frag:72-82):
 * at pixel centre f, p = f - rect.xy, uv = mat_inv(p) / rect.zw, sampled with the sampler the
 * pattern's filter and extend select (NEAREST / LINEAR; CLAMP_TO_BORDER with transparent black,
 * REPEAT, MIRRORED_REPEAT, CLAMP_TO_EDGE) and multiplied by the draw's opacity */
typedef struct {
    const void *rgba;     /* device pointer of the source surface's image (RGBA8 as stored) */
    uint32_t width, height;
    uint32_t extend;      /* vkvg_extend_t */
    uint32_t filter;      /* vkvg_filter_t: BILINEAR and BEST sample linearly, others nearest */
    float    rect[4];     /* the context's push-constant source rectangle: x, y, width, height */
    float    mat_inv[6];  /* the context's push-constant inverse matrix (xx yx xy yy x0 y0) */
} vkh_surface_source;

typedef struct {
    void     *rgba;       /* device pointer, width*height*4 bytes, premultiplied RGBA8 */
    uint32_t  width, height;
    uint32_t *sig;        /* optional device pointer to width*height uint32 coverage signatures */
    uint32_t *stencil;    /* optional device pointer to width*height uint32 stencil words; needed
                           * by the stencil draws, and then every draw of the target is clipped */
    uint32_t  stencil_load; /* 1: the frame starts from the plane's contents; 0: from zero */
    uint32_t  reserved;
} vkh_target;

typedef struct {
    const vkh_item         *items;    uint32_t n_items;
    const vkh_subpath      *subpaths; uint32_t n_subpaths;
    const vkh_draw         *draws;    uint32_t n_draws;
    const vkh_stroke_style *styles;   uint32_t n_styles;
    const float            *dashes;   uint32_t n_dashes;
    const vkh_gradient     *grads;    uint32_t n_grads;
    const vkh_target       *targets;  uint32_t n_targets;
    const vkh_surface_source *sources; uint32_t n_sources;
} vkh_frame;

/* per-frame counters (sizes of the intermediate buffers the last render used) */
typedef struct {
    uint64_t points, stroke_vertices, stroke_triangles, edges;
    uint64_t path_tiles, segments, commands, tiles;
    uint32_t tile_size;
    float    ms_total;      /* host wall time of the last vkvg_hip_execute, ms */
    float    ms_fine;       /* device time of the fine (coverage + composite) kernel, ms */
    float    ms_device;     /* device time from the first to the last kernel of the frame, ms */
    float    ms_stage[6];   /* flatten, geometry (stroke + edges), tile setup, binning, coarse, fine */
    uint32_t graph;         /* 1 when the frame ran as a captured graph (stage times are then 0) */
    uint32_t fine_words;    /* fine batch capacity in LDS words (commands per batch = words / tile^2) */
} vkh_stats;

/* one autotuner measurement */
typedef struct {
    uint32_t tile, fine_words;
    float    ms;            /* best device time of the measured executions */
} vkh_tune_entry;
typedef struct {
    uint32_t       n, best;
    vkh_tune_entry e[16];
} vkh_tune;

typedef struct vkh_engine vkh_engine;

/* Engine: device buffers, stream and events for one HIP device. Returns NULL on failure. */
vkvg_public vkh_engine *vkvg_hip_engine_create(int device_ordinal);
vkvg_public void        vkvg_hip_engine_destroy(vkh_engine *eng);
/* tile size in pixels (8, 16 or 32); 0 restores the default. Returns 0 on success. */
vkvg_public int         vkvg_hip_set_tile_size(vkh_engine *eng, uint32_t tile);
/* Optional external stream (a hipStream_t passed as void*); NULL = the engine's own. */
vkvg_public void        vkvg_hip_set_stream(vkh_engine *eng, void *stream);
/* Fine batch capacity in LDS words (2048, 4096 or 8192; 0 = 4096): commands staged per batch =
 * words / tile^2. Applies to the next execution. */
vkvg_public int         vkvg_hip_set_fine_words(vkh_engine *eng, uint32_t words);
/* Re-tiles the resident frame (8, 16 or 32 px) without uploading it again. */
vkvg_public int         vkvg_hip_retile(vkh_engine *eng, uint32_t tile);
/* Graph replay: with on != 0, vkvg_hip_execute captures the resident frame's launch sequence
 * after its first (sizing) execution and replays it as one hipGraph launch afterwards. */
vkvg_public void        vkvg_hip_set_graph(vkh_engine *eng, int on);
/* Times every tile size x fine capacity on the resident frame (reps executions each, best time
 * kept), applies the fastest pair and reports all of them. */
vkvg_public int         vkvg_hip_autotune(vkh_engine *eng, uint32_t reps, vkh_tune *out);

/* device images */
vkvg_public int  vkvg_hip_image_alloc(vkh_engine *eng, uint32_t width, uint32_t height, void **dev_ptr);
vkvg_public void vkvg_hip_image_free(vkh_engine *eng, void *dev_ptr);
vkvg_public int  vkvg_hip_image_clear(vkh_engine *eng, void *dev_ptr, size_t bytes);
vkvg_public int  vkvg_hip_image_download(vkh_engine *eng, const void *dev_ptr, void *host, size_t bytes);
/* clears the stencil byte of a surface's stencil plane (one 32-bit word per pixel; the bytes
 * holding saved clip levels are kept) */
vkvg_public int  vkvg_hip_stencil_clear(vkh_engine *eng, void *dev_ptr, size_t npixels);
vkvg_public int  vkvg_hip_image_upload(vkh_engine *eng, void *dev_ptr, const void *host, size_t bytes);
/* fused BGRA/RGBA swizzle-free readback with vkvg's truncating unpremultiply
 * (src/vkvg_surface.c:441-452) on the device, then copy to host */
vkvg_public int  vkvg_hip_image_read_unpremultiplied(vkh_engine *eng, const void *dev_ptr, uint32_t width,
                                                     uint32_t height, void *host);
/* The same readback without allocating anything: no counterpart in the reference; for callers that
 * cannot allocate device memory (serving loops, allocation-free harnesses). Stages through a static
 * 64 MiB device array (at most 4096 x 4096 pixels; larger images return -1 with the reason in
 * vkvg_hip_last_error). Every call on one GPU shares that array: calls that overlap on one GPU race. */
vkvg_public int  vkvg_hip_image_read_unpremultiplied_noalloc(vkh_engine *eng, const void *dev_ptr, uint32_t width,
                                                             uint32_t height, void *host);

/* Render one frame (any number of draws over any number of targets) and wait for it:
 * vkvg_hip_upload followed by vkvg_hip_execute. Returns 0 on success, < 0 on error. */
vkvg_public int  vkvg_hip_render(vkh_engine *eng, const vkh_frame *frame);
/* Copies a frame into device memory (one transfer); it stays resident until the next upload. */
vkvg_public int  vkvg_hip_upload(vkh_engine *eng, const vkh_frame *frame);
/* Runs the kernel pipeline on the resident frame (re-executable; the frame's targets are
 * composited again each time). */
vkvg_public int  vkvg_hip_execute(vkh_engine *eng);
vkvg_public void vkvg_hip_get_stats(vkh_engine *eng, vkh_stats *out);
vkvg_public const char *vkvg_hip_last_error(vkh_engine *eng);

/* Host side, no device needed: widens ext = {x_min, y_min, x_max, y_max} by the points that a cubic
 * item flattens to (its subdivision points and, unless VKH_ITEM_NO_END is set, its end point), with
 * the comparisons of _vkvg_path_extents (src/vkvg_context_internal.c:1879-1917). The subdivision is
 * the flattening kernels' own code compiled for the host. Used by vkvg_path_extents
 * (src/vkvg_context.c:684-696). */
vkvg_public void vkvg_hip_item_extents(const vkh_item *item, float ext[4]);

#ifdef __cplusplus
}
#endif
#endif
