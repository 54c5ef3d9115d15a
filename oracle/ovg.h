/*
 * ovg — CPU oracle for the vkvg fill/stroke/composite path.
 *
 * TEST INFRASTRUCTURE ONLY. Nothing in the shipped library links, loads or calls
 * this code; only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use
 * it, and only as the checker (or the timed CPU baseline).
 *
 * It restates, in plain C, the reference's host geometry and the Vulkan fixed-function
 * semantics that vkvg relies on:
 *   - path encoding, AGG bezier flattening, arcs, rectangles, ellipses
 *       (src/vkvg_context.c:332-639, src/vkvg_context_internal.c:136-252, :1313-1580)
 *   - stroke joins / caps / dashes through literal vertex and index caches
 *       (src/vkvg_context.c:822-948, src/vkvg_context_internal.c:924-1282)
 *   - even-odd fills as stencil-INVERT triangle fans + scissored cover
 *       (src/vkvg_context_internal.c:1583-1655, :1919-1952; src/vkvg_device_internal.c:226-246)
 *   - non-zero fills: convex fan fast path, otherwise the [winding != 0] region that the
 *       absent glutess tessellation covers (src/vkvg_context_internal.c:1720-1793)
 *   - the vertex transform (shaders/vkvg_main.vert:59-81), 1/256 sub-pixel snap,
 *       pixel-centre sampling and the top-left tie rule, premultiplied OVER in UNORM8
 *   - linear / radial gradients exactly as shaders/vkvg_main.frag:83-144
 *   - readback with truncating unpremultiply (src/vkvg_surface.c:441-452)
 *
 * Every rasterized draw also folds its per-pixel coverage count k into a 32-bit
 * signature image:  sig[p] += k * ovg_draw_key(draw_index), so two renderers agree on
 * every per-draw 1-sample mask exactly when their signature images are equal.
 */
#ifndef OVG_H
#define OVG_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Scene command stream: a flat uint32 array, opcode followed by fixed operands
 * (floats are stored bit-cast). The same stream drives the oracle and the HIP library. */
enum {
    OVG_OP_END = 0,
    OVG_OP_MOVE_TO = 1,          /* x y */
    OVG_OP_LINE_TO = 2,          /* x y */
    OVG_OP_CURVE_TO = 3,         /* x1 y1 x2 y2 x3 y3 */
    OVG_OP_CLOSE_PATH = 4,
    OVG_OP_ARC = 5,              /* xc yc r a1 a2 */
    OVG_OP_ARC_NEG = 6,          /* xc yc r a1 a2 */
    OVG_OP_RECTANGLE = 7,        /* x y w h */
    OVG_OP_FILL = 8,
    OVG_OP_STROKE = 9,
    OVG_OP_FILL_PRESERVE = 10,
    OVG_OP_STROKE_PRESERVE = 11,
    OVG_OP_PAINT = 12,
    OVG_OP_SET_SOURCE_RGBA = 13, /* r g b a */
    OVG_OP_SET_SOURCE_RGB = 14,  /* r g b */
    OVG_OP_SET_LINE_WIDTH = 15,  /* w */
    OVG_OP_SET_LINE_JOIN = 16,   /* u32 */
    OVG_OP_SET_LINE_CAP = 17,    /* u32 */
    OVG_OP_SET_MITER_LIMIT = 18, /* m */
    OVG_OP_SET_FILL_RULE = 19,   /* u32 */
    OVG_OP_SET_DASH = 20,        /* u32 n, offset, n floats */
    OVG_OP_SET_OPACITY = 21,     /* o */
    OVG_OP_SET_OPERATOR = 22,    /* u32 */
    OVG_OP_LINEAR = 23,          /* x0 y0 x1 y1 u32 nstops, nstops*(off r g b a) */
    OVG_OP_RADIAL = 24,          /* cx0 cy0 r0 cx1 cy1 r1 u32 nstops, nstops*(off r g b a) */
    OVG_OP_NEW_PATH = 25,
    OVG_OP_NEW_SUB_PATH = 26,
    OVG_OP_REL_LINE_TO = 27,     /* dx dy */
    OVG_OP_REL_MOVE_TO = 28,     /* dx dy */
    OVG_OP_QUADRATIC_TO = 29,    /* x1 y1 x2 y2 */
    OVG_OP_TRANSLATE = 30,       /* dx dy */
    OVG_OP_SCALE = 31,           /* sx sy */
    OVG_OP_ROTATE = 32,          /* a */
    OVG_OP_IDENTITY_MATRIX = 33,
    OVG_OP_CLEAR = 34,
    OVG_OP_ELLIPSE = 35,         /* rx ry x y rot */
    OVG_OP_REL_CURVE_TO = 36,    /* 6 floats */
    OVG_OP_SET_MATRIX = 37,      /* xx yx xy yy x0 y0 */
    OVG_OP_ELLIPTIC_ARC_TO = 38, /* x y u32 large u32 sweep rx ry phi */
    OVG_OP_ROUNDED_RECT = 39,    /* x y w h r */
    OVG_OP_SET_SOURCE_COLOR = 40,/* u32 */
    OVG_OP_CLIP = 41,
    OVG_OP_CLIP_PRESERVE = 42,
    OVG_OP_RESET_CLIP = 43,
    OVG_OP_SAVE = 44,
    OVG_OP_RESTORE = 45,
    OVG_OP_FLUSH = 46,           /* vkvg_flush: a frame boundary, no drawing state changes */
    OVG_OP_SET_SOURCE_SURFACE = 47, /* u32 source index, x, y */
    OVG_OP_SOURCE_PATTERN = 48,  /* u32 source index, u32 extend, u32 filter, u32 has_matrix, 6 floats:
                                  * pattern for the surface, set_source, destroy */
    OVG_OP_ROUNDED_RECT2 = 49,   /* x y w h rx ry */
    OVG_OP_LINEAR_MATRIX = 50,   /* 6 floats (pattern matrix), then the operands of OVG_OP_LINEAR */
    OVG_OP_RADIAL_MATRIX = 51,   /* 6 floats (pattern matrix), then the operands of OVG_OP_RADIAL */
    OVG_OP_COUNT
};

typedef struct {
    uint32_t width, height;
    uint8_t *rgba;     /* premultiplied RGBA8, row-major, as stored on the device */
    uint8_t *stencil;  /* S8: bit0 FILL, bit1 CLIP, bits 2..7 saved CLIP bits */
    uint32_t *sig;     /* per-pixel coverage signature */
    int32_t *acc;      /* scratch for the non-zero winding pass */
    uint32_t draw_counter;
    uint64_t tri_count;   /* triangles rasterized (stats) */
    uint64_t frag_count;  /* fragments blended (stats) */
} ovg_surface;

/* a source surface of a stream: premultiplied RGBA8 as stored, row-major */
typedef struct {
    const uint8_t *rgba;
    uint32_t width, height;
} ovg_source;

/* key mixed into the signature for the n-th rasterized draw (a 32-bit bijective hash of n) */
uint32_t ovg_draw_key(uint32_t n);

/* Render one command stream onto a fresh transparent surface of the given size.
 * out_rgba: w*h*4 premultiplied bytes (device layout), out_sig: w*h uint32 (may be NULL).
 * Returns the context status (0 = success). */
int ovg_render_stream(uint32_t width, uint32_t height, const uint32_t *words, size_t nwords,
                      uint8_t *out_rgba, uint32_t *out_sig, uint64_t *out_stats);
/* the same with source surfaces that OVG_OP_SET_SOURCE_SURFACE / OVG_OP_SOURCE_PATTERN index */
int ovg_render_stream_src(uint32_t width, uint32_t height, const uint32_t *words, size_t nwords,
                          const ovg_source *sources, uint32_t n_sources, uint8_t *out_rgba, uint32_t *out_sig,
                          uint64_t *out_stats);

/* The same, rasterizing only the rows [row0, row1): the band's rows equal the full render's (other
 * rows are left as they are); bands of one stream render concurrently on separate threads. */
int ovg_render_stream_rows(uint32_t width, uint32_t height, const uint32_t *words, size_t nwords, uint32_t row0,
                           uint32_t row1, uint8_t *out_rgba, uint32_t *out_sig, uint64_t *out_stats);

/* Readback as vkvg_surface_write_to_memory: RGBA with truncating unpremultiply. */
void ovg_unpremultiply(const uint8_t *premul, uint8_t *out, size_t npixels);

/* Low-level rasterizer entry points, exported so the tests can pin them against
 * brute force and against each other. Vertices are framebuffer-space floats. */
void ovg_snap_vertex(float x, float y, uint32_t w, uint32_t h, const float *mat6, int32_t *fx, int32_t *fy);
int  ovg_sample_in_triangle(const int32_t *fx, const int32_t *fy, int32_t sx, int32_t sy);
/* coverage of one snapped triangle over a w*h grid, scanline method; adds 1 into cov[] */
void ovg_raster_triangle_scan(const int32_t *fx, const int32_t *fy, uint32_t w, uint32_t h, int32_t *cov);
/* same, brute force per sample */
void ovg_raster_triangle_brute(const int32_t *fx, const int32_t *fy, uint32_t w, uint32_t h, int32_t *cov);
/* winding of a closed polygon (snapped) at every sample, scanline crossing method */
void ovg_winding_scan(const int32_t *px, const int32_t *py, uint32_t n, uint32_t w, uint32_t h, int32_t *wind);
/* winding as the signed sum of fan-triangle coverage (independent formulation) */
void ovg_winding_fan(const int32_t *px, const int32_t *py, uint32_t n, uint32_t w, uint32_t h, int32_t *wind);

/* AGG flattening of one cubic as vkvg's _curve_to does (start point excluded, end point
 * included); returns the number of points written (<= cap). */
uint32_t ovg_flatten_cubic(float x0, float y0, float x1, float y1, float x2, float y2, float x3, float y3,
                           float tolerance, float *out_xy, uint32_t cap);

/* The same subdivision started at a given recursion level: level >= 1 applies the flatness
 * tests to the given control points themselves (level 0 always subdivides once). */
uint32_t ovg_flatten_cubic_level(float x0, float y0, float x1, float y1, float x2, float y2, float x3, float y3,
                                 float tolerance, unsigned level, float *out_xy, uint32_t cap);

/* Geometry of the first fill or stroke of a stream, before rasterization: the finished path
 * (points and the path-record array with its CLOSED / CURVES / CONVEX bits) and, for a stroke,
 * the vertex positions and triangle indices the stroke emits. Counts are always the full
 * counts; at most `cap_*` entries are written. */
typedef struct {
    float    *points;   uint32_t cap_points, n_points;   /* x, y pairs */
    uint32_t *pathes;   uint32_t cap_pathes, n_pathes;
    float    *verts;    uint32_t cap_verts,  n_verts;    /* x, y pairs */
    uint32_t *inds;     uint32_t cap_inds,   n_inds;
} ovg_geometry;
int ovg_geometry_dump(uint32_t width, uint32_t height, const uint32_t *words, size_t nwords, ovg_geometry *out);

/* H-blend: one premultiplied OVER step in UNORM8. */
void ovg_blend_over(uint8_t *d, const uint8_t *s);

#ifdef __cplusplus
}
#endif
#endif
