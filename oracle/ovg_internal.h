/* Internal declarations of the CPU oracle (test infrastructure only). */
#ifndef OVG_INTERNAL_H
#define OVG_INTERNAL_H

#include "ovg.h"
#include <float.h>
#include <math.h>
#include <stdbool.h>
#include <stdlib.h>
#include <string.h>

/* float constants of src/vkvg_internal.h:42-47 */
#define OVG_PIF   3.14159265358979323846f
#define OVG_PIF_2 1.57079632679489661923f
#define OVG_2_PIF 0.63661977236758134308f /* 2/pi, used where 2*pi was meant (reproduced) */

/* path element bits, src/vkvg_internal.h:62-68 */
#define P_CLOSED   0x80000000u
#define P_CURVES   0x40000000u
#define P_CONVEX   0x20000000u
#define P_MASK     0x1FFFFFFFu

#define EQUF(a, b) (fabsf((a) - (b)) <= FLT_EPSILON)

/* libm: the oracle evaluates the float transcendental functions in double and rounds
 * once (the HIP kernels do the same). Build with -DOVG_GLIBC_FLOAT_LIBM to use glibc's
 * float entry points instead, as the reference does; the tests count the differences. */
#ifdef OVG_GLIBC_FLOAT_LIBM
#define L_ATAN2F(y, x) atan2f((y), (x))
#define L_COSF(a)      cosf(a)
#define L_SINF(a)      sinf(a)
#define L_ACOSF(a)     acosf(a)
#define L_POWF(a, b)   powf((a), (b))
#else
#define L_ATAN2F(y, x) ((float)atan2((double)(y), (double)(x)))
#define L_COSF(a)      ((float)cos((double)(a)))
#define L_SINF(a)      ((float)sin((double)(a)))
#define L_ACOSF(a)     ((float)acos((double)(a)))
#define L_POWF(a, b)   ((float)pow((double)(a), (double)(b)))
#endif

typedef struct { float x, y; } v2;
typedef struct { float xx, yx, xy, yy, x0, y0; } mat6;
typedef struct { float x, y; uint32_t color; } vtx;

enum { SRC_SOLID = 0, SRC_SURFACE = 1, SRC_LINEAR = 2, SRC_RADIAL = 3 };

/* a surface pattern as defined (vkvg_pattern_create_for_surface + extend / filter / matrix) */
typedef struct {
    uint32_t index;       /* into the stream's sources */
    uint32_t extend, filter;
    bool     hasMatrix;
    mat6     matrix;
} ovg_surf_pat;
enum { OP_CLEAR = 0, OP_SOURCE = 1, OP_OVER = 2, OP_DIFFERENCE = 3 };

typedef struct {
    int      type;          /* SRC_LINEAR / SRC_RADIAL */
    float    colors[16][4]; /* premultiplied at add time */
    float    stops[16];
    float    cp[2][4];
    uint32_t count;
    bool     hasMatrix;     /* vkvg_pattern_set_matrix */
    mat6     matrix;
} ovg_grad;

/* vkvg_clip_state_t, src/vkvg_context_internal.h:94-99 */
enum { OVG_CLIP_NONE = 0, OVG_CLIP_CLEAR = 1, OVG_CLIP_CLIP = 2, OVG_CLIP_SAVED = 6 };

typedef struct octx_s {
    int      status;
    ovg_surface *surf;

    /* path storage, same layout as the reference's points[] / pathes[] */
    v2       *points;
    uint32_t  pointCount, sizePoints;
    uint32_t *pathes;
    uint32_t  pathPtr, sizePathes, segmentPtr, subpathCount;
    bool      simpleConvex;

    /* graphic state */
    float    lineWidth, miterLimit, dashOffset, opacity;
    uint32_t dashCount;
    float   *dashes;
    int      lineCap, lineJoin, fillRule, op;
    uint32_t curColor;
    mat6     mat;
    int      srcType;       /* SRC_SOLID or gradient */
    ovg_grad grad;          /* control points already transformed at set_source time */
    uint32_t patId;         /* identity of the current pattern, 0 for a solid colour */
    ovg_grad gradDef;       /* a gradient pattern as defined (before the CTM), for vkvg_restore */
    ovg_surf_pat surfPat;   /* a surface pattern (srcType SRC_SURFACE) */
    /* push constants a surface source reads (src/vkvg_context_internal.h:74-81) */
    float    srcRect[4];    /* x, y, width, height */
    mat6     matInv;        /* inverse of the CTM (times a surface pattern's matrix) */
    const ovg_source *sources;
    uint32_t nSources;

    /* clip and save/restore state (src/vkvg_context_internal.h:87-130) */
    int       clipState;    /* OVG_CLIP_* */
    uint32_t  curSavBit;    /* stencil save levels in use */
    struct osave *saved;    /* vkvg_save stack */
    uint8_t **spills;       /* whole-stencil copies taken every 6 save levels */
    uint32_t  nSpills;

    /* vertex / index caches used by stroke and fill geometry */
    vtx      *vc;
    uint32_t  vertCount, sizeV;
    uint32_t *ic;
    uint32_t  indCount, sizeI;
} octx;

/* one vkvg_save record (vkvg_context_save_t, src/vkvg_context_internal.h:101-130) */
struct osave {
    struct osave *next;
    int      clipState;
    float    dashOffset;
    uint32_t dashCount;
    float   *dashes;
    float    lineWidth, miterLimit, opacity;
    int      op, lineCap, fillRule;
    mat6     mat, matInv;
    float    srcRect[4];
    uint32_t curColor, patId;
    int      patType;
    ovg_grad gradDef;
    ovg_surf_pat surfPat;
};

/* geometry (ovg_geom.c) */
void octx_init(octx *c, ovg_surface *s);
void octx_free(octx *c);
void o_move_to(octx *c, float x, float y);
void o_line_to(octx *c, float x, float y);
void o_rel_line_to(octx *c, float dx, float dy);
void o_rel_move_to(octx *c, float dx, float dy);
void o_curve_to(octx *c, float x1, float y1, float x2, float y2, float x3, float y3);
void o_rel_curve_to(octx *c, float x1, float y1, float x2, float y2, float x3, float y3);
void o_quadratic_to(octx *c, float x1, float y1, float x2, float y2);
void o_close_path(octx *c);
void o_new_path(octx *c);
void o_new_sub_path(octx *c);
void o_arc(octx *c, float xc, float yc, float r, float a1, float a2);
void o_arc_negative(octx *c, float xc, float yc, float r, float a1, float a2);
int  o_rectangle(octx *c, float x, float y, float w, float h);
int  o_rounded_rectangle(octx *c, float x, float y, float w, float h, float radius);
void o_ellipse(octx *c, float rx, float ry, float x, float y, float rot);
void o_elliptic_arc_to(octx *c, float x2, float y2, bool large, bool sweep, float rx, float ry, float phi);
void o_finish_path(octx *c);
void o_clear_path(octx *c);
/* builds the stroke triangles of the current path into vc/ic (indices local to 0) */
void o_stroke_geometry(octx *c);
void o_set_dash(octx *c, const float *d, uint32_t n, float off);

/* matrix (ovg_geom.c), src/vkvg_matrix.c */
void  m_multiply(mat6 *r, const mat6 *a, const mat6 *b);
void  m_transform_point(const mat6 *m, float *x, float *y);
void  m_transform_distance(const mat6 *m, float *x, float *y);
int   m_invert(mat6 *m);
void  m_get_scale(const mat6 *m, float *sx, float *sy);

/* rows [ovg_row_lo, ovg_row_hi) that the rasterizing loops visit (the whole surface unless a band
 * render narrows them); per thread, so bands render concurrently */
extern __thread int32_t ovg_row_lo, ovg_row_hi;

/* raster (ovg_raster.c) */
void r_fill(octx *c);
void r_stroke(octx *c);
void r_paint(octx *c);
void r_clear(octx *c);
void r_clip(octx *c);
/* clip state machine and save/restore (ovg_state.c) */
void o_clip_preserve(octx *c);
void o_reset_clip(octx *c);
void o_clear(octx *c);
void o_save(octx *c);
void o_restore(octx *c);
void o_set_gradient(octx *c, const ovg_grad *def);
void o_set_surface_pattern(octx *c, const ovg_surf_pat *p);
void o_update_mat_inv(octx *c);
void o_state_free(octx *c);
uint32_t o_rgbaf(float r, float g, float b, float a);

#endif
