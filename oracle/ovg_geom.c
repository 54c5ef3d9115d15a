/*
 * CPU oracle — path encoding, flattening and stroke geometry (test infrastructure only).
 *
 * Restates the reference's host geometry with its float evaluation order kept:
 *   path bookkeeping        src/vkvg_context_internal.c:136-238, src/vkvg_context.c:332-393,504-522
 *   AGG cubic flattening   src/vkvg_context_internal.c:1305-1461, src/vkvg_context.c:541-577
 *   arcs / rect / ellipse  src/vkvg_context.c:394-503,620-639,1605-1639, src/vkvg_context_internal.c:245-252,1473-1580
 *   stroke                 src/vkvg_context.c:822-948, src/vkvg_context_internal.c:924-1282
 *   matrices               src/vkvg_matrix.c:107-229
 * Compile with -ffp-contract=off: no fused multiply-add may change a rounding.
 */
#include "ovg_internal.h"

/* ---------------------------------------------------------------- vec2 helpers
 * (src/vectors.h:92-165, operand order preserved) */
static inline v2 V(float x, float y) { v2 r = {x, y}; return r; }
static inline v2 vadd(v2 a, v2 b) { return V(a.x + b.x, a.y + b.y); }
static inline v2 vsub(v2 a, v2 b) { return V(a.x - b.x, a.y - b.y); }
static inline v2 vmul(v2 a, float m) { return V(a.x * m, a.y * m); }
static inline v2 vdiv(v2 a, float m) { return V(a.x / m, a.y / m); }
static inline v2 vperp(v2 a) { return V(a.y, -a.x); }
static inline float vlen(v2 v) { return sqrtf(v.x * v.x + v.y * v.y); }
static inline v2 vnorm(v2 a) { float m = sqrtf(a.x * a.x + a.y * a.y); return V(a.x / m, a.y / m); }
static inline float vdot(v2 a, v2 b) { return (a.x * b.x) + (a.y * b.y); }
static inline bool vequ(v2 a, v2 b) { return EQUF(a.x, b.x) & EQUF(a.y, b.y); }
static inline v2 vline_norm(v2 a, v2 b) {
    v2 d = {b.x - a.x, b.y - a.y};
    float md = sqrtf(d.x * d.x + d.y * d.y);
    d.x /= md;
    d.y /= md;
    return d;
}

/* ---------------------------------------------------------------- matrices */
void m_multiply(mat6 *res, const mat6 *a, const mat6 *b) {
    mat6 r;
    r.xx = a->xx * b->xx + a->yx * b->xy;
    r.yx = a->xx * b->yx + a->yx * b->yy;
    r.xy = a->xy * b->xx + a->yy * b->xy;
    r.yy = a->xy * b->yx + a->yy * b->yy;
    r.x0 = a->x0 * b->xx + a->y0 * b->xy + b->x0;
    r.y0 = a->x0 * b->yx + a->y0 * b->yy + b->y0;
    *res = r;
}
void m_transform_distance(const mat6 *m, float *dx, float *dy) {
    float nx = (m->xx * *dx + m->xy * *dy);
    float ny = (m->yx * *dx + m->yy * *dy);
    *dx = nx;
    *dy = ny;
}
void m_transform_point(const mat6 *m, float *x, float *y) {
    m_transform_distance(m, x, y);
    *x += m->x0;
    *y += m->y0;
}
int m_invert(mat6 *m) {
    if (m->xy == 0. && m->yx == 0.) {
        m->x0 = -m->x0;
        m->y0 = -m->y0;
        if (m->xx != 1.f) {
            if (m->xx == 0.)
                return 1;
            m->xx = 1.f / m->xx;
            m->x0 *= m->xx;
        }
        if (m->yy != 1.f) {
            if (m->yy == 0.)
                return 1;
            m->yy = 1.f / m->yy;
            m->y0 *= m->yy;
        }
        return 0;
    }
    float det = m->xx * m->yy - m->yx * m->xy;
    if (!(det * det >= 0.) || det == 0)
        return 1;
    float a = m->xx, b = m->yx, cc = m->xy, d = m->yy, tx = m->x0, ty = m->y0;
    mat6 adj = {d, -b, -cc, a, cc * ty - d * tx, b * tx - a * ty};
    float s = 1 / det;
    adj.xx *= s; adj.yx *= s; adj.xy *= s; adj.yy *= s; adj.x0 *= s; adj.y0 *= s;
    *m = adj;
    return 0;
}
void m_get_scale(const mat6 *m, float *sx, float *sy) {
    *sx = (float)sqrt(m->xx * m->xx + m->xy * m->xy);
    *sy = (float)sqrt(m->yx * m->yx + m->yy * m->yy);
}

/* ---------------------------------------------------------------- context */
void octx_init(octx *c, ovg_surface *s) {
    memset(c, 0, sizeof(*c));
    c->surf = s;
    c->sizePoints = 1024;
    c->points = (v2 *)malloc(c->sizePoints * sizeof(v2));
    c->sizePathes = 64;
    c->pathes = (uint32_t *)calloc(c->sizePathes, sizeof(uint32_t));
    c->sizeV = 4096;
    c->vc = (vtx *)malloc(c->sizeV * sizeof(vtx));
    c->sizeI = 4096 * 6;
    c->ic = (uint32_t *)malloc(c->sizeI * sizeof(uint32_t));
    c->lineWidth = 1.f;
    c->miterLimit = 10.f;
    c->op = OP_OVER;
    c->fillRule = 1; /* NON_ZERO */
    c->opacity = 1.f;
    c->curColor = 0xff000000u;
    mat6 id = {1, 0, 0, 1, 0, 0};
    c->mat = id;
    c->srcType = SRC_SOLID;
}
void octx_free(octx *c) {
    free(c->points);
    free(c->pathes);
    free(c->vc);
    free(c->ic);
    free(c->dashes);
}

static void ensure_pathes(octx *c) {
    if (c->sizePathes - c->pathPtr - c->segmentPtr > 8)
        return;
    uint32_t ns = c->sizePathes * 2;
    c->pathes = (uint32_t *)realloc(c->pathes, ns * sizeof(uint32_t));
    memset(c->pathes + c->sizePathes, 0, (ns - c->sizePathes) * sizeof(uint32_t));
    c->sizePathes = ns;
}
static void ensure_points(octx *c) {
    if (c->sizePoints - 8 > c->pointCount)
        return;
    c->sizePoints *= 2;
    c->points = (v2 *)realloc(c->points, c->sizePoints * sizeof(v2));
}
static void ensure_verts(octx *c, uint32_t add) {
    if (c->vertCount + add + 8 < c->sizeV)
        return;
    while (c->vertCount + add + 8 >= c->sizeV)
        c->sizeV *= 2;
    c->vc = (vtx *)realloc(c->vc, c->sizeV * sizeof(vtx));
}
static void ensure_inds(octx *c, uint32_t add) {
    if (c->indCount + add + 8 < c->sizeI)
        return;
    while (c->indCount + add + 8 >= c->sizeI)
        c->sizeI *= 2;
    c->ic = (uint32_t *)realloc(c->ic, c->sizeI * sizeof(uint32_t));
}

static inline bool path_empty(octx *c) { return c->pathes[c->pathPtr] == 0; }
static inline v2 cur_pos(octx *c) { return c->points[c->pointCount - 1]; }

/* _add_point, src/vkvg_context_internal.c:221-238 */
static void add_point(octx *c, float x, float y) {
    ensure_points(c);
    if (isnan(x) || isnan(y))
        return;
    c->points[c->pointCount++] = V(x, y);
    c->pathes[c->pathPtr]++;
    if (c->segmentPtr > 0)
        c->pathes[c->pathPtr + c->segmentPtr]++;
}
/* _set_curve_start, :136-153 */
static void set_curve_start(octx *c) {
    if (c->segmentPtr > 0) {
        if ((c->pathes[c->pathPtr + c->segmentPtr] & P_MASK) > 0)
            c->segmentPtr++;
    } else {
        if (c->pathes[c->pathPtr] > 0) {
            c->pathes[c->pathPtr + 1] = c->pathes[c->pathPtr];
            c->segmentPtr = 2;
        } else
            c->segmentPtr = 1;
    }
    ensure_pathes(c);
    c->pathes[c->pathPtr + c->segmentPtr] = 0;
}
/* _set_curve_end, :155-162 */
static void set_curve_end(octx *c) {
    c->pathes[c->pathPtr + c->segmentPtr] |= P_CURVES;
    c->segmentPtr++;
    ensure_pathes(c);
    c->pathes[c->pathPtr + c->segmentPtr] = 0;
}
/* _finish_path, :163-197 */
void o_finish_path(octx *c) {
    if (c->pathes[c->pathPtr] == 0)
        return;
    if ((c->pathes[c->pathPtr] & P_MASK) < 2) {
        c->pointCount -= c->pathes[c->pathPtr];
        c->pathes[c->pathPtr] = 0;
        c->segmentPtr = 0;
        return;
    }
    if (c->pathPtr == 0 && c->simpleConvex)
        c->pathes[0] |= P_CONVEX;
    if (c->segmentPtr > 0) {
        c->pathes[c->pathPtr] |= P_CURVES;
        if ((c->pathes[c->pathPtr + c->segmentPtr] & P_CURVES) == 0 &&
            (c->pathes[c->pathPtr + c->segmentPtr] & P_MASK) > 0)
            c->segmentPtr++;
        c->pathPtr += c->segmentPtr;
    } else
        c->pathPtr++;
    ensure_pathes(c);
    c->pathes[c->pathPtr] = 0;
    c->segmentPtr = 0;
    c->subpathCount++;
    c->simpleConvex = false;
}
/* _clear_path, :199-206 */
void o_clear_path(octx *c) {
    c->pathPtr = 0;
    c->pathes[0] = 0;
    c->pointCount = 0;
    c->segmentPtr = 0;
    c->subpathCount = 0;
    c->simpleConvex = false;
}
/* _remove_last_point, :207-219 */
static void remove_last_point(octx *c) {
    c->pathes[c->pathPtr]--;
    c->pointCount--;
    if (c->segmentPtr > 0) {
        if (!c->pathes[c->pathPtr + c->segmentPtr])
            c->segmentPtr--;
        c->pathes[c->pathPtr + c->segmentPtr]--;
        if ((c->pathes[c->pathPtr + c->segmentPtr] & P_MASK) == 0)
            c->pathes[c->pathPtr + c->segmentPtr] = 0;
        else if (c->pathes[c->pathPtr + c->segmentPtr] & P_CURVES)
            c->segmentPtr++;
    }
}
/* _get_arc_step, :245-252 */
static float arc_step(octx *c, float radius) {
    float sx, sy;
    m_get_scale(&c->mat, &sx, &sy);
    float r = radius * fabsf(fmaxf(sx, sy));
    if (r < 30.0f)
        return fminf(OVG_PIF / 3.f, OVG_PIF / r);
    return fminf(OVG_PIF / 3.f, OVG_PIF / (r * 0.4f));
}
/* _line_to, :1463-1472 */
static void line_to_i(octx *c, float x, float y) {
    v2 p = {x, y};
    if (!path_empty(c)) {
        if (vequ(cur_pos(c), p))
            return;
    }
    add_point(c, x, y);
    c->simpleConvex = false;
}

/* ---------------------------------------------------------------- AGG flattening
 * _recursive_bezier, src/vkvg_context_internal.c:1313-1461. Constants are doubles in the
 * reference (CURVE_COLLINEARITY_EPSILON 1.7, M_CUSP_LIMIT 0.01, M_ANGLE_TOLERANCE 0.01),
 * so the comparisons that use them directly promote the float operand to double. */
typedef void (*emit_fn)(void *user, float x, float y);

static void rbez(emit_fn emit, void *u, float tol, float x1, float y1, float x2, float y2, float x3, float y3,
                 float x4, float y4, unsigned level) {
    if (level > 100)
        return;
    float x12 = (x1 + x2) / 2, y12 = (y1 + y2) / 2;
    float x23 = (x2 + x3) / 2, y23 = (y2 + y3) / 2;
    float x34 = (x3 + x4) / 2, y34 = (y3 + y4) / 2;
    float x123 = (x12 + x23) / 2, y123 = (y12 + y23) / 2;
    float x234 = (x23 + x34) / 2, y234 = (y23 + y34) / 2;
    float x1234 = (x123 + x234) / 2, y1234 = (y123 + y234) / 2;

    if (level > 0) {
        float dx = x4 - x1, dy = y4 - y1;
        float d2 = fabsf(((x2 - x4) * dy - (y2 - y4) * dx));
        float d3 = fabsf(((x3 - x4) * dy - (y3 - y4) * dx));
        float da1, da2;
        if ((double)d2 > 1.7 && (double)d3 > 1.7) {
            if ((d2 + d3) * (d2 + d3) <= (dx * dx + dy * dy) * tol) {
                float a23 = L_ATAN2F(y3 - y2, x3 - x2);
                da1 = fabsf(a23 - L_ATAN2F(y2 - y1, x2 - x1));
                da2 = fabsf(L_ATAN2F(y4 - y3, x4 - x3) - a23);
                if (da1 >= OVG_PIF)
                    da1 = OVG_2_PIF - da1;
                if (da2 >= OVG_PIF)
                    da2 = OVG_2_PIF - da2;
                if (da1 + da2 < (float)0.01) {
                    emit(u, x1234, y1234);
                    return;
                }
                if ((double)da1 > 0.01) {
                    emit(u, x2, y2);
                    return;
                }
                if ((double)da2 > 0.01) {
                    emit(u, x3, y3);
                    return;
                }
            }
        } else {
            if ((double)d2 > 1.7) {
                if (d2 * d2 <= tol * (dx * dx + dy * dy)) {
                    da1 = fabsf(L_ATAN2F(y3 - y2, x3 - x2) - L_ATAN2F(y2 - y1, x2 - x1));
                    if (da1 >= OVG_PIF)
                        da1 = OVG_2_PIF - da1;
                    if ((double)da1 < 0.01) {
                        emit(u, x2, y2);
                        emit(u, x3, y3);
                        return;
                    }
                    if ((double)da1 > 0.01) {
                        emit(u, x2, y2);
                        return;
                    }
                }
            } else if ((double)d3 > 1.7) {
                if (d3 * d3 <= tol * (dx * dx + dy * dy)) {
                    da1 = fabsf(L_ATAN2F(y4 - y3, x4 - x3) - L_ATAN2F(y3 - y2, x3 - x2));
                    if (da1 >= OVG_PIF)
                        da1 = OVG_2_PIF - da1;
                    if ((double)da1 < 0.01) {
                        emit(u, x2, y2);
                        emit(u, x3, y3);
                        return;
                    }
                    if ((double)da1 > 0.01) {
                        emit(u, x3, y3);
                        return;
                    }
                }
            } else {
                dx = x1234 - (x1 + x4) / 2;
                dy = y1234 - (y1 + y4) / 2;
                if (dx * dx + dy * dy <= tol) {
                    emit(u, x1234, y1234);
                    return;
                }
            }
        }
    }
    rbez(emit, u, tol, x1, y1, x12, y12, x123, y123, x1234, y1234, level + 1);
    rbez(emit, u, tol, x1234, y1234, x234, y234, x34, y34, x4, y4, level + 1);
}

static void emit_ctx(void *u, float x, float y) { add_point((octx *)u, x, y); }

typedef struct { float *out; uint32_t n, cap; } flat_buf;
static void emit_buf(void *u, float x, float y) {
    flat_buf *b = (flat_buf *)u;
    if (b->n < b->cap) {
        b->out[2 * b->n] = x;
        b->out[2 * b->n + 1] = y;
    }
    b->n++;
}
uint32_t ovg_flatten_cubic_level(float x0, float y0, float x1, float y1, float x2, float y2, float x3, float y3,
                                 float tolerance, unsigned level, float *out_xy, uint32_t cap) {
    flat_buf b = {out_xy, 0, cap};
    rbez(emit_buf, &b, tolerance, x0, y0, x1, y1, x2, y2, x3, y3, level);
    emit_buf(&b, x3, y3);
    return b.n;
}
uint32_t ovg_flatten_cubic(float x0, float y0, float x1, float y1, float x2, float y2, float x3, float y3,
                           float tolerance, float *out_xy, uint32_t cap) {
    return ovg_flatten_cubic_level(x0, y0, x1, y1, x2, y2, x3, y3, tolerance, 0, out_xy, cap);
}

/* _curve_to, src/vkvg_context.c:541-566 */
static void curve_to_i(octx *c, float x1, float y1, float x2, float y2, float x3, float y3) {
    if (EQUF(x1, x2) && EQUF(x2, x3) && EQUF(y1, y2) && EQUF(y2, y3)) {
        if (path_empty(c) || (EQUF(cur_pos(c).x, x1) && EQUF(cur_pos(c).y, y1)))
            return;
    }
    c->simpleConvex = false;
    set_curve_start(c);
    if (path_empty(c))
        add_point(c, x1, y1);
    v2 cp = cur_pos(c);
    float sx = 1, sy = 1;
    m_get_scale(&c->mat, &sx, &sy);
    float tol = (float)fabs(0.25f / fmaxf(sx, sy));
    rbez(emit_ctx, c, tol, cp.x, cp.y, x1, y1, x2, y2, x3, y3, 0);
    add_point(c, x3, y3);
    set_curve_end(c);
}

/* ---------------------------------------------------------------- public path API */
void o_new_sub_path(octx *c) { o_finish_path(c); }
void o_new_path(octx *c) { o_clear_path(c); }
void o_close_path(octx *c) {
    if (c->pathes[c->pathPtr] & P_CLOSED)
        return;
    if (c->pathes[c->pathPtr] < 3)
        return;
    if (vequ(c->points[c->pointCount - 1], c->points[c->pointCount - c->pathes[c->pathPtr]])) {
        if (c->pathes[c->pathPtr] < 4)
            return;
        remove_last_point(c);
    }
    c->pathes[c->pathPtr] |= P_CLOSED;
    o_finish_path(c);
}
void o_rel_line_to(octx *c, float dx, float dy) {
    if (path_empty(c))
        add_point(c, 0, 0);
    v2 cp = cur_pos(c);
    line_to_i(c, cp.x + dx, cp.y + dy);
}
void o_line_to(octx *c, float x, float y) { line_to_i(c, x, y); }
void o_move_to(octx *c, float x, float y) {
    o_finish_path(c);
    add_point(c, x, y);
}
void o_rel_move_to(octx *c, float x, float y) {
    if (path_empty(c))
        add_point(c, 0, 0);
    v2 cp = cur_pos(c);
    o_finish_path(c);
    add_point(c, cp.x + x, cp.y + y);
}
void o_arc(octx *c, float xc, float yc, float radius, float a1, float a2) {
    while (a2 < a1)
        a2 += 2.f * OVG_PIF;
    if (a2 - a1 > 2.f * OVG_PIF)
        a2 = a1 + 2.f * OVG_PIF;
    v2 v = {L_COSF(a1) * radius + xc, L_SINF(a1) * radius + yc};
    float step = arc_step(c, radius);
    float a = a1;
    if (path_empty(c)) {
        set_curve_start(c);
        add_point(c, v.x, v.y);
        c->simpleConvex = !c->pathPtr;
    } else {
        line_to_i(c, v.x, v.y);
        set_curve_start(c);
        c->simpleConvex = false;
    }
    a += step;
    if (EQUF(a2, a1))
        return;
    while (a < a2) {
        v.x = L_COSF(a) * radius + xc;
        v.y = L_SINF(a) * radius + yc;
        add_point(c, v.x, v.y);
        a += step;
    }
    if (EQUF(a2 - a1, OVG_PIF * 2.f)) {
        set_curve_end(c);
        o_close_path(c);
        return;
    }
    a = a2;
    v.x = L_COSF(a) * radius + xc;
    v.y = L_SINF(a) * radius + yc;
    add_point(c, v.x, v.y);
    set_curve_end(c);
}
void o_arc_negative(octx *c, float xc, float yc, float radius, float a1, float a2) {
    while (a2 > a1)
        a2 -= 2.f * OVG_PIF;
    if (a1 - a2 > a1 + 2.f * OVG_PIF) /* reproduced as written in the reference */
        a2 = a1 - 2.f * OVG_PIF;
    v2 v = {L_COSF(a1) * radius + xc, L_SINF(a1) * radius + yc};
    float step = arc_step(c, radius);
    float a = a1;
    if (path_empty(c)) {
        set_curve_start(c);
        add_point(c, v.x, v.y);
        c->simpleConvex = !c->pathPtr;
    } else {
        line_to_i(c, v.x, v.y);
        set_curve_start(c);
        c->simpleConvex = false;
    }
    a -= step;
    if (EQUF(a2, a1))
        return;
    while (a > a2) {
        v.x = L_COSF(a) * radius + xc;
        v.y = L_SINF(a) * radius + yc;
        add_point(c, v.x, v.y);
        a -= step;
    }
    if (EQUF(a1 - a2, OVG_PIF * 2.f)) {
        set_curve_end(c);
        o_close_path(c);
        return;
    }
    a = a2;
    v.x = L_COSF(a) * radius + xc;
    v.y = L_SINF(a) * radius + yc;
    add_point(c, v.x, v.y);
    set_curve_end(c);
}
void o_curve_to(octx *c, float x1, float y1, float x2, float y2, float x3, float y3) {
    curve_to_i(c, x1, y1, x2, y2, x3, y3);
}
void o_rel_curve_to(octx *c, float x1, float y1, float x2, float y2, float x3, float y3) {
    if (path_empty(c)) {
        c->status = 4; /* VKVG_STATUS_NO_CURRENT_POINT */
        return;
    }
    v2 cp = cur_pos(c);
    curve_to_i(c, cp.x + x1, cp.y + y1, cp.x + x2, cp.y + y2, cp.x + x3, cp.y + y3);
}
void o_quadratic_to(octx *c, float x1, float y1, float x2, float y2) {
    float x0, y0;
    if (path_empty(c)) {
        x0 = x1;
        y0 = y1;
    } else {
        v2 cp = cur_pos(c);
        x0 = cp.x;
        y0 = cp.y;
    }
    const double f = 2.0 / 3.0;
    curve_to_i(c, (float)(x0 + (x1 - x0) * f), (float)(y0 + (y1 - y0) * f), (float)(x2 + (x1 - x2) * f),
               (float)(y2 + (y1 - y2) * f), x2, y2);
}
int o_rectangle(octx *c, float x, float y, float w, float h) {
    o_finish_path(c);
    if (w <= 0 || h <= 0)
        return 14; /* VKVG_STATUS_INVALID_RECT */
    add_point(c, x, y);
    add_point(c, x + w, y);
    add_point(c, x + w, y + h);
    add_point(c, x, y + h);
    c->pathes[c->pathPtr] |= (P_CLOSED | P_CONVEX);
    o_finish_path(c);
    return 0;
}
int o_rounded_rectangle(octx *c, float x, float y, float w, float h, float radius) {
    o_finish_path(c);
    if (w <= 0 || h <= 0)
        return 14;
    if ((radius > w / 2.0f) || (radius > h / 2.0f))
        radius = (float)fmin(w / 2.0f, h / 2.0f);
    o_move_to(c, x, y + radius);
    o_arc(c, x + radius, y + radius, radius, OVG_PIF, -OVG_PIF_2);
    o_line_to(c, x + w - radius, y);
    o_arc(c, x + w - radius, y + radius, radius, -OVG_PIF_2, 0);
    o_line_to(c, x + w, y + h - radius);
    o_arc(c, x + w - radius, y + h - radius, radius, 0, OVG_PIF_2);
    o_line_to(c, x + radius, y + h);
    o_arc(c, x + radius, y + h - radius, radius, OVG_PIF_2, OVG_PIF);
    o_line_to(c, x, y + radius);
    o_close_path(c);
    return 0;
}
/* vkvg_ellipse, src/vkvg_context.c:1605-1639 */
void o_ellipse(octx *c, float radiusX, float radiusY, float x, float y, float rot) {
    float wtt = radiusX * 4 / 3;
    float dx1 = L_SINF(rot) * radiusY, dy1 = L_COSF(rot) * radiusY;
    float dx2 = L_COSF(rot) * wtt, dy2 = L_SINF(rot) * wtt;
    float tcx = x - dx1, tcy = y + dy1;
    float trx = tcx + dx2, try_ = tcy + dy2;
    float tlx = tcx - dx2, tly = tcy - dy2;
    float bcx = x + dx1, bcy = y - dy1;
    float brx = bcx + dx2, bry = bcy + dy2;
    float blx = bcx - dx2, bly = bcy - dy2;
    o_finish_path(c);
    add_point(c, bcx, bcy);
    curve_to_i(c, brx, bry, trx, try_, tcx, tcy);
    curve_to_i(c, tlx, tly, blx, bly, bcx, bcy);
    c->pathes[c->pathPtr] |= P_CLOSED;
    o_finish_path(c);
}
/* _elliptic_arc, src/vkvg_context_internal.c:1473-1580 */
void o_elliptic_arc_to(octx *c, float x2, float y2, bool largeArc, bool ccw, float _rx, float _ry, float phi) {
    float x1 = 0, y1 = 0;
    if (!path_empty(c)) {
        x1 = cur_pos(c).x;
        y1 = cur_pos(c).y;
    }
    if (_rx == 0 || _ry == 0) {
        if (path_empty(c))
            o_move_to(c, x1, y1);
        o_line_to(c, x2, y2);
        return;
    }
    float rx = fabsf(_rx), ry = fabsf(_ry);
    float cph = L_COSF(phi), sph = L_SINF(phi);
    v2 p = {(x1 - x2) / 2, (y1 - y2) / 2};
    v2 p1 = {(cph * p.x) + (sph * p.y), (-sph * p.x) + (cph * p.y)};
    double lambda = L_POWF(p1.x, 2) / L_POWF(rx, 2) + L_POWF(p1.y, 2) / L_POWF(ry, 2);
    if (lambda > 1) {
        lambda = sqrtf((float)lambda);
        rx *= lambda;
        ry *= lambda;
    }
    p = V(rx * p1.y / ry, -ry * p1.x / rx);
    float num = L_POWF(rx, 2) * L_POWF(ry, 2) - L_POWF(rx, 2) * L_POWF(p1.y, 2) - L_POWF(ry, 2) * L_POWF(p1.x, 2);
    float den = L_POWF(rx, 2) * L_POWF(p1.y, 2) + L_POWF(ry, 2) * L_POWF(p1.x, 2);
    v2 cp = vmul(p, sqrtf(fabsf(num / den)));
    if (largeArc == ccw) {
        cp.x = -cp.x;
        cp.y = -cp.y;
    }
    p = V((x1 + x2) / 2, (y1 + y2) / 2);
    v2 ctr = vadd(V((cph * cp.x) + (-sph * cp.y), (sph * cp.x) + (cph * cp.y)), p);
    v2 u = {1.f, 0};
    v2 v = {(p1.x - cp.x) / rx, (p1.y - cp.y) / ry};
    double sa = L_ACOSF(vdot(u, v) / (fabsf(vlen(v)) * fabsf(vlen(u))));
    if (isnan(sa))
        sa = OVG_PIF;
    if (u.x * v.y - u.y * v.x < 0)
        sa = -sa;
    u = v;
    v = V((-p1.x - cp.x) / rx, (-p1.y - cp.y) / ry);
    double dt = L_ACOSF(vdot(u, v) / (fabsf(vlen(v)) * fabsf(vlen(u))));
    if (isnan(dt))
        dt = OVG_PIF;
    if (u.x * v.y - u.y * v.x < 0)
        dt = -dt;
    if (ccw) {
        if (dt < 0)
            dt += OVG_PIF * 2.0;
    } else if (dt > 0)
        dt -= OVG_PIF * 2.0;
    double theta = sa, ea = sa + dt;
    float step = fmaxf(0.001f, fminf(OVG_PIF, arc_step(c, fminf(rx, ry)) * 0.1f));
    p = V(rx * L_COSF((float)theta), ry * L_SINF((float)theta));
    v2 xy = vadd(V((cph * p.x) + (-sph * p.y), (sph * p.x) + (cph * p.y)), ctr);
    if (path_empty(c)) {
        set_curve_start(c);
        add_point(c, xy.x, xy.y);
        c->simpleConvex = !c->pathPtr;
    } else {
        line_to_i(c, xy.x, xy.y);
        set_curve_start(c);
        c->simpleConvex = false;
    }
    set_curve_start(c);
    if (sa < ea) {
        theta += step;
        while (theta < ea) {
            p = V(rx * L_COSF((float)theta), ry * L_SINF((float)theta));
            xy = vadd(V((cph * p.x) + (-sph * p.y), (sph * p.x) + (cph * p.y)), ctr);
            add_point(c, xy.x, xy.y);
            theta += step;
        }
    } else {
        theta -= step;
        while (theta > ea) {
            p = V(rx * L_COSF((float)theta), ry * L_SINF((float)theta));
            xy = vadd(V((cph * p.x) + (-sph * p.y), (sph * p.x) + (cph * p.y)), ctr);
            add_point(c, xy.x, xy.y);
            theta -= step;
        }
    }
    p = V(rx * L_COSF((float)ea), ry * L_SINF((float)ea));
    xy = vadd(V((cph * p.x) + (-sph * p.y), (sph * p.x) + (cph * p.y)), ctr);
    add_point(c, xy.x, xy.y);
    set_curve_end(c);
}
void o_set_dash(octx *c, const float *d, uint32_t n, float off) {
    free(c->dashes);
    c->dashes = NULL;
    c->dashCount = n;
    c->dashOffset = off;
    if (!n)
        return;
    c->dashes = (float *)malloc(n * sizeof(float));
    memcpy(c->dashes, d, n * sizeof(float));
}

/* ---------------------------------------------------------------- stroke geometry */
typedef struct {
    uint32_t iL, iR, cp, firstIdx;
    float hw, lhMax, arcStep;
} strk;
typedef struct {
    bool dashOn;
    uint32_t curDash;
    float curDashOffset, totDashLength;
    v2 normal;
} dashc;

static inline void addv(octx *c, v2 p) {
    ensure_verts(c, 1);
    c->vc[c->vertCount].x = p.x;
    c->vc[c->vertCount].y = p.y;
    c->vc[c->vertCount].color = c->curColor;
    c->vertCount++;
}
static inline void addt(octx *c, uint32_t a, uint32_t b, uint32_t d) {
    ensure_inds(c, 3);
    c->ic[c->indCount++] = a;
    c->ic[c->indCount++] = b;
    c->ic[c->indCount++] = d;
}
static inline void add_rect_tris(octx *c, uint32_t i) {
    addt(c, i, i + 2, i + 1);
    addt(c, i + 1, i + 2, i + 3);
}

/* _build_vb_step, src/vkvg_context_internal.c:924-1163 */
static bool vb_step(octx *c, strk *s, bool isCurve) {
    v2 p0 = c->points[s->cp];
    v2 v0 = vsub(p0, c->points[s->iL]);
    v2 v1 = vsub(c->points[s->iR], p0);
    float lv0 = vlen(v0), lv1 = vlen(v1);
    if (lv0 < FLT_EPSILON || lv1 < FLT_EPSILON)
        return false;
    v2 v0n = vdiv(v0, lv0), v1n = vdiv(v1, lv1);
    float dot = vdot(v0n, v1n);
    float det = v0n.x * v1n.y - v0n.y * v1n.x;
    if (EQUF(dot, 1.0f))
        return false;
    uint32_t idx = c->vertCount;
    if (EQUF(dot, -1.0f)) {
        v2 vPerp = vmul(vperp(v0n), s->hw);
        addv(c, vadd(p0, vPerp));
        addv(c, vsub(p0, vPerp));
        addt(c, idx, idx + 1, idx + 2);
        addt(c, idx, idx + 2, idx + 3);
        return true;
    }
    v2 bisec_n = vnorm(vadd(v0n, v1n));
    float alpha = L_ACOSF(dot);
    if (det < 0)
        alpha = -alpha;
    float halfAlpha = alpha / 2.f;
    float cosHalfAlpha = L_COSF(halfAlpha);
    float lh = s->hw / cosHalfAlpha;
    v2 bisec_n_perp = vperp(bisec_n);
    float rlh = lh;
    if (dot < 0.f)
        rlh = fminf(rlh, fminf(lv0, lv1));
    v2 bisec = vmul(bisec_n_perp, rlh);
    v2 in_pos, out_pos;
    if (rlh < lh) {
        v2 vnPerp = (lv0 < lv1) ? vperp(v1n) : vperp(v0n);
        v2 vHwPerp = vmul(vnPerp, s->hw);
        float lbc = cosHalfAlpha * rlh; /* stored in a double, always used as float */
        if (det < 0.f) {
            in_pos = vadd(vadd(vmul(vnPerp, -lbc), vadd(p0, bisec)), vHwPerp);
            out_pos = vsub(p0, vmul(bisec_n_perp, lh));
        } else {
            in_pos = vsub(vadd(vmul(vnPerp, lbc), vsub(p0, bisec)), vHwPerp);
            out_pos = vadd(p0, vmul(bisec_n_perp, lh));
        }
    } else {
        if (det < 0.0) {
            in_pos = vadd(p0, bisec);
            out_pos = vsub(p0, bisec);
        } else {
            in_pos = vsub(p0, bisec);
            out_pos = vadd(p0, bisec);
        }
    }
    int join = c->lineJoin;
    if (isCurve)
        join = (dot < 0.8f) ? 1 /*ROUND*/ : 0 /*MITER*/;

    if (join == 0) {
        if (lh > s->lhMax) {
            float x = (lh - s->lhMax) * cosHalfAlpha;
            v2 bisecPerp = vmul(bisec_n, x);
            bisec = vmul(bisec_n_perp, s->lhMax);
            if (det < 0) {
                addv(c, in_pos);
                v2 p = vsub(p0, bisec);
                addv(c, vsub(p, bisecPerp));
                addv(c, vadd(p, bisecPerp));
                addt(c, idx, idx + 2, idx + 1);
                addt(c, idx + 2, idx + 4, idx);
                addt(c, idx, idx + 3, idx + 4);
                return true;
            } else {
                v2 p = vadd(p0, bisec);
                addv(c, vsub(p, bisecPerp));
                addv(c, in_pos);
                addv(c, vadd(p, bisecPerp));
                addt(c, idx, idx + 2, idx + 1);
                addt(c, idx + 2, idx + 3, idx + 1);
                addt(c, idx + 1, idx + 3, idx + 4);
                return false;
            }
        } else {
            if (det < 0) {
                addv(c, in_pos);
                addv(c, out_pos);
            } else {
                addv(c, out_pos);
                addv(c, in_pos);
            }
            add_rect_tris(c, idx);
            return false;
        }
    } else {
        v2 vp = vperp(v0n);
        if (det < 0) {
            if (dot < 0 && rlh < lh)
                addv(c, in_pos);
            else
                addv(c, vadd(p0, bisec));
            addv(c, vsub(p0, vmul(vp, s->hw)));
        } else {
            addv(c, vadd(p0, vmul(vp, s->hw)));
            if (dot < 0 && rlh < lh)
                addv(c, in_pos);
            else
                addv(c, vsub(p0, bisec));
        }
        if (join == 2) { /* BEVEL */
            if (det < 0) {
                addt(c, idx, idx + 2, idx + 1);
                addt(c, idx + 2, idx + 4, idx + 0);
                addt(c, idx, idx + 3, idx + 4);
            } else {
                addt(c, idx, idx + 2, idx + 1);
                addt(c, idx + 2, idx + 3, idx + 1);
                addt(c, idx + 1, idx + 3, idx + 4);
            }
        } else if (join == 1) { /* ROUND */
            if (!s->arcStep)
                s->arcStep = arc_step(c, s->hw);
            float a = L_ACOSF(vp.x);
            if (vp.y < 0)
                a = -a;
            if (det < 0) {
                a += OVG_PIF;
                float a1 = a + alpha;
                a -= s->arcStep;
                while (a > a1) {
                    addv(c, V(L_COSF(a) * s->hw + p0.x, L_SINF(a) * s->hw + p0.y));
                    a -= s->arcStep;
                }
            } else {
                float a1 = a + alpha;
                a += s->arcStep;
                while (a < a1) {
                    addv(c, V(L_COSF(a) * s->hw + p0.x, L_SINF(a) * s->hw + p0.y));
                    a += s->arcStep;
                }
            }
            uint32_t p0Idx = c->vertCount;
            addt(c, idx, idx + 2, idx + 1);
            if (det < 0) {
                for (uint32_t p = idx + 2; p < p0Idx; p++)
                    addt(c, p, p + 1, idx);
                addt(c, p0Idx, p0Idx + 2, idx);
                addt(c, idx, p0Idx + 1, p0Idx + 2);
            } else {
                for (uint32_t p = idx + 2; p < p0Idx; p++)
                    addt(c, p, p + 1, idx + 1);
                addt(c, p0Idx, p0Idx + 1, idx + 1);
                addt(c, idx + 1, p0Idx + 1, p0Idx + 2);
            }
        }
        vp = vmul(vperp(v1n), s->hw);
        if (det < 0)
            addv(c, vsub(p0, vp));
        else
            addv(c, vadd(p0, vp));
    }
    return (det < 0);
}

/* _draw_stoke_cap, src/vkvg_context_internal.c:1165-1239 */
static void stroke_cap(octx *c, strk *s, v2 p0, v2 n, bool isStart) {
    uint32_t firstIdx = c->vertCount;
    if (isStart) {
        v2 vhw = vmul(n, s->hw);
        if (c->lineCap == 2)
            p0 = vsub(p0, vhw);
        vhw = vperp(vhw);
        if (c->lineCap == 1) {
            if (!s->arcStep)
                s->arcStep = arc_step(c, s->hw);
            float a = L_ACOSF(n.x) + OVG_PIF_2;
            if (n.y < 0)
                a = OVG_PIF - a;
            float a1 = a + OVG_PIF;
            a += s->arcStep;
            while (a < a1) {
                addv(c, V(L_COSF(a) * s->hw + p0.x, L_SINF(a) * s->hw + p0.y));
                a += s->arcStep;
            }
            uint32_t p0Idx = c->vertCount;
            for (uint32_t p = firstIdx; p < p0Idx; p++)
                addt(c, p0Idx + 1, p, p + 1);
            firstIdx = p0Idx;
        }
        addv(c, vadd(p0, vhw));
        addv(c, vsub(p0, vhw));
        add_rect_tris(c, firstIdx);
    } else {
        v2 vhw = vmul(n, s->hw);
        if (c->lineCap == 2)
            p0 = vadd(p0, vhw);
        vhw = vperp(vhw);
        addv(c, vadd(p0, vhw));
        addv(c, vsub(p0, vhw));
        firstIdx = c->vertCount;
        if (c->lineCap == 1) {
            if (!s->arcStep)
                s->arcStep = arc_step(c, s->hw);
            float a = L_ACOSF(n.x) + OVG_PIF_2;
            if (n.y < 0)
                a = OVG_PIF - a;
            float a1 = a - OVG_PIF;
            a -= s->arcStep;
            while (a > a1) {
                addv(c, V(L_COSF(a) * s->hw + p0.x, L_SINF(a) * s->hw + p0.y));
                a -= s->arcStep;
            }
            uint32_t p0Idx = c->vertCount - 1;
            for (uint32_t p = firstIdx - 1; p < p0Idx; p++)
                addt(c, p + 1, p, firstIdx - 2);
        }
    }
}

/* _draw_dashed_segment, :1240-1264 */
static void dashed_segment(octx *c, strk *s, dashc *dc, bool isCurve) {
    v2 p = c->points[s->cp];
    v2 pR = c->points[s->iR];
    if (!dc->dashOn)
        vb_step(c, s, isCurve);
    v2 d = vsub(pR, p);
    dc->normal = vnorm(d);
    float segLen = vlen(d);
    while (dc->curDashOffset < segLen) {
        v2 p0 = vadd(p, vmul(dc->normal, dc->curDashOffset));
        stroke_cap(c, s, p0, dc->normal, dc->dashOn);
        dc->dashOn ^= true;
        dc->curDashOffset += c->dashes[dc->curDash];
        if (++dc->curDash == c->dashCount)
            dc->curDash = 0;
    }
    dc->curDashOffset -= segLen;
    dc->curDashOffset = fmodf(dc->curDashOffset, dc->totDashLength);
}
/* _draw_segment, :1265-1282 (the 1/3-of-UINT32 batch split never triggers) */
static void draw_segment(octx *c, strk *s, dashc *dc, bool isCurve) {
    s->iR = s->cp + 1;
    if (c->dashCount > 0)
        dashed_segment(c, s, dc, isCurve);
    else
        vb_step(c, s, isCurve);
    s->iL = s->cp++;
}

/* _stroke_preserve, src/vkvg_context.c:822-948 (path already finished by the caller) */
void o_stroke_geometry(octx *c) {
    strk s;
    memset(&s, 0, sizeof(s));
    s.hw = c->lineWidth * 0.5f;
    s.lhMax = c->miterLimit * c->lineWidth;
    uint32_t ptrPath = 0;
    c->vertCount = c->indCount = 0;

    while (ptrPath < c->pathPtr) {
        uint32_t ptrSegment = 0, lastSegIdx = 0;
        uint32_t firstIdx = s.cp;
        uint32_t count = c->pathes[ptrPath] & P_MASK;
        uint32_t lastIdx = s.cp + count - 1;
        dashc dc;
        memset(&dc, 0, sizeof(dc));
        bool curved_path = (c->pathes[ptrPath] & P_CURVES) != 0;
        bool closed = (c->pathes[ptrPath] & P_CLOSED) != 0;
        if (curved_path) {
            ptrSegment = 1;
            lastSegIdx = s.cp + (c->pathes[ptrPath + ptrSegment] & P_MASK) - 1;
        }
        s.firstIdx = c->vertCount;
        if (c->dashCount > 0) {
            dc.dashOn = true;
            dc.curDash = 0;
            dc.totDashLength = 0;
            bool neg = false;
            for (uint32_t i = 0; i < c->dashCount; i++) {
                dc.totDashLength += c->dashes[i];
                neg |= !(c->dashes[i] >= 0.f);
            }
            /* the reference rejects a zero total; negative entries and totals below 1/1024 px,
             * on which its dash walk would run (almost) forever, are rejected the same way */
            if (dc.totDashLength == 0 || neg || !(dc.totDashLength >= 1.0f / 1024.0f)) {
                c->status = 13; /* VKVG_STATUS_INVALID_DASH */
                return;
            }
            dc.curDashOffset = fmodf(c->dashOffset, dc.totDashLength);
            s.iL = lastIdx;
        } else if (closed) {
            s.iL = lastIdx;
        } else {
            stroke_cap(c, &s, c->points[s.cp], vline_norm(c->points[s.cp], c->points[s.cp + 1]), true);
            s.iL = s.cp++;
        }
        bool walked = false;
        if (curved_path) {
            while (s.cp < lastIdx) {
                walked = true;
                bool curved = (c->pathes[ptrPath + ptrSegment] & P_CURVES) != 0;
                if (lastSegIdx == lastIdx)
                    lastSegIdx--;
                while (s.cp <= lastSegIdx)
                    draw_segment(c, &s, &dc, curved);
                ptrSegment++;
                uint32_t cnt = c->pathes[ptrPath + ptrSegment] & P_MASK;
                lastSegIdx = s.cp + cnt - 1;
                if (lastSegIdx == lastIdx && cnt == 1) {
                    ptrSegment++;
                    break;
                }
            }
        } else {
            while (s.cp < lastIdx)
                draw_segment(c, &s, &dc, false);
        }
        if (c->dashCount > 0) {
            if (closed) {
                s.iR = firstIdx;
                dashed_segment(c, &s, &dc, false);
                s.iL++;
                s.cp++;
            }
            if (!dc.dashOn) {
                int32_t prevDash = (int32_t)dc.curDash - 1;
                if (prevDash < 0)
                    dc.curDash = c->dashCount - 1;
                /* the reference then reads dashes[-1] (undefined); read the wrapped entry */
                float dprev = c->dashes[prevDash < 0 ? (int32_t)c->dashCount - 1 : prevDash];
                float m = fminf(dprev - dc.curDashOffset, c->dashes[dc.curDash]);
                v2 p = vsub(c->points[s.iR], vmul(dc.normal, m));
                stroke_cap(c, &s, p, dc.normal, false);
            }
        } else if (closed) {
            s.iR = firstIdx;
            bool inverse = vb_step(c, &s, false);
            uint32_t *inds = &c->ic[c->indCount - 6];
            uint32_t ii = s.firstIdx;
            if (inverse) {
                inds[1] = ii + 1;
                inds[4] = ii + 1;
                inds[5] = ii;
            } else {
                inds[1] = ii;
                inds[4] = ii;
                inds[5] = ii + 1;
            }
            s.cp++;
        } else {
            stroke_cap(c, &s, c->points[s.cp], vline_norm(c->points[s.cp - 1], c->points[s.cp]), false);
        }
        s.cp = firstIdx + count;
        if (curved_path && !walked) {
            /* Deviation: an open, solid, curved sub-path of two points never enters the segment
             * walk above, so the reference (src/vkvg_context.c:935-938) advances by one record
             * and strokes the curve-segment record as a further sub-path made of the points that
             * follow in its point buffer: leftovers of earlier paths or uninitialised memory.
             * Here the sub-path's segment records are skipped the way the fill walks them
             * (src/vkvg_context_internal.c:1645-1652). */
            uint32_t q = ptrPath + 1, tot = 0;
            while (tot < count && q < c->pathPtr)
                tot += c->pathes[q++] & P_MASK;
            ptrPath = q;
        } else if (ptrSegment > 0)
            ptrPath += ptrSegment;
        else
            ptrPath++;
    }
}
