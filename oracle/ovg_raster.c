/*
 * CPU oracle — Vulkan fixed-function semantics used by vkvg (test infrastructure only).
 *
 * Vertex path:  shaders/vkvg_main.vert:59-81 (p = mat * pos, ndc = p*2/size - 1), then the
 *               viewport transform x_f = ndc*(W/2) + W/2 and a 1/256 sub-pixel snap with
 *               round-to-nearest-even (hypothesis H-snap; lavapipe's subPixelPrecisionBits = 8).
 * Sampling:     one sample at the pixel centre, top-left tie rule in y-down framebuffer
 *               coordinates (hypothesis H-tie; Vulkan spec "Polygon rasterization").
 * Stencil:      src/vkvg_device_internal.c:226-246 — poly-fill INVERTs FILL where CLIP==0,
 *               pipe_OVER passes where CLIP==0 (or FILL==1 for the even-odd cover) and zeroes FILL.
 * Colour:       vertex colour UNORM8 -> float, gradient math of shaders/vkvg_main.frag:83-144,
 *               c *= opacity (:150-154), float -> UNORM8 with round-to-nearest-even (H-conv),
 *               premultiplied OVER in 8-bit: d' = sat(s + round(d*(255-sa)/255)) (H-blend).
 * Fills:        even-odd = fan INVERT + scissored cover (src/vkvg_context_internal.c:1583-1655,
 *               :1919-1952, src/vkvg_context.c:796-821); non-zero = convex fan fast path, else the
 *               region [winding != 0] that glutess triangulates (:1720-1793; glutess itself is not
 *               in the reference tree, so its exact triangles are not reproduced).
 */
#include "ovg_internal.h"

/* ------------------------------------------------------------------ helpers */
static inline int64_t fdiv(int64_t a, int64_t b) { /* floor(a/b), b != 0 */
    int64_t q = a / b, r = a % b;
    if (r != 0 && ((r < 0) != (b < 0)))
        q--;
    return q;
}
static inline int64_t cdiv(int64_t a, int64_t b) { return -fdiv(-a, b); }

uint32_t ovg_draw_key(uint32_t n) {
    uint32_t x = n + 0x9E3779B9u;
    x ^= x >> 16;
    x *= 0x85EBCA6Bu;
    x ^= x >> 13;
    x *= 0xC2B2AE35u;
    x ^= x >> 16;
    return x | 1u;
}

/* CreateRgbaf, src/vkvg_context_internal.h:60-66 (VKVG_PREMULT_ALPHA) */
uint32_t o_rgbaf(float r, float g, float b, float a) {
    return ((((uint32_t)(a * 255.0f) & 0xFF) << 24) | (((uint32_t)(b * a * 255.0f) & 0xFF) << 16) |
            (((uint32_t)(g * a * 255.0f) & 0xFF) << 8) | ((uint32_t)(r * a * 255.0f) & 0xFF));
}

static inline uint32_t mul255(uint32_t a, uint32_t b) {
    uint32_t t = a * b + 128u;
    return (t + (t >> 8)) >> 8;
}
void ovg_blend_over(uint8_t *d, const uint8_t *s) {
    uint32_t ia = 255u - s[3];
    for (int i = 0; i < 4; i++) {
        uint32_t v = s[i] + mul255(d[i], ia);
        d[i] = (uint8_t)(v > 255u ? 255u : v);
    }
}
static inline void blend_op(int op, uint8_t *d, const uint8_t *s) {
    if (op == OP_CLEAR) {
        d[0] = d[1] = d[2] = d[3] = 0;
    } else if (op == OP_DIFFERENCE) { /* BLEND_OP_SUBTRACT: s - d*(1-sa), clamped at 0 */
        uint32_t ia = 255u - s[3];
        for (int i = 0; i < 4; i++) {
            int v = (int)s[i] - (int)mul255(d[i], ia);
            d[i] = (uint8_t)(v < 0 ? 0 : v);
        }
    } else {
        ovg_blend_over(d, s);
    }
}
static inline uint8_t f2u8(float f) {
    if (!(f > 0.f))
        return 0;
    if (f >= 1.f)
        return 255;
    return (uint8_t)rintf(f * 255.0f);
}

/* ------------------------------------------------------------------ vertex transform */
void ovg_snap_vertex(float x, float y, uint32_t w, uint32_t h, const float *m, int32_t *fx, int32_t *fy) {
    float px = (m[0] * x + m[2] * y) + m[4];
    float py = (m[1] * x + m[3] * y) + m[5];
    float sw = (float)w, sh = (float)h;
    float nx = (px * 2.0f) / sw - 1.0f;
    float ny = (py * 2.0f) / sh - 1.0f;
    float hw = sw * 0.5f, hh = sh * 0.5f;
    float X = nx * hw + hw;
    float Y = ny * hh + hh;
    /* supported coordinate range: |x|,|y| <= 32768 px (24.8 fixed point) */
    if (isnan(X) || isnan(Y)) {
        *fx = INT32_MIN;
        *fy = INT32_MIN;
        return;
    }
    X = fminf(fmaxf(X, -32768.0f), 32768.0f);
    Y = fminf(fmaxf(Y, -32768.0f), 32768.0f);
    *fx = (int32_t)rintf(X * 256.0f);
    *fy = (int32_t)rintf(Y * 256.0f);
}
static inline bool vinvalid(int32_t x) { return x == INT32_MIN; }

/* ------------------------------------------------------------------ triangle coverage
 * A sample S is covered when every edge function, oriented positive inside, is > 0, or is
 * == 0 on a top edge (horizontal, interior below in y-down) or a left edge (interior to +x). */
static inline int orient_fix(int32_t *x, int32_t *y) {
    int64_t a2 = (int64_t)(x[1] - x[0]) * (y[2] - y[0]) - (int64_t)(y[1] - y[0]) * (x[2] - x[0]);
    if (a2 == 0)
        return 0;
    if (a2 < 0) {
        int32_t t = x[1]; x[1] = x[2]; x[2] = t;
        t = y[1]; y[1] = y[2]; y[2] = t;
        return -1;
    }
    return 1;
}
int ovg_sample_in_triangle(const int32_t *fx, const int32_t *fy, int32_t sx, int32_t sy) {
    int32_t x[3] = {fx[0], fx[1], fx[2]}, y[3] = {fy[0], fy[1], fy[2]};
    if (vinvalid(x[0]) || vinvalid(x[1]) || vinvalid(x[2]))
        return 0;
    if (!orient_fix(x, y))
        return 0;
    for (int e = 0; e < 3; e++) {
        int a = e, b = (e + 1) % 3;
        int64_t dx = x[b] - x[a], dy = y[b] - y[a];
        int64_t E = dx * ((int64_t)sy - y[a]) - dy * ((int64_t)sx - x[a]);
        if (E > 0)
            continue;
        if (E < 0)
            return 0;
        bool top = (dy == 0 && dx > 0);  /* horizontal, interior below */
        bool left = (dy < 0);            /* interior towards +x */
        if (!(top || left))
            return 0;
    }
    return 1;
}
void ovg_raster_triangle_brute(const int32_t *fx, const int32_t *fy, uint32_t w, uint32_t h, int32_t *cov) {
    for (uint32_t r = 0; r < h; r++)
        for (uint32_t c = 0; c < w; c++)
            if (ovg_sample_in_triangle(fx, fy, 256 * (int32_t)c + 128, 256 * (int32_t)r + 128))
                cov[(size_t)r * w + c] += 1;
}
/* per-row span of a positively oriented triangle; returns false when the row is empty */
static bool tri_row_span(const int32_t *x, const int32_t *y, int64_t sy, int64_t *lo, int64_t *hi) {
    int64_t clo = INT64_MIN / 4, chi = INT64_MAX / 4;
    for (int e = 0; e < 3; e++) {
        int a = e, b = (e + 1) % 3;
        int64_t dx = x[b] - x[a], dy = y[b] - y[a];
        /* E(c) = K - D*c with sample x = 256c + 128 */
        int64_t K = dx * (sy - y[a]) - dy * (128 - (int64_t)x[a]);
        int64_t D = 256 * dy;
        if (dy == 0) {
            if (!(K > 0 || (K == 0 && dx > 0)))
                return false;
        } else if (dy > 0) { /* not a tie edge: K - D c > 0 */
            int64_t m = cdiv(K, D) - 1;
            if (m < chi)
                chi = m;
        } else { /* left edge, tie inside: K - D c >= 0 */
            int64_t m = cdiv(K, D);
            if (m > clo)
                clo = m;
        }
    }
    *lo = clo;
    *hi = chi;
    return clo <= chi;
}
typedef struct { int32_t x0, y0, x1, y1; } bbox_t;
__thread int32_t ovg_row_lo = 0, ovg_row_hi = INT32_MAX;
static inline void band_rows(int64_t *r0, int64_t *r1) {
    if (*r0 < ovg_row_lo)
        *r0 = ovg_row_lo;
    if (*r1 > (int64_t)ovg_row_hi - 1)
        *r1 = (int64_t)ovg_row_hi - 1;
}
static void raster_tri(const int32_t *fx, const int32_t *fy, uint32_t w, uint32_t h, int32_t *cov, bbox_t *bb) {
    int32_t x[3] = {fx[0], fx[1], fx[2]}, y[3] = {fy[0], fy[1], fy[2]};
    if (vinvalid(x[0]) || vinvalid(x[1]) || vinvalid(x[2]))
        return;
    if (!orient_fix(x, y))
        return;
    int32_t ymin = y[0], ymax = y[0];
    for (int i = 1; i < 3; i++) {
        if (y[i] < ymin) ymin = y[i];
        if (y[i] > ymax) ymax = y[i];
    }
    int64_t r0 = cdiv((int64_t)ymin - 128, 256), r1 = fdiv((int64_t)ymax - 128, 256);
    if (r0 < 0) r0 = 0;
    if (r1 > (int64_t)h - 1) r1 = (int64_t)h - 1;
    band_rows(&r0, &r1);
    for (int64_t r = r0; r <= r1; r++) {
        int64_t lo, hi;
        if (!tri_row_span(x, y, 256 * r + 128, &lo, &hi))
            continue;
        if (lo < 0) lo = 0;
        if (hi > (int64_t)w - 1) hi = (int64_t)w - 1;
        if (lo > hi)
            continue;
        int32_t *row = cov + (size_t)r * w;
        for (int64_t c = lo; c <= hi; c++)
            row[c] += 1;
        if (bb) {
            if ((int32_t)lo < bb->x0) bb->x0 = (int32_t)lo;
            if ((int32_t)hi > bb->x1) bb->x1 = (int32_t)hi;
            if ((int32_t)r < bb->y0) bb->y0 = (int32_t)r;
            if ((int32_t)r > bb->y1) bb->y1 = (int32_t)r;
        }
    }
}
void ovg_raster_triangle_scan(const int32_t *fx, const int32_t *fy, uint32_t w, uint32_t h, int32_t *cov) {
    raster_tri(fx, fy, w, h, cov, NULL);
}

/* ------------------------------------------------------------------ winding
 * Ray to -x from the sample perturbed by (+eps, +eps^2); an edge counts +1 when it runs
 * towards +y. An edge crosses sample row sy iff ytop <= sy < ybot; the sample lies right of
 * it iff its column c >= ceil((dyb*(xtop-128) + (xbot-xtop)*(sy-ytop)) / (256*dyb)). */
typedef struct { int32_t xa, ya, xb, yb; int dir; } edge_t;
static void winding_edges(const edge_t *E, size_t ne, uint32_t w, uint32_t h, int32_t *wind, bbox_t *bb) {
    int32_t *diff = (int32_t *)calloc(w + 1, sizeof(int32_t));
    int32_t ymin = INT32_MAX, ymax = INT32_MIN;
    for (size_t i = 0; i < ne; i++) {
        if (E[i].ya < ymin) ymin = E[i].ya;
        if (E[i].yb > ymax) ymax = E[i].yb;
    }
    if (ne == 0) {
        free(diff);
        return;
    }
    int64_t r0 = cdiv((int64_t)ymin - 128, 256), r1 = fdiv((int64_t)ymax - 128, 256);
    if (r0 < 0) r0 = 0;
    if (r1 > (int64_t)h - 1) r1 = (int64_t)h - 1;
    band_rows(&r0, &r1);
    /* rows are independent: sort-free difference-array sweep per row */
    for (int64_t r = r0; r <= r1; r++) {
        int64_t sy = 256 * r + 128;
        bool any = false;
        int64_t cmin = w, cmax = 0;
        for (size_t i = 0; i < ne; i++) {
            const edge_t *e = &E[i];
            if (!(e->ya <= sy && sy < e->yb))
                continue;
            int64_t dyb = (int64_t)e->yb - e->ya;
            int64_t N = dyb * ((int64_t)e->xa - 128) + ((int64_t)e->xb - e->xa) * (sy - e->ya);
            int64_t c = cdiv(N, 256 * dyb);
            if (c < 0) c = 0;
            if (c > (int64_t)w) c = w;
            diff[c] += e->dir;
            if (c < cmin) cmin = c;
            if (c > cmax) cmax = c;
            any = true;
        }
        if (!any)
            continue;
        int32_t run = 0;
        int32_t *row = wind + (size_t)r * w;
        for (int64_t c = cmin; c < (int64_t)w; c++) {
            run += diff[c];
            diff[c] = 0;
            if (run != 0) {
                row[c] += run;
                if (bb) {
                    if ((int32_t)c < bb->x0) bb->x0 = (int32_t)c;
                    if ((int32_t)c > bb->x1) bb->x1 = (int32_t)c;
                    if ((int32_t)r < bb->y0) bb->y0 = (int32_t)r;
                    if ((int32_t)r > bb->y1) bb->y1 = (int32_t)r;
                }
            } else if (c >= cmax)
                break;
        }
        diff[w] = 0;
        for (int64_t c = cmin; c <= cmax && c < (int64_t)w; c++)
            diff[c] = 0;
    }
    free(diff);
}
static void push_edge(edge_t **E, size_t *n, size_t *cap, int32_t x0, int32_t y0, int32_t x1, int32_t y1) {
    if (vinvalid(x0) || vinvalid(x1) || y0 == y1)
        return;
    if (*n == *cap) {
        *cap = *cap ? *cap * 2 : 1024;
        *E = (edge_t *)realloc(*E, *cap * sizeof(edge_t));
    }
    edge_t e;
    if (y0 < y1) {
        e.xa = x0; e.ya = y0; e.xb = x1; e.yb = y1; e.dir = 1;
    } else {
        e.xa = x1; e.ya = y1; e.xb = x0; e.yb = y0; e.dir = -1;
    }
    (*E)[(*n)++] = e;
}
void ovg_winding_scan(const int32_t *px, const int32_t *py, uint32_t n, uint32_t w, uint32_t h, int32_t *wind) {
    edge_t *E = NULL;
    size_t ne = 0, cap = 0;
    for (uint32_t i = 0; i < n; i++) {
        uint32_t j = (i + 1) % n;
        push_edge(&E, &ne, &cap, px[i], py[i], px[j], py[j]);
    }
    winding_edges(E, ne, w, h, wind, NULL);
    free(E);
}
void ovg_winding_fan(const int32_t *px, const int32_t *py, uint32_t n, uint32_t w, uint32_t h, int32_t *wind) {
    int32_t *cov = (int32_t *)calloc((size_t)w * h, sizeof(int32_t));
    for (uint32_t i = 1; i + 1 < n; i++) {
        int32_t tx[3] = {px[0], px[i], px[i + 1]}, ty[3] = {py[0], py[i], py[i + 1]};
        int64_t a2 = (int64_t)(tx[1] - tx[0]) * (ty[2] - ty[0]) - (int64_t)(ty[1] - ty[0]) * (tx[2] - tx[0]);
        if (a2 == 0)
            continue;
        memset(cov, 0, (size_t)w * h * sizeof(int32_t));
        ovg_raster_triangle_brute(tx, ty, w, h, cov);
        int s = a2 > 0 ? -1 : 1; /* positive area (y-down) winds -1 under the +y-edge convention */
        for (size_t k = 0; k < (size_t)w * h; k++)
            wind[k] += s * cov[k];
    }
    free(cov);
}

/* ------------------------------------------------------------------ sources */
static inline float o_clampf(float x) { return fminf(fmaxf(x, 0.f), 1.f); }
static inline float smoothstepf(float e0, float e1, float x) {
    float t = o_clampf((x - e0) / (e1 - e0));
    return t * t * (3.f - 2.f * t);
}
static inline void mix4(float *c, const float *d, float a) {
    for (int i = 0; i < 4; i++)
        c[i] = c[i] * (1.f - a) + d[i] * a;
}
/* shaders/vkvg_main.frag:83-144, p = gl_FragCoord.xy / size */
static void grad_color(const octx *c, float fx, float fy, float *col) {
    const ovg_grad *g = &c->grad;
    float W = (float)c->surf->width, H = (float)c->surf->height;
    float px = fx / W, py = fy / H;
    if (g->type == SRC_LINEAR) {
        float dist = 1;
        float p0x = g->cp[0][0] / W, p0y = g->cp[0][1] / H;
        float p1x = g->cp[0][2] / W, p1y = g->cp[0][3] / H;
        float dx = p1x - p0x, dy = p1y - p0y;
        float l = sqrtf(dx * dx + dy * dy);
        float ux = dx / l, uy = dy / l; /* normalize(v) = v / length(v) */
        if (uy == 0) {
            if (ux < 0)
                dist = -(px - p0x) / l;
            else
                dist = (px - p0x) / l;
        } else {
            float m = -ux / uy;
            float bb = p0y - m * p0x;
            dist = ((py - m * px - bb) / sqrtf(1 + m * m)) / l;
            if (uy < 0)
                dist = -dist;
        }
        for (int i = 0; i < 4; i++)
            col[i] = g->colors[0][i];
        mix4(col, g->colors[1], smoothstepf(g->stops[0], g->stops[1], dist));
        for (int i = 1; i < (int)g->count - 1; ++i)
            mix4(col, g->colors[i + 1], smoothstepf(g->stops[i], g->stops[i + 1], dist));
    } else {
        float c0x = g->cp[0][0] / W, c0y = g->cp[0][1] / H;
        float c1x = g->cp[1][0] / W, c1y = g->cp[1][1] / H;
        float r0 = g->cp[0][2] / W, r1 = g->cp[1][2] / W;
        float gradLength = 1.0f;
        float diffx = c0x - c1x, diffy = c0y - c1y;
        float rx = px - c0x, ry = py - c0y;
        float rl = sqrtf(rx * rx + ry * ry);
        float rdx = rx / rl, rdy = ry / rl;
        float a = rdx * rdx + rdy * rdy;
        float b = 2.0f * (rdx * diffx + rdy * diffy);
        float cc = (diffx * diffx + diffy * diffy) - r1 * r1;
        float disc = b * b - 4.0f * a * cc;
        if (disc >= 0.0f) {
            float t = (-b + sqrtf(fabsf(disc))) / (2.0f * a);
            float prx = c0x + rdx * t, pry = c0y + rdy * t;
            float ddx = prx - c0x, ddy = pry - c0y;
            gradLength = sqrtf(ddx * ddx + ddy * ddy) - r0;
        }
        float qx = px - c0x, qy = py - c0y;
        float grad = (sqrtf(qx * qx + qy * qy) - r0) / gradLength;
        for (int i = 0; i < 4; i++)
            col[i] = g->colors[0][i];
        mix4(col, g->colors[1], smoothstepf(g->stops[0], g->stops[1], grad));
        for (int i = 2; i < (int)g->count; i++)
            mix4(col, g->colors[i], smoothstepf(g->stops[i - 1], g->stops[i], grad));
    }
}
/* ---- surface sources: texture(source, uv) of shaders/vkvg_main.frag:72-82 with the sampler that
 * src/vkvg_context_internal.c:730-755 creates, evaluated with the Vulkan specification's formulas
 * for unnormalized texel coordinates (u' = u * width; NEAREST: floor(u'); LINEAR: floor(u' - 0.5)
 * and its fraction, the four texels weighted ((w00 t00 + w10 t10) + w01 t01) + w11 t11) */
enum { EXT_NONE = 0, EXT_REPEAT = 1, EXT_REFLECT = 2, EXT_PAD = 3 };
static int64_t floor_index(float x) { return (int64_t)fminf(fmaxf(floorf(x), -1073741824.f), 1073741824.f); }
/* address modes: NONE = CLAMP_TO_BORDER, REPEAT, REFLECT = MIRRORED_REPEAT, PAD = CLAMP_TO_EDGE;
 * -1 for the border */
static int64_t wrap_index(int64_t i, int64_t n, uint32_t extend) {
    switch (extend) {
    case EXT_REPEAT:
        return ((i % n) + n) % n;
    case EXT_REFLECT: {
        int64_t t = ((i % (2 * n)) + 2 * n) % (2 * n);
        return t < n ? t : 2 * n - 1 - t;
    }
    case EXT_PAD:
        return i < 0 ? 0 : (i >= n ? n - 1 : i);
    default:
        return (i < 0 || i >= n) ? -1 : i;
    }
}
static void texel(const ovg_source *src, int64_t i, int64_t j, uint32_t extend, float *t) {
    int64_t wi = wrap_index(i, src->width, extend), wj = wrap_index(j, src->height, extend);
    if (wi < 0 || wj < 0) { /* border colour: transparent black */
        t[0] = t[1] = t[2] = t[3] = 0.f;
        return;
    }
    const uint8_t *q = src->rgba + 4 * ((size_t)wj * src->width + (size_t)wi);
    for (int k = 0; k < 4; k++)
        t[k] = (float)q[k] / 255.0f;
}
static void surface_color(const octx *c, float fx, float fy, float *col) {
    const ovg_surf_pat *sp = &c->surfPat;
    const ovg_source *src = &c->sources[sp->index];
    const mat6 *m = &c->matInv;
    const float px = fx - c->srcRect[0], py = fy - c->srcRect[1];
    float u = m->xx * px + m->xy * py + m->x0;
    float v = m->yx * px + m->yy * py + m->y0;
    u = u / c->srcRect[2];
    v = v / c->srcRect[3];
    const float uu = u * (float)src->width, vv = v * (float)src->height;
    if (sp->filter == 2 || sp->filter == 4) { /* VKVG_FILTER_BEST, VKVG_FILTER_BILINEAR: LINEAR */
        const float a = uu - 0.5f, b = vv - 0.5f;
        const float fa = floorf(a), fb = floorf(b);
        const float al = a - fa, be = b - fb;
        const int64_t i0 = floor_index(a), j0 = floor_index(b);
        float t00[4], t10[4], t01[4], t11[4];
        texel(src, i0, j0, sp->extend, t00);
        texel(src, i0 + 1, j0, sp->extend, t10);
        texel(src, i0, j0 + 1, sp->extend, t01);
        texel(src, i0 + 1, j0 + 1, sp->extend, t11);
        const float w00 = (1.f - al) * (1.f - be), w10 = al * (1.f - be), w01 = (1.f - al) * be, w11 = al * be;
        for (int k = 0; k < 4; k++)
            col[k] = ((w00 * t00[k] + w10 * t10[k]) + w01 * t01[k]) + w11 * t11[k];
    } else {
        texel(src, floor_index(uu), floor_index(vv), sp->extend, col);
    }
}

static void src_at(const octx *c, uint32_t color, int x, int y, uint8_t *s) {
    float col[4];
    if (c->srcType == SRC_SURFACE) {
        surface_color(c, (float)x + 0.5f, (float)y + 0.5f, col);
    } else if (c->srcType == SRC_SOLID) {
        col[0] = (float)(color & 0xFF) / 255.0f;
        col[1] = (float)((color >> 8) & 0xFF) / 255.0f;
        col[2] = (float)((color >> 16) & 0xFF) / 255.0f;
        col[3] = (float)((color >> 24) & 0xFF) / 255.0f;
    } else {
        grad_color(c, (float)x + 0.5f, (float)y + 0.5f, col);
    }
    for (int i = 0; i < 4; i++)
        s[i] = f2u8(col[i] * c->opacity);
}

/* ------------------------------------------------------------------ draw emission */
static void composite_counts(octx *c, uint32_t color, int32_t *cov, bbox_t bb, bool once) {
    ovg_surface *S = c->surf;
    uint32_t key = ovg_draw_key(S->draw_counter);
    if (bb.x0 > bb.x1)
        return;
    for (int32_t r = bb.y0; r <= bb.y1; r++) {
        for (int32_t x = bb.x0; x <= bb.x1; x++) {
            size_t i = (size_t)r * S->width + x;
            int32_t k = cov[i];
            if (!k)
                continue;
            cov[i] = 0;
            if (once)
                k = 1;
            if (k < 0)
                k = -k; /* only reachable for the triangle-count path, where k >= 0 */
            if (S->stencil[i] & 2)
                continue; /* clipped (CLIP bit set) */
            uint8_t s[4];
            src_at(c, color, x, r, s);
            for (int32_t t = 0; t < k; t++)
                blend_op(c->op, S->rgba + 4 * i, s);
            S->stencil[i] &= (uint8_t)~1u;
            S->sig[i] += (uint32_t)k * key;
            S->frag_count += (uint64_t)k;
        }
    }
}

static void snap_all(octx *c, int32_t *X, int32_t *Y) {
    const float m[6] = {c->mat.xx, c->mat.yx, c->mat.xy, c->mat.yy, c->mat.x0, c->mat.y0};
    for (uint32_t i = 0; i < c->vertCount; i++)
        ovg_snap_vertex(c->vc[i].x, c->vc[i].y, c->surf->width, c->surf->height, m, &X[i], &Y[i]);
}

/* triangles in ic[] over vertices vc[] composite once per covering triangle */
static void draw_triangles(octx *c) {
    ovg_surface *S = c->surf;
    int32_t *X = (int32_t *)malloc((c->vertCount + 1) * sizeof(int32_t));
    int32_t *Y = (int32_t *)malloc((c->vertCount + 1) * sizeof(int32_t));
    snap_all(c, X, Y);
    bbox_t bb = {INT32_MAX, INT32_MAX, INT32_MIN, INT32_MIN};
    for (uint32_t t = 0; t + 2 < c->indCount; t += 3) {
        uint32_t i0 = c->ic[t], i1 = c->ic[t + 1], i2 = c->ic[t + 2];
        if (i0 >= c->vertCount || i1 >= c->vertCount || i2 >= c->vertCount)
            continue; /* dangling forward reference (never drawn by a valid stream) */
        int32_t tx[3] = {X[i0], X[i1], X[i2]}, ty[3] = {Y[i0], Y[i1], Y[i2]};
        raster_tri(tx, ty, S->width, S->height, S->acc, &bb);
        S->tri_count++;
    }
    composite_counts(c, c->curColor, S->acc, bb, false);
    free(X);
    free(Y);
}

/* collect the subpaths with more than 2 points: (first point, count) pairs */
static uint32_t list_subpaths(octx *c, uint32_t *first, uint32_t *cnt, uint32_t cap) {
    uint32_t ptrPath = 0, firstPt = 0, n = 0;
    while (ptrPath < c->pathPtr) {
        uint32_t count = c->pathes[ptrPath] & P_MASK;
        if (count > 2 && n < cap) {
            first[n] = firstPt;
            cnt[n] = count;
            n++;
        }
        firstPt += count;
        if (c->pathes[ptrPath] & P_CURVES) {
            ptrPath++;
            uint32_t tot = 0;
            while (tot < count)
                tot += (c->pathes[ptrPath++] & P_MASK);
        } else
            ptrPath++;
    }
    return n;
}

/* _poly_fill, src/vkvg_context_internal.c:1583-1655 with polyFillOpState
 * (src/vkvg_device_internal.c:226-232): one triangle fan per subpath INVERTs the FILL bit where
 * CLIP == 0. Accumulates the CTM-transformed point bounds (starting from `bounds`) and the pixel
 * bbox of the fans. */
static void eo_fans(octx *c, const uint32_t *first, const uint32_t *cnt, uint32_t nsp, float *bounds, bbox_t *bb) {
    ovg_surface *S = c->surf;
    const float m[6] = {c->mat.xx, c->mat.yx, c->mat.xy, c->mat.yy, c->mat.x0, c->mat.y0};
    for (uint32_t s = 0; s < nsp; s++) {
        uint32_t n = cnt[s];
        int32_t *X = (int32_t *)malloc(n * sizeof(int32_t)), *Y = (int32_t *)malloc(n * sizeof(int32_t));
        for (uint32_t i = 0; i < n; i++) {
            v2 p = c->points[first[s] + i];
            ovg_snap_vertex(p.x, p.y, S->width, S->height, m, &X[i], &Y[i]);
            float bx = p.x, by = p.y;
            m_transform_point(&c->mat, &bx, &by);
            if (bx < bounds[0]) bounds[0] = bx;
            if (bx > bounds[2]) bounds[2] = bx;
            if (by < bounds[1]) bounds[1] = by;
            if (by > bounds[3]) bounds[3] = by;
        }
        for (uint32_t i = 1; i + 1 < n; i++) {
            int32_t tx[3] = {X[0], X[i], X[i + 1]}, ty[3] = {Y[0], Y[i], Y[i + 1]};
            bbox_t tb = {INT32_MAX, INT32_MAX, INT32_MIN, INT32_MIN};
            raster_tri(tx, ty, S->width, S->height, S->acc, &tb);
            S->tri_count++;
            if (tb.x0 > tb.x1)
                continue;
            for (int32_t r = tb.y0; r <= tb.y1; r++)
                for (int32_t x = tb.x0; x <= tb.x1; x++) {
                    size_t k = (size_t)r * S->width + x;
                    if (S->acc[k]) {
                        if (!(S->stencil[k] & 2))
                            S->stencil[k] ^= 1u;
                        S->acc[k] = 0;
                    }
                }
            if (tb.x0 < bb->x0) bb->x0 = tb.x0;
            if (tb.x1 > bb->x1) bb->x1 = tb.x1;
            if (tb.y0 < bb->y0) bb->y0 = tb.y0;
            if (tb.y1 > bb->y1) bb->y1 = tb.y1;
        }
        free(X);
        free(Y);
    }
}

/* the region [winding != 0] over all contours (what glutess triangulates,
 * src/vkvg_context_internal.c:1747-1793), added into S->acc */
static void nz_winding(octx *c, const uint32_t *first, const uint32_t *cnt, uint32_t nsp, bbox_t *bb) {
    ovg_surface *S = c->surf;
    const float m[6] = {c->mat.xx, c->mat.yx, c->mat.xy, c->mat.yy, c->mat.x0, c->mat.y0};
    edge_t *E = NULL;
    size_t ne = 0, ecap = 0;
    for (uint32_t s = 0; s < nsp; s++) {
        uint32_t n = cnt[s];
        int32_t *X = (int32_t *)malloc(n * sizeof(int32_t)), *Y = (int32_t *)malloc(n * sizeof(int32_t));
        for (uint32_t i = 0; i < n; i++) {
            v2 p = c->points[first[s] + i];
            ovg_snap_vertex(p.x, p.y, S->width, S->height, m, &X[i], &Y[i]);
        }
        for (uint32_t i = 0; i < n; i++) {
            uint32_t j = (i + 1) % n;
            push_edge(&E, &ne, &ecap, X[i], Y[i], X[j], Y[j]);
        }
        free(X);
        free(Y);
    }
    winding_edges(E, ne, S->width, S->height, S->acc, bb);
    free(E);
}

static bool convex_fast_path(const octx *c) { return c->pathPtr == 1 && (c->pathes[0] & P_CONVEX); }

void r_fill(octx *c) {
    ovg_surface *S = c->surf;
    o_finish_path(c);
    if (!c->pathPtr)
        return;
    uint32_t cap = c->pathPtr + 1;
    uint32_t *first = (uint32_t *)malloc(cap * sizeof(uint32_t)), *cnt = (uint32_t *)malloc(cap * sizeof(uint32_t));
    uint32_t nsp = list_subpaths(c, first, cnt, cap);

    if (c->fillRule == 0) {
        /* EVEN_ODD: fans INVERT FILL, bounds start as src/vkvg_context.c:806 has them */
        float bounds[4] = {FLT_MAX, FLT_MAX, FLT_MIN, FLT_MIN};
        bbox_t bb = {INT32_MAX, INT32_MAX, INT32_MIN, INT32_MIN};
        eo_fans(c, first, cnt, nsp, bounds, &bb);
        const float xMin = bounds[0], yMin = bounds[1], xMax = bounds[2], yMax = bounds[3];
        /* cover: full-screen triangle scissored to the bounds, passes where FILL == 1 */
        int32_t sx = (int32_t)fmaxf(xMin, 0), sy = (int32_t)fmaxf(yMin, 0);
        int32_t sw = (int32_t)fmaxf(xMax - (float)(int32_t)xMin + 1, 1);
        int32_t sh = (int32_t)fmaxf(yMax - (float)(int32_t)yMin + 1, 1);
        int64_t x0 = sx, y0 = sy, x1 = (int64_t)sx + sw - 1, y1 = (int64_t)sy + sh - 1;
        if (x1 > (int64_t)S->width - 1) x1 = (int64_t)S->width - 1;
        if (y1 > (int64_t)S->height - 1) y1 = (int64_t)S->height - 1;
        band_rows(&y0, &y1);
        uint32_t key = ovg_draw_key(S->draw_counter);
        for (int64_t r = y0; r <= y1; r++)
            for (int64_t x = x0; x <= x1; x++) {
                size_t k = (size_t)r * S->width + x;
                if ((S->stencil[k] & 1) == 1) {
                    uint8_t s[4];
                    src_at(c, c->curColor, (int)x, (int)r, s);
                    blend_op(c->op, S->rgba + 4 * k, s);
                    S->stencil[k] &= (uint8_t)~1u;
                    S->sig[k] += key;
                    S->frag_count++;
                }
            }
    } else if (convex_fast_path(c)) {
        /* convex fast path: fan (0, i-1, i) drawn with pipe_OVER */
        uint32_t n = c->pathes[0] & P_MASK;
        c->vertCount = c->indCount = 0;
        for (uint32_t i = 0; i < n; i++) {
            c->vc[c->vertCount].x = c->points[i].x;
            c->vc[c->vertCount].y = c->points[i].y;
            c->vc[c->vertCount].color = c->curColor;
            c->vertCount++;
            if (c->vertCount + 8 >= c->sizeV) {
                c->sizeV *= 2;
                c->vc = (vtx *)realloc(c->vc, c->sizeV * sizeof(vtx));
            }
        }
        for (uint32_t i = 2; i < n; i++) {
            if (c->indCount + 8 >= c->sizeI) {
                c->sizeI *= 2;
                c->ic = (uint32_t *)realloc(c->ic, c->sizeI * sizeof(uint32_t));
            }
            c->ic[c->indCount++] = 0;
            c->ic[c->indCount++] = i - 1;
            c->ic[c->indCount++] = i;
        }
        draw_triangles(c);
    } else {
        /* NON_ZERO: the region with winding != 0 over all contours, composited once */
        bbox_t bb = {INT32_MAX, INT32_MAX, INT32_MIN, INT32_MIN};
        nz_winding(c, first, cnt, nsp, &bb);
        composite_counts(c, c->curColor, S->acc, bb, true);
    }
    S->draw_counter++;
    free(first);
    free(cnt);
}

/* _clip_preserve's stencil passes, src/vkvg_context.c:754-795 with clipingOpState
 * (src/vkvg_device_internal.c:233-239):
 *   1. FILL := coverage of the path where CLIP == 0 (even-odd: the _poly_fill fans; non-zero:
 *      pipelineClipping with reference FILL, compare mask CLIP, write mask FILL over the
 *      _fill_non_zero triangles, whose fail op ZEROes FILL where CLIP == 1);
 *   2. a full-screen quad with reference CLIP, compare mask FILL, write mask FILL|CLIP: where
 *      FILL == 0 the pass op REPLACEs the two bits with CLIP, elsewhere the fail op ZEROes them.
 * The clip therefore only ever shrinks: CLIP' = CLIP | !coverage. The pixels that keep CLIP == 0
 * fold the draw's key into the signature (the 1-sample mask of the clip). */
void r_clip(octx *c) {
    ovg_surface *S = c->surf;
    o_finish_path(c);
    if (!c->pathPtr)
        return;
    uint32_t cap = c->pathPtr + 1;
    uint32_t *first = (uint32_t *)malloc(cap * sizeof(uint32_t)), *cnt = (uint32_t *)malloc(cap * sizeof(uint32_t));
    uint32_t nsp = list_subpaths(c, first, cnt, cap);
    const size_t np = (size_t)S->width * S->height;
    if (c->fillRule == 0) {
        float bounds[4] = {FLT_MAX, FLT_MAX, FLT_MIN, FLT_MIN}; /* not used: no scissor */
        bbox_t bb = {INT32_MAX, INT32_MAX, INT32_MIN, INT32_MIN};
        eo_fans(c, first, cnt, nsp, bounds, &bb);
    } else {
        bbox_t bb = {INT32_MAX, INT32_MAX, INT32_MIN, INT32_MIN};
        if (convex_fast_path(c)) {
            const float m[6] = {c->mat.xx, c->mat.yx, c->mat.xy, c->mat.yy, c->mat.x0, c->mat.y0};
            uint32_t n = c->pathes[0] & P_MASK;
            int32_t *X = (int32_t *)malloc((n + 1) * sizeof(int32_t)), *Y = (int32_t *)malloc((n + 1) * sizeof(int32_t));
            for (uint32_t i = 0; i < n; i++)
                ovg_snap_vertex(c->points[i].x, c->points[i].y, S->width, S->height, m, &X[i], &Y[i]);
            for (uint32_t i = 2; i < n; i++) {
                int32_t tx[3] = {X[0], X[i - 1], X[i]}, ty[3] = {Y[0], Y[i - 1], Y[i]};
                raster_tri(tx, ty, S->width, S->height, S->acc, &bb);
                S->tri_count++;
            }
            free(X);
            free(Y);
        } else {
            nz_winding(c, first, cnt, nsp, &bb);
        }
        if (bb.x0 <= bb.x1)
            for (int32_t r = bb.y0; r <= bb.y1; r++)
                for (int32_t x = bb.x0; x <= bb.x1; x++) {
                    size_t k = (size_t)r * S->width + x;
                    if (S->acc[k] && !(S->stencil[k] & 2))
                        S->stencil[k] |= 1u; /* REPLACE with reference FILL, write mask FILL */
                    S->acc[k] = 0;
                }
    }
    uint32_t key = ovg_draw_key(S->draw_counter);
    for (size_t k = 0; k < np; k++) {
        if (S->stencil[k] & 1) {
            S->stencil[k] &= (uint8_t)~3u;
            S->sig[k] += key;
        } else {
            S->stencil[k] = (uint8_t)((S->stencil[k] & ~3u) | 2u);
        }
    }
    S->draw_counter++;
    c->clipState = OVG_CLIP_CLIP;
    free(first);
    free(cnt);
}

void r_stroke(octx *c) {
    o_finish_path(c);
    if (!c->pathPtr)
        return;
    o_stroke_geometry(c);
    if (c->status)
        return;
    draw_triangles(c);
    c->surf->draw_counter++;
}

void r_paint(octx *c) {
    o_finish_path(c);
    if (c->pathPtr) {
        r_fill(c);
        o_clear_path(c);
        return;
    }
    ovg_surface *S = c->surf;
    uint32_t key = ovg_draw_key(S->draw_counter);
    int64_t pr0 = 0, pr1 = (int64_t)S->height - 1;
    band_rows(&pr0, &pr1);
    for (int64_t r = pr0; r <= pr1; r++)
        for (uint32_t x = 0; x < S->width; x++) {
            size_t k = (size_t)r * S->width + x;
            if (S->stencil[k] & 2)
                continue;
            uint8_t s[4];
            src_at(c, c->curColor, (int)x, (int)r, s);
            blend_op(c->op, S->rgba + 4 * k, s);
            S->stencil[k] &= (uint8_t)~1u;
            S->sig[k] += key;
        }
    S->frag_count += (uint64_t)S->width * S->height;
    S->draw_counter++;
}

void r_clear(octx *c) {
    ovg_surface *S = c->surf;
    size_t n = (size_t)S->width * S->height;
    memset(S->rgba, 0, n * 4);
    memset(S->stencil, 0, n);
    uint32_t key = ovg_draw_key(S->draw_counter);
    for (size_t k = 0; k < n; k++)
        S->sig[k] += key;
    S->draw_counter++;
}

/* src/vkvg_surface.c:441-452: p = (uchar)((double)p / ((double)a / 255)), truncating */
void ovg_unpremultiply(const uint8_t *in, uint8_t *out, size_t npix) {
    for (size_t i = 0; i < npix; i++) {
        const uint8_t *p = in + 4 * i;
        uint8_t *q = out + 4 * i;
        double alpha = (double)p[3] / 255.f;
        for (int k = 0; k < 3; k++) {
            double v = (double)p[k] / alpha;
            /* a == 0 divides by zero: x86 converts the NaN/inf to INT_MIN, whose low byte is 0 */
            q[k] = (p[3] == 0) ? 0 : (uint8_t)(int32_t)v;
        }
        q[3] = p[3];
    }
}
