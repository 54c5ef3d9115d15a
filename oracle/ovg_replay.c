/*
 * CPU oracle — scene command stream interpreter (test infrastructure only).
 * Maps each opcode onto the oracle's restatement of the corresponding vkvg entry point
 * (src/vkvg_context.c) and src/vkvg_pattern.c for gradients.
 */
#include "ovg_internal.h"

static inline float F(const uint32_t *w) {
    float f;
    memcpy(&f, w, 4);
    return f;
}

/* vkvg_pattern_edit_radial, src/vkvg_pattern.c:95-119 */
static void edit_radial(ovg_grad *g, float cx0, float cy0, float radius0, float cx1, float cy1, float radius1) {
    float c0x = cx0, c0y = cy0, c1x = cx1, c1y = cy1;
    if (radius0 > radius1 - 1.0f)
        radius0 = radius1 - 1.0f;
    float ux = c0x - c1x, uy = c0y - c1y;
    float l = sqrtf(ux * ux + uy * uy);
    if (l + radius0 + 1.0f >= radius1) {
        float vx = ux / l, vy = uy / l;
        float k = radius1 - radius0 - 1.0f;
        c0x = c1x + vx * k;
        c0y = c1y + vy * k;
    }
    g->cp[0][0] = c0x; g->cp[0][1] = c0y; g->cp[0][2] = radius0; g->cp[0][3] = 0;
    g->cp[1][0] = c1x; g->cp[1][1] = c1y; g->cp[1][2] = radius1; g->cp[1][3] = 0;
}
int ovg_render_stream(uint32_t width, uint32_t height, const uint32_t *w, size_t nw, uint8_t *out_rgba,
                      uint32_t *out_sig, uint64_t *out_stats) {
    return ovg_render_stream_src(width, height, w, nw, NULL, 0, out_rgba, out_sig, out_stats);
}

/* Copies the finished path (and, for a stroke, the stroke's vertex and index caches) of the
 * first fill / stroke of a stream into `d`, without rasterizing it. */
static void dump_geometry(octx *c, ovg_geometry *d, bool stroke) {
    o_finish_path(c);
    d->n_points = c->pointCount;
    for (uint32_t k = 0; k < c->pointCount && k < d->cap_points; k++) {
        d->points[2 * k] = c->points[k].x;
        d->points[2 * k + 1] = c->points[k].y;
    }
    d->n_pathes = c->pathPtr;
    for (uint32_t k = 0; k < c->pathPtr && k < d->cap_pathes; k++)
        d->pathes[k] = c->pathes[k];
    d->n_verts = d->n_inds = 0;
    if (!stroke)
        return;
    o_stroke_geometry(c);
    d->n_verts = c->vertCount;
    for (uint32_t k = 0; k < c->vertCount && k < d->cap_verts; k++) {
        d->verts[2 * k] = c->vc[k].x;
        d->verts[2 * k + 1] = c->vc[k].y;
    }
    d->n_inds = c->indCount;
    for (uint32_t k = 0; k < c->indCount && k < d->cap_inds; k++)
        d->inds[k] = c->ic[k];
}

static int run_stream(uint32_t width, uint32_t height, const uint32_t *w, size_t nw, const ovg_source *sources,
                      uint32_t n_sources, uint8_t *out_rgba, uint32_t *out_sig, uint64_t *out_stats,
                      ovg_geometry *dump) {
    ovg_surface S;
    memset(&S, 0, sizeof(S));
    S.width = width;
    S.height = height;
    size_t np = (size_t)width * height;
    S.rgba = (uint8_t *)calloc(np * 4, 1);
    S.stencil = (uint8_t *)calloc(np, 1);
    S.sig = (uint32_t *)calloc(np, sizeof(uint32_t));
    S.acc = (int32_t *)calloc(np, sizeof(int32_t));
    octx c;
    octx_init(&c, &S);
    c.sources = sources;
    c.nSources = n_sources;
    c.srcRect[3] = 1.f; /* {.a = 1}, src/vkvg_context.c:29 */
    c.matInv = c.mat;
    uint32_t n_patterns = 0;

    size_t i = 0;
    while (i < nw && !c.status) {
        uint32_t op = w[i++];
        const uint32_t *a = w + i;
        switch (op) {
        case OVG_OP_END: i = nw; break;
        case OVG_OP_MOVE_TO: o_move_to(&c, F(a), F(a + 1)); i += 2; break;
        case OVG_OP_LINE_TO: o_line_to(&c, F(a), F(a + 1)); i += 2; break;
        case OVG_OP_CURVE_TO: o_curve_to(&c, F(a), F(a + 1), F(a + 2), F(a + 3), F(a + 4), F(a + 5)); i += 6; break;
        case OVG_OP_REL_CURVE_TO:
            o_rel_curve_to(&c, F(a), F(a + 1), F(a + 2), F(a + 3), F(a + 4), F(a + 5));
            i += 6;
            break;
        case OVG_OP_CLOSE_PATH: o_close_path(&c); break;
        case OVG_OP_ARC: o_arc(&c, F(a), F(a + 1), F(a + 2), F(a + 3), F(a + 4)); i += 5; break;
        case OVG_OP_ARC_NEG: o_arc_negative(&c, F(a), F(a + 1), F(a + 2), F(a + 3), F(a + 4)); i += 5; break;
        case OVG_OP_RECTANGLE: o_rectangle(&c, F(a), F(a + 1), F(a + 2), F(a + 3)); i += 4; break;
        case OVG_OP_ROUNDED_RECT: o_rounded_rectangle(&c, F(a), F(a + 1), F(a + 2), F(a + 3), F(a + 4)); i += 5; break;
        case OVG_OP_ELLIPSE: o_ellipse(&c, F(a), F(a + 1), F(a + 2), F(a + 3), F(a + 4)); i += 5; break;
        case OVG_OP_ELLIPTIC_ARC_TO:
            o_elliptic_arc_to(&c, F(a), F(a + 1), a[2] != 0, a[3] != 0, F(a + 4), F(a + 5), F(a + 6));
            i += 7;
            break;
        case OVG_OP_NEW_PATH: o_new_path(&c); break;
        case OVG_OP_NEW_SUB_PATH: o_new_sub_path(&c); break;
        case OVG_OP_REL_LINE_TO: o_rel_line_to(&c, F(a), F(a + 1)); i += 2; break;
        case OVG_OP_REL_MOVE_TO: o_rel_move_to(&c, F(a), F(a + 1)); i += 2; break;
        case OVG_OP_QUADRATIC_TO: o_quadratic_to(&c, F(a), F(a + 1), F(a + 2), F(a + 3)); i += 4; break;
        case OVG_OP_FILL:
        case OVG_OP_FILL_PRESERVE:
        case OVG_OP_STROKE:
        case OVG_OP_STROKE_PRESERVE:
            if (dump) {
                dump_geometry(&c, dump, op == OVG_OP_STROKE || op == OVG_OP_STROKE_PRESERVE);
                i = nw;
                break;
            }
            if (op == OVG_OP_FILL || op == OVG_OP_FILL_PRESERVE)
                r_fill(&c);
            else
                r_stroke(&c);
            if (op == OVG_OP_FILL || op == OVG_OP_STROKE)
                o_clear_path(&c);
            break;
        case OVG_OP_PAINT: r_paint(&c); break;
        case OVG_OP_CLEAR: o_clear(&c); break;
        case OVG_OP_CLIP: o_clip_preserve(&c); o_clear_path(&c); break;
        case OVG_OP_CLIP_PRESERVE: o_clip_preserve(&c); break;
        case OVG_OP_RESET_CLIP: o_reset_clip(&c); break;
        case OVG_OP_SAVE: o_save(&c); break;
        case OVG_OP_RESTORE: o_restore(&c); break;
        case OVG_OP_FLUSH: break; /* frame boundary: no state change */
        case OVG_OP_ROUNDED_RECT2: { /* vkvg_rounded_rectangle2, src/vkvg_context.c:665-684 */
            const float x = F(a), y = F(a + 1), w2 = F(a + 2), h2 = F(a + 3), rx = F(a + 4), ry = F(a + 5);
            o_move_to(&c, x + rx, y);
            o_line_to(&c, x + w2 - rx, y);
            o_elliptic_arc_to(&c, x + w2, y + ry, false, true, rx, ry, 0);
            o_line_to(&c, x + w2, y + h2 - ry);
            o_elliptic_arc_to(&c, x + w2 - rx, y + h2, false, true, rx, ry, 0);
            o_line_to(&c, x + rx, y + h2);
            o_elliptic_arc_to(&c, x, y + h2 - ry, false, true, rx, ry, 0);
            o_line_to(&c, x, y + ry);
            o_elliptic_arc_to(&c, x + rx, y, false, true, rx, ry, 0);
            o_close_path(&c);
            i += 6;
            break;
        }
        case OVG_OP_SET_SOURCE_SURFACE: /* src/vkvg_context.c:1025-1033 */
        case OVG_OP_SOURCE_PATTERN: {
            if (a[0] >= c.nSources) {
                c.status = 7; /* VKVG_STATUS_INVALID_INDEX */
                break;
            }
            ovg_surf_pat sp;
            memset(&sp, 0, sizeof(sp));
            sp.index = a[0];
            if (op == OVG_OP_SET_SOURCE_SURFACE) {
                c.srcRect[0] = F(a + 1);
                c.srcRect[1] = F(a + 2);
                i += 3;
            } else {
                sp.extend = a[1];
                sp.filter = a[2];
                sp.hasMatrix = a[3] != 0;
                mat6 pm = {F(a + 4), F(a + 5), F(a + 6), F(a + 7), F(a + 8), F(a + 9)};
                sp.matrix = pm;
                i += 10;
            }
            c.patId = ++n_patterns;
            o_set_surface_pattern(&c, &sp);
            break;
        }
        case OVG_OP_SET_SOURCE_RGBA:
            c.curColor = o_rgbaf(F(a), F(a + 1), F(a + 2), F(a + 3));
            c.srcType = SRC_SOLID;
            c.patId = 0;
            i += 4;
            break;
        case OVG_OP_SET_SOURCE_RGB:
            c.curColor = o_rgbaf(F(a), F(a + 1), F(a + 2), 1);
            c.srcType = SRC_SOLID;
            c.patId = 0;
            i += 3;
            break;
        case OVG_OP_SET_SOURCE_COLOR: c.curColor = a[0]; c.srcType = SRC_SOLID; c.patId = 0; i += 1; break;
        case OVG_OP_SET_LINE_WIDTH: c.lineWidth = F(a); i += 1; break;
        case OVG_OP_SET_LINE_JOIN: c.lineJoin = (int)a[0]; i += 1; break;
        case OVG_OP_SET_LINE_CAP: c.lineCap = (int)a[0]; i += 1; break;
        case OVG_OP_SET_MITER_LIMIT: c.miterLimit = F(a); i += 1; break;
        case OVG_OP_SET_FILL_RULE: c.fillRule = (int)a[0]; i += 1; break;
        case OVG_OP_SET_DASH: {
            uint32_t n = a[0];
            float off = F(a + 1);
            float d[64];
            uint32_t nn = n > 64 ? 64 : n;
            for (uint32_t k = 0; k < nn; k++)
                d[k] = F(a + 2 + k);
            o_set_dash(&c, d, nn, off);
            i += 2 + n;
            break;
        }
        case OVG_OP_SET_OPACITY: c.opacity = F(a); i += 1; break;
        case OVG_OP_SET_OPERATOR: c.op = (int)a[0]; i += 1; break;
        case OVG_OP_LINEAR:
        case OVG_OP_RADIAL:
        case OVG_OP_LINEAR_MATRIX:
        case OVG_OP_RADIAL_MATRIX: {
            ovg_grad g;
            memset(&g, 0, sizeof(g));
            if (op == OVG_OP_LINEAR_MATRIX || op == OVG_OP_RADIAL_MATRIX) { /* vkvg_pattern_set_matrix */
                g.hasMatrix = true;
                g.matrix = (mat6){F(a), F(a + 1), F(a + 2), F(a + 3), F(a + 4), F(a + 5)};
                a += 6;
                i += 6;
                op = op == OVG_OP_LINEAR_MATRIX ? OVG_OP_LINEAR : OVG_OP_RADIAL;
            }
            uint32_t ncp = (op == OVG_OP_LINEAR) ? 4 : 6;
            if (op == OVG_OP_LINEAR) {
                g.type = SRC_LINEAR;
                g.cp[0][0] = F(a); g.cp[0][1] = F(a + 1); g.cp[0][2] = F(a + 2); g.cp[0][3] = F(a + 3);
            } else {
                g.type = SRC_RADIAL;
                edit_radial(&g, F(a), F(a + 1), F(a + 2), F(a + 3), F(a + 4), F(a + 5));
            }
            uint32_t ns = a[ncp];
            const uint32_t *st = a + ncp + 1;
            for (uint32_t k = 0; k < ns; k++) {
                if (g.count >= 16)
                    break; /* the reference writes past its 16-entry arrays here */
                float off = F(st + 5 * k), r = F(st + 5 * k + 1), gg = F(st + 5 * k + 2), b = F(st + 5 * k + 3),
                      al = F(st + 5 * k + 4);
                g.colors[g.count][0] = al * r;
                g.colors[g.count][1] = al * gg;
                g.colors[g.count][2] = al * b;
                g.colors[g.count][3] = al;
                g.stops[g.count] = off;
                g.count++;
            }
            c.patId = ++n_patterns; /* every gradient op sets a new pattern */
            c.gradDef = g;
            o_set_gradient(&c, &g);
            i += ncp + 1 + 5 * ns;
            break;
        }
        case OVG_OP_TRANSLATE: {
            mat6 t = {1, 0, 0, 1, F(a), F(a + 1)};
            m_multiply(&c.mat, &t, &c.mat);
            i += 2;
            o_update_mat_inv(&c);
            break;
        }
        case OVG_OP_SCALE: {
            mat6 t = {F(a), 0, 0, F(a + 1), 0, 0};
            m_multiply(&c.mat, &t, &c.mat);
            i += 2;
            o_update_mat_inv(&c);
            break;
        }
        case OVG_OP_ROTATE: {
            float s = L_SINF(F(a)), co = L_COSF(F(a));
            mat6 t = {co, s, -s, co, 0, 0};
            m_multiply(&c.mat, &t, &c.mat);
            i += 1;
            o_update_mat_inv(&c);
            break;
        }
        case OVG_OP_IDENTITY_MATRIX: {
            mat6 t = {1, 0, 0, 1, 0, 0};
            c.mat = t;
            o_update_mat_inv(&c);
            break;
        }
        case OVG_OP_SET_MATRIX: {
            mat6 t = {F(a), F(a + 1), F(a + 2), F(a + 3), F(a + 4), F(a + 5)};
            c.mat = t;
            i += 6;
            o_update_mat_inv(&c);
            break;
        }
        default:
            c.status = 6; /* VKVG_STATUS_INVALID_STATUS: unknown opcode */
            break;
        }
    }
    int status = c.status;
    if (out_rgba)
        memcpy(out_rgba, S.rgba, np * 4);
    if (out_sig)
        memcpy(out_sig, S.sig, np * sizeof(uint32_t));
    if (out_stats) {
        out_stats[0] = S.draw_counter;
        out_stats[1] = S.tri_count;
        out_stats[2] = S.frag_count;
    }
    o_state_free(&c);
    octx_free(&c);
    free(S.rgba);
    free(S.stencil);
    free(S.sig);
    free(S.acc);
    return status;
}

int ovg_render_stream_src(uint32_t width, uint32_t height, const uint32_t *w, size_t nw, const ovg_source *sources,
                          uint32_t n_sources, uint8_t *out_rgba, uint32_t *out_sig, uint64_t *out_stats) {
    return run_stream(width, height, w, nw, sources, n_sources, out_rgba, out_sig, out_stats, NULL);
}

int ovg_render_stream_rows(uint32_t width, uint32_t height, const uint32_t *w, size_t nw, uint32_t row0,
                           uint32_t row1, uint8_t *out_rgba, uint32_t *out_sig, uint64_t *out_stats) {
    ovg_row_lo = (int32_t)row0;
    ovg_row_hi = (int32_t)row1;
    int st = run_stream(width, height, w, nw, NULL, 0, out_rgba, out_sig, out_stats, NULL);
    ovg_row_lo = 0;
    ovg_row_hi = INT32_MAX;
    return st;
}

int ovg_geometry_dump(uint32_t width, uint32_t height, const uint32_t *w, size_t nw, ovg_geometry *out) {
    return run_stream(width, height, w, nw, NULL, 0, NULL, NULL, NULL, out);
}
