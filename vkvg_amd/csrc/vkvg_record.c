/*
 * Recording of a context's API calls and their replay (include/vkvg.h:1960-1970; the reference's
 * src/recording/vkvg_record.c and vkvg_record_internal.c).
 *
 * Between vkvg_start_recording and vkvg_stop_recording a context executes nothing: each call that
 * the reference records is appended as a command code plus its operands; vkvg_replay calls the
 * public entry points again, in order, on any context. The store is one command array and one
 * byte buffer; vkvg_recording_get_command returns a command's code and the byte offset of its
 * operands in the buffer that vkvg_recording_get_data returns, as the reference does.
 *
 * Operand layout (all packed, little-endian, unaligned): floats as IEEE binary32, enums and counts
 * as uint32, booleans as one byte each, patterns and surfaces as the handle (the recording holds a
 * reference until it is destroyed), matrices as the 24-byte vkvg_matrix_t, text NUL-terminated.
 *
 * Where the reference's recording loses calls, this one keeps them (each is a replay defect there):
 *   vkvg_set_miter_limit is recorded as VKVG_CMD_SET_LINE_WIDTH (src/vkvg_context.c:1050) -> here
 *     VKVG_CMD_SET_MITER_LIMIT;
 *   VKVG_CMD_ARC_NEG replays as vkvg_arc (src/recording/vkvg_record_internal.c:329-331) -> here as
 *     vkvg_arc_negative;
 *   VKVG_CMD_SET_DASH stores no operands (:160-161) but its replay reads a count, an offset and the
 *     dashes (:308-309) -> here those operands are stored;
 *   VKVG_CMD_RESET_CLIP is recorded (src/vkvg_context.c:723) and never replayed -> replayed here;
 *   vkvg_ellipse and vkvg_set_opacity are executed on the recording context instead of recorded ->
 *     recorded here (VKVG_CMD_ELLIPSE, VKVG_CMD_SET_OPACITY).
 */
#include <stdarg.h>
#include <stdlib.h>
#include <string.h>

#include "vkvg_record.h"

typedef struct {
    uint16_t cmd;
    size_t   dataOffset;
} rec_cmd;

struct _vkvg_recording_t {
    rec_cmd *cmds;
    uint32_t count, cap;
    size_t   size, reserved;
    char    *buf;
};

vkvg_rec *vkvg_rec_new(void) { return (vkvg_rec *)calloc(1, sizeof(vkvg_rec)); }

void vkvg_rec_free(vkvg_rec *rec) {
    if (!rec)
        return;
    for (uint32_t i = 0; i < rec->count; i++) {
        const char *d = rec->buf + rec->cmds[i].dataOffset;
        if (rec->cmds[i].cmd == VKVG_CMD_SET_SOURCE) {
            VkvgPattern p;
            memcpy(&p, d, sizeof(p));
            vkvg_pattern_destroy(p);
        } else if (rec->cmds[i].cmd == VKVG_CMD_SET_SOURCE_SURFACE) {
            VkvgSurface s;
            memcpy(&s, d + 2 * sizeof(float), sizeof(s));
            vkvg_surface_destroy(s);
        }
    }
    free(rec->cmds);
    free(rec->buf);
    free(rec);
}

void vkvg_rec_push(vkvg_rec *rec, uint32_t cmd, const void *data, size_t n) {
    if (rec->count == rec->cap) {
        uint32_t cap = rec->cap ? rec->cap * 2 : 64;
        rec_cmd *p = (rec_cmd *)realloc(rec->cmds, cap * sizeof(rec_cmd));
        if (!p)
            return;
        rec->cmds = p;
        rec->cap = cap;
    }
    if (rec->size + n > rec->reserved) {
        size_t cap = rec->reserved ? rec->reserved : 1024;
        while (cap < rec->size + n)
            cap *= 2;
        char *p = (char *)realloc(rec->buf, cap);
        if (!p)
            return;
        rec->buf = p;
        rec->reserved = cap;
    }
    rec->cmds[rec->count].cmd = (uint16_t)cmd;
    rec->cmds[rec->count].dataOffset = rec->size;
    rec->count++;
    if (n)
        memcpy(rec->buf + rec->size, data, n);
    rec->size += n;
}

void vkvg_rec_floats(vkvg_rec *rec, uint32_t cmd, uint32_t n, ...) {
    float v[8];
    va_list ap;
    va_start(ap, n);
    for (uint32_t i = 0; i < n && i < 8; i++)
        v[i] = (float)va_arg(ap, double);
    va_end(ap);
    vkvg_rec_push(rec, cmd, v, n * sizeof(float));
}
void vkvg_rec_u32(vkvg_rec *rec, uint32_t cmd, uint32_t v) { vkvg_rec_push(rec, cmd, &v, sizeof(v)); }
void vkvg_rec_elliptic(vkvg_rec *rec, uint32_t cmd, float x, float y, float rx, float ry, float phi, bool large,
                       bool sweep) {
    char d[5 * sizeof(float) + 2];
    float f[5] = {x, y, rx, ry, phi};
    memcpy(d, f, sizeof(f));
    d[20] = (char)large;
    d[21] = (char)sweep;
    vkvg_rec_push(rec, cmd, d, sizeof(d));
}
void vkvg_rec_dash(vkvg_rec *rec, const float *dashes, uint32_t n, float offset) {
    size_t bytes = 8 + (size_t)n * sizeof(float);
    char *d = (char *)malloc(bytes);
    if (!d)
        return;
    memcpy(d, &n, 4);
    memcpy(d + 4, &offset, 4);
    if (n && dashes)
        memcpy(d + 8, dashes, (size_t)n * sizeof(float));
    else if (n)
        memset(d + 8, 0, (size_t)n * sizeof(float));
    vkvg_rec_push(rec, VKVG_CMD_SET_DASH, d, bytes);
    free(d);
}
void vkvg_rec_pattern(vkvg_rec *rec, VkvgPattern pat) {
    vkvg_pattern_reference(pat);
    vkvg_rec_push(rec, VKVG_CMD_SET_SOURCE, &pat, sizeof(pat));
}
void vkvg_rec_surface(vkvg_rec *rec, VkvgSurface surf, float x, float y) {
    char d[2 * sizeof(float) + sizeof(VkvgSurface)];
    vkvg_surface_reference(surf);
    memcpy(d, &x, 4);
    memcpy(d + 4, &y, 4);
    memcpy(d + 8, &surf, sizeof(surf));
    vkvg_rec_push(rec, VKVG_CMD_SET_SOURCE_SURFACE, d, sizeof(d));
}
void vkvg_rec_text(vkvg_rec *rec, uint32_t cmd, const char *text) {
    const char *t = text ? text : "";
    vkvg_rec_push(rec, cmd, t, strlen(t) + 1);
}

/* ---------------------------------------------------------------- public API */
void vkvg_start_recording(VkvgContext ctx) {
    if (vkvg_status(ctx))
        return;
    vkvg_rec **slot = vkvg_ctx_recording(ctx);
    vkvg_rec_free(*slot);
    *slot = vkvg_rec_new();
}
VkvgRecording vkvg_stop_recording(VkvgContext ctx) {
    if (vkvg_status(ctx))
        return NULL;
    vkvg_rec **slot = vkvg_ctx_recording(ctx);
    vkvg_rec *rec = *slot;
    *slot = NULL;
    if (rec && !rec->count) { /* an empty recording is none */
        vkvg_rec_free(rec);
        return NULL;
    }
    return rec;
}
uint32_t vkvg_recording_get_count(VkvgRecording rec) { return rec ? rec->count : 0; }
void *vkvg_recording_get_data(VkvgRecording rec) { return rec ? rec->buf : NULL; }
void vkvg_recording_get_command(VkvgRecording rec, uint32_t cmdIndex, uint32_t *cmd, void **dataOffset) {
    if (!rec)
        return;
    uint32_t c = 0;
    void *off = NULL;
    if (cmdIndex < rec->count) {
        c = rec->cmds[cmdIndex].cmd;
        off = (void *)rec->cmds[cmdIndex].dataOffset;
    }
    if (cmd)
        *cmd = c;
    if (dataOffset)
        *dataOffset = off;
}
void vkvg_recording_destroy(VkvgRecording rec) { vkvg_rec_free(rec); }

static void replay_one(VkvgContext ctx, const vkvg_rec *rec, uint32_t i) {
    const char *d = rec->buf + rec->cmds[i].dataOffset;
    float f[8] = {0};
    uint32_t u = 0;
    size_t avail = (i + 1 < rec->count ? rec->cmds[i + 1].dataOffset : rec->size) - rec->cmds[i].dataOffset;
    memcpy(f, d, avail < sizeof(f) ? avail : sizeof(f));
    memcpy(&u, f, sizeof(u));
    switch (rec->cmds[i].cmd) {
    case VKVG_CMD_SAVE: vkvg_save(ctx); break;
    case VKVG_CMD_RESTORE: vkvg_restore(ctx); break;
    case VKVG_CMD_NEW_PATH: vkvg_new_path(ctx); break;
    case VKVG_CMD_NEW_SUB_PATH: vkvg_new_sub_path(ctx); break;
    case VKVG_CMD_CLOSE_PATH: vkvg_close_path(ctx); break;
    case VKVG_CMD_MOVE_TO: vkvg_move_to(ctx, f[0], f[1]); break;
    case VKVG_CMD_LINE_TO: vkvg_line_to(ctx, f[0], f[1]); break;
    case VKVG_CMD_REL_MOVE_TO: vkvg_rel_move_to(ctx, f[0], f[1]); break;
    case VKVG_CMD_REL_LINE_TO: vkvg_rel_line_to(ctx, f[0], f[1]); break;
    case VKVG_CMD_RECTANGLE: vkvg_rectangle(ctx, f[0], f[1], f[2], f[3]); break;
    case VKVG_CMD_ARC: vkvg_arc(ctx, f[0], f[1], f[2], f[3], f[4]); break;
    case VKVG_CMD_ARC_NEG: vkvg_arc_negative(ctx, f[0], f[1], f[2], f[3], f[4]); break;
    case VKVG_CMD_ELLIPSE: vkvg_ellipse(ctx, f[0], f[1], f[2], f[3], f[4]); break;
    case VKVG_CMD_CURVE_TO: vkvg_curve_to(ctx, f[0], f[1], f[2], f[3], f[4], f[5]); break;
    case VKVG_CMD_REL_CURVE_TO: vkvg_rel_curve_to(ctx, f[0], f[1], f[2], f[3], f[4], f[5]); break;
    case VKVG_CMD_QUADRATIC_TO: vkvg_quadratic_to(ctx, f[0], f[1], f[2], f[3]); break;
    case VKVG_CMD_REL_QUADRATIC_TO: vkvg_rel_quadratic_to(ctx, f[0], f[1], f[2], f[3]); break;
    case VKVG_CMD_ELLIPTICAL_ARC_TO:
        vkvg_elliptic_arc_to(ctx, f[0], f[1], d[20] != 0, d[21] != 0, f[2], f[3], f[4]);
        break;
    case VKVG_CMD_REL_ELLIPTICAL_ARC_TO:
        vkvg_rel_elliptic_arc_to(ctx, f[0], f[1], d[20] != 0, d[21] != 0, f[2], f[3], f[4]);
        break;
    case VKVG_CMD_SET_LINE_WIDTH: vkvg_set_line_width(ctx, f[0]); break;
    case VKVG_CMD_SET_MITER_LIMIT: vkvg_set_miter_limit(ctx, f[0]); break;
    case VKVG_CMD_SET_LINE_JOIN: vkvg_set_line_join(ctx, (vkvg_line_join_t)u); break;
    case VKVG_CMD_SET_LINE_CAP: vkvg_set_line_cap(ctx, (vkvg_line_cap_t)u); break;
    case VKVG_CMD_SET_OPERATOR: vkvg_set_operator(ctx, (vkvg_operator_t)u); break;
    case VKVG_CMD_SET_FILL_RULE: vkvg_set_fill_rule(ctx, (vkvg_fill_rule_t)u); break;
    case VKVG_CMD_SET_OPACITY: vkvg_set_opacity(ctx, f[0]); break;
    case VKVG_CMD_SET_DASH: {
        float *dashes = u ? (float *)malloc((size_t)u * sizeof(float)) : NULL;
        if (u && !dashes)
            break;
        if (u)
            memcpy(dashes, d + 8, (size_t)u * sizeof(float));
        vkvg_set_dash(ctx, dashes, u, f[1]);
        free(dashes);
        break;
    }
    case VKVG_CMD_TRANSLATE: vkvg_translate(ctx, f[0], f[1]); break;
    case VKVG_CMD_SCALE: vkvg_scale(ctx, f[0], f[1]); break;
    case VKVG_CMD_ROTATE: vkvg_rotate(ctx, f[0]); break;
    case VKVG_CMD_IDENTITY_MATRIX: vkvg_identity_matrix(ctx); break;
    case VKVG_CMD_TRANSFORM:
    case VKVG_CMD_SET_MATRIX: {
        vkvg_matrix_t m;
        memcpy(&m, d, sizeof(m));
        if (rec->cmds[i].cmd == VKVG_CMD_TRANSFORM)
            vkvg_transform(ctx, &m);
        else
            vkvg_set_matrix(ctx, &m);
        break;
    }
    case VKVG_CMD_PAINT: vkvg_paint(ctx); break;
    case VKVG_CMD_FILL: vkvg_fill(ctx); break;
    case VKVG_CMD_STROKE: vkvg_stroke(ctx); break;
    case VKVG_CMD_CLIP: vkvg_clip(ctx); break;
    case VKVG_CMD_RESET_CLIP: vkvg_reset_clip(ctx); break;
    case VKVG_CMD_CLEAR: vkvg_clear(ctx); break;
    case VKVG_CMD_FILL_PRESERVE: vkvg_fill_preserve(ctx); break;
    case VKVG_CMD_STROKE_PRESERVE: vkvg_stroke_preserve(ctx); break;
    case VKVG_CMD_CLIP_PRESERVE: vkvg_clip_preserve(ctx); break;
    case VKVG_CMD_SET_SOURCE_RGB: vkvg_set_source_rgb(ctx, f[0], f[1], f[2]); break;
    case VKVG_CMD_SET_SOURCE_RGBA: vkvg_set_source_rgba(ctx, f[0], f[1], f[2], f[3]); break;
    case VKVG_CMD_SET_SOURCE_COLOR: vkvg_set_source_color(ctx, u); break;
    case VKVG_CMD_SET_SOURCE: {
        VkvgPattern p;
        memcpy(&p, d, sizeof(p));
        vkvg_set_source(ctx, p);
        break;
    }
    case VKVG_CMD_SET_SOURCE_SURFACE: {
        VkvgSurface s;
        memcpy(&s, d + 8, sizeof(s));
        vkvg_set_source_surface(ctx, s, f[0], f[1]);
        break;
    }
    case VKVG_CMD_SET_FONT_SIZE: vkvg_set_font_size(ctx, u); break;
    case VKVG_CMD_SET_FONT_FACE: vkvg_select_font_face(ctx, d); break;
    case VKVG_CMD_SET_FONT_PATH: vkvg_load_font_from_path(ctx, d, NULL); break;
    case VKVG_CMD_SHOW_TEXT: vkvg_show_text(ctx, d); break;
    default: break;
    }
}
void vkvg_replay(VkvgContext ctx, VkvgRecording rec) {
    if (!rec)
        return;
    for (uint32_t i = 0; i < rec->count; i++)
        replay_one(ctx, rec, i);
}
void vkvg_replay_command(VkvgContext ctx, VkvgRecording rec, uint32_t cmdIndex) {
    if (rec && cmdIndex < rec->count)
        replay_one(ctx, rec, cmdIndex);
}
