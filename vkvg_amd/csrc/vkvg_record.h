/* Command recording of a context (include/vkvg.h:1960-1970): internal interface between the
 * recorder of the API calls (vkvg_api.c) and the recording store (vkvg_record.c). */
#ifndef VKVG_RECORD_H
#define VKVG_RECORD_H

#include "vkvg.h"

/* Command codes: the values of the reference (src/recording/vkvg_record_internal.h:28-99), so that a
 * client decoding vkvg_recording_get_command sees the same numbers. ELLIPSE uses the code the
 * reference reserves for it (commented out there); SET_OPACITY has no code in the reference. */
#define VKVG_CMD_SAVE               0x0001
#define VKVG_CMD_RESTORE            0x0002
#define VKVG_CMD_PATH_COMMANDS      0x0100
#define VKVG_CMD_DRAW_COMMANDS      0x0200
#define VKVG_CMD_RELATIVE_COMMANDS  (0x0400 | VKVG_CMD_PATH_COMMANDS)
#define VKVG_CMD_PATHPROPS_COMMANDS (0x1000 | VKVG_CMD_PATH_COMMANDS)
#define VKVG_CMD_PRESERVE_COMMANDS  (0x0400 | VKVG_CMD_DRAW_COMMANDS)
#define VKVG_CMD_PATTERN_COMMANDS   0x0800
#define VKVG_CMD_TRANSFORM_COMMANDS 0x2000
#define VKVG_CMD_TEXT_COMMANDS      0x4000

#define VKVG_CMD_NEW_PATH           (0x0001 | VKVG_CMD_PATH_COMMANDS)
#define VKVG_CMD_NEW_SUB_PATH       (0x0002 | VKVG_CMD_PATH_COMMANDS)
#define VKVG_CMD_CLOSE_PATH         (0x0003 | VKVG_CMD_PATH_COMMANDS)
#define VKVG_CMD_MOVE_TO            (0x0004 | VKVG_CMD_PATH_COMMANDS)
#define VKVG_CMD_LINE_TO            (0x0005 | VKVG_CMD_PATH_COMMANDS)
#define VKVG_CMD_RECTANGLE          (0x0006 | VKVG_CMD_PATH_COMMANDS)
#define VKVG_CMD_ARC                (0x0007 | VKVG_CMD_PATH_COMMANDS)
#define VKVG_CMD_ARC_NEG            (0x0008 | VKVG_CMD_PATH_COMMANDS)
#define VKVG_CMD_ELLIPSE            (0x0009 | VKVG_CMD_PATH_COMMANDS)
#define VKVG_CMD_CURVE_TO           (0x000A | VKVG_CMD_PATH_COMMANDS)
#define VKVG_CMD_QUADRATIC_TO       (0x000B | VKVG_CMD_PATH_COMMANDS)
#define VKVG_CMD_ELLIPTICAL_ARC_TO  (0x000C | VKVG_CMD_PATH_COMMANDS)

#define VKVG_CMD_SET_LINE_WIDTH     (0x0001 | VKVG_CMD_PATHPROPS_COMMANDS)
#define VKVG_CMD_SET_MITER_LIMIT    (0x0002 | VKVG_CMD_PATHPROPS_COMMANDS)
#define VKVG_CMD_SET_LINE_JOIN      (0x0003 | VKVG_CMD_PATHPROPS_COMMANDS)
#define VKVG_CMD_SET_LINE_CAP       (0x0004 | VKVG_CMD_PATHPROPS_COMMANDS)
#define VKVG_CMD_SET_OPERATOR       (0x0005 | VKVG_CMD_PATHPROPS_COMMANDS)
#define VKVG_CMD_SET_FILL_RULE      (0x0006 | VKVG_CMD_PATHPROPS_COMMANDS)
#define VKVG_CMD_SET_DASH           (0x0007 | VKVG_CMD_PATHPROPS_COMMANDS)
#define VKVG_CMD_SET_OPACITY        (0x0008 | VKVG_CMD_PATHPROPS_COMMANDS)

#define VKVG_CMD_TRANSLATE          (0x0001 | VKVG_CMD_TRANSFORM_COMMANDS)
#define VKVG_CMD_ROTATE             (0x0002 | VKVG_CMD_TRANSFORM_COMMANDS)
#define VKVG_CMD_SCALE              (0x0003 | VKVG_CMD_TRANSFORM_COMMANDS)
#define VKVG_CMD_TRANSFORM          (0x0004 | VKVG_CMD_TRANSFORM_COMMANDS)
#define VKVG_CMD_IDENTITY_MATRIX    (0x0005 | VKVG_CMD_TRANSFORM_COMMANDS)
#define VKVG_CMD_SET_MATRIX         (0x0006 | VKVG_CMD_TRANSFORM_COMMANDS)

#define VKVG_CMD_SET_FONT_SIZE      (0x0001 | VKVG_CMD_TEXT_COMMANDS)
#define VKVG_CMD_SET_FONT_FACE      (0x0002 | VKVG_CMD_TEXT_COMMANDS)
#define VKVG_CMD_SET_FONT_PATH      (0x0003 | VKVG_CMD_TEXT_COMMANDS)
#define VKVG_CMD_SHOW_TEXT          (0x0004 | VKVG_CMD_TEXT_COMMANDS)

#define VKVG_CMD_REL_MOVE_TO           (VKVG_CMD_MOVE_TO | VKVG_CMD_RELATIVE_COMMANDS)
#define VKVG_CMD_REL_LINE_TO           (VKVG_CMD_LINE_TO | VKVG_CMD_RELATIVE_COMMANDS)
#define VKVG_CMD_REL_CURVE_TO          (VKVG_CMD_CURVE_TO | VKVG_CMD_RELATIVE_COMMANDS)
#define VKVG_CMD_REL_QUADRATIC_TO      (VKVG_CMD_QUADRATIC_TO | VKVG_CMD_RELATIVE_COMMANDS)
#define VKVG_CMD_REL_ELLIPTICAL_ARC_TO (VKVG_CMD_ELLIPTICAL_ARC_TO | VKVG_CMD_RELATIVE_COMMANDS)

#define VKVG_CMD_PAINT              (0x0001 | VKVG_CMD_DRAW_COMMANDS)
#define VKVG_CMD_FILL               (0x0002 | VKVG_CMD_DRAW_COMMANDS)
#define VKVG_CMD_STROKE             (0x0003 | VKVG_CMD_DRAW_COMMANDS)
#define VKVG_CMD_CLIP               (0x0004 | VKVG_CMD_DRAW_COMMANDS)
#define VKVG_CMD_RESET_CLIP         (0x0005 | VKVG_CMD_DRAW_COMMANDS)
#define VKVG_CMD_CLEAR              (0x0006 | VKVG_CMD_DRAW_COMMANDS)
#define VKVG_CMD_FILL_PRESERVE      (VKVG_CMD_FILL | VKVG_CMD_PRESERVE_COMMANDS)
#define VKVG_CMD_STROKE_PRESERVE    (VKVG_CMD_STROKE | VKVG_CMD_PRESERVE_COMMANDS)
#define VKVG_CMD_CLIP_PRESERVE      (VKVG_CMD_CLIP | VKVG_CMD_PRESERVE_COMMANDS)

#define VKVG_CMD_SET_SOURCE_RGB     (0x0001 | VKVG_CMD_PATTERN_COMMANDS)
#define VKVG_CMD_SET_SOURCE_RGBA    (0x0002 | VKVG_CMD_PATTERN_COMMANDS)
#define VKVG_CMD_SET_SOURCE_COLOR   (0x0003 | VKVG_CMD_PATTERN_COMMANDS)
#define VKVG_CMD_SET_SOURCE         (0x0004 | VKVG_CMD_PATTERN_COMMANDS)
#define VKVG_CMD_SET_SOURCE_SURFACE (0x0005 | VKVG_CMD_PATTERN_COMMANDS)

typedef struct _vkvg_recording_t vkvg_rec;

vkvg_rec *vkvg_rec_new(void);
void      vkvg_rec_free(vkvg_rec *rec);
/* appends a command with `n` bytes of operands */
void      vkvg_rec_push(vkvg_rec *rec, uint32_t cmd, const void *data, size_t n);
void      vkvg_rec_floats(vkvg_rec *rec, uint32_t cmd, uint32_t n, ...); /* n doubles (promoted floats) */
void      vkvg_rec_u32(vkvg_rec *rec, uint32_t cmd, uint32_t v);
void      vkvg_rec_elliptic(vkvg_rec *rec, uint32_t cmd, float x, float y, float rx, float ry, float phi, bool large,
                            bool sweep);
void      vkvg_rec_dash(vkvg_rec *rec, const float *dashes, uint32_t n, float offset);
void      vkvg_rec_pattern(vkvg_rec *rec, VkvgPattern pat);              /* takes a reference */
void      vkvg_rec_surface(vkvg_rec *rec, VkvgSurface surf, float x, float y); /* takes a reference */
void      vkvg_rec_text(vkvg_rec *rec, uint32_t cmd, const char *text);

/* the context's recording slot (vkvg_api.c) */
vkvg_rec **vkvg_ctx_recording(VkvgContext ctx);

#endif
