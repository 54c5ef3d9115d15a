/*
 * A C client of the drop-in boundary: compiled against include/vkvg.h with gcc and linked with
 * libvkvg.so, as a program written for the reference would be.
 *
 *   abi_client          checks that need no GPU: struct layouts of the Vulkan-typed ABI, NULL and
 *                       error handles, the matrix vectors of gunit_tests/matrices.cpp:4-70
 *   abi_client device   the expectations of the reference's unit tests that need a device:
 *                       gunit_tests/device.cpp:4-41, context.cpp:30-218, surface.cpp:37-119, and
 *                       the recording API (include/vkvg.h:1960-1970)
 *
 * Prints one line per failed expectation; the exit status is the number of failures.
 */
#include <math.h>
#include <stddef.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "vkvg.h"

/* x86-64 layout of the reference's create-info with Vulkan's types (VkSampleCountFlags = uint32_t,
 * handles are pointers): a client built against either header passes the same bytes */
_Static_assert(sizeof(vkvg_device_create_info_t) == 48, "create-info size");
_Static_assert(offsetof(vkvg_device_create_info_t, samples) == 0, "samples");
_Static_assert(offsetof(vkvg_device_create_info_t, deferredResolve) == 4, "deferredResolve");
_Static_assert(offsetof(vkvg_device_create_info_t, inst) == 8, "inst");
_Static_assert(offsetof(vkvg_device_create_info_t, phy) == 16, "phy");
_Static_assert(offsetof(vkvg_device_create_info_t, vkdev) == 24, "vkdev");
_Static_assert(offsetof(vkvg_device_create_info_t, qFamIdx) == 32, "qFamIdx");
_Static_assert(offsetof(vkvg_device_create_info_t, qIndex) == 36, "qIndex");
_Static_assert(offsetof(vkvg_device_create_info_t, threadAware) == 40, "threadAware");
_Static_assert(sizeof(vkvg_matrix_t) == 24 && offsetof(vkvg_matrix_t, y0) == 20, "matrix");
_Static_assert(sizeof(vkvg_color_t) == 16, "colour");
_Static_assert(sizeof(vkvg_text_extents_t) == 24 && sizeof(vkvg_font_extents_t) == 20, "text metrics");
_Static_assert(sizeof(vkvg_glyph_info_t) == 20, "glyph info");
_Static_assert(sizeof(VkFormat) == 4 && sizeof(VkSampleCountFlags) == 4, "Vulkan scalar types");
_Static_assert(VKVG_STATUS_INVALID_DEVICE_CREATE_INFO == 17 && VKVG_STATUS_IN_CACHE == 21, "status values");
_Static_assert(VKVG_FILL_RULE_EVEN_ODD == 0 && VKVG_FILL_RULE_NON_ZERO == 1, "fill rule values");
_Static_assert(VKVG_LINE_JOIN_BEVEL == 2 && VKVG_LINE_CAP_SQUARE == 2, "join / cap values");
_Static_assert(VKVG_OPERATOR_OVER == 2 && VKVG_OPERATOR_DIFFERENCE == 3, "operator values");

static int failures = 0;
#define EXPECT(cond)                                                                                                   \
    do {                                                                                                               \
        if (!(cond)) {                                                                                                 \
            printf("FAIL %s:%d: %s\n", __FILE__, __LINE__, #cond);                                                     \
            failures++;                                                                                                \
        }                                                                                                              \
    } while (0)
#define EXPECT_FEQ(a, b) EXPECT(fabsf((float)(a) - (float)(b)) <= 4 * 1.1920929e-7f * fmaxf(1.f, fabsf((float)(b))))

static void no_device_checks(void) {
    EXPECT(vkvg_device_status(NULL) == VKVG_STATUS_NULL_POINTER);
    VkvgDevice d = vkvg_device_create(NULL);
    EXPECT(d != NULL && vkvg_device_status(d) == VKVG_STATUS_INVALID_DEVICE_CREATE_INFO);
    EXPECT(vkvg_device_get_reference_count(d) == 0);
    vkvg_device_destroy(d);
    EXPECT(vkvg_surface_status(NULL) == VKVG_STATUS_NULL_POINTER);
    VkvgSurface s = vkvg_surface_create(NULL, 0, 0);
    EXPECT(vkvg_surface_status(s) == VKVG_STATUS_DEVICE_ERROR);
    EXPECT(vkvg_surface_get_reference_count(NULL) == 0 && vkvg_surface_get_reference_count(s) == 0);
    EXPECT(vkvg_surface_get_width(NULL) == 0 && vkvg_surface_get_height(NULL) == 0);
    EXPECT(vkvg_surface_get_vk_image(NULL) == NULL && vkvg_surface_get_vk_format(NULL) == VK_FORMAT_UNDEFINED);
    EXPECT(vkvg_surface_write_to_memory(NULL, NULL) == VKVG_STATUS_INVALID_STATUS);
    EXPECT(vkvg_surface_write_to_png(NULL, NULL) == VKVG_STATUS_INVALID_STATUS);
    EXPECT(vkvg_status(NULL) == VKVG_STATUS_NULL_POINTER);
    VkvgContext c = vkvg_create(NULL);
    EXPECT(vkvg_status(c) == VKVG_STATUS_INVALID_SURFACE && vkvg_get_reference_count(c) == 0);
    EXPECT(vkvg_pattern_status(NULL) == VKVG_STATUS_NULL_POINTER);
    EXPECT(vkvg_recording_get_count(NULL) == 0 && vkvg_recording_get_data(NULL) == NULL);

    /* gunit_tests/matrices.cpp:4-70 */
    vkvg_matrix_t m;
    float sx, sy;
    vkvg_matrix_init_scale(&m, 2.1f, 1.5f);
    vkvg_matrix_get_scale(&m, &sx, &sy);
    EXPECT_FEQ(sx, 2.1f);
    EXPECT_FEQ(sy, 1.5f);
    EXPECT(vkvg_matrix_invert(&m) == VKVG_STATUS_SUCCESS);
    EXPECT_FEQ(m.xx, 1.f / 2.1f);
    EXPECT_FEQ(m.yy, 1.f / 1.5f);
    vkvg_matrix_t z = {0, 0, 0, 0, 0, 0};
    EXPECT(vkvg_matrix_invert(&z) == VKVG_STATUS_INVALID_MATRIX);
    vkvg_matrix_t sing = {1, 1, 0, 0, 0, 0};
    EXPECT(vkvg_matrix_invert(&sing) == VKVG_STATUS_INVALID_MATRIX);
    vkvg_matrix_init_translate(&m, 2.1f, 1.5f);
    EXPECT(vkvg_matrix_invert(&m) == VKVG_STATUS_SUCCESS);
    EXPECT_FEQ(m.x0, -2.1f);
    EXPECT_FEQ(m.y0, -1.5f);
    vkvg_matrix_t id = VKVG_IDENTITY_MATRIX;
    float px = 3, py = 4;
    vkvg_matrix_transform_point(&id, &px, &py);
    EXPECT(px == 3 && py == 4);
}

#define EXPECT_CP(ctx, x, y)                                                                                           \
    do {                                                                                                               \
        float a_ = -1, b_ = -1;                                                                                        \
        vkvg_get_current_point(ctx, &a_, &b_);                                                                         \
        EXPECT_FEQ(a_, x);                                                                                             \
        EXPECT_FEQ(b_, y);                                                                                             \
    } while (0)
#define EXPECT_NO_CP(ctx)                                                                                              \
    do {                                                                                                               \
        EXPECT(!vkvg_has_current_point(ctx));                                                                          \
        EXPECT_CP(ctx, 0, 0);                                                                                          \
    } while (0)

static void null_context_calls(VkvgContext ctx) {
    /* gunit_tests/context.cpp:77-170: every entry point is a no-op on a context in error */
    vkvg_new_path(ctx);
    vkvg_close_path(ctx);
    vkvg_new_sub_path(ctx);
    vkvg_path_extents(ctx, NULL, NULL, NULL, NULL);
    vkvg_get_current_point(ctx, NULL, NULL);
    vkvg_line_to(ctx, 0, 0);
    vkvg_rel_line_to(ctx, 0, 0);
    vkvg_move_to(ctx, 0, 0);
    vkvg_rel_move_to(ctx, 0, 0);
    vkvg_arc(ctx, 0, 0, 0, 0, 0);
    vkvg_arc_negative(ctx, 0, 0, 0, 0, 0);
    vkvg_curve_to(ctx, 0, 0, 0, 0, 0, 0);
    vkvg_rel_curve_to(ctx, 0, 0, 0, 0, 0, 0);
    vkvg_quadratic_to(ctx, 0, 0, 0, 0);
    vkvg_rel_quadratic_to(ctx, 0, 0, 0, 0);
    vkvg_rectangle(ctx, 0, 0, 0, 0);
    vkvg_rounded_rectangle(ctx, 0, 0, 0, 0, 0);
    vkvg_rounded_rectangle2(ctx, 0, 0, 0, 0, 0, 0);
    vkvg_ellipse(ctx, 0, 0, 0, 0, 0);
    vkvg_elliptic_arc_to(ctx, 0, 0, 0, 0, 0, 0, 0);
    vkvg_rel_elliptic_arc_to(ctx, 0, 0, 0, 0, 0, 0, 0);
    vkvg_stroke(ctx);
    vkvg_stroke_preserve(ctx);
    vkvg_fill(ctx);
    vkvg_fill_preserve(ctx);
    vkvg_paint(ctx);
    vkvg_clear(ctx);
    vkvg_reset_clip(ctx);
    vkvg_clip(ctx);
    vkvg_clip_preserve(ctx);
    vkvg_set_opacity(ctx, 0);
    EXPECT(vkvg_get_opacity(ctx) == 0);
    vkvg_set_source_color(ctx, 0);
    vkvg_set_source_rgb(ctx, 0, 0, 0);
    vkvg_set_source_rgba(ctx, 0, 0, 0, 0);
    vkvg_set_source_surface(ctx, NULL, 0, 0);
    vkvg_set_source(ctx, NULL);
    vkvg_set_line_width(ctx, 0);
    vkvg_set_miter_limit(ctx, 0);
    EXPECT(vkvg_get_miter_limit(ctx) == 0);
    vkvg_set_line_cap(ctx, VKVG_LINE_CAP_ROUND);
    vkvg_set_line_join(ctx, VKVG_LINE_JOIN_MITER);
    vkvg_set_operator(ctx, VKVG_OPERATOR_OVER);
    vkvg_set_fill_rule(ctx, VKVG_FILL_RULE_EVEN_ODD);
    vkvg_set_dash(ctx, NULL, 0, 0);
    vkvg_get_dash(ctx, NULL, NULL, NULL);
    EXPECT(vkvg_get_line_width(ctx) == 0);
    EXPECT(vkvg_get_line_cap(ctx) == VKVG_LINE_CAP_BUTT);
    EXPECT(vkvg_get_line_join(ctx) == VKVG_LINE_JOIN_MITER);
    EXPECT(vkvg_get_operator(ctx) == VKVG_OPERATOR_OVER);
    EXPECT(vkvg_get_fill_rule(ctx) == VKVG_FILL_RULE_NON_ZERO);
    (void)vkvg_get_source(ctx);
    (void)vkvg_get_target(ctx);
    EXPECT(!vkvg_has_current_point(ctx));
    vkvg_save(ctx);
    vkvg_restore(ctx);
    vkvg_translate(ctx, 0, 0);
    vkvg_scale(ctx, 0, 0);
    vkvg_rotate(ctx, 0);
    vkvg_transform(ctx, NULL);
    vkvg_set_matrix(ctx, NULL);
    vkvg_get_matrix(ctx, NULL);
    vkvg_identity_matrix(ctx);
    vkvg_select_font_face(ctx, NULL);
    vkvg_load_font_from_path(ctx, NULL, NULL);
    vkvg_load_font_from_memory(ctx, NULL, 0, NULL);
    vkvg_set_font_size(ctx, 0);
    vkvg_show_text(ctx, NULL);
    vkvg_text_extents(ctx, NULL, NULL);
    vkvg_font_extents(ctx, NULL);
    vkvg_text_run_create(ctx, NULL);
    vkvg_text_run_create_with_length(ctx, NULL, 10);
    vkvg_show_text_run(ctx, NULL);
    vkvg_start_recording(ctx);
    EXPECT(vkvg_stop_recording(ctx) == NULL);
}

static int read_pixels(VkvgSurface s, uint32_t *buf) { return vkvg_surface_write_to_memory(s, (unsigned char *)buf); }

static void draw_scene(VkvgContext ctx) {
    /* a little of everything that records: paths, arcs, an ellipse, styles, a dash, a gradient,
     * a transform, save / restore, clip */
    vkvg_set_source_rgb(ctx, 1, 1, 1);
    vkvg_paint(ctx);
    vkvg_save(ctx);
    vkvg_translate(ctx, 6, 4);
    vkvg_set_line_width(ctx, 5);
    vkvg_set_miter_limit(ctx, 1.2f);
    vkvg_set_line_join(ctx, VKVG_LINE_JOIN_MITER);
    vkvg_set_source_rgba(ctx, 0.8f, 0.1f, 0.2f, 0.7f);
    vkvg_move_to(ctx, 10, 10);
    vkvg_line_to(ctx, 60, 14);
    vkvg_rel_line_to(ctx, -30, 30);
    vkvg_stroke(ctx);
    vkvg_restore(ctx);
    const float dashes[2] = {6, 3};
    vkvg_set_dash(ctx, dashes, 2, 1.5f);
    vkvg_set_line_cap(ctx, VKVG_LINE_CAP_ROUND);
    vkvg_set_line_width(ctx, 3);
    vkvg_arc_negative(ctx, 70, 60, 18, 3.0f, 0.4f);
    vkvg_stroke(ctx);
    vkvg_set_dash(ctx, NULL, 0, 0);
    VkvgPattern pat = vkvg_pattern_create_radial(40, 70, 2, 45, 75, 30);
    vkvg_pattern_add_color_stop(pat, 0, 0, 0.5f, 1, 1);
    vkvg_pattern_add_color_stop(pat, 1, 1, 1, 0, 0.6f);
    vkvg_set_source(ctx, pat);
    vkvg_pattern_destroy(pat);
    vkvg_ellipse(ctx, 25, 15, 40, 75, 0.3f);
    vkvg_set_fill_rule(ctx, VKVG_FILL_RULE_EVEN_ODD);
    vkvg_fill(ctx);
    vkvg_set_opacity(ctx, 0.5f);
    vkvg_rectangle(ctx, 50, 5, 40, 30);
    vkvg_clip(ctx);
    vkvg_set_source_rgb(ctx, 0, 0.6f, 0.2f);
    vkvg_rounded_rectangle(ctx, 40, 0, 40, 40, 8);
    vkvg_fill(ctx);
    vkvg_reset_clip(ctx);
    vkvg_set_source_color(ctx, 0xff204080u);
    vkvg_move_to(ctx, 5, 90);
    vkvg_curve_to(ctx, 30, 60, 60, 120, 95, 85);
    vkvg_quadratic_to(ctx, 60, 70, 20, 95);
    vkvg_close_path(ctx);
    vkvg_fill(ctx);
}

static void device_checks(void) {
    vkvg_device_create_info_t info;
    memset(&info, 0, sizeof(info));
    /* gunit_tests/device.cpp:17-41 */
    VkvgDevice dev = vkvg_device_create(&info);
    EXPECT(vkvg_device_status(dev) == VKVG_STATUS_SUCCESS);
    EXPECT(vkvg_device_get_reference_count(dev) == 1);
    vkvg_device_reference(dev);
    EXPECT(vkvg_device_get_reference_count(dev) == 2);
    vkvg_device_destroy(dev);
    EXPECT(vkvg_device_status(dev) == VKVG_STATUS_SUCCESS && vkvg_device_get_reference_count(dev) == 1);
    vkvg_device_set_context_cache_size(dev, 0);

    /* gunit_tests/surface.cpp:37-105 */
    VkvgSurface surf = vkvg_surface_create(dev, 512, 256);
    EXPECT(vkvg_surface_status(surf) == VKVG_STATUS_SUCCESS);
    EXPECT(vkvg_device_get_reference_count(dev) == 2);
    EXPECT(vkvg_surface_get_reference_count(surf) == 1);
    vkvg_surface_reference(surf);
    EXPECT(vkvg_surface_get_reference_count(surf) == 2);
    vkvg_surface_destroy(surf);
    EXPECT(vkvg_surface_get_reference_count(surf) == 1);
    EXPECT(vkvg_surface_get_width(surf) == 512 && vkvg_surface_get_height(surf) == 256);
    EXPECT(vkvg_surface_get_vk_format(surf) == VK_FORMAT_B8G8R8A8_UNORM && vkvg_surface_get_vk_image(surf) != NULL);
    EXPECT(vkvg_surface_write_to_memory(surf, NULL) == VKVG_STATUS_WRITE_ERROR);
    EXPECT(vkvg_surface_write_to_png(surf, NULL) == VKVG_STATUS_WRITE_ERROR);
    vkvg_surface_destroy(surf);
    EXPECT(vkvg_device_get_reference_count(dev) == 1);

    /* gunit_tests/context.cpp:30-75 */
    surf = vkvg_surface_create(dev, 512, 512);
    VkvgContext ctx = vkvg_create(NULL);
    EXPECT(vkvg_status(ctx) == VKVG_STATUS_INVALID_SURFACE);
    null_context_calls(ctx);
    ctx = vkvg_create(surf);
    EXPECT(vkvg_status(ctx) == VKVG_STATUS_SUCCESS);
    EXPECT(vkvg_get_reference_count(ctx) == 1);
    EXPECT(vkvg_surface_get_reference_count(surf) == 2);
    EXPECT(vkvg_device_get_reference_count(dev) == 2);
    vkvg_reference(ctx);
    EXPECT(vkvg_get_reference_count(ctx) == 2);
    vkvg_destroy(ctx);
    EXPECT(vkvg_get_reference_count(ctx) == 1);
    vkvg_destroy(ctx);
    EXPECT(vkvg_surface_get_reference_count(surf) == 1 && vkvg_device_get_reference_count(dev) == 2);

    /* gunit_tests/context.cpp:184-218: current point */
    ctx = vkvg_create(surf);
    EXPECT_NO_CP(ctx);
    vkvg_new_path(ctx);
    vkvg_close_path(ctx);
    vkvg_new_sub_path(ctx);
    float a = 1, b = 1, c = 1, d = 1;
    vkvg_path_extents(ctx, &a, &b, &c, &d);
    EXPECT(a == 0 && b == 0 && c == 0 && d == 0);
    EXPECT_NO_CP(ctx);
    vkvg_line_to(ctx, 50, 10);
    EXPECT_CP(ctx, 50, 10);
    vkvg_move_to(ctx, 10, 50);
    EXPECT(vkvg_has_current_point(ctx));
    EXPECT_CP(ctx, 10, 50);
    vkvg_line_to(ctx, 50, 10);
    EXPECT_CP(ctx, 50, 10);
    vkvg_rel_line_to(ctx, 10, 10);
    EXPECT_CP(ctx, 60, 20);
    vkvg_close_path(ctx);
    EXPECT_NO_CP(ctx);
    vkvg_line_to(ctx, 50, 10);
    vkvg_rel_line_to(ctx, 10, 10);
    vkvg_new_sub_path(ctx);
    EXPECT_NO_CP(ctx);
    vkvg_line_to(ctx, 50, 10);
    vkvg_rel_line_to(ctx, 10, 10);
    vkvg_new_path(ctx);
    EXPECT_NO_CP(ctx);
    /* text entry points link and draw nothing */
    vkvg_text_extents_t te;
    memset(&te, 0xff, sizeof(te));
    vkvg_text_extents(ctx, "abc", &te);
    EXPECT(te.width == 0 && te.x_advance == 0);
    vkvg_show_text(ctx, "abc");
    EXPECT(vkvg_status(ctx) == VKVG_STATUS_SUCCESS);
    vkvg_destroy(ctx);
    vkvg_surface_destroy(surf);

    /* surface.cpp:107-119 SurfClear */
    surf = vkvg_surface_create(dev, 8, 8);
    ctx = vkvg_create(surf);
    vkvg_set_source_rgb(ctx, 1, 0, 0);
    vkvg_paint(ctx);
    vkvg_destroy(ctx);
    vkvg_surface_clear(surf);
    uint32_t px[64];
    EXPECT(read_pixels(surf, px) == VKVG_STATUS_SUCCESS);
    for (int i = 0; i < 64; i++)
        EXPECT(px[i] == 0);
    vkvg_surface_destroy(surf);

    /* recording: nothing is drawn while recording; the replay draws what direct calls draw */
    const uint32_t W = 100, H = 100;
    VkvgSurface direct = vkvg_surface_create(dev, W, H), recd = vkvg_surface_create(dev, W, H),
                replayed = vkvg_surface_create(dev, W, H);
    ctx = vkvg_create(direct);
    draw_scene(ctx);
    vkvg_destroy(ctx);
    ctx = vkvg_create(recd);
    vkvg_start_recording(ctx);
    draw_scene(ctx);
    VkvgRecording rec = vkvg_stop_recording(ctx);
    EXPECT(rec != NULL);
    EXPECT(vkvg_status(ctx) == VKVG_STATUS_SUCCESS);
    vkvg_destroy(ctx);
    uint32_t n = vkvg_recording_get_count(rec);
    EXPECT(n == 45); /* the calls of draw_scene; rounded_rectangle records as its ten path calls */
    uint32_t cmd = 0;
    void *off = NULL;
    vkvg_recording_get_command(rec, 0, &cmd, &off);
    EXPECT(cmd == 0x0801 && off == (void *)0);          /* SET_SOURCE_RGB: r, g, b */
    vkvg_recording_get_command(rec, 5, &cmd, &off);
    EXPECT(cmd == 0x1102);                               /* SET_MITER_LIMIT */
    const char *data = (const char *)vkvg_recording_get_data(rec);
    float lim;
    memcpy(&lim, data + (size_t)off, 4);
    EXPECT(lim == 1.2f);
    vkvg_recording_get_command(rec, 8, &cmd, &off);
    EXPECT(cmd == 0x0104);                               /* MOVE_TO 10, 10 */
    float mv[2];
    memcpy(mv, data + (size_t)off, 8);
    EXPECT(mv[0] == 10 && mv[1] == 10);
    vkvg_recording_get_command(rec, n, &cmd, &off);
    EXPECT(cmd == 0 && off == NULL);
    ctx = vkvg_create(replayed);
    vkvg_replay(ctx, rec);
    vkvg_destroy(ctx);
    uint32_t *p0 = (uint32_t *)malloc(W * H * 4), *p1 = (uint32_t *)malloc(W * H * 4),
             *p2 = (uint32_t *)malloc(W * H * 4);
    EXPECT(read_pixels(direct, p0) == 0 && read_pixels(recd, p1) == 0 && read_pixels(replayed, p2) == 0);
    uint32_t drawn = 0, differ = 0, nonzero_rec = 0;
    for (uint32_t i = 0; i < W * H; i++) {
        drawn += p0[i] != 0;
        differ += p0[i] != p2[i];
        nonzero_rec += p1[i] != 0;
    }
    EXPECT(drawn == W * H);
    EXPECT(differ == 0);
    EXPECT(nonzero_rec == 0);
    /* one command at a time gives the same image again */
    ctx = vkvg_create(replayed);
    vkvg_clear(ctx);
    for (uint32_t i = 0; i < n; i++)
        vkvg_replay_command(ctx, rec, i);
    vkvg_destroy(ctx);
    EXPECT(read_pixels(replayed, p2) == 0);
    differ = 0;
    for (uint32_t i = 0; i < W * H; i++)
        differ += p0[i] != p2[i];
    EXPECT(differ == 0);
    vkvg_recording_destroy(rec);
    free(p0);
    free(p1);
    free(p2);
    vkvg_surface_destroy(direct);
    vkvg_surface_destroy(recd);
    vkvg_surface_destroy(replayed);
    EXPECT(vkvg_device_get_reference_count(dev) == 1);
    vkvg_device_destroy(dev);
}

int main(int argc, char **argv) {
    no_device_checks();
    if (argc > 1 && !strcmp(argv[1], "device"))
        device_checks();
    printf("%s: %d failure(s)\n", argc > 1 ? argv[1] : "no device", failures);
    return failures;
}
