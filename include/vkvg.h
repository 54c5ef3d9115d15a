/*
 * vkvg.h — public C API of the MI355X-native vkvg backend.
 *
 * Drop-in replacement for include/vkvg.h of the reference (jpbruyere/vkvg): every function
 * declared here has the reference's name, signature, argument meaning and error behaviour,
 * and every enum keeps the reference's values, so client code recompiles and relinks
 * unchanged. Each declaration cites the reference line it replaces.
 *
 * The reference header includes <vulkan/vulkan.h> for a handful of types that appear in
 * signatures (include/vkvg.h:69, :555-564, :801-807). This backend has no Vulkan; when
 * vulkan.h has not been included, those types are defined here with identical size and
 * representation (dispatchable handles and non-dispatchable handles are pointers on
 * 64-bit targets, VkFormat is a 32-bit enum, VkSampleCountFlags is a 32-bit mask).
 *
 * Semantics summary (see the reference documentation for details):
 *  - Objects are reference counted; constructors never return NULL but, on failure,
 *    the address of a static status value (include/vkvg.h:294-329).
 *  - Every entry point first checks the object's status and returns silently when it is
 *    not VKVG_STATUS_SUCCESS; errors are sticky.
 *  - Drawing is recorded on the host and rendered by HIP kernels on the GPU at
 *    vkvg_flush / vkvg_destroy / surface readback.
 */
#ifndef VKVG_H
#define VKVG_H

#ifdef __cplusplus
extern "C" {
#endif

#include <math.h>
#include <stdbool.h>
#include <stdint.h>

#ifndef vkvg_public
#define vkvg_public __attribute__((visibility("default")))
#endif

#ifndef VK_VERSION_1_0
typedef uint32_t VkSampleCountFlags;
typedef struct VkInstance_T *VkInstance;
typedef struct VkPhysicalDevice_T *VkPhysicalDevice;
typedef struct VkDevice_T *VkDevice;
typedef struct VkImage_T *VkImage;
typedef struct VkPhysicalDeviceFeatures VkPhysicalDeviceFeatures;
typedef enum VkFormat {
    VK_FORMAT_UNDEFINED = 0,
    VK_FORMAT_R8G8B8A8_UNORM = 37,
    VK_FORMAT_B8G8R8A8_UNORM = 44,
    VK_FORMAT_MAX_ENUM = 0x7FFFFFFF
} VkFormat;
#define VK_SAMPLE_COUNT_1_BIT 1u
#endif

/* include/vkvg.h:87-101 */
#define VKVG_LOG_ERR        0x00000001
#define VKVG_LOG_DEBUG      0x00000002
#define VKVG_LOG_INFO_PTS   0x00000004
#define VKVG_LOG_INFO_PATH  0x00000008
#define VKVG_LOG_INFO_CMD   0x00000010
#define VKVG_LOG_INFO_VBO   0x00000020
#define VKVG_LOG_INFO_IBO   0x00000040
#define VKVG_LOG_INFO_VAO   (VKVG_LOG_INFO_VBO | VKVG_LOG_INFO_IBO)
#define VKVG_LOG_THREAD     0x00000080
#define VKVG_LOG_DBG_ARRAYS 0x00001000
#define VKVG_LOG_STROKE     0x00010000
#define VKVG_LOG_FULL       0xffffffff

/* include/vkvg.h:125-150 */
typedef enum {
    VKVG_STATUS_SUCCESS = 0,
    VKVG_STATUS_NO_MEMORY,
    VKVG_STATUS_NULL_POINTER,
    VKVG_STATUS_INVALID_RESTORE,
    VKVG_STATUS_NO_CURRENT_POINT,
    VKVG_STATUS_INVALID_MATRIX,
    VKVG_STATUS_INVALID_STATUS,
    VKVG_STATUS_INVALID_INDEX,
    VKVG_STATUS_WRITE_ERROR,
    VKVG_STATUS_PATTERN_TYPE_MISMATCH,
    VKVG_STATUS_PATTERN_INVALID_GRADIENT,
    VKVG_STATUS_INVALID_FORMAT,
    VKVG_STATUS_FILE_NOT_FOUND,
    VKVG_STATUS_INVALID_DASH,
    VKVG_STATUS_INVALID_RECT,
    VKVG_STATUS_TIMEOUT,
    VKVG_STATUS_DEVICE_ERROR,
    VKVG_STATUS_INVALID_DEVICE_CREATE_INFO,
    VKVG_STATUS_INVALID_IMAGE,
    VKVG_STATUS_INVALID_SURFACE,
    VKVG_STATUS_INVALID_FONT,
    VKVG_STATUS_IN_CACHE,
    VKVG_STATUS_ENUM_MAX = 0x7FFFFFFF
} vkvg_status_t;

typedef enum { VKVG_HORIZONTAL = 0, VKVG_VERTICAL = 1 } vkvg_direction_t;                           /* :152 */
typedef enum { VKVG_FORMAT_ARGB32, VKVG_FORMAT_RGB24, VKVG_FORMAT_A8, VKVG_FORMAT_A1 } vkvg_format_t; /* :154 */

typedef enum { VKVG_EXTEND_NONE, VKVG_EXTEND_REPEAT, VKVG_EXTEND_REFLECT, VKVG_EXTEND_PAD } vkvg_extend_t; /* :161 */

typedef enum {
    VKVG_FILTER_FAST,
    VKVG_FILTER_GOOD,
    VKVG_FILTER_BEST,
    VKVG_FILTER_NEAREST,
    VKVG_FILTER_BILINEAR,
    VKVG_FILTER_GAUSSIAN,
} vkvg_filter_t; /* :168-175 */

typedef enum {
    VKVG_PATTERN_TYPE_SOLID,
    VKVG_PATTERN_TYPE_SURFACE,
    VKVG_PATTERN_TYPE_LINEAR,
    VKVG_PATTERN_TYPE_RADIAL,
    VKVG_PATTERN_TYPE_MESH,
    VKVG_PATTERN_TYPE_RASTER_SOURCE,
} vkvg_pattern_type_t; /* :182-189 */

typedef enum { VKVG_LINE_CAP_BUTT, VKVG_LINE_CAP_ROUND, VKVG_LINE_CAP_SQUARE } vkvg_line_cap_t;    /* :197-201 */
typedef enum { VKVG_LINE_JOIN_MITER, VKVG_LINE_JOIN_ROUND, VKVG_LINE_JOIN_BEVEL } vkvg_line_join_t; /* :208-212 */
typedef enum { VKVG_FILL_RULE_EVEN_ODD, VKVG_FILL_RULE_NON_ZERO } vkvg_fill_rule_t;                /* :222-225 */

typedef struct {
    float r, g, b, a;
} vkvg_color_t; /* :227-232 */

/* text metrics and runs (:240-294): fonts are outside this backend's scope; the text entry points
 * below exist so that clients link, and draw nothing */
typedef struct {
    float ascent, descent, height, max_x_advance, max_y_advance;
} vkvg_font_extents_t; /* :240-248 */
typedef struct {
    float x_bearing, y_bearing, width, height, x_advance, y_advance;
} vkvg_text_extents_t; /* :256-265 */
typedef struct _glyph_info_t {
    int32_t  x_advance, y_advance, x_offset, y_offset;
    uint32_t codepoint;
} vkvg_glyph_info_t; /* :272-280 */
typedef struct _vkvg_text_run_t *VkvgText; /* :294 */

typedef struct _vkvg_context_t *VkvgContext; /* :302 */
typedef struct _vkvg_surface_t *VkvgSurface; /* :314 */
typedef struct _vkvg_device_t *VkvgDevice;   /* :321 */
typedef struct _vkvg_pattern_t *VkvgPattern; /* :329 */

/* :377-384 */
typedef struct {
    float xx, yx, xy, yy, x0, y0;
} vkvg_matrix_t;

#define VKVG_IDENTITY_MATRIX ((vkvg_matrix_t){1, 0, 0, 1, 0, 0})

/* :855-891 */
typedef enum _vkvg_operator {
    VKVG_OPERATOR_CLEAR,
    VKVG_OPERATOR_SOURCE,
    VKVG_OPERATOR_OVER,
    VKVG_OPERATOR_DIFFERENCE,
    VKVG_OPERATOR_MAX,
} vkvg_operator_t;

/* :555-564 */
typedef struct {
    VkSampleCountFlags samples;
    bool               deferredResolve;
    VkInstance         inst;
    VkPhysicalDevice   phy;
    VkDevice           vkdev;
    uint32_t           qFamIdx;
    uint32_t           qIndex;
    bool               threadAware;
} vkvg_device_create_info_t;

/* ---- matrices (include/vkvg.h:391-514) */
vkvg_public void vkvg_matrix_init_identity(vkvg_matrix_t *matrix);
vkvg_public void vkvg_matrix_init(vkvg_matrix_t *matrix, float xx, float yx, float xy, float yy, float x0, float y0);
vkvg_public void vkvg_matrix_init_translate(vkvg_matrix_t *matrix, float tx, float ty);
vkvg_public void vkvg_matrix_init_scale(vkvg_matrix_t *matrix, float sx, float sy);
vkvg_public void vkvg_matrix_init_rotate(vkvg_matrix_t *matrix, float radians);
vkvg_public void vkvg_matrix_translate(vkvg_matrix_t *matrix, float tx, float ty);
vkvg_public void vkvg_matrix_scale(vkvg_matrix_t *matrix, float sx, float sy);
vkvg_public void vkvg_matrix_rotate(vkvg_matrix_t *matrix, float radians);
vkvg_public void vkvg_matrix_multiply(vkvg_matrix_t *result, const vkvg_matrix_t *a, const vkvg_matrix_t *b);
vkvg_public void vkvg_matrix_transform_distance(const vkvg_matrix_t *matrix, float *dx, float *dy);
vkvg_public void vkvg_matrix_transform_point(const vkvg_matrix_t *matrix, float *x, float *y);
vkvg_public vkvg_status_t vkvg_matrix_invert(vkvg_matrix_t *matrix);
vkvg_public void vkvg_matrix_get_scale(const vkvg_matrix_t *matrix, float *sx, float *sy);

/* ---- device (include/vkvg.h:594-664) */
vkvg_public VkvgDevice    vkvg_device_create(vkvg_device_create_info_t *info); /* :594 */
vkvg_public void          vkvg_device_destroy(VkvgDevice dev);                 /* :640 */
vkvg_public vkvg_status_t vkvg_device_status(VkvgDevice dev);                  /* :648 */
vkvg_public VkvgDevice    vkvg_device_reference(VkvgDevice dev);               /* :656 */
vkvg_public uint32_t      vkvg_device_get_reference_count(VkvgDevice dev);     /* :664 */
vkvg_public void          vkvg_device_set_dpy(VkvgDevice dev, int hdpy, int vdpy);   /* :673 */
vkvg_public void          vkvg_device_get_dpy(VkvgDevice dev, int *hdpy, int *vdpy); /* :683 */
/* :577: contexts are not cached by this backend (creating one allocates no device memory); the
 * count is kept and read back by nothing */
vkvg_public void          vkvg_device_set_context_cache_size(VkvgDevice dev, uint32_t maxCount);
/* Vulkan set-up helpers (include/vkvg.h:692-710): this backend needs no Vulkan instance or device,
 * so it requires no extension and no feature chain */
vkvg_public void          vkvg_get_required_instance_extensions(const char **pExtensions, uint32_t *pExtCount); /* :692 */
vkvg_public vkvg_status_t vkvg_get_required_device_extensions(VkPhysicalDevice phy, const char **pExtensions,
                                                              uint32_t *pExtCount); /* :701 */
vkvg_public const void   *vkvg_get_device_requirements(VkPhysicalDeviceFeatures *pEnabledFeatures); /* :710 */

/* ---- surface (include/vkvg.h:725-845) */
vkvg_public VkvgSurface   vkvg_surface_create(VkvgDevice dev, uint32_t width, uint32_t height); /* :725 */
vkvg_public VkvgSurface   vkvg_surface_create_from_image(VkvgDevice dev, const char *filePath); /* :735 */
vkvg_public VkvgSurface   vkvg_surface_create_from_bitmap(VkvgDevice dev, unsigned char *img, uint32_t width,
                                                          uint32_t height); /* :752 */
vkvg_public vkvg_status_t vkvg_surface_status(VkvgSurface surf);                 /* :761 */
vkvg_public VkvgSurface   vkvg_surface_reference(VkvgSurface surf);              /* :767 */
vkvg_public uint32_t      vkvg_surface_get_reference_count(VkvgSurface surf);    /* :773 */
vkvg_public void          vkvg_surface_destroy(VkvgSurface surf);                /* :778 */
vkvg_public void          vkvg_surface_clear(VkvgSurface surf);                  /* :795 */
vkvg_public VkImage       vkvg_surface_get_vk_image(VkvgSurface surf);           /* :801: returns the HBM image pointer */
vkvg_public VkFormat      vkvg_surface_get_vk_format(VkvgSurface surf);          /* :807 */
vkvg_public uint32_t      vkvg_surface_get_width(VkvgSurface surf);              /* :813 */
vkvg_public uint32_t      vkvg_surface_get_height(VkvgSurface surf);             /* :819 */
vkvg_public vkvg_status_t vkvg_surface_write_to_png(VkvgSurface surf, const char *path);                 /* :826 */
vkvg_public vkvg_status_t vkvg_surface_write_to_memory(VkvgSurface surf, unsigned char *const bitmap);  /* :833 */
vkvg_public void          vkvg_surface_resolve(VkvgSurface surf);                /* :845 */

/* ---- context (include/vkvg.h:906-952) */
vkvg_public VkvgContext   vkvg_create(VkvgSurface surf);                   /* :906 */
vkvg_public void          vkvg_destroy(VkvgContext ctx);                   /* :913 */
vkvg_public vkvg_status_t vkvg_status(VkvgContext ctx);                    /* :921 */
vkvg_public const char   *vkvg_status_to_string(vkvg_status_t status);     /* :930 */
vkvg_public VkvgContext   vkvg_reference(VkvgContext ctx);                 /* :936 */
vkvg_public uint32_t      vkvg_get_reference_count(VkvgContext ctx);       /* :942 */
vkvg_public void          vkvg_flush(VkvgContext ctx);                     /* :952 */

/* ---- path (include/vkvg.h:961-1235) */
vkvg_public void vkvg_new_path(VkvgContext ctx);                                                 /* :961 */
vkvg_public void vkvg_close_path(VkvgContext ctx);                                               /* :970 */
vkvg_public void vkvg_new_sub_path(VkvgContext ctx);                                             /* :980 */
vkvg_public void vkvg_get_current_point(VkvgContext ctx, float *x, float *y);                    /* :998 */
vkvg_public void vkvg_path_extents(VkvgContext ctx, float *const x1, float *const y1, float *const x2,
                                   float *const y2);                                             /* :989 */
vkvg_public void vkvg_line_to(VkvgContext ctx, float x, float y);                                /* :1009 */
vkvg_public void vkvg_rel_line_to(VkvgContext ctx, float dx, float dy);                          /* :1020 */
vkvg_public void vkvg_move_to(VkvgContext ctx, float x, float y);                                /* :1033 */
vkvg_public void vkvg_rel_move_to(VkvgContext ctx, float x, float y);                            /* :1045 */
vkvg_public void vkvg_arc(VkvgContext ctx, float xc, float yc, float radius, float a1, float a2); /* :1071 */
vkvg_public void vkvg_arc_negative(VkvgContext ctx, float xc, float yc, float radius, float a1, float a2); /* :1089 */
vkvg_public void vkvg_curve_to(VkvgContext ctx, float x1, float y1, float x2, float y2, float x3, float y3); /* :1106 */
vkvg_public void vkvg_rel_curve_to(VkvgContext ctx, float x1, float y1, float x2, float y2, float x3, float y3); /* :1123 */
vkvg_public void vkvg_quadratic_to(VkvgContext ctx, float x1, float y1, float x2, float y2);     /* :1135 */
vkvg_public void vkvg_rel_quadratic_to(VkvgContext ctx, float x1, float y1, float x2, float y2); /* :1145 */
vkvg_public vkvg_status_t vkvg_rectangle(VkvgContext ctx, float x, float y, float w, float h);  /* :1157 */
vkvg_public vkvg_status_t vkvg_rounded_rectangle(VkvgContext ctx, float x, float y, float w, float h,
                                                 float radius); /* :1170 */
vkvg_public void vkvg_rounded_rectangle2(VkvgContext ctx, float x, float y, float w, float h, float rx,
                                         float ry); /* :1184 */
vkvg_public void vkvg_ellipse(VkvgContext ctx, float radiusX, float radiusY, float x, float y,
                              float rotationAngle); /* :1198 */
vkvg_public void vkvg_elliptic_arc_to(VkvgContext ctx, float x, float y, bool large_arc_flag, bool sweep_flag,
                                      float rx, float ry, float phi); /* :1219 */
vkvg_public void vkvg_rel_elliptic_arc_to(VkvgContext ctx, float x, float y, bool large_arc_flag, bool sweep_flag,
                                          float rx, float ry, float phi); /* :1235 */
vkvg_public bool vkvg_has_current_point(VkvgContext ctx); /* :1555 */

/* ---- painting (include/vkvg.h:1246-1291) */
vkvg_public void vkvg_stroke(VkvgContext ctx);          /* :1246 */
vkvg_public void vkvg_stroke_preserve(VkvgContext ctx); /* :1254 */
vkvg_public void vkvg_fill(VkvgContext ctx);            /* :1263 */
vkvg_public void vkvg_fill_preserve(VkvgContext ctx);   /* :1271 */
vkvg_public void vkvg_paint(VkvgContext ctx);           /* :1280 */
vkvg_public void vkvg_clear(VkvgContext ctx);           /* :1291 */

/* ---- clip and graphic-state stack (include/vkvg.h:1299-1325, :1565-1574) */
vkvg_public void vkvg_reset_clip(VkvgContext ctx);      /* :1299 */
vkvg_public void vkvg_clip(VkvgContext ctx);            /* :1317 */
vkvg_public void vkvg_clip_preserve(VkvgContext ctx);   /* :1325 */
vkvg_public void vkvg_save(VkvgContext ctx);            /* :1565 */
vkvg_public void vkvg_restore(VkvgContext ctx);         /* :1574 */

/* ---- state (include/vkvg.h:1334-1527) */
vkvg_public void  vkvg_set_opacity(VkvgContext ctx, float opacity);                            /* :1334 */
vkvg_public float vkvg_get_opacity(VkvgContext ctx);                                           /* :1335 */
vkvg_public void  vkvg_set_source_color(VkvgContext ctx, uint32_t c);                          /* :1350 */
vkvg_public void  vkvg_set_source_rgba(VkvgContext ctx, float r, float g, float b, float a);   /* :1361 */
vkvg_public void  vkvg_set_source_rgb(VkvgContext ctx, float r, float g, float b);             /* :1376 */
vkvg_public void  vkvg_set_source_surface(VkvgContext ctx, VkvgSurface surf, float x, float y); /* :1386 */
vkvg_public void  vkvg_set_source(VkvgContext ctx, VkvgPattern pat);                           /* :1394 */
vkvg_public void  vkvg_set_line_width(VkvgContext ctx, float width);                           /* :1404 */
vkvg_public void  vkvg_set_miter_limit(VkvgContext ctx, float limit);                          /* :1421 */
vkvg_public float vkvg_get_miter_limit(VkvgContext ctx);                                       /* :1430 */
vkvg_public void  vkvg_set_line_cap(VkvgContext ctx, vkvg_line_cap_t cap);                     /* :1438 */
vkvg_public void  vkvg_set_line_join(VkvgContext ctx, vkvg_line_join_t join);                  /* :1446 */
vkvg_public void  vkvg_set_operator(VkvgContext ctx, vkvg_operator_t op);                      /* :1453 */
vkvg_public void  vkvg_set_fill_rule(VkvgContext ctx, vkvg_fill_rule_t fr);                    /* :1460 */
vkvg_public void  vkvg_set_dash(VkvgContext ctx, const float *dashes, uint32_t num_dashes, float offset); /* :1473 */
vkvg_public void  vkvg_get_dash(VkvgContext ctx, const float *dashes, uint32_t *num_dashes, float *offset); /* :1485 */
vkvg_public float vkvg_get_line_width(VkvgContext ctx);                                        /* :1494 */
vkvg_public vkvg_line_cap_t  vkvg_get_line_cap(VkvgContext ctx);                               /* :1501 */
vkvg_public vkvg_line_join_t vkvg_get_line_join(VkvgContext ctx);                              /* :1508 */
vkvg_public vkvg_operator_t  vkvg_get_operator(VkvgContext ctx);                               /* :1515 */
vkvg_public vkvg_fill_rule_t vkvg_get_fill_rule(VkvgContext ctx);                              /* :1527 */
vkvg_public VkvgPattern      vkvg_get_source(VkvgContext ctx);                                 /* :1536 */
vkvg_public VkvgSurface      vkvg_get_target(VkvgContext ctx);                                 /* :1545 */

/* ---- transforms (include/vkvg.h:1584-1637) */
vkvg_public void vkvg_translate(VkvgContext ctx, float dx, float dy);                  /* :1584 */
vkvg_public void vkvg_scale(VkvgContext ctx, float sx, float sy);                      /* :1594 */
vkvg_public void vkvg_rotate(VkvgContext ctx, float radians);                          /* :1602 */
vkvg_public void vkvg_transform(VkvgContext ctx, const vkvg_matrix_t *matrix);         /* :1611 */
vkvg_public void vkvg_set_matrix(VkvgContext ctx, const vkvg_matrix_t *matrix);        /* :1620 */
vkvg_public void vkvg_get_matrix(VkvgContext ctx, vkvg_matrix_t *const matrix);        /* :1629 */
vkvg_public void vkvg_identity_matrix(VkvgContext ctx);                                /* :1637 */

/* ---- patterns (include/vkvg.h:1763-1953) */
vkvg_public vkvg_status_t vkvg_pattern_status(VkvgPattern pat);                                    /* :1763 */
vkvg_public VkvgPattern   vkvg_pattern_reference(VkvgPattern pat);                                 /* :1771 */
vkvg_public uint32_t      vkvg_pattern_get_reference_count(VkvgPattern pat);                       /* :1779 */
vkvg_public VkvgPattern   vkvg_pattern_create_for_surface(VkvgSurface surf);                       /* :1790 */
vkvg_public VkvgPattern   vkvg_pattern_create_linear(float x0, float y0, float x1, float y1);      /* :1803 */
vkvg_public vkvg_status_t vkvg_pattern_edit_linear(VkvgPattern pat, float x0, float y0, float x1, float y1); /* :1815 */
vkvg_public vkvg_status_t vkvg_pattern_get_linear_points(VkvgPattern pat, float *x0, float *y0, float *x1,
                                                         float *y1); /* :1827 */
vkvg_public VkvgPattern   vkvg_pattern_create_radial(float cx0, float cy0, float radius0, float cx1, float cy1,
                                                     float radius1); /* :1842 */
vkvg_public vkvg_status_t vkvg_pattern_edit_radial(VkvgPattern pat, float cx0, float cy0, float radius0, float cx1,
                                                   float cy1, float radius1); /* :1858 */
vkvg_public vkvg_status_t vkvg_pattern_get_color_stop_count(VkvgPattern pat, uint32_t *count); /* :1869 */
vkvg_public vkvg_status_t vkvg_pattern_get_color_stop_rgba(VkvgPattern pat, uint32_t index, float *offset, float *r,
                                                           float *g, float *b, float *a); /* :1886 */
vkvg_public void          vkvg_pattern_destroy(VkvgPattern pat);                          /* :1896 */
vkvg_public vkvg_status_t vkvg_pattern_add_color_stop(VkvgPattern pat, float offset, float r, float g, float b,
                                                      float a); /* :1910 */
vkvg_public void          vkvg_pattern_set_extend(VkvgPattern pat, vkvg_extend_t extend);  /* :1921 */
vkvg_public void          vkvg_pattern_set_filter(VkvgPattern pat, vkvg_filter_t filter);  /* :1928 */
vkvg_public vkvg_extend_t vkvg_pattern_get_extend(VkvgPattern pat);                        /* :1935 */
vkvg_public vkvg_filter_t vkvg_pattern_get_filter(VkvgPattern pat);                        /* :1942 */
vkvg_public vkvg_pattern_type_t vkvg_pattern_get_type(VkvgPattern pat);                    /* :1951 */
vkvg_public void vkvg_pattern_set_matrix(VkvgPattern pat, const vkvg_matrix_t *matrix);    /* :1952 */
vkvg_public void vkvg_pattern_get_matrix(VkvgPattern pat, vkvg_matrix_t *matrix);          /* :1953 */

/* ---- recording (include/vkvg.h:1960-1970; the reference builds it with VKVG_RECORDING, which this
 * header defines). Between start and stop a context records its calls instead of executing them;
 * vkvg_replay calls them again on any context. */
#ifndef VKVG_RECORDING
#define VKVG_RECORDING 1
#endif
typedef struct _vkvg_recording_t *VkvgRecording;
vkvg_public void          vkvg_start_recording(VkvgContext ctx);
vkvg_public VkvgRecording vkvg_stop_recording(VkvgContext ctx);
vkvg_public void          vkvg_replay(VkvgContext ctx, VkvgRecording rec);
vkvg_public void          vkvg_replay_command(VkvgContext ctx, VkvgRecording rec, uint32_t cmdIndex);
vkvg_public void          vkvg_recording_get_command(VkvgRecording rec, uint32_t cmdIndex, uint32_t *cmd,
                                                     void **dataOffset);
vkvg_public uint32_t      vkvg_recording_get_count(VkvgRecording rec);
vkvg_public void         *vkvg_recording_get_data(VkvgRecording rec);
vkvg_public void          vkvg_recording_destroy(VkvgRecording rec);

/* ---- text (include/vkvg.h:1645-1745): no font support; these draw nothing, leave the context's
 * status unchanged, report zero extents and zero glyphs, and vkvg_text_run_create returns NULL */
vkvg_public void     vkvg_select_font_face(VkvgContext ctx, const char *name);                    /* :1645 */
vkvg_public void     vkvg_load_font_from_path(VkvgContext ctx, const char *path, const char *name); /* :1653 */
vkvg_public void     vkvg_load_font_from_memory(VkvgContext ctx, unsigned char *fontBuffer, long fontBufferByteSize,
                                                const char *name);                                /* :1662 */
vkvg_public void     vkvg_set_font_size(VkvgContext ctx, uint32_t size);                          /* :1670 */
vkvg_public void     vkvg_show_text(VkvgContext ctx, const char *utf8);                           /* :1680 */
vkvg_public void     vkvg_text_extents(VkvgContext ctx, const char *utf8, vkvg_text_extents_t *extents); /* :1688 */
vkvg_public void     vkvg_font_extents(VkvgContext ctx, vkvg_font_extents_t *extents);            /* :1695 */
vkvg_public VkvgText vkvg_text_run_create(VkvgContext ctx, const char *text);                     /* :1705 */
vkvg_public VkvgText vkvg_text_run_create_with_length(VkvgContext ctx, const char *text, uint32_t length); /* :1714 */
vkvg_public void     vkvg_text_run_destroy(VkvgText textRun);                                     /* :1720 */
vkvg_public void     vkvg_show_text_run(VkvgContext ctx, VkvgText textRun);                       /* :1727 */
vkvg_public void     vkvg_text_run_get_extents(VkvgText textRun, vkvg_text_extents_t *extents);   /* :1734 */
vkvg_public uint32_t vkvg_text_run_get_glyph_count(VkvgText textRun);                             /* :1740 */
vkvg_public void     vkvg_text_run_get_glyph_position(VkvgText textRun, uint32_t index,
                                                      vkvg_glyph_info_t *pGlyphInfo);             /* :1745 */

#ifdef __cplusplus
}
#endif
#endif
