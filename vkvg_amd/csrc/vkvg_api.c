/*
 * vkvg_api.c — the vkvg C API of the MI355X backend: host recorder + objects.
 *
 * Path calls are recorded as *items* (explicit points and unflattened cubics) with the
 * reference's path bookkeeping kept exactly: duplicate-point removal (src/vkvg_context.c
 * _line_to, src/vkvg_context_internal.c:1463-1472), sub-path finishing and discard
 * (:163-197), close_path's closing-point removal (src/vkvg_context.c:350-373), the curve
 * segments that force stroke joins (:136-160) and the convex/empty rules. Flattening,
 * stroking and rasterization run on the GPU at flush time through include/vkvg_hip.h.
 *
 * A point count is needed on the host only for the close/finish tests (< 2, < 3, < 4
 * points). A finite cubic always flattens to at least 2 interior points plus its end
 * point (the first AGG level always subdivides and every finite branch ends in an
 * emitting leaf), so a sub-path holding a cubic always passes those tests. Cubics with
 * non-finite control points are ignored (the reference would recurse 2^100 times).
 */
#define _GNU_SOURCE
#include <float.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <zlib.h>

#include "vkvg.h"
#include "vkvg_amd.h"
#include "vkvg_hip.h"
#include "vkvg_record.h"

#define PIF   3.14159265358979323846f
#define PIF_2 1.57079632679489661923f
#define EQUF(a, b) (fabsf((a) - (b)) <= FLT_EPSILON)
/* transcendental functions evaluated in double, rounded once (as the GPU kernels do) */
#define D_COSF(a) ((float)cos((double)(a)))
#define D_SINF(a) ((float)sin((double)(a)))
#define D_ACOSF(a) ((float)acos((double)(a)))
#define D_POWF(a, b) ((float)pow((double)(a), (double)(b)))

/* ---------------------------------------------------------------------------- growable arrays */
#define VEC(T) struct { T *p; uint32_t n, cap; }
#define VEC_RESERVE(v, want)                                                                                           \
    do {                                                                                                               \
        if ((want) > (v).cap) {                                                                                        \
            uint32_t nc_ = (v).cap ? (v).cap : 64;                                                                     \
            while (nc_ < (want))                                                                                       \
                nc_ *= 2;                                                                                              \
            void *np_ = realloc((v).p, (size_t)nc_ * sizeof(*(v).p));                                                  \
            if (np_) {                                                                                                 \
                (v).p = np_;                                                                                           \
                (v).cap = nc_;                                                                                         \
            }                                                                                                          \
        }                                                                                                              \
    } while (0)
#define VEC_PUSH(v, val)                                                                                               \
    do {                                                                                                               \
        VEC_RESERVE(v, (v).n + 1);                                                                                     \
        (v).p[(v).n++] = (val);                                                                                        \
    } while (0)

typedef struct { float x, y; } v2;

/* ---------------------------------------------------------------------------- objects */
struct _vkvg_device_t {
    vkvg_status_t status;
    uint32_t      references;
    vkh_engine   *eng;
    int           hdpy, vdpy;
    uint32_t      ctx_cache_size; /* vkvg_device_set_context_cache_size: kept, contexts are not cached */
};

struct _vkvg_surface_t {
    vkvg_status_t status;
    uint32_t      references;
    VkvgDevice    dev;
    uint32_t      width, height;
    void         *img; /* device RGBA8 premultiplied */
    uint32_t     *sig; /* device coverage signatures, optional */
    uint32_t     *stencil; /* device stencil words, allocated by the first clip */
    uint32_t      draw_counter;
};

typedef struct {
    uint32_t count;
    float    colors[16][4];
    float    stops[16];
    float    cp[2][4];
} grad_data;

struct _vkvg_pattern_t {
    vkvg_status_t       status;
    uint32_t            references;
    vkvg_pattern_type_t type;
    vkvg_extend_t       extend;
    vkvg_filter_t       filter;
    vkvg_matrix_t       matrix;
    bool                hasMatrix;
    grad_data           g;
    VkvgSurface         surf; /* SURFACE patterns: the referenced source surface */
};

typedef struct {
    VEC(vkh_item) items;
    VEC(vkh_subpath) sps;
    VEC(vkh_draw) draws;
    VEC(vkh_stroke_style) styles;
    VEC(float) dashes;
    VEC(vkh_gradient) grads;
    VEC(vkh_surface_source) srcs;
    VEC(VkvgSurface) src_refs; /* source surfaces referenced until the frame has rendered */
} frame_t;

struct _vkvg_context_t {
    vkvg_status_t status;
    uint32_t      references;
    VkvgSurface   surf;
    vkvg_rec     *recording; /* non-NULL between vkvg_start_recording and vkvg_stop_recording */
    bool          deferred; /* batch mode: frames are collected, not rendered */

    /* graphic state */
    float            lineWidth, miterLimit, opacity, dashOffset;
    float           *dashes;
    uint32_t         dashCount;
    vkvg_line_cap_t  lineCap;
    vkvg_line_join_t lineJoin;
    vkvg_fill_rule_t fillRule;
    vkvg_operator_t  op;
    uint32_t         curColor;
    vkvg_matrix_t    mat;
    VkvgPattern      pattern;  /* referenced source pattern or NULL */
    int32_t          cur_grad; /* index in frame.grads, -1 for solid */
    /* the push constants a surface source reads (src/vkvg_context_internal.h:74-81): the source
     * rectangle and the inverse matrix, updated as the reference updates them */
    float            src_rect[4];
    vkvg_matrix_t    mat_inv;

    /* path recording */
    frame_t  f;
    uint32_t path_sp_begin;   /* first committed subpath of the current path */
    uint32_t path_item_begin; /* first item of the current path */
    bool     path_drawn;      /* a draw references the current path */
    uint32_t cur_first_item;  /* first item of the open subpath */
    uint32_t cur_count;       /* point count of the open subpath (exact without cubics) */
    bool     cur_has_cubic;
    bool     cur_closed, cur_convex;
    v2       cur_first, cur_last;
    bool     simpleConvex;
    uint32_t seg_ptr;         /* 0 = not segmented */
    uint32_t seg_first_item;  /* first item of the current segment */
    uint32_t seg_count;

    /* clip and save/restore (src/vkvg_context_internal.h:87-130, :218-224) */
    int              clip_state;  /* CLIP_STATE_* */
    uint32_t         sav_bit;     /* stencil save levels in use (curSavBit) */
    struct ctx_save *saved;       /* vkvg_save stack */
    int              stencil_use; /* STENCIL_*: how this context's frames use the surface's plane */
};

/* vkvg_clip_state_t, src/vkvg_context_internal.h:94-99 */
enum { CLIP_STATE_NONE = 0, CLIP_STATE_CLEAR = 1, CLIP_STATE_CLIP = 2, CLIP_STATE_SAVED = 6 };
/* the stencil plane starts from zero in the first frame of a context that clips (a context's first
 * render pass clears the stencil, src/vkvg_context.c:42-49); later frames load it */
enum { STENCIL_UNUSED = 0, STENCIL_FIRST_FRAME = 1, STENCIL_LOADED = 2 };

/* vkvg_context_save_t, src/vkvg_context_internal.h:101-125 */
struct ctx_save {
    struct ctx_save *next;
    int              clip_state;
    float            dashOffset;
    uint32_t         dashCount;
    float           *dashes;
    float            lineWidth, miterLimit, opacity;
    vkvg_operator_t  op;
    vkvg_line_cap_t  lineCap;
    vkvg_fill_rule_t fillRule;
    vkvg_matrix_t    mat, mat_inv;
    float            src_rect[4];
    uint32_t         curColor;
    VkvgPattern      pattern; /* referenced */
};

static vkvg_status_t s_no_mem = VKVG_STATUS_NO_MEMORY;
static vkvg_status_t s_null = VKVG_STATUS_NULL_POINTER;
static vkvg_status_t s_dev_err = VKVG_STATUS_DEVICE_ERROR;
static vkvg_status_t s_inv_surf = VKVG_STATUS_INVALID_SURFACE;
static vkvg_status_t s_inv_dev_ci = VKVG_STATUS_INVALID_DEVICE_CREATE_INFO;

/* ---------------------------------------------------------------------------- matrices
 * src/vkvg_matrix.c:107-229 (cairo's formulas) */
void vkvg_matrix_init(vkvg_matrix_t *m, float xx, float yx, float xy, float yy, float x0, float y0) {
    m->xx = xx;
    m->yx = yx;
    m->xy = xy;
    m->yy = yy;
    m->x0 = x0;
    m->y0 = y0;
}
void vkvg_matrix_init_identity(vkvg_matrix_t *m) { vkvg_matrix_init(m, 1, 0, 0, 1, 0, 0); }
void vkvg_matrix_init_translate(vkvg_matrix_t *m, float tx, float ty) { vkvg_matrix_init(m, 1, 0, 0, 1, tx, ty); }
void vkvg_matrix_init_scale(vkvg_matrix_t *m, float sx, float sy) { vkvg_matrix_init(m, sx, 0, 0, sy, 0, 0); }
void vkvg_matrix_init_rotate(vkvg_matrix_t *m, float radians) {
    float s = D_SINF(radians), c = D_COSF(radians);
    vkvg_matrix_init(m, c, s, -s, c, 0, 0);
}
void vkvg_matrix_multiply(vkvg_matrix_t *result, const vkvg_matrix_t *a, const vkvg_matrix_t *b) {
    vkvg_matrix_t r;
    r.xx = a->xx * b->xx + a->yx * b->xy;
    r.yx = a->xx * b->yx + a->yx * b->yy;
    r.xy = a->xy * b->xx + a->yy * b->xy;
    r.yy = a->xy * b->yx + a->yy * b->yy;
    r.x0 = a->x0 * b->xx + a->y0 * b->xy + b->x0;
    r.y0 = a->x0 * b->yx + a->y0 * b->yy + b->y0;
    *result = r;
}
void vkvg_matrix_translate(vkvg_matrix_t *m, float tx, float ty) {
    vkvg_matrix_t t;
    vkvg_matrix_init_translate(&t, tx, ty);
    vkvg_matrix_multiply(m, &t, m);
}
void vkvg_matrix_scale(vkvg_matrix_t *m, float sx, float sy) {
    vkvg_matrix_t t;
    vkvg_matrix_init_scale(&t, sx, sy);
    vkvg_matrix_multiply(m, &t, m);
}
void vkvg_matrix_rotate(vkvg_matrix_t *m, float radians) {
    vkvg_matrix_t t;
    vkvg_matrix_init_rotate(&t, radians);
    vkvg_matrix_multiply(m, &t, m);
}
void vkvg_matrix_transform_distance(const vkvg_matrix_t *m, float *dx, float *dy) {
    float nx = (m->xx * *dx + m->xy * *dy);
    float ny = (m->yx * *dx + m->yy * *dy);
    *dx = nx;
    *dy = ny;
}
void vkvg_matrix_transform_point(const vkvg_matrix_t *m, float *x, float *y) {
    vkvg_matrix_transform_distance(m, x, y);
    *x += m->x0;
    *y += m->y0;
}
vkvg_status_t vkvg_matrix_invert(vkvg_matrix_t *m) {
    if (m->xy == 0. && m->yx == 0.) {
        m->x0 = -m->x0;
        m->y0 = -m->y0;
        if (m->xx != 1.f) {
            if (m->xx == 0.)
                return VKVG_STATUS_INVALID_MATRIX;
            m->xx = 1.f / m->xx;
            m->x0 *= m->xx;
        }
        if (m->yy != 1.f) {
            if (m->yy == 0.)
                return VKVG_STATUS_INVALID_MATRIX;
            m->yy = 1.f / m->yy;
            m->y0 *= m->yy;
        }
        return VKVG_STATUS_SUCCESS;
    }
    float det = m->xx * m->yy - m->yx * m->xy;
    if (!(det * det >= 0.) || det == 0)
        return VKVG_STATUS_INVALID_MATRIX;
    float a = m->xx, b = m->yx, c = m->xy, d = m->yy, tx = m->x0, ty = m->y0;
    vkvg_matrix_init(m, d, -b, -c, a, c * ty - d * tx, b * tx - a * ty);
    float s = 1 / det;
    m->xx *= s;
    m->yx *= s;
    m->xy *= s;
    m->yy *= s;
    m->x0 *= s;
    m->y0 *= s;
    return VKVG_STATUS_SUCCESS;
}
void vkvg_matrix_get_scale(const vkvg_matrix_t *m, float *sx, float *sy) {
    *sx = (float)sqrt(m->xx * m->xx + m->xy * m->xy);
    *sy = (float)sqrt(m->yx * m->yx + m->yy * m->yy);
}

/* ---------------------------------------------------------------------------- device */
/* src/vkvg_device.c:286-291: a NULL create-info is an error; its samples and Vulkan handles have no
 * meaning here (1 sample, HIP device 0) */
VkvgDevice vkvg_device_create(vkvg_device_create_info_t *info) {
    if (!info)
        return (VkvgDevice)&s_inv_dev_ci;
    VkvgDevice dev = (VkvgDevice)calloc(1, sizeof(struct _vkvg_device_t));
    if (!dev)
        return (VkvgDevice)&s_no_mem;
    dev->eng = vkvg_hip_engine_create(0);
    if (!dev->eng) {
        free(dev);
        return (VkvgDevice)&s_dev_err;
    }
    dev->references = 1;
    dev->hdpy = dev->vdpy = 96;
    return dev;
}
void vkvg_get_required_instance_extensions(const char **pExtensions, uint32_t *pExtCount) {
    (void)pExtensions;
    if (pExtCount)
        *pExtCount = 0;
}
vkvg_status_t vkvg_get_required_device_extensions(VkPhysicalDevice phy, const char **pExtensions, uint32_t *pExtCount) {
    (void)phy;
    (void)pExtensions;
    if (pExtCount)
        *pExtCount = 0;
    return VKVG_STATUS_SUCCESS;
}
const void *vkvg_get_device_requirements(VkPhysicalDeviceFeatures *pEnabledFeatures) {
    (void)pEnabledFeatures;
    return NULL;
}
vkvg_status_t vkvg_device_status(VkvgDevice dev) { return !dev ? VKVG_STATUS_NULL_POINTER : dev->status; }
void vkvg_device_set_context_cache_size(VkvgDevice dev, uint32_t maxCount) {
    if (!vkvg_device_status(dev))
        dev->ctx_cache_size = maxCount;
}
VkvgDevice vkvg_device_reference(VkvgDevice dev) {
    if (!vkvg_device_status(dev))
        dev->references++;
    return dev;
}
uint32_t vkvg_device_get_reference_count(VkvgDevice dev) { return vkvg_device_status(dev) ? 0 : dev->references; }
void vkvg_device_destroy(VkvgDevice dev) {
    if (vkvg_device_status(dev))
        return;
    if (--dev->references > 0)
        return;
    vkvg_hip_engine_destroy(dev->eng);
    free(dev);
}
void vkvg_device_set_dpy(VkvgDevice dev, int hdpy, int vdpy) {
    if (vkvg_device_status(dev))
        return;
    dev->hdpy = hdpy;
    dev->vdpy = vdpy;
}
void vkvg_device_get_dpy(VkvgDevice dev, int *hdpy, int *vdpy) {
    if (vkvg_device_status(dev))
        return;
    *hdpy = dev->hdpy;
    *vdpy = dev->vdpy;
}
void vkvg_amd_device_get_stats(VkvgDevice dev, vkh_stats *out) {
    if (!vkvg_device_status(dev))
        vkvg_hip_get_stats(dev->eng, out);
}
int vkvg_amd_device_set_tile_size(VkvgDevice dev, uint32_t tile) {
    return vkvg_device_status(dev) ? -1 : vkvg_hip_set_tile_size(dev->eng, tile);
}
vkh_engine *vkvg_amd_device_engine(VkvgDevice dev) { return vkvg_device_status(dev) ? NULL : dev->eng; }

/* ---------------------------------------------------------------------------- surface */
VkvgSurface vkvg_surface_create(VkvgDevice dev, uint32_t width, uint32_t height) {
    if (vkvg_device_status(dev))
        return (VkvgSurface)&s_dev_err;
    if (width == 0 || height == 0 || width > 32768 || height > 32768)
        return (VkvgSurface)&s_inv_surf;
    VkvgSurface s = (VkvgSurface)calloc(1, sizeof(struct _vkvg_surface_t));
    if (!s)
        return (VkvgSurface)&s_no_mem;
    s->dev = dev;
    s->width = width;
    s->height = height;
    if (vkvg_hip_image_alloc(dev->eng, width, height, &s->img)) {
        free(s);
        return (VkvgSurface)&s_dev_err;
    }
    s->references = 1;
    vkvg_device_reference(dev);
    return s;
}
VkvgSurface vkvg_surface_create_from_bitmap(VkvgDevice dev, unsigned char *img, uint32_t width, uint32_t height) {
    VkvgSurface s = vkvg_surface_create(dev, width, height);
    if (vkvg_surface_status(s) || !img)
        return s;
    /* the reference blits the bytes into the surface and paints them through a surface source onto
     * the cleared image (src/vkvg_surface.c:82-179): nearest samples at pixel centres, opacity 1,
     * OVER onto zero, so the surface holds the bytes as given (used as premultiplied from then on) */
    vkvg_hip_image_upload(dev->eng, s->img, img, (size_t)width * height * 4);
    return s;
}
/* vkvg_surface_create_from_image, src/vkvg_surface.c:181-194: the decoded file (PNG here) through
 * vkvg_surface_create_from_bitmap; a file that cannot be read or decoded gives a NULL_POINTER surface */
VkvgSurface vkvg_surface_create_from_image(VkvgDevice dev, const char *filePath) {
    if (vkvg_device_status(dev))
        return (VkvgSurface)&s_dev_err;
    FILE *f = filePath ? fopen(filePath, "rb") : NULL;
    if (!f)
        return (VkvgSurface)&s_null;
    uint8_t *buf = NULL;
    size_t n = 0, cap = 0, got;
    do {
        if (n == cap) {
            cap = cap ? cap * 2 : 1u << 16;
            uint8_t *p = (uint8_t *)realloc(buf, cap);
            if (!p) {
                free(buf);
                fclose(f);
                return (VkvgSurface)&s_no_mem;
            }
            buf = p;
        }
        got = fread(buf + n, 1, cap - n, f);
        n += got;
    } while (got > 0);
    fclose(f);
    uint32_t w = 0, h = 0;
    uint8_t *img = vkvg_amd_png_decode(buf, n, &w, &h);
    free(buf);
    if (!img)
        return (VkvgSurface)&s_null;
    VkvgSurface s = vkvg_surface_create_from_bitmap(dev, img, w, h);
    vkvg_amd_png_free(img);
    return s;
}
vkvg_status_t vkvg_surface_status(VkvgSurface surf) { return !surf ? VKVG_STATUS_NULL_POINTER : surf->status; }
VkvgSurface vkvg_surface_reference(VkvgSurface surf) {
    if (!vkvg_surface_status(surf))
        surf->references++;
    return surf;
}
uint32_t vkvg_surface_get_reference_count(VkvgSurface surf) {
    return vkvg_surface_status(surf) ? 0 : surf->references;
}
void vkvg_surface_destroy(VkvgSurface surf) {
    if (vkvg_surface_status(surf))
        return;
    if (--surf->references > 0)
        return;
    VkvgDevice dev = surf->dev;
    vkvg_hip_image_free(dev->eng, surf->img);
    if (surf->sig)
        vkvg_hip_image_free(dev->eng, surf->sig);
    if (surf->stencil)
        vkvg_hip_image_free(dev->eng, surf->stencil);
    free(surf);
    vkvg_device_destroy(dev);
}
/* vkvg_surface_clear, src/vkvg_surface.c:32-36: colour and stencil cleared. The stencil plane's
 * word per pixel keeps the saved clip levels a context spilled in its upper bytes (context-owned
 * images in the reference), so only the stencil byte is cleared. */
void vkvg_surface_clear(VkvgSurface surf) {
    if (vkvg_surface_status(surf))
        return;
    vkvg_hip_image_clear(surf->dev->eng, surf->img, (size_t)surf->width * surf->height * 4);
    if (surf->stencil)
        vkvg_hip_stencil_clear(surf->dev->eng, surf->stencil, (size_t)surf->width * surf->height);
}
VkImage vkvg_surface_get_vk_image(VkvgSurface surf) {
    return vkvg_surface_status(surf) ? NULL : (VkImage)surf->img;
}
VkFormat vkvg_surface_get_vk_format(VkvgSurface surf) {
    return vkvg_surface_status(surf) ? VK_FORMAT_UNDEFINED : VK_FORMAT_B8G8R8A8_UNORM;
}
uint32_t vkvg_surface_get_width(VkvgSurface surf) { return vkvg_surface_status(surf) ? 0 : surf->width; }
uint32_t vkvg_surface_get_height(VkvgSurface surf) { return vkvg_surface_status(surf) ? 0 : surf->height; }
void vkvg_surface_resolve(VkvgSurface surf) { (void)surf; /* single-sample surfaces need no resolve */ }
void *vkvg_amd_surface_device_ptr(VkvgSurface surf) { return vkvg_surface_status(surf) ? NULL : surf->img; }

vkvg_status_t vkvg_surface_write_to_memory(VkvgSurface surf, unsigned char *const bitmap) {
    if (vkvg_surface_status(surf))
        return VKVG_STATUS_INVALID_STATUS;
    if (!bitmap)
        return VKVG_STATUS_WRITE_ERROR;
    if (vkvg_hip_image_read_unpremultiplied(surf->dev->eng, surf->img, surf->width, surf->height, bitmap))
        return VKVG_STATUS_DEVICE_ERROR;
    return VKVG_STATUS_SUCCESS;
}
vkvg_status_t vkvg_amd_surface_read_premultiplied(VkvgSurface surf, unsigned char *out) {
    if (vkvg_surface_status(surf))
        return VKVG_STATUS_INVALID_STATUS;
    if (!out)
        return VKVG_STATUS_WRITE_ERROR;
    if (vkvg_hip_image_download(surf->dev->eng, surf->img, out, (size_t)surf->width * surf->height * 4))
        return VKVG_STATUS_DEVICE_ERROR;
    return VKVG_STATUS_SUCCESS;
}
vkvg_status_t vkvg_amd_surface_enable_signature(VkvgSurface surf) {
    if (vkvg_surface_status(surf))
        return VKVG_STATUS_INVALID_STATUS;
    if (surf->sig)
        return VKVG_STATUS_SUCCESS;
    void *p = NULL;
    if (vkvg_hip_image_alloc(surf->dev->eng, surf->width, surf->height, &p))
        return VKVG_STATUS_DEVICE_ERROR;
    surf->sig = (uint32_t *)p;
    return VKVG_STATUS_SUCCESS;
}
vkvg_status_t vkvg_amd_surface_read_signature(VkvgSurface surf, uint32_t *out) {
    if (vkvg_surface_status(surf))
        return VKVG_STATUS_INVALID_STATUS;
    if (!surf->sig || !out)
        return VKVG_STATUS_WRITE_ERROR;
    if (vkvg_hip_image_download(surf->dev->eng, surf->sig, out, (size_t)surf->width * surf->height * 4))
        return VKVG_STATUS_DEVICE_ERROR;
    return VKVG_STATUS_SUCCESS;
}

/* minimal PNG encoder (RGBA8, zlib deflate) */
static void png_u32(unsigned char *p, uint32_t v) {
    p[0] = (unsigned char)(v >> 24);
    p[1] = (unsigned char)(v >> 16);
    p[2] = (unsigned char)(v >> 8);
    p[3] = (unsigned char)v;
}
static int png_chunk(FILE *fp, const char *type, const unsigned char *data, uint32_t len) {
    unsigned char hdr[8];
    png_u32(hdr, len);
    memcpy(hdr + 4, type, 4);
    uLong crc = crc32(0L, (const Bytef *)type, 4);
    if (len)
        crc = crc32(crc, data, len);
    unsigned char c[4];
    png_u32(c, (uint32_t)crc);
    return fwrite(hdr, 1, 8, fp) == 8 && (len == 0 || fwrite(data, 1, len, fp) == len) && fwrite(c, 1, 4, fp) == 4;
}
vkvg_status_t vkvg_surface_write_to_png(VkvgSurface surf, const char *path) {
    if (vkvg_surface_status(surf))
        return VKVG_STATUS_INVALID_STATUS;
    if (!path)
        return VKVG_STATUS_WRITE_ERROR;
    uint32_t w = surf->width, h = surf->height;
    unsigned char *img = (unsigned char *)malloc((size_t)w * h * 4);
    unsigned char *raw = (unsigned char *)malloc((size_t)(w * 4 + 1) * h);
    if (!img || !raw) {
        free(img);
        free(raw);
        return VKVG_STATUS_NO_MEMORY;
    }
    vkvg_status_t st = vkvg_surface_write_to_memory(surf, img);
    if (st) {
        free(img);
        free(raw);
        return st;
    }
    for (uint32_t y = 0; y < h; y++) {
        raw[(size_t)y * (w * 4 + 1)] = 0;
        memcpy(raw + (size_t)y * (w * 4 + 1) + 1, img + (size_t)y * w * 4, (size_t)w * 4);
    }
    uLongf zlen = compressBound((uLong)((size_t)(w * 4 + 1) * h));
    unsigned char *z = (unsigned char *)malloc(zlen);
    FILE *fp = fopen(path, "wb");
    int ok = z && fp && compress2(z, &zlen, raw, (uLong)((size_t)(w * 4 + 1) * h), 6) == Z_OK;
    if (ok) {
        static const unsigned char sigb[8] = {0x89, 'P', 'N', 'G', '\r', '\n', 0x1a, '\n'};
        unsigned char ihdr[13];
        png_u32(ihdr, w);
        png_u32(ihdr + 4, h);
        ihdr[8] = 8;
        ihdr[9] = 6;
        ihdr[10] = ihdr[11] = ihdr[12] = 0;
        ok = fwrite(sigb, 1, 8, fp) == 8 && png_chunk(fp, "IHDR", ihdr, 13) && png_chunk(fp, "IDAT", z, (uint32_t)zlen) &&
             png_chunk(fp, "IEND", NULL, 0);
    }
    if (fp)
        fclose(fp);
    free(z);
    free(img);
    free(raw);
    return ok ? VKVG_STATUS_SUCCESS : VKVG_STATUS_WRITE_ERROR;
}

/* ---------------------------------------------------------------------------- patterns
 * src/vkvg_pattern.c */
static VkvgPattern pattern_new(vkvg_pattern_type_t type) {
    VkvgPattern p = (VkvgPattern)calloc(1, sizeof(struct _vkvg_pattern_t));
    if (!p)
        return (VkvgPattern)&s_null;
    p->type = type;
    p->extend = VKVG_EXTEND_NONE;
    p->references = 1;
    return p;
}
vkvg_status_t vkvg_pattern_status(VkvgPattern pat) { return !pat ? VKVG_STATUS_NULL_POINTER : pat->status; }
VkvgPattern vkvg_pattern_reference(VkvgPattern pat) {
    if (!vkvg_pattern_status(pat))
        pat->references++;
    return pat;
}
uint32_t vkvg_pattern_get_reference_count(VkvgPattern pat) { return vkvg_pattern_status(pat) ? 0 : pat->references; }
vkvg_status_t vkvg_pattern_edit_linear(VkvgPattern pat, float x0, float y0, float x1, float y1) {
    if (vkvg_pattern_status(pat))
        return vkvg_pattern_status(pat);
    if (pat->type != VKVG_PATTERN_TYPE_LINEAR)
        return VKVG_STATUS_PATTERN_TYPE_MISMATCH;
    pat->g.cp[0][0] = x0;
    pat->g.cp[0][1] = y0;
    pat->g.cp[0][2] = x1;
    pat->g.cp[0][3] = y1;
    return VKVG_STATUS_SUCCESS;
}
VkvgPattern vkvg_pattern_create_linear(float x0, float y0, float x1, float y1) {
    VkvgPattern p = pattern_new(VKVG_PATTERN_TYPE_LINEAR);
    if (!vkvg_pattern_status(p))
        vkvg_pattern_edit_linear(p, x0, y0, x1, y1);
    return p;
}
vkvg_status_t vkvg_pattern_get_linear_points(VkvgPattern pat, float *x0, float *y0, float *x1, float *y1) {
    if (vkvg_pattern_status(pat))
        return vkvg_pattern_status(pat);
    if (pat->type != VKVG_PATTERN_TYPE_LINEAR)
        return VKVG_STATUS_PATTERN_TYPE_MISMATCH;
    *x0 = pat->g.cp[0][0];
    *y0 = pat->g.cp[0][1];
    *x1 = pat->g.cp[0][2];
    *y1 = pat->g.cp[0][3];
    return VKVG_STATUS_SUCCESS;
}
/* radius and focus clamping of src/vkvg_pattern.c:95-119 */
vkvg_status_t vkvg_pattern_edit_radial(VkvgPattern pat, float cx0, float cy0, float radius0, float cx1, float cy1,
                                       float radius1) {
    if (vkvg_pattern_status(pat))
        return vkvg_pattern_status(pat);
    if (pat->type != VKVG_PATTERN_TYPE_RADIAL)
        return VKVG_STATUS_PATTERN_TYPE_MISMATCH;
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
    float *a = pat->g.cp[0], *b = pat->g.cp[1];
    a[0] = c0x;
    a[1] = c0y;
    a[2] = radius0;
    a[3] = 0;
    b[0] = c1x;
    b[1] = c1y;
    b[2] = radius1;
    b[3] = 0;
    return VKVG_STATUS_SUCCESS;
}
VkvgPattern vkvg_pattern_create_radial(float cx0, float cy0, float radius0, float cx1, float cy1, float radius1) {
    VkvgPattern p = pattern_new(VKVG_PATTERN_TYPE_RADIAL);
    if (!vkvg_pattern_status(p))
        vkvg_pattern_edit_radial(p, cx0, cy0, radius0, cx1, cy1, radius1);
    return p;
}
vkvg_status_t vkvg_pattern_add_color_stop(VkvgPattern pat, float offset, float r, float g, float b, float a) {
    if (vkvg_pattern_status(pat))
        return vkvg_pattern_status(pat);
    if (pat->type == VKVG_PATTERN_TYPE_SURFACE || pat->type == VKVG_PATTERN_TYPE_SOLID)
        return VKVG_STATUS_PATTERN_TYPE_MISMATCH;
    if (pat->g.count >= 16)
        return VKVG_STATUS_INVALID_INDEX; /* the reference writes past its 16-entry arrays */
    float *c = pat->g.colors[pat->g.count];
    c[0] = a * r;
    c[1] = a * g;
    c[2] = a * b;
    c[3] = a;
    pat->g.stops[pat->g.count] = offset;
    pat->g.count++;
    return VKVG_STATUS_SUCCESS;
}
vkvg_status_t vkvg_pattern_get_color_stop_count(VkvgPattern pat, uint32_t *count) {
    if (vkvg_pattern_status(pat))
        return vkvg_pattern_status(pat);
    if (pat->type == VKVG_PATTERN_TYPE_SURFACE || pat->type == VKVG_PATTERN_TYPE_SOLID)
        return VKVG_STATUS_PATTERN_TYPE_MISMATCH;
    *count = pat->g.count;
    return VKVG_STATUS_SUCCESS;
}
vkvg_status_t vkvg_pattern_get_color_stop_rgba(VkvgPattern pat, uint32_t index, float *offset, float *r, float *g,
                                               float *b, float *a) {
    if (vkvg_pattern_status(pat))
        return vkvg_pattern_status(pat);
    if (pat->type == VKVG_PATTERN_TYPE_SURFACE || pat->type == VKVG_PATTERN_TYPE_SOLID)
        return VKVG_STATUS_PATTERN_TYPE_MISMATCH;
    if (index >= pat->g.count)
        return VKVG_STATUS_INVALID_INDEX;
    *offset = pat->g.stops[index];
    *r = pat->g.colors[index][0];
    *g = pat->g.colors[index][1];
    *b = pat->g.colors[index][2];
    *a = pat->g.colors[index][3];
    return VKVG_STATUS_SUCCESS;
}
void vkvg_pattern_set_extend(VkvgPattern pat, vkvg_extend_t e) {
    if (!vkvg_pattern_status(pat))
        pat->extend = e;
}
void vkvg_pattern_set_filter(VkvgPattern pat, vkvg_filter_t f) {
    if (!vkvg_pattern_status(pat))
        pat->filter = f;
}
vkvg_extend_t vkvg_pattern_get_extend(VkvgPattern pat) { return vkvg_pattern_status(pat) ? 0 : pat->extend; }
vkvg_filter_t vkvg_pattern_get_filter(VkvgPattern pat) { return vkvg_pattern_status(pat) ? 0 : pat->filter; }
vkvg_pattern_type_t vkvg_pattern_get_type(VkvgPattern pat) { return vkvg_pattern_status(pat) ? 0 : pat->type; }
void vkvg_pattern_set_matrix(VkvgPattern pat, const vkvg_matrix_t *m) {
    if (vkvg_pattern_status(pat) || !m)
        return;
    pat->matrix = *m;
    pat->hasMatrix = true;
}
void vkvg_pattern_get_matrix(VkvgPattern pat, vkvg_matrix_t *m) {
    if (vkvg_pattern_status(pat) || !m)
        return;
    if (pat->hasMatrix)
        *m = pat->matrix;
    else
        vkvg_matrix_init_identity(m);
}
void vkvg_pattern_destroy(VkvgPattern pat) {
    if (vkvg_pattern_status(pat))
        return;
    if (--pat->references > 0)
        return;
    if (pat->surf)
        vkvg_surface_destroy(pat->surf);
    free(pat);
}
/* vkvg_pattern_create_for_surface, src/vkvg_pattern.c:28-49: extend NONE, filter FAST */
VkvgPattern vkvg_pattern_create_for_surface(VkvgSurface surf) {
    if (vkvg_surface_status(surf))
        return (VkvgPattern)&s_null;
    VkvgPattern p = pattern_new(VKVG_PATTERN_TYPE_SURFACE);
    if (vkvg_pattern_status(p))
        return p;
    p->surf = vkvg_surface_reference(surf);
    return p;
}

/* ---------------------------------------------------------------------------- context */
static void frame_release_sources(frame_t *f) {
    for (uint32_t i = 0; i < f->src_refs.n; i++)
        vkvg_surface_destroy(f->src_refs.p[i]);
    f->src_refs.n = 0;
}
static void frame_free(frame_t *f) {
    frame_release_sources(f);
    free(f->src_refs.p);
    free(f->items.p);
    free(f->sps.p);
    free(f->draws.p);
    free(f->styles.p);
    free(f->dashes.p);
    free(f->grads.p);
    free(f->srcs.p);
    memset(f, 0, sizeof(*f));
}

static void reset_open_subpath(VkvgContext c) {
    c->cur_first_item = c->f.items.n;
    c->cur_count = 0;
    c->cur_has_cubic = false;
    c->cur_closed = false;
    c->cur_convex = false;
    c->seg_ptr = 0;
    c->seg_first_item = c->f.items.n;
    c->seg_count = 0;
}

VkvgContext vkvg_create(VkvgSurface surf) {
    if (vkvg_surface_status(surf))
        return (VkvgContext)&s_inv_surf;
    if (vkvg_device_status(surf->dev))
        return (VkvgContext)&s_dev_err;
    VkvgContext c = (VkvgContext)calloc(1, sizeof(struct _vkvg_context_t));
    if (!c)
        return (VkvgContext)&s_no_mem;
    c->references = 1;
    c->surf = surf;
    vkvg_surface_reference(surf);
    /* defaults of src/vkvg_context.c:24-69 */
    c->lineWidth = 1.f;
    c->miterLimit = 10.f;
    c->opacity = 1.f;
    c->op = VKVG_OPERATOR_OVER;
    c->fillRule = VKVG_FILL_RULE_NON_ZERO;
    c->lineCap = VKVG_LINE_CAP_BUTT;
    c->lineJoin = VKVG_LINE_JOIN_MITER;
    c->curColor = 0xff000000u;
    vkvg_matrix_init_identity(&c->mat);
    vkvg_matrix_init_identity(&c->mat_inv);
    c->src_rect[3] = 1.f; /* {.a = 1}, src/vkvg_context.c:29 */
    c->cur_grad = -1;
    reset_open_subpath(c);
    return c;
}
vkvg_status_t vkvg_status(VkvgContext ctx) { return !ctx ? VKVG_STATUS_NULL_POINTER : ctx->status; }
vkvg_rec **vkvg_ctx_recording(VkvgContext ctx) { return &ctx->recording; }
VkvgContext vkvg_reference(VkvgContext ctx) {
    if (!vkvg_status(ctx))
        ctx->references++;
    return ctx;
}
uint32_t vkvg_get_reference_count(VkvgContext ctx) { return vkvg_status(ctx) ? 0 : ctx->references; }
VkvgSurface vkvg_get_target(VkvgContext ctx) { return vkvg_status(ctx) ? NULL : ctx->surf; }

/* keeps the current path alive across a frame boundary */
static void frame_reset_keep_path(VkvgContext c) {
    frame_t *f = &c->f;
    uint32_t ib = c->path_item_begin, sb = c->path_sp_begin;
    uint32_t ni = f->items.n - ib, ns = f->sps.n - sb;
    if (ib)
        memmove(f->items.p, f->items.p + ib, (size_t)ni * sizeof(vkh_item));
    if (sb)
        memmove(f->sps.p, f->sps.p + sb, (size_t)ns * sizeof(vkh_subpath));
    for (uint32_t i = 0; i < ns; i++)
        f->sps.p[i].first_item -= ib;
    f->items.n = ni;
    f->sps.n = ns;
    c->cur_first_item -= ib;
    c->seg_first_item = c->seg_first_item >= ib ? c->seg_first_item - ib : 0;
    c->path_item_begin = 0;
    c->path_sp_begin = 0;
    c->path_drawn = false;
    f->draws.n = 0;
    f->styles.n = 0;
    f->dashes.n = 0;
    f->srcs.n = 0;
    frame_release_sources(f);
    /* keep the current gradient source */
    if (c->cur_grad >= 0) {
        vkh_gradient g = f->grads.p[c->cur_grad];
        f->grads.n = 0;
        VEC_PUSH(f->grads, g);
        c->cur_grad = 0;
    } else
        f->grads.n = 0;
}

static void fill_frame(VkvgContext c, vkh_frame *fr, vkh_target *t) {
    frame_t *f = &c->f;
    t->rgba = c->surf->img;
    t->width = c->surf->width;
    t->height = c->surf->height;
    t->sig = c->surf->sig;
    t->stencil = c->stencil_use ? c->surf->stencil : NULL;
    t->stencil_load = c->stencil_use == STENCIL_LOADED;
    t->reserved = 0;
    for (uint32_t i = 0; i < f->draws.n; i++)
        f->draws.p[i].target = 0;
    fr->items = f->items.p;
    fr->n_items = f->items.n;
    fr->subpaths = f->sps.p;
    fr->n_subpaths = f->sps.n;
    fr->draws = f->draws.p;
    fr->n_draws = f->draws.n;
    fr->styles = f->styles.p;
    fr->n_styles = f->styles.n;
    fr->dashes = f->dashes.p;
    fr->n_dashes = f->dashes.n;
    fr->grads = f->grads.p;
    fr->n_grads = f->grads.n;
    fr->sources = f->srcs.p;
    fr->n_sources = f->srcs.n;
    fr->targets = t;
    fr->n_targets = 1;
}

/* upload_only: copy the frame to the device without running the pipeline */
static void submit_frame_ex(VkvgContext c, bool upload_only) {
    frame_t *f = &c->f;
    if (f->draws.n == 0 || c->deferred)
        return;
    vkh_target t;
    vkh_frame fr;
    fill_frame(c, &fr, &t);
    vkh_engine *eng = c->surf->dev->eng;
    int r = upload_only ? vkvg_hip_upload(eng, &fr) : vkvg_hip_render(eng, &fr);
    if (r) {
        fprintf(stderr, "vkvg: render failed: %s\n", vkvg_hip_last_error(eng));
        c->status = VKVG_STATUS_DEVICE_ERROR;
        return;
    }
    if (c->stencil_use)
        c->stencil_use = STENCIL_LOADED;
    frame_reset_keep_path(c);
}
static void submit_frame(VkvgContext c) { submit_frame_ex(c, false); }

vkvg_status_t vkvg_amd_upload(VkvgContext ctx) {
    if (vkvg_status(ctx))
        return vkvg_status(ctx);
    submit_frame_ex(ctx, true);
    return ctx->status;
}
int vkvg_amd_execute(VkvgDevice dev) {
    if (vkvg_device_status(dev))
        return -1;
    return vkvg_hip_execute(dev->eng);
}

void vkvg_flush(VkvgContext ctx) {
    if (vkvg_status(ctx))
        return;
    submit_frame(ctx);
}

void vkvg_destroy(VkvgContext ctx) {
    if (vkvg_status(ctx))
        return;
    if (--ctx->references > 0)
        return;
    submit_frame(ctx);
    while (ctx->saved) {
        struct ctx_save *sv = ctx->saved;
        ctx->saved = sv->next;
        if (sv->pattern)
            vkvg_pattern_destroy(sv->pattern);
        free(sv->dashes);
        free(sv);
    }
    if (ctx->pattern)
        vkvg_pattern_destroy(ctx->pattern);
    vkvg_rec_free(ctx->recording); /* a recording never stopped (src/vkvg_context.c:261-264) */
    free(ctx->dashes);
    frame_free(&ctx->f);
    vkvg_surface_destroy(ctx->surf);
    free(ctx);
}

/* ---- path bookkeeping */
static inline bool path_empty(VkvgContext c) { return c->cur_count == 0; }
static inline bool no_committed_subpath(VkvgContext c) { return c->f.sps.n == c->path_sp_begin; }

/* _add_point, src/vkvg_context_internal.c:221-238 */
static void add_point(VkvgContext c, float x, float y) {
    if (isnan(x) || isnan(y))
        return;
    vkh_item it;
    memset(&it, 0, sizeof(it));
    it.p[0] = x;
    it.p[1] = y;
    VEC_PUSH(c->f.items, it);
    v2 p = {x, y};
    if (c->cur_count == 0)
        c->cur_first = p;
    c->cur_last = p;
    c->cur_count++;
    if (c->seg_ptr > 0)
        c->seg_count++;
}
/* _set_curve_start, :136-153: a new segment starts unless the current one is empty */
static void set_curve_start(VkvgContext c) {
    if (c->seg_ptr > 0) {
        if (c->seg_count > 0) {
            c->seg_ptr++;
            c->seg_first_item = c->f.items.n;
            c->seg_count = 0;
        }
    } else {
        c->seg_ptr = 1;
        c->seg_first_item = c->f.items.n;
        c->seg_count = 0;
    }
}
/* _set_curve_end, :155-162: the current segment is a curve */
static void set_curve_end(VkvgContext c) {
    for (uint32_t i = c->seg_first_item; i < c->f.items.n; i++)
        c->f.items.p[i].flags |= VKH_ITEM_CURVED;
    c->seg_ptr++;
    c->seg_first_item = c->f.items.n;
    c->seg_count = 0;
}
/* _finish_path, :163-197 */
static void finish_path(VkvgContext c) {
    if (c->cur_count == 0)
        return;
    if (!c->cur_has_cubic && c->cur_count < 2) {
        c->f.items.n = c->cur_first_item;
        reset_open_subpath(c);
        return;
    }
    vkh_subpath sp;
    sp.first_item = c->cur_first_item;
    sp.n_items = c->f.items.n - c->cur_first_item;
    sp.flags = c->cur_closed ? VKH_SP_CLOSED : 0u;
    sp.reserved = 0;
    VEC_PUSH(c->f.sps, sp);
    c->simpleConvex = false;
    reset_open_subpath(c);
}
/* _clear_path, :199-206 */
static void clear_path(VkvgContext c) {
    if (!c->path_drawn) {
        c->f.items.n = c->path_item_begin;
        c->f.sps.n = c->path_sp_begin;
    }
    c->path_item_begin = c->f.items.n;
    c->path_sp_begin = c->f.sps.n;
    c->path_drawn = false;
    c->simpleConvex = false;
    reset_open_subpath(c);
}
/* _remove_last_point, :207-219 */
static void remove_last_point(VkvgContext c) {
    vkh_item *last = &c->f.items.p[c->f.items.n - 1];
    if (last->flags & VKH_ITEM_CUBIC)
        last->flags |= VKH_ITEM_NO_END;
    else if ((last->flags & (VKH_ITEM_ARC | VKH_ITEM_EARC)) && (last->flags >> VKH_ITEM_COUNT_SHIFT) > 1)
        last->flags -= 1u << VKH_ITEM_COUNT_SHIFT; /* the arc's last interior point */
    else
        c->f.items.n--;
    c->cur_count--;
}
static float arc_step(VkvgContext c, float radius) {
    float sx, sy;
    vkvg_matrix_get_scale(&c->mat, &sx, &sy);
    float r = radius * fabsf(fmaxf(sx, sy));
    if (r < 30.0f)
        return fminf(PIF / 3.f, PIF / r);
    return fminf(PIF / 3.f, PIF / (r * 0.4f));
}
static void line_to_i(VkvgContext c, float x, float y) {
    if (!path_empty(c)) {
        if (EQUF(c->cur_last.x, x) && EQUF(c->cur_last.y, y))
            return;
    }
    add_point(c, x, y);
    c->simpleConvex = false;
}
/* _curve_to, src/vkvg_context.c:541-566; the cubic itself is flattened on the GPU */
static void curve_to_i(VkvgContext c, float x1, float y1, float x2, float y2, float x3, float y3) {
    if (EQUF(x1, x2) && EQUF(x2, x3) && EQUF(y1, y2) && EQUF(y2, y3)) {
        if (path_empty(c) || (EQUF(c->cur_last.x, x1) && EQUF(c->cur_last.y, y1)))
            return;
    }
    if (!isfinite(x1) || !isfinite(y1) || !isfinite(x2) || !isfinite(y2) || !isfinite(x3) || !isfinite(y3))
        return;
    c->simpleConvex = false;
    set_curve_start(c);
    if (path_empty(c))
        add_point(c, x1, y1);
    v2 cp = c->cur_last;
    float sx = 1, sy = 1;
    vkvg_matrix_get_scale(&c->mat, &sx, &sy);
    float tol = (float)fabs(0.25f / fmaxf(sx, sy));
    vkh_item it;
    memset(&it, 0, sizeof(it));
    it.p[0] = cp.x;
    it.p[1] = cp.y;
    it.p[2] = x1;
    it.p[3] = y1;
    it.p[4] = x2;
    it.p[5] = y2;
    it.p[6] = x3;
    it.p[7] = y3;
    it.flags = VKH_ITEM_CUBIC;
    it.tol = tol;
    VEC_PUSH(c->f.items, it);
    c->cur_count += 3; /* >= 2 interior points + end point */
    c->cur_has_cubic = true;
    if (c->seg_ptr > 0)
        c->seg_count += 3;
    c->cur_last.x = x3;
    c->cur_last.y = y3;
    set_curve_end(c);
}

void vkvg_new_sub_path(VkvgContext ctx) {
    if (vkvg_status(ctx))
        return;
    if (ctx->recording) { /* recorded, not executed */
        vkvg_rec_push(ctx->recording, VKVG_CMD_NEW_SUB_PATH, NULL, 0);
        return;
    }
    finish_path(ctx);
}
void vkvg_new_path(VkvgContext ctx) {
    if (vkvg_status(ctx))
        return;
    if (ctx->recording) { /* recorded, not executed */
        vkvg_rec_push(ctx->recording, VKVG_CMD_NEW_PATH, NULL, 0);
        return;
    }
    clear_path(ctx);
}
void vkvg_close_path(VkvgContext ctx) {
    if (vkvg_status(ctx))
        return;
    if (ctx->recording) { /* recorded, not executed */
        vkvg_rec_push(ctx->recording, VKVG_CMD_CLOSE_PATH, NULL, 0);
        return;
    }
    if (ctx->cur_closed)
        return;
    if (ctx->cur_count < 3)
        return;
    if (EQUF(ctx->cur_last.x, ctx->cur_first.x) && EQUF(ctx->cur_last.y, ctx->cur_first.y)) {
        if (ctx->cur_count < 4)
            return;
        remove_last_point(ctx);
    }
    ctx->cur_closed = true;
    finish_path(ctx);
}
void vkvg_rel_line_to(VkvgContext ctx, float dx, float dy) {
    if (vkvg_status(ctx))
        return;
    if (ctx->recording) { /* recorded, not executed */
        vkvg_rec_floats(ctx->recording, VKVG_CMD_REL_LINE_TO, 2, dx, dy);
        return;
    }
    if (path_empty(ctx))
        add_point(ctx, 0, 0);
    v2 cp = ctx->cur_last;
    line_to_i(ctx, cp.x + dx, cp.y + dy);
}
void vkvg_line_to(VkvgContext ctx, float x, float y) {
    if (vkvg_status(ctx))
        return;
    if (ctx->recording) { /* recorded, not executed */
        vkvg_rec_floats(ctx->recording, VKVG_CMD_LINE_TO, 2, x, y);
        return;
    }
    line_to_i(ctx, x, y);
}
void vkvg_move_to(VkvgContext ctx, float x, float y) {
    if (vkvg_status(ctx))
        return;
    if (ctx->recording) { /* recorded, not executed */
        vkvg_rec_floats(ctx->recording, VKVG_CMD_MOVE_TO, 2, x, y);
        return;
    }
    finish_path(ctx);
    add_point(ctx, x, y);
}
void vkvg_rel_move_to(VkvgContext ctx, float x, float y) {
    if (vkvg_status(ctx))
        return;
    if (ctx->recording) { /* recorded, not executed */
        vkvg_rec_floats(ctx->recording, VKVG_CMD_REL_MOVE_TO, 2, x, y);
        return;
    }
    if (path_empty(ctx))
        add_point(ctx, 0, 0);
    v2 cp = ctx->cur_last;
    finish_path(ctx);
    add_point(ctx, cp.x + x, cp.y + y);
}
bool vkvg_has_current_point(VkvgContext ctx) {
    if (vkvg_status(ctx))
        return false;
    return !path_empty(ctx);
}
void vkvg_get_current_point(VkvgContext ctx, float *x, float *y) {
    if (vkvg_status(ctx) || !x || !y)
        return;
    if (path_empty(ctx)) {
        *x = *y = 0;
        return;
    }
    *x = ctx->cur_last.x;
    *y = ctx->cur_last.y;
}
/* vkvg_path_extents, src/vkvg_context.c:684-696 -> _vkvg_path_extents (untransformed),
 * src/vkvg_context_internal.c:1879-1917: the bounds of every point of the current path, curves as
 * flattened, in user space. The maxima start at FLT_MIN as in the reference, so a path left of or
 * above the origin reports FLT_MIN as its x2 or y2. */
void vkvg_path_extents(VkvgContext ctx, float *const x1, float *const y1, float *const x2, float *const y2) {
    if (vkvg_status(ctx) || !x1 || !y1 || !x2 || !y2)
        return;
    finish_path(ctx);
    if (no_committed_subpath(ctx)) {
        *x1 = *x2 = *y1 = *y2 = 0;
        return;
    }
    float ext[4] = {FLT_MAX, FLT_MAX, FLT_MIN, FLT_MIN};
    for (uint32_t s = ctx->path_sp_begin; s < ctx->f.sps.n; s++) {
        const vkh_subpath *sp = &ctx->f.sps.p[s];
        for (uint32_t k = 0; k < sp->n_items; k++)
            vkvg_hip_item_extents(&ctx->f.items.p[sp->first_item + k], ext);
    }
    *x1 = ext[0];
    *y1 = ext[1];
    *x2 = ext[2];
    *y2 = ext[3];
}
/* vkvg_arc / vkvg_arc_negative, src/vkvg_context.c:394-503: arc points are generated here
 * with the reference's sequential angle accumulation */
/* Coordinates small enough that every point of the arc is finite (the reference skips NaN points,
 * src/vkvg_context_internal.c:224-227; arcs with larger or non-finite operands keep that test by
 * evaluating their points here). */
static bool arc_on_device(float x, float y, float r) { return fabsf(x) < 1e30f && fabsf(y) < 1e30f && fabsf(r) < 1e30f; }
/* n interior points of a circular arc from angle a (step s) as one item; the current point
 * becomes the last of them, at angle a_last */
static void push_arc_item(VkvgContext c, float xc, float yc, float r, float a, float s, uint32_t n, float a_last) {
    vkh_item it;
    memset(&it, 0, sizeof(it));
    it.p[0] = xc;
    it.p[1] = yc;
    it.p[2] = r;
    it.p[3] = a;
    it.p[4] = s;
    it.flags = VKH_ITEM_ARC | (n << VKH_ITEM_COUNT_SHIFT);
    VEC_PUSH(c->f.items, it);
    c->cur_last.x = D_COSF(a_last) * r + xc;
    c->cur_last.y = D_SINF(a_last) * r + yc;
    c->cur_count += n;
    if (c->seg_ptr > 0)
        c->seg_count += n;
}
/* the same for an elliptic arc: theta is a double advanced by the float step s */
static void push_earc_item(VkvgContext c, double th, float s, float rx, float ry, float cph, float sph, v2 ctr,
                           uint32_t n, double th_last) {
    vkh_item it;
    memset(&it, 0, sizeof(it));
    memcpy(it.p, &th, sizeof(th));
    it.p[2] = s;
    it.p[3] = rx;
    it.p[4] = ry;
    it.p[5] = cph;
    it.p[6] = sph;
    it.p[7] = ctr.x;
    it.tol = ctr.y;
    it.flags = VKH_ITEM_EARC | (n << VKH_ITEM_COUNT_SHIFT);
    VEC_PUSH(c->f.items, it);
    const float ex = rx * D_COSF((float)th_last), ey = ry * D_SINF((float)th_last);
    c->cur_last.x = ((cph * ex) + (-sph * ey)) + ctr.x;
    c->cur_last.y = ((sph * ex) + (cph * ey)) + ctr.y;
    c->cur_count += n;
    if (c->seg_ptr > 0)
        c->seg_count += n;
}
static void arc_impl(VkvgContext c, float xc, float yc, float radius, float a1, float a2, bool neg) {
    if (!neg) {
        while (a2 < a1)
            a2 += 2.f * PIF;
        if (a2 - a1 > 2.f * PIF)
            a2 = a1 + 2.f * PIF;
    } else {
        while (a2 > a1)
            a2 -= 2.f * PIF;
        if (a1 - a2 > a1 + 2.f * PIF)
            a2 = a1 - 2.f * PIF;
    }
    v2 v = {D_COSF(a1) * radius + xc, D_SINF(a1) * radius + yc};
    float step = arc_step(c, radius);
    float a = a1;
    if (path_empty(c)) {
        set_curve_start(c);
        add_point(c, v.x, v.y);
        c->simpleConvex = no_committed_subpath(c);
    } else {
        line_to_i(c, v.x, v.y);
        set_curve_start(c);
        c->simpleConvex = false;
    }
    if (!neg)
        a += step;
    else
        a -= step;
    if (EQUF(a2, a1))
        return;
    if (!(step > 0.f) || a == a1) /* an angle the step cannot advance (the reference loops forever) */
        return;
    if (arc_on_device(xc, yc, radius)) {
        /* the interior points are an arc item evaluated by the flattening kernels; here the loop
         * only counts them (float additions, the reference's recurrence) */
        const float s = neg ? -step : step;
        float a_first = a, a_last = a;
        uint32_t n = 0;
        while (neg ? (a > a2) : (a < a2)) {
            a_last = a;
            if (++n == VKH_ITEM_MAX_COUNT) {
                push_arc_item(c, xc, yc, radius, a_first, s, n, a_last);
                n = 0;
                a_first = a + s;
            }
            if (a + s == a) /* the angle stopped advancing (the reference loops forever) */
                break;
            a += s;
        }
        if (n)
            push_arc_item(c, xc, yc, radius, a_first, s, n, a_last);
    } else if (!neg) {
        while (a < a2) {
            add_point(c, D_COSF(a) * radius + xc, D_SINF(a) * radius + yc);
            a += step;
        }
    } else {
        while (a > a2) {
            add_point(c, D_COSF(a) * radius + xc, D_SINF(a) * radius + yc);
            a -= step;
        }
    }
    if (EQUF(neg ? a1 - a2 : a2 - a1, PIF * 2.f)) {
        set_curve_end(c);
        vkvg_close_path(c);
        return;
    }
    a = a2;
    add_point(c, D_COSF(a) * radius + xc, D_SINF(a) * radius + yc);
    set_curve_end(c);
}
void vkvg_arc(VkvgContext ctx, float xc, float yc, float radius, float a1, float a2) {
    if (vkvg_status(ctx))
        return;
    if (ctx->recording) { /* recorded, not executed */
        vkvg_rec_floats(ctx->recording, VKVG_CMD_ARC, 5, xc, yc, radius, a1, a2);
        return;
    }
    arc_impl(ctx, xc, yc, radius, a1, a2, false);
}
void vkvg_arc_negative(VkvgContext ctx, float xc, float yc, float radius, float a1, float a2) {
    if (vkvg_status(ctx))
        return;
    if (ctx->recording) { /* recorded, not executed */
        vkvg_rec_floats(ctx->recording, VKVG_CMD_ARC_NEG, 5, xc, yc, radius, a1, a2);
        return;
    }
    arc_impl(ctx, xc, yc, radius, a1, a2, true);
}
void vkvg_curve_to(VkvgContext ctx, float x1, float y1, float x2, float y2, float x3, float y3) {
    if (vkvg_status(ctx))
        return;
    if (ctx->recording) { /* recorded, not executed */
        vkvg_rec_floats(ctx->recording, VKVG_CMD_CURVE_TO, 6, x1, y1, x2, y2, x3, y3);
        return;
    }
    curve_to_i(ctx, x1, y1, x2, y2, x3, y3);
}
void vkvg_rel_curve_to(VkvgContext ctx, float x1, float y1, float x2, float y2, float x3, float y3) {
    if (vkvg_status(ctx))
        return;
    if (ctx->recording) { /* recorded, not executed */
        vkvg_rec_floats(ctx->recording, VKVG_CMD_REL_CURVE_TO, 6, x1, y1, x2, y2, x3, y3);
        return;
    }
    if (path_empty(ctx)) {
        ctx->status = VKVG_STATUS_NO_CURRENT_POINT;
        return;
    }
    v2 cp = ctx->cur_last;
    curve_to_i(ctx, cp.x + x1, cp.y + y1, cp.x + x2, cp.y + y2, cp.x + x3, cp.y + y3);
}
static void quadratic_to_i(VkvgContext c, float x1, float y1, float x2, float y2) {
    float x0, y0;
    if (path_empty(c)) {
        x0 = x1;
        y0 = y1;
    } else {
        x0 = c->cur_last.x;
        y0 = c->cur_last.y;
    }
    const double f = 2.0 / 3.0;
    curve_to_i(c, (float)(x0 + (x1 - x0) * f), (float)(y0 + (y1 - y0) * f), (float)(x2 + (x1 - x2) * f),
               (float)(y2 + (y1 - y2) * f), x2, y2);
}
void vkvg_quadratic_to(VkvgContext ctx, float x1, float y1, float x2, float y2) {
    if (vkvg_status(ctx))
        return;
    if (ctx->recording) { /* recorded, not executed */
        vkvg_rec_floats(ctx->recording, VKVG_CMD_QUADRATIC_TO, 4, x1, y1, x2, y2);
        return;
    }
    quadratic_to_i(ctx, x1, y1, x2, y2);
}
void vkvg_rel_quadratic_to(VkvgContext ctx, float x1, float y1, float x2, float y2) {
    if (vkvg_status(ctx))
        return;
    if (ctx->recording) { /* recorded, not executed */
        vkvg_rec_floats(ctx->recording, VKVG_CMD_REL_QUADRATIC_TO, 4, x1, y1, x2, y2);
        return;
    }
    v2 cp = ctx->cur_last;
    quadratic_to_i(ctx, cp.x + x1, cp.y + y1, cp.x + x2, cp.y + y2);
}
vkvg_status_t vkvg_rectangle(VkvgContext ctx, float x, float y, float w, float h) {
    if (vkvg_status(ctx))
        return ctx ? ctx->status : VKVG_STATUS_NULL_POINTER;
    if (ctx->recording) { /* recorded, not executed */
        vkvg_rec_floats(ctx->recording, VKVG_CMD_RECTANGLE, 4, x, y, w, h);
        return VKVG_STATUS_SUCCESS;
    }
    finish_path(ctx);
    if (w <= 0 || h <= 0)
        return VKVG_STATUS_INVALID_RECT;
    add_point(ctx, x, y);
    add_point(ctx, x + w, y);
    add_point(ctx, x + w, y + h);
    add_point(ctx, x, y + h);
    ctx->cur_closed = true;
    ctx->cur_convex = true;
    finish_path(ctx);
    return VKVG_STATUS_SUCCESS;
}
vkvg_status_t vkvg_rounded_rectangle(VkvgContext ctx, float x, float y, float w, float h, float radius) {
    if (vkvg_status(ctx))
        return ctx ? ctx->status : VKVG_STATUS_NULL_POINTER;
    finish_path(ctx);
    if (w <= 0 || h <= 0)
        return VKVG_STATUS_INVALID_RECT;
    if ((radius > w / 2.0f) || (radius > h / 2.0f))
        radius = (float)fmin(w / 2.0f, h / 2.0f);
    vkvg_move_to(ctx, x, y + radius);
    vkvg_arc(ctx, x + radius, y + radius, radius, PIF, -PIF_2);
    vkvg_line_to(ctx, x + w - radius, y);
    vkvg_arc(ctx, x + w - radius, y + radius, radius, -PIF_2, 0);
    vkvg_line_to(ctx, x + w, y + h - radius);
    vkvg_arc(ctx, x + w - radius, y + h - radius, radius, 0, PIF_2);
    vkvg_line_to(ctx, x + radius, y + h);
    vkvg_arc(ctx, x + radius, y + h - radius, radius, PIF_2, PIF);
    vkvg_line_to(ctx, x, y + radius);
    vkvg_close_path(ctx);
    return VKVG_STATUS_SUCCESS;
}
/* vkvg_ellipse, src/vkvg_context.c:1605-1639 */
/* vkvg_rounded_rectangle2, src/vkvg_context.c:665-684: corners as elliptic arcs */
void vkvg_rounded_rectangle2(VkvgContext ctx, float x, float y, float w, float h, float rx, float ry) {
    if (vkvg_status(ctx))
        return;
    vkvg_move_to(ctx, x + rx, y);
    vkvg_line_to(ctx, x + w - rx, y);
    vkvg_elliptic_arc_to(ctx, x + w, y + ry, false, true, rx, ry, 0);
    vkvg_line_to(ctx, x + w, y + h - ry);
    vkvg_elliptic_arc_to(ctx, x + w - rx, y + h, false, true, rx, ry, 0);
    vkvg_line_to(ctx, x + rx, y + h);
    vkvg_elliptic_arc_to(ctx, x, y + h - ry, false, true, rx, ry, 0);
    vkvg_line_to(ctx, x, y + ry);
    vkvg_elliptic_arc_to(ctx, x + rx, y, false, true, rx, ry, 0);
    vkvg_close_path(ctx);
}
void vkvg_ellipse(VkvgContext ctx, float radiusX, float radiusY, float x, float y, float rot) {
    if (vkvg_status(ctx))
        return;
    if (ctx->recording) { /* recorded, not executed */
        vkvg_rec_floats(ctx->recording, VKVG_CMD_ELLIPSE, 5, radiusX, radiusY, x, y, rot);
        return;
    }
    float wtt = radiusX * 4 / 3;
    float dx1 = D_SINF(rot) * radiusY, dy1 = D_COSF(rot) * radiusY;
    float dx2 = D_COSF(rot) * wtt, dy2 = D_SINF(rot) * wtt;
    float tcx = x - dx1, tcy = y + dy1;
    float trx = tcx + dx2, try_ = tcy + dy2;
    float tlx = tcx - dx2, tly = tcy - dy2;
    float bcx = x + dx1, bcy = y - dy1;
    float brx = bcx + dx2, bry = bcy + dy2;
    float blx = bcx - dx2, bly = bcy - dy2;
    finish_path(ctx);
    add_point(ctx, bcx, bcy);
    curve_to_i(ctx, brx, bry, trx, try_, tcx, tcy);
    curve_to_i(ctx, tlx, tly, blx, bly, bcx, bcy);
    ctx->cur_closed = true;
    finish_path(ctx);
}
/* _elliptic_arc, src/vkvg_context_internal.c:1473-1580 (points generated here) */
static void elliptic_arc(VkvgContext c, float x1, float y1, float x2, float y2, bool largeArc, bool ccw, float _rx,
                         float _ry, float phi) {
    if (_rx == 0 || _ry == 0) {
        if (path_empty(c))
            vkvg_move_to(c, x1, y1);
        vkvg_line_to(c, x2, y2);
        return;
    }
    float rx = fabsf(_rx), ry = fabsf(_ry);
    float cph = D_COSF(phi), sph = D_SINF(phi);
    v2 p = {(x1 - x2) / 2, (y1 - y2) / 2};
    v2 p1 = {(cph * p.x) + (sph * p.y), (-sph * p.x) + (cph * p.y)};
    double lambda = D_POWF(p1.x, 2) / D_POWF(rx, 2) + D_POWF(p1.y, 2) / D_POWF(ry, 2);
    if (lambda > 1) {
        lambda = sqrtf((float)lambda);
        rx *= lambda;
        ry *= lambda;
    }
    p.x = rx * p1.y / ry;
    p.y = -ry * p1.x / rx;
    float num = D_POWF(rx, 2) * D_POWF(ry, 2) - D_POWF(rx, 2) * D_POWF(p1.y, 2) - D_POWF(ry, 2) * D_POWF(p1.x, 2);
    float den = D_POWF(rx, 2) * D_POWF(p1.y, 2) + D_POWF(ry, 2) * D_POWF(p1.x, 2);
    float k = sqrtf(fabsf(num / den));
    v2 cp = {p.x * k, p.y * k};
    if (largeArc == ccw) {
        cp.x = -cp.x;
        cp.y = -cp.y;
    }
    p.x = (x1 + x2) / 2;
    p.y = (y1 + y2) / 2;
    v2 ctr = {((cph * cp.x) + (-sph * cp.y)) + p.x, ((sph * cp.x) + (cph * cp.y)) + p.y};
    v2 u = {1.f, 0}, v = {(p1.x - cp.x) / rx, (p1.y - cp.y) / ry};
    float lv = sqrtf(v.x * v.x + v.y * v.y), lu = sqrtf(u.x * u.x + u.y * u.y);
    double sa = D_ACOSF(((u.x * v.x) + (u.y * v.y)) / (fabsf(lv) * fabsf(lu)));
    if (isnan(sa))
        sa = PIF;
    if (u.x * v.y - u.y * v.x < 0)
        sa = -sa;
    u = v;
    v.x = (-p1.x - cp.x) / rx;
    v.y = (-p1.y - cp.y) / ry;
    lv = sqrtf(v.x * v.x + v.y * v.y);
    lu = sqrtf(u.x * u.x + u.y * u.y);
    double dt = D_ACOSF(((u.x * v.x) + (u.y * v.y)) / (fabsf(lv) * fabsf(lu)));
    if (isnan(dt))
        dt = PIF;
    if (u.x * v.y - u.y * v.x < 0)
        dt = -dt;
    if (ccw) {
        if (dt < 0)
            dt += PIF * 2.0;
    } else if (dt > 0)
        dt -= PIF * 2.0;
    double theta = sa, ea = sa + dt;
    float step = fmaxf(0.001f, fminf(PIF, arc_step(c, fminf(rx, ry)) * 0.1f));
    float ex = rx * D_COSF((float)theta), ey = ry * D_SINF((float)theta);
    v2 xy = {((cph * ex) + (-sph * ey)) + ctr.x, ((sph * ex) + (cph * ey)) + ctr.y};
    if (path_empty(c)) {
        set_curve_start(c);
        add_point(c, xy.x, xy.y);
        c->simpleConvex = no_committed_subpath(c);
    } else {
        line_to_i(c, xy.x, xy.y);
        set_curve_start(c);
        c->simpleConvex = false;
    }
    set_curve_start(c);
    if (arc_on_device(ctr.x, ctr.y, rx) && fabsf(ry) < 1e30f && fabsf(cph) <= 1.f && fabsf(sph) <= 1.f && step > 0.f &&
        isfinite(sa) && isfinite(ea)) {
        /* the interior points as an item evaluated by the flattening kernels, counted here with the
         * reference's double additions */
        const float s = sa < ea ? step : -step;
        theta += s;
        double th_first = theta, th_last = theta;
        uint32_t n = 0;
        while (sa < ea ? (theta < ea) : (theta > ea)) {
            th_last = theta;
            if (++n == VKH_ITEM_MAX_COUNT) {
                push_earc_item(c, th_first, s, rx, ry, cph, sph, ctr, n, th_last);
                n = 0;
                th_first = theta + s;
            }
            if (theta + s == theta) /* the angle stopped advancing (the reference loops forever) */
                break;
            theta += s;
        }
        if (n)
            push_earc_item(c, th_first, s, rx, ry, cph, sph, ctr, n, th_last);
    } else if (sa < ea) {
        theta += step;
        while (theta < ea) {
            ex = rx * D_COSF((float)theta);
            ey = ry * D_SINF((float)theta);
            add_point(c, ((cph * ex) + (-sph * ey)) + ctr.x, ((sph * ex) + (cph * ey)) + ctr.y);
            theta += step;
        }
    } else {
        theta -= step;
        while (theta > ea) {
            ex = rx * D_COSF((float)theta);
            ey = ry * D_SINF((float)theta);
            add_point(c, ((cph * ex) + (-sph * ey)) + ctr.x, ((sph * ex) + (cph * ey)) + ctr.y);
            theta -= step;
        }
    }
    ex = rx * D_COSF((float)ea);
    ey = ry * D_SINF((float)ea);
    add_point(c, ((cph * ex) + (-sph * ey)) + ctr.x, ((sph * ex) + (cph * ey)) + ctr.y);
    set_curve_end(c);
}
void vkvg_elliptic_arc_to(VkvgContext ctx, float x2, float y2, bool large, bool sweep, float rx, float ry,
                          float phi) {
    if (vkvg_status(ctx))
        return;
    if (ctx->recording) { /* recorded, not executed */
        vkvg_rec_elliptic(ctx->recording, VKVG_CMD_ELLIPTICAL_ARC_TO, x2, y2, rx, ry, phi, large, sweep);
        return;
    }
    float x1 = 0, y1 = 0;
    vkvg_get_current_point(ctx, &x1, &y1);
    elliptic_arc(ctx, x1, y1, x2, y2, large, sweep, rx, ry, phi);
}
void vkvg_rel_elliptic_arc_to(VkvgContext ctx, float x2, float y2, bool large, bool sweep, float rx, float ry,
                              float phi) {
    if (vkvg_status(ctx))
        return;
    if (ctx->recording) { /* recorded, not executed */
        vkvg_rec_elliptic(ctx->recording, VKVG_CMD_REL_ELLIPTICAL_ARC_TO, x2, y2, rx, ry, phi, large, sweep);
        return;
    }
    float x1 = 0, y1 = 0;
    vkvg_get_current_point(ctx, &x1, &y1);
    elliptic_arc(ctx, x1, y1, x2 + x1, y2 + y1, large, sweep, rx, ry, phi);
}

/* ---- draws */
static vkh_draw *new_draw(VkvgContext c, uint32_t kind) {
    vkh_draw d;
    memset(&d, 0, sizeof(d));
    d.kind = kind;
    d.op = (uint32_t)c->op;
    d.color = c->curColor;
    d.grad = c->cur_grad;
    d.opacity = c->opacity;
    d.mat[0] = c->mat.xx;
    d.mat[1] = c->mat.yx;
    d.mat[2] = c->mat.xy;
    d.mat[3] = c->mat.yy;
    d.mat[4] = c->mat.x0;
    d.mat[5] = c->mat.y0;
    d.first_subpath = c->path_sp_begin;
    d.n_subpaths = c->f.sps.n - c->path_sp_begin;
    d.style = -1;
    d.target = 0;
    d.key_index = c->surf->draw_counter++;
    d.source = -1;
    if (c->pattern && c->pattern->type == VKVG_PATTERN_TYPE_SURFACE) {
        /* the sampler state of the pattern and the push constants current at this draw */
        vkh_surface_source src;
        memset(&src, 0, sizeof(src));
        const VkvgSurface ss = c->pattern->surf;
        src.rgba = ss->img;
        src.width = ss->width;
        src.height = ss->height;
        src.extend = (uint32_t)c->pattern->extend;
        src.filter = (uint32_t)c->pattern->filter;
        memcpy(src.rect, c->src_rect, sizeof(src.rect));
        const vkvg_matrix_t *m = &c->mat_inv;
        src.mat_inv[0] = m->xx;
        src.mat_inv[1] = m->yx;
        src.mat_inv[2] = m->xy;
        src.mat_inv[3] = m->yy;
        src.mat_inv[4] = m->x0;
        src.mat_inv[5] = m->y0;
        if (!c->f.srcs.n || memcmp(&c->f.srcs.p[c->f.srcs.n - 1], &src, sizeof(src))) {
            VEC_PUSH(c->f.srcs, src);
            /* the frame reads the image when it renders: keep the surface alive until then */
            VEC_PUSH(c->f.src_refs, vkvg_surface_reference(ss));
        }
        d.source = (int32_t)(c->f.srcs.n - 1);
        d.grad = -1;
    }
    VEC_PUSH(c->f.draws, d);
    return &c->f.draws.p[c->f.draws.n - 1];
}
/* _fill_preserve, src/vkvg_context.c:796-821 */
static void fill_preserve_i(VkvgContext c) {
    finish_path(c);
    if (no_committed_subpath(c))
        return;
    new_draw(c, c->fillRule == VKVG_FILL_RULE_EVEN_ODD ? VKH_DRAW_FILL_EO : VKH_DRAW_FILL_NZ);
    c->path_drawn = true;
}
/* _stroke_preserve, src/vkvg_context.c:822-948 */
static void stroke_preserve_i(VkvgContext c) {
    finish_path(c);
    if (no_committed_subpath(c))
        return;
    vkh_stroke_style st;
    memset(&st, 0, sizeof(st));
    st.hw = c->lineWidth * 0.5f;
    st.lhMax = c->miterLimit * c->lineWidth;
    /* raised to VKH_MIN_ARC_STEP where the reference's arc loops would not end (huge or
     * negative widths); every width that renders in finite time keeps the reference's step */
    st.arcStep = fmaxf(arc_step(c, st.hw), VKH_MIN_ARC_STEP);
    st.join = (uint32_t)c->lineJoin;
    st.cap = (uint32_t)c->lineCap;
    st.dash_first = c->f.dashes.n;
    st.dash_count = c->dashCount;
    st.dash_offset = c->dashOffset;
    if (c->dashCount > 0) {
        float tot = 0;
        bool neg = false;
        for (uint32_t i = 0; i < c->dashCount; i++) {
            tot += c->dashes[i];
            neg |= !(c->dashes[i] >= 0.f);
        }
        /* the reference rejects a zero total (src/vkvg_context.c:860-863); negative or NaN
         * entries and totals below 1/1024 px would make its dash walk run (almost) forever */
        if (tot == 0 || neg || !(tot >= 1.0f / 1024.0f)) {
            c->status = VKVG_STATUS_INVALID_DASH;
            return;
        }
        for (uint32_t i = 0; i < c->dashCount; i++)
            VEC_PUSH(c->f.dashes, c->dashes[i]);
    }
    VEC_PUSH(c->f.styles, st);
    vkh_draw *d = new_draw(c, VKH_DRAW_STROKE);
    d->style = (int32_t)(c->f.styles.n - 1);
    c->path_drawn = true;
}
void vkvg_fill_preserve(VkvgContext ctx) {
    if (vkvg_status(ctx))
        return;
    if (ctx->recording) { /* recorded, not executed */
        vkvg_rec_push(ctx->recording, VKVG_CMD_FILL_PRESERVE, NULL, 0);
        return;
    }
    fill_preserve_i(ctx);
}
void vkvg_fill(VkvgContext ctx) {
    if (vkvg_status(ctx))
        return;
    if (ctx->recording) { /* recorded, not executed */
        vkvg_rec_push(ctx->recording, VKVG_CMD_FILL, NULL, 0);
        return;
    }
    fill_preserve_i(ctx);
    clear_path(ctx);
}
void vkvg_stroke_preserve(VkvgContext ctx) {
    if (vkvg_status(ctx))
        return;
    if (ctx->recording) { /* recorded, not executed */
        vkvg_rec_push(ctx->recording, VKVG_CMD_STROKE_PRESERVE, NULL, 0);
        return;
    }
    stroke_preserve_i(ctx);
}
void vkvg_stroke(VkvgContext ctx) {
    if (vkvg_status(ctx))
        return;
    if (ctx->recording) { /* recorded, not executed */
        vkvg_rec_push(ctx->recording, VKVG_CMD_STROKE, NULL, 0);
        return;
    }
    stroke_preserve_i(ctx);
    clear_path(ctx);
}
void vkvg_paint(VkvgContext ctx) {
    if (vkvg_status(ctx))
        return;
    if (ctx->recording) { /* recorded, not executed */
        vkvg_rec_push(ctx->recording, VKVG_CMD_PAINT, NULL, 0);
        return;
    }
    finish_path(ctx);
    if (!no_committed_subpath(ctx)) {
        vkvg_fill(ctx);
        return;
    }
    new_draw(ctx, VKH_DRAW_PAINT);
}
static void set_solid(VkvgContext ctx);
static int prev_clip_state(VkvgContext c) { return c->saved ? c->saved->clip_state : CLIP_STATE_CLEAR; }

/* vkvg_clear, src/vkvg_context.c:734-753: colour and stencil cleared */
void vkvg_clear(VkvgContext ctx) {
    if (vkvg_status(ctx))
        return;
    if (ctx->recording) { /* recorded, not executed */
        vkvg_rec_push(ctx->recording, VKVG_CMD_CLEAR, NULL, 0);
        return;
    }
    ctx->clip_state = prev_clip_state(ctx) == CLIP_STATE_CLEAR ? CLIP_STATE_NONE : CLIP_STATE_CLEAR;
    new_draw(ctx, VKH_DRAW_CLEAR);
}

/* ---- clip and save/restore */
/* a draw that only rewrites the stencil plane: no geometry and no signature key; nothing to do
 * while this context has not clipped (its stencil is still all zero) */
static void stencil_draw(VkvgContext c, uint32_t kind, uint32_t param) {
    if (!c->stencil_use)
        return;
    vkh_draw d;
    memset(&d, 0, sizeof(d));
    d.kind = kind;
    d.op = VKVG_OPERATOR_OVER;
    d.grad = -1;
    d.opacity = 1.f;
    d.mat[0] = d.mat[3] = 1.f;
    d.style = (int32_t)param;
    d.source = -1;
    VEC_PUSH(c->f.draws, d);
}
/* _clip_preserve, src/vkvg_context.c:754-795 */
static void clip_preserve_i(VkvgContext c) {
    finish_path(c);
    if (no_committed_subpath(c))
        return;
    if (!c->surf->stencil) {
        void *p = NULL;
        if (vkvg_hip_image_alloc(c->surf->dev->eng, c->surf->width, c->surf->height, &p)) {
            c->status = VKVG_STATUS_NO_MEMORY;
            return;
        }
        c->surf->stencil = (uint32_t *)p;
    }
    if (!c->stencil_use)
        c->stencil_use = STENCIL_FIRST_FRAME;
    vkh_draw *d = new_draw(c, c->fillRule == VKVG_FILL_RULE_EVEN_ODD ? VKH_DRAW_CLIP_EO : VKH_DRAW_CLIP_NZ);
    d->grad = -1;
    d->source = -1;
    c->path_drawn = true;
    c->clip_state = CLIP_STATE_CLIP;
}
void vkvg_clip_preserve(VkvgContext ctx) {
    if (vkvg_status(ctx))
        return;
    if (ctx->recording) { /* recorded, not executed */
        vkvg_rec_push(ctx->recording, VKVG_CMD_CLIP_PRESERVE, NULL, 0);
        return;
    }
    clip_preserve_i(ctx);
}
void vkvg_clip(VkvgContext ctx) {
    if (vkvg_status(ctx))
        return;
    if (ctx->recording) { /* recorded, not executed */
        vkvg_rec_push(ctx->recording, VKVG_CMD_CLIP, NULL, 0);
        return;
    }
    clip_preserve_i(ctx);
    clear_path(ctx);
}
/* vkvg_reset_clip, src/vkvg_context.c:719-733 */
void vkvg_reset_clip(VkvgContext ctx) {
    if (vkvg_status(ctx))
        return;
    if (ctx->recording) { /* recorded, not executed */
        vkvg_rec_push(ctx->recording, VKVG_CMD_RESET_CLIP, NULL, 0);
        return;
    }
    if (ctx->clip_state == CLIP_STATE_CLEAR)
        return;
    ctx->clip_state = prev_clip_state(ctx) == CLIP_STATE_CLEAR ? CLIP_STATE_NONE : CLIP_STATE_CLEAR;
    stencil_draw(ctx, VKH_DRAW_CLIP_RESET, 0);
}
/* vkvg_save, src/vkvg_context.c:1251-1375 */
void vkvg_save(VkvgContext ctx) {
    if (vkvg_status(ctx))
        return;
    if (ctx->recording) { /* recorded, not executed */
        vkvg_rec_push(ctx->recording, VKVG_CMD_SAVE, NULL, 0);
        return;
    }
    struct ctx_save *sv = (struct ctx_save *)calloc(1, sizeof(struct ctx_save));
    if (!sv) {
        ctx->status = VKVG_STATUS_NO_MEMORY;
        return;
    }
    if (ctx->clip_state == CLIP_STATE_CLIP) {
        sv->clip_state = CLIP_STATE_SAVED;
        if (ctx->sav_bit > 0 && ctx->sav_bit % 6 == 0) { /* the reference copies the stencil to a new image */
            const uint32_t idx = ctx->sav_bit / 6 - 1;
            if (idx >= VKH_STENCIL_SPILLS) { /* more nested clip saves than the plane holds */
                free(sv);
                ctx->status = VKVG_STATUS_NO_MEMORY;
                return;
            }
            stencil_draw(ctx, VKH_DRAW_STENCIL_SPILL, idx);
        }
        stencil_draw(ctx, VKH_DRAW_CLIP_SAVE, ctx->sav_bit % 6);
        ctx->sav_bit++;
    } else if (ctx->clip_state == CLIP_STATE_NONE) {
        sv->clip_state = prev_clip_state(ctx) & 0x03;
    } else {
        sv->clip_state = CLIP_STATE_CLEAR;
    }
    sv->dashOffset = ctx->dashOffset;
    sv->dashCount = ctx->dashCount;
    if (ctx->dashCount) {
        sv->dashes = (float *)malloc(sizeof(float) * ctx->dashCount);
        if (!sv->dashes) {
            free(sv);
            ctx->status = VKVG_STATUS_NO_MEMORY;
            return;
        }
        memcpy(sv->dashes, ctx->dashes, sizeof(float) * ctx->dashCount);
    }
    sv->lineWidth = ctx->lineWidth;
    sv->miterLimit = ctx->miterLimit;
    sv->op = ctx->op;
    sv->lineCap = ctx->lineCap;
    sv->fillRule = ctx->fillRule;
    sv->mat = ctx->mat; /* the push constants: CTM, inverse, source rectangle, opacity */
    sv->mat_inv = ctx->mat_inv;
    memcpy(sv->src_rect, ctx->src_rect, sizeof(sv->src_rect));
    sv->opacity = ctx->opacity;
    if (ctx->pattern)
        sv->pattern = vkvg_pattern_reference(ctx->pattern);
    else
        sv->curColor = ctx->curColor;
    sv->next = ctx->saved;
    ctx->saved = sv;
}
/* vkvg_restore, src/vkvg_context.c:1376-1512 */
void vkvg_restore(VkvgContext ctx) {
    if (vkvg_status(ctx))
        return;
    if (ctx->recording) { /* recorded, not executed */
        vkvg_rec_push(ctx->recording, VKVG_CMD_RESTORE, NULL, 0);
        return;
    }
    struct ctx_save *sv = ctx->saved;
    if (!sv) {
        ctx->status = VKVG_STATUS_INVALID_RESTORE;
        return;
    }
    ctx->saved = sv->next;
    ctx->mat = sv->mat;
    ctx->mat_inv = sv->mat_inv;
    memcpy(ctx->src_rect, sv->src_rect, sizeof(ctx->src_rect));
    ctx->opacity = sv->opacity;
    if (ctx->clip_state != CLIP_STATE_NONE) {
        if (ctx->clip_state == CLIP_STATE_CLIP && sv->clip_state == CLIP_STATE_CLEAR)
            stencil_draw(ctx, VKH_DRAW_CLIP_RESET, 0);
        else if (ctx->sav_bit > 0)
            stencil_draw(ctx, VKH_DRAW_CLIP_RESTORE, (ctx->sav_bit - 1) % 6);
        /* with no level saved the reference's pass copies CLIP onto itself */
    }
    if (sv->clip_state == CLIP_STATE_SAVED) {
        ctx->sav_bit--;
        if (ctx->sav_bit > 0 && ctx->sav_bit % 6 == 0) /* the saved stencil image comes back */
            stencil_draw(ctx, VKH_DRAW_STENCIL_UNSPILL, ctx->sav_bit / 6 - 1);
    }
    ctx->clip_state = CLIP_STATE_NONE;
    ctx->dashOffset = sv->dashOffset;
    free(ctx->dashes);
    ctx->dashes = sv->dashes; /* ownership moves back */
    ctx->dashCount = sv->dashCount;
    sv->dashes = NULL;
    ctx->lineWidth = sv->lineWidth;
    ctx->miterLimit = sv->miterLimit;
    ctx->op = sv->op;
    ctx->lineCap = sv->lineCap;
    ctx->lineJoin = VKVG_LINE_JOIN_MITER; /* read from a save field vkvg_save never writes, :1492 */
    ctx->fillRule = sv->fillRule;
    if (sv->pattern) {
        if (sv->pattern != ctx->pattern) /* _update_cur_pattern under the restored CTM */
            vkvg_set_source(ctx, sv->pattern);
        vkvg_pattern_destroy(sv->pattern);
    } else {
        ctx->curColor = sv->curColor;
        set_solid(ctx);
    }
    free(sv);
}

/* ---- state */
void vkvg_set_opacity(VkvgContext ctx, float opacity) {
    if (vkvg_status(ctx))
        return;
    if (ctx->recording) { /* recorded, not executed */
        vkvg_rec_floats(ctx->recording, VKVG_CMD_SET_OPACITY, 1, opacity);
        return;
    }
    ctx->opacity = opacity;
}
float vkvg_get_opacity(VkvgContext ctx) { return vkvg_status(ctx) ? 0 : ctx->opacity; }
static void set_solid(VkvgContext ctx) {
    if (ctx->pattern) {
        vkvg_pattern_destroy(ctx->pattern);
        ctx->pattern = NULL;
    }
    ctx->cur_grad = -1;
}
/* CreateRgbaf, src/vkvg_context_internal.h:60-66 (premultiplied, truncating) */
static uint32_t rgbaf(float r, float g, float b, float a) {
    return ((((uint32_t)(a * 255.0f) & 0xFF) << 24) | (((uint32_t)(b * a * 255.0f) & 0xFF) << 16) |
            (((uint32_t)(g * a * 255.0f) & 0xFF) << 8) | ((uint32_t)(r * a * 255.0f) & 0xFF));
}
void vkvg_set_source_color(VkvgContext ctx, uint32_t c) {
    if (vkvg_status(ctx))
        return;
    if (ctx->recording) { /* recorded, not executed */
        vkvg_rec_u32(ctx->recording, VKVG_CMD_SET_SOURCE_COLOR, c);
        return;
    }
    ctx->curColor = c;
    set_solid(ctx);
}
void vkvg_set_source_rgba(VkvgContext ctx, float r, float g, float b, float a) {
    if (vkvg_status(ctx))
        return;
    if (ctx->recording) { /* recorded, not executed */
        vkvg_rec_floats(ctx->recording, VKVG_CMD_SET_SOURCE_RGBA, 4, r, g, b, a);
        return;
    }
    ctx->curColor = rgbaf(r, g, b, a);
    set_solid(ctx);
}
void vkvg_set_source_rgb(VkvgContext ctx, float r, float g, float b) {
    if (vkvg_status(ctx))
        return;
    if (ctx->recording) { /* recorded, not executed */
        vkvg_rec_floats(ctx->recording, VKVG_CMD_SET_SOURCE_RGB, 3, r, g, b);
        return;
    }
    ctx->curColor = rgbaf(r, g, b, 1);
    set_solid(ctx);
}
/* _update_cur_pattern for gradients, src/vkvg_context_internal.c:774-826 */
void vkvg_set_source(VkvgContext ctx, VkvgPattern pat) {
    if (vkvg_status(ctx) || vkvg_pattern_status(pat))
        return;
    if (ctx->recording) { /* recorded, not executed */
        vkvg_rec_pattern(ctx->recording, pat);
        return;
    }
    if (pat->type != VKVG_PATTERN_TYPE_LINEAR && pat->type != VKVG_PATTERN_TYPE_RADIAL &&
        pat->type != VKVG_PATTERN_TYPE_SURFACE) {
        ctx->status = VKVG_STATUS_PATTERN_TYPE_MISMATCH;
        return;
    }
    vkvg_pattern_reference(pat);
    if (ctx->pattern)
        vkvg_pattern_destroy(ctx->pattern);
    ctx->pattern = pat;
    if (pat->type == VKVG_PATTERN_TYPE_SURFACE) {
        /* src/vkvg_context_internal.c:705-772: the source size, and the pattern matrix (not
         * inverted) multiplied into the inverse matrix */
        ctx->cur_grad = -1;
        ctx->src_rect[2] = (float)pat->surf->width;
        ctx->src_rect[3] = (float)pat->surf->height;
        if (pat->hasMatrix)
            vkvg_matrix_multiply(&ctx->mat_inv, &ctx->mat_inv, &pat->matrix);
        return;
    }
    /* the target's size stored in the unused source field (:782-786) */
    ctx->src_rect[0] = (float)ctx->surf->width;
    ctx->src_rect[1] = (float)ctx->surf->height;
    ctx->src_rect[2] = ctx->src_rect[3] = 0.f;
    if (pat->g.count < 2) {
        ctx->status = VKVG_STATUS_PATTERN_INVALID_GRADIENT;
        return;
    }
    vkh_gradient g;
    memset(&g, 0, sizeof(g));
    g.type = (uint32_t)pat->type;
    g.count = pat->g.count;
    memcpy(g.colors, pat->g.colors, sizeof(g.colors));
    memcpy(g.stops, pat->g.stops, sizeof(g.stops));
    memcpy(g.cp, pat->g.cp, sizeof(g.cp));
    vkvg_matrix_t pm;
    bool hasm = false;
    if (pat->hasMatrix) {
        pm = pat->matrix;
        if (vkvg_matrix_invert(&pm) != VKVG_STATUS_SUCCESS)
            vkvg_matrix_init_identity(&pm);
        hasm = true;
        vkvg_matrix_transform_point(&pm, &g.cp[0][0], &g.cp[0][1]);
    }
    vkvg_matrix_transform_point(&ctx->mat, &g.cp[0][0], &g.cp[0][1]);
    if (pat->type == VKVG_PATTERN_TYPE_LINEAR) {
        if (hasm)
            vkvg_matrix_transform_point(&pm, &g.cp[0][2], &g.cp[0][3]);
        vkvg_matrix_transform_point(&ctx->mat, &g.cp[0][2], &g.cp[0][3]);
    } else {
        if (hasm)
            vkvg_matrix_transform_point(&pm, &g.cp[1][0], &g.cp[1][1]);
        vkvg_matrix_transform_point(&ctx->mat, &g.cp[1][0], &g.cp[1][1]);
        if (hasm) {
            vkvg_matrix_transform_distance(&pm, &g.cp[0][2], &g.cp[0][3]);
            vkvg_matrix_transform_distance(&pm, &g.cp[1][2], &g.cp[0][3]);
        }
        vkvg_matrix_transform_distance(&ctx->mat, &g.cp[0][2], &g.cp[0][3]);
        vkvg_matrix_transform_distance(&ctx->mat, &g.cp[1][2], &g.cp[0][3]); /* as written in the reference */
    }
    VEC_PUSH(ctx->f.grads, g);
    ctx->cur_grad = (int32_t)(ctx->f.grads.n - 1);
}
/* vkvg_set_source_surface, src/vkvg_context.c:1025-1033 */
void vkvg_set_source_surface(VkvgContext ctx, VkvgSurface surf, float x, float y) {
    if (vkvg_status(ctx) || vkvg_surface_status(surf))
        return;
    if (ctx->recording) { /* recorded, not executed */
        vkvg_rec_surface(ctx->recording, surf, x, y);
        return;
    }
    ctx->src_rect[0] = x;
    ctx->src_rect[1] = y;
    VkvgPattern p = vkvg_pattern_create_for_surface(surf);
    if (vkvg_pattern_status(p)) {
        ctx->status = VKVG_STATUS_NO_MEMORY;
        return;
    }
    vkvg_set_source(ctx, p);
    vkvg_pattern_destroy(p); /* the context keeps the only reference */
}
VkvgPattern vkvg_get_source(VkvgContext ctx) {
    if (vkvg_status(ctx))
        return NULL;
    vkvg_pattern_reference(ctx->pattern);
    return ctx->pattern;
}
void vkvg_set_line_width(VkvgContext ctx, float width) {
    if (vkvg_status(ctx))
        return;
    if (ctx->recording) /* recorded, not executed */
        vkvg_rec_floats(ctx->recording, VKVG_CMD_SET_LINE_WIDTH, 1, width);
    else
        ctx->lineWidth = width;
}
float vkvg_get_line_width(VkvgContext ctx) { return vkvg_status(ctx) ? 0 : ctx->lineWidth; }
void vkvg_set_miter_limit(VkvgContext ctx, float limit) {
    if (vkvg_status(ctx))
        return;
    if (ctx->recording) /* recorded, not executed */
        vkvg_rec_floats(ctx->recording, VKVG_CMD_SET_MITER_LIMIT, 1, limit);
    else
        ctx->miterLimit = limit;
}
float vkvg_get_miter_limit(VkvgContext ctx) { return vkvg_status(ctx) ? 0 : ctx->miterLimit; }
void vkvg_set_line_cap(VkvgContext ctx, vkvg_line_cap_t cap) {
    if (vkvg_status(ctx))
        return;
    if (ctx->recording) /* recorded, not executed */
        vkvg_rec_u32(ctx->recording, VKVG_CMD_SET_LINE_CAP, (uint32_t)cap);
    else
        ctx->lineCap = cap;
}
vkvg_line_cap_t vkvg_get_line_cap(VkvgContext ctx) { return vkvg_status(ctx) ? 0 : ctx->lineCap; }
void vkvg_set_line_join(VkvgContext ctx, vkvg_line_join_t join) {
    if (vkvg_status(ctx))
        return;
    if (ctx->recording) /* recorded, not executed */
        vkvg_rec_u32(ctx->recording, VKVG_CMD_SET_LINE_JOIN, (uint32_t)join);
    else
        ctx->lineJoin = join;
}
vkvg_line_join_t vkvg_get_line_join(VkvgContext ctx) { return vkvg_status(ctx) ? 0 : ctx->lineJoin; }
void vkvg_set_operator(VkvgContext ctx, vkvg_operator_t op) {
    if (vkvg_status(ctx))
        return;
    if (ctx->recording) /* recorded, not executed */
        vkvg_rec_u32(ctx->recording, VKVG_CMD_SET_OPERATOR, (uint32_t)op);
    else
        ctx->op = op;
}
vkvg_operator_t vkvg_get_operator(VkvgContext ctx) { return vkvg_status(ctx) ? VKVG_OPERATOR_OVER : ctx->op; }
void vkvg_set_fill_rule(VkvgContext ctx, vkvg_fill_rule_t fr) {
    if (vkvg_status(ctx))
        return;
    if (ctx->recording) /* recorded, not executed */
        vkvg_rec_u32(ctx->recording, VKVG_CMD_SET_FILL_RULE, (uint32_t)fr);
    else
        ctx->fillRule = fr;
}
vkvg_fill_rule_t vkvg_get_fill_rule(VkvgContext ctx) {
    return vkvg_status(ctx) ? VKVG_FILL_RULE_NON_ZERO : ctx->fillRule;
}
void vkvg_set_dash(VkvgContext ctx, const float *dashes, uint32_t num_dashes, float offset) {
    if (vkvg_status(ctx))
        return;
    if (ctx->recording) { /* recorded, not executed */
        vkvg_rec_dash(ctx->recording, dashes, num_dashes, offset);
        return;
    }
    free(ctx->dashes);
    ctx->dashes = NULL;
    ctx->dashCount = num_dashes;
    ctx->dashOffset = offset;
    if (num_dashes == 0)
        return;
    ctx->dashes = (float *)malloc(sizeof(float) * num_dashes);
    if (!ctx->dashes || !dashes) {
        ctx->dashCount = 0;
        return;
    }
    memcpy(ctx->dashes, dashes, sizeof(float) * num_dashes);
}
void vkvg_get_dash(VkvgContext ctx, const float *dashes, uint32_t *num_dashes, float *offset) {
    if (vkvg_status(ctx))
        return;
    if (num_dashes)
        *num_dashes = ctx->dashCount;
    if (offset)
        *offset = ctx->dashOffset;
    if (ctx->dashCount == 0 || dashes == NULL)
        return;
    memcpy((float *)dashes, ctx->dashes, sizeof(float) * ctx->dashCount);
}

/* ---- transforms, src/vkvg_context.c:1514-1577 */
/* _set_mat_inv_and_vkCmdPush, src/vkvg_context_internal.c:672-676: the inverse of a singular
 * matrix stays the matrix itself (vkvg_matrix_invert leaves it untouched) */
static void update_mat_inv(VkvgContext ctx) {
    ctx->mat_inv = ctx->mat;
    vkvg_matrix_invert(&ctx->mat_inv);
}
void vkvg_translate(VkvgContext ctx, float dx, float dy) {
    if (vkvg_status(ctx))
        return;
    if (ctx->recording) { /* recorded, not executed */
        vkvg_rec_floats(ctx->recording, VKVG_CMD_TRANSLATE, 2, dx, dy);
        return;
    }
    vkvg_matrix_translate(&ctx->mat, dx, dy);
    update_mat_inv(ctx);
}
void vkvg_scale(VkvgContext ctx, float sx, float sy) {
    if (vkvg_status(ctx))
        return;
    if (ctx->recording) { /* recorded, not executed */
        vkvg_rec_floats(ctx->recording, VKVG_CMD_SCALE, 2, sx, sy);
        return;
    }
    vkvg_matrix_scale(&ctx->mat, sx, sy);
    update_mat_inv(ctx);
}
void vkvg_rotate(VkvgContext ctx, float radians) {
    if (vkvg_status(ctx))
        return;
    if (ctx->recording) { /* recorded, not executed */
        vkvg_rec_floats(ctx->recording, VKVG_CMD_ROTATE, 1, radians);
        return;
    }
    vkvg_matrix_rotate(&ctx->mat, radians);
    update_mat_inv(ctx);
}
void vkvg_transform(VkvgContext ctx, const vkvg_matrix_t *matrix) {
    if (vkvg_status(ctx) || !matrix)
        return;
    if (ctx->recording) { /* recorded, not executed */
        vkvg_rec_push(ctx->recording, VKVG_CMD_TRANSFORM, matrix, sizeof(*matrix));
        return;
    }
    vkvg_matrix_t res;
    vkvg_matrix_multiply(&res, &ctx->mat, matrix);
    ctx->mat = res;
    update_mat_inv(ctx);
}
void vkvg_identity_matrix(VkvgContext ctx) {
    if (vkvg_status(ctx))
        return;
    if (ctx->recording) { /* recorded, not executed */
        vkvg_rec_push(ctx->recording, VKVG_CMD_IDENTITY_MATRIX, NULL, 0);
        return;
    }
    vkvg_matrix_init_identity(&ctx->mat);
    update_mat_inv(ctx);
}
void vkvg_set_matrix(VkvgContext ctx, const vkvg_matrix_t *matrix) {
    if (vkvg_status(ctx) || !matrix)
        return;
    if (ctx->recording) { /* recorded, not executed */
        vkvg_rec_push(ctx->recording, VKVG_CMD_SET_MATRIX, matrix, sizeof(*matrix));
        return;
    }
    ctx->mat = *matrix;
    update_mat_inv(ctx);
}
void vkvg_get_matrix(VkvgContext ctx, vkvg_matrix_t *const matrix) {
    if (!vkvg_status(ctx) && matrix)
        *matrix = ctx->mat;
}

const char *vkvg_status_to_string(vkvg_status_t status) {
    switch (status) {
    case VKVG_STATUS_SUCCESS: return "no error has occurred";
    case VKVG_STATUS_NO_MEMORY: return "out of memory";
    case VKVG_STATUS_NULL_POINTER: return "NULL pointer";
    case VKVG_STATUS_INVALID_RESTORE: return "vkvg_restore() without matching vkvg_save()";
    case VKVG_STATUS_NO_CURRENT_POINT: return "no current point defined";
    case VKVG_STATUS_INVALID_MATRIX: return "invalid matrix (not invertible)";
    case VKVG_STATUS_INVALID_STATUS: return "invalid value for an input vkvg_status_t";
    case VKVG_STATUS_INVALID_INDEX: return "invalid index passed to getter";
    case VKVG_STATUS_WRITE_ERROR: return "error while writing to output stream";
    case VKVG_STATUS_PATTERN_TYPE_MISMATCH: return "the pattern type is not appropriate for the operation";
    case VKVG_STATUS_PATTERN_INVALID_GRADIENT: return "the stops count is zero";
    case VKVG_STATUS_INVALID_FORMAT: return "invalid value for an input vkvg_format_t";
    case VKVG_STATUS_FILE_NOT_FOUND: return "file not found";
    case VKVG_STATUS_INVALID_DASH: return "invalid value for a dash setting";
    case VKVG_STATUS_INVALID_RECT: return "a rectangle has the height or width equal to 0";
    case VKVG_STATUS_TIMEOUT: return "waiting for a device operation to finish resulted in a timeout";
    case VKVG_STATUS_DEVICE_ERROR: return "the initialization of the device resulted in an error";
    case VKVG_STATUS_INVALID_IMAGE: return "invalid image";
    case VKVG_STATUS_INVALID_SURFACE: return "invalid surface";
    case VKVG_STATUS_INVALID_FONT: return "unresolved font name";
    default: return "<unknown error status>";
    }
}

/* ---------------------------------------------------------------------------- replay */
static inline float F(const uint32_t *w) {
    float f;
    memcpy(&f, w, 4);
    return f;
}
/* operand counts per opcode (fixed part); SET_DASH and the gradients carry a count */
static int stream_step(VkvgContext c, const uint32_t *w, size_t avail, size_t *used, const VkvgSurface *srcs,
                       uint32_t n_srcs) {
    uint32_t op = w[0];
    const uint32_t *a = w + 1;
    size_t need = 1;
#define NEED(k)                                                                                                        \
    do {                                                                                                               \
        need = 1 + (k);                                                                                                \
        if (need > avail)                                                                                              \
            return -1;                                                                                                 \
    } while (0)
    switch (op) {
    case 0: *used = avail; return 1;
    case 1: NEED(2); vkvg_move_to(c, F(a), F(a + 1)); break;
    case 2: NEED(2); vkvg_line_to(c, F(a), F(a + 1)); break;
    case 3: NEED(6); vkvg_curve_to(c, F(a), F(a + 1), F(a + 2), F(a + 3), F(a + 4), F(a + 5)); break;
    case 4: vkvg_close_path(c); break;
    case 5: NEED(5); vkvg_arc(c, F(a), F(a + 1), F(a + 2), F(a + 3), F(a + 4)); break;
    case 6: NEED(5); vkvg_arc_negative(c, F(a), F(a + 1), F(a + 2), F(a + 3), F(a + 4)); break;
    case 7: NEED(4); vkvg_rectangle(c, F(a), F(a + 1), F(a + 2), F(a + 3)); break;
    case 8: vkvg_fill(c); break;
    case 9: vkvg_stroke(c); break;
    case 10: vkvg_fill_preserve(c); break;
    case 11: vkvg_stroke_preserve(c); break;
    case 12: vkvg_paint(c); break;
    case 13: NEED(4); vkvg_set_source_rgba(c, F(a), F(a + 1), F(a + 2), F(a + 3)); break;
    case 14: NEED(3); vkvg_set_source_rgb(c, F(a), F(a + 1), F(a + 2)); break;
    case 15: NEED(1); vkvg_set_line_width(c, F(a)); break;
    case 16: NEED(1); vkvg_set_line_join(c, (vkvg_line_join_t)a[0]); break;
    case 17: NEED(1); vkvg_set_line_cap(c, (vkvg_line_cap_t)a[0]); break;
    case 18: NEED(1); vkvg_set_miter_limit(c, F(a)); break;
    case 19: NEED(1); vkvg_set_fill_rule(c, (vkvg_fill_rule_t)a[0]); break;
    case 20: {
        NEED(2);
        uint32_t n = a[0];
        NEED(2 + (size_t)n);
        float d[64];
        uint32_t nn = n > 64 ? 64 : n;
        for (uint32_t k = 0; k < nn; k++)
            d[k] = F(a + 2 + k);
        vkvg_set_dash(c, d, nn, F(a + 1));
        break;
    }
    case 21: NEED(1); vkvg_set_opacity(c, F(a)); break;
    case 22: NEED(1); vkvg_set_operator(c, (vkvg_operator_t)a[0]); break;
    case 23:
    case 24:
    case 50:
    case 51: {
        /* 50 / 51: a pattern matrix (6 floats) first, then the operands of 23 / 24 */
        const size_t pre = op >= 50 ? 6 : 0;
        const bool linear = op == 23 || op == 50;
        const size_t ncp = linear ? 4 : 6;
        NEED(pre + ncp + 1);
        const uint32_t *g = a + pre;
        uint32_t ns = g[ncp];
        NEED(pre + ncp + 1 + 5 * (size_t)ns);
        VkvgPattern p = linear ? vkvg_pattern_create_linear(F(g), F(g + 1), F(g + 2), F(g + 3))
                               : vkvg_pattern_create_radial(F(g), F(g + 1), F(g + 2), F(g + 3), F(g + 4), F(g + 5));
        const uint32_t *st = g + ncp + 1;
        for (uint32_t k = 0; k < ns; k++)
            vkvg_pattern_add_color_stop(p, F(st + 5 * k), F(st + 5 * k + 1), F(st + 5 * k + 2), F(st + 5 * k + 3),
                                        F(st + 5 * k + 4));
        if (pre) {
            vkvg_matrix_t m = {F(a), F(a + 1), F(a + 2), F(a + 3), F(a + 4), F(a + 5)};
            vkvg_pattern_set_matrix(p, &m);
        }
        vkvg_set_source(c, p);
        vkvg_pattern_destroy(p);
        break;
    }
    case 25: vkvg_new_path(c); break;
    case 26: vkvg_new_sub_path(c); break;
    case 27: NEED(2); vkvg_rel_line_to(c, F(a), F(a + 1)); break;
    case 28: NEED(2); vkvg_rel_move_to(c, F(a), F(a + 1)); break;
    case 29: NEED(4); vkvg_quadratic_to(c, F(a), F(a + 1), F(a + 2), F(a + 3)); break;
    case 30: NEED(2); vkvg_translate(c, F(a), F(a + 1)); break;
    case 31: NEED(2); vkvg_scale(c, F(a), F(a + 1)); break;
    case 32: NEED(1); vkvg_rotate(c, F(a)); break;
    case 33: vkvg_identity_matrix(c); break;
    case 34: vkvg_clear(c); break;
    case 35: NEED(5); vkvg_ellipse(c, F(a), F(a + 1), F(a + 2), F(a + 3), F(a + 4)); break;
    case 36: NEED(6); vkvg_rel_curve_to(c, F(a), F(a + 1), F(a + 2), F(a + 3), F(a + 4), F(a + 5)); break;
    case 37: {
        NEED(6);
        vkvg_matrix_t m = {F(a), F(a + 1), F(a + 2), F(a + 3), F(a + 4), F(a + 5)};
        vkvg_set_matrix(c, &m);
        break;
    }
    case 38: NEED(7); vkvg_elliptic_arc_to(c, F(a), F(a + 1), a[2] != 0, a[3] != 0, F(a + 4), F(a + 5), F(a + 6)); break;
    case 39: NEED(5); vkvg_rounded_rectangle(c, F(a), F(a + 1), F(a + 2), F(a + 3), F(a + 4)); break;
    case 40: NEED(1); vkvg_set_source_color(c, a[0]); break;
    case 41: vkvg_clip(c); break;
    case 42: vkvg_clip_preserve(c); break;
    case 43: vkvg_reset_clip(c); break;
    case 44: vkvg_save(c); break;
    case 45: vkvg_restore(c); break;
    case 46: vkvg_flush(c); break;
    case 47:
        NEED(3);
        if (a[0] >= n_srcs) {
            c->status = VKVG_STATUS_INVALID_INDEX;
            break;
        }
        vkvg_set_source_surface(c, srcs[a[0]], F(a + 1), F(a + 2));
        break;
    case 49: NEED(6); vkvg_rounded_rectangle2(c, F(a), F(a + 1), F(a + 2), F(a + 3), F(a + 4), F(a + 5)); break;
    case 48: {
        NEED(10);
        if (a[0] >= n_srcs) {
            c->status = VKVG_STATUS_INVALID_INDEX;
            break;
        }
        VkvgPattern p = vkvg_pattern_create_for_surface(srcs[a[0]]);
        vkvg_pattern_set_extend(p, (vkvg_extend_t)a[1]);
        vkvg_pattern_set_filter(p, (vkvg_filter_t)a[2]);
        if (a[3]) {
            vkvg_matrix_t m = {F(a + 4), F(a + 5), F(a + 6), F(a + 7), F(a + 8), F(a + 9)};
            vkvg_pattern_set_matrix(p, &m);
        }
        vkvg_set_source(c, p);
        vkvg_pattern_destroy(p);
        break;
    }
    default: return -1;
    }
#undef NEED
    *used = need;
    return 0;
}
vkvg_status_t vkvg_amd_replay_sources(VkvgContext ctx, const uint32_t *words, size_t n, const VkvgSurface *sources,
                                      uint32_t n_sources) {
    if (vkvg_status(ctx))
        return vkvg_status(ctx);
    size_t i = 0;
    while (i < n && !ctx->status) {
        size_t used = 0;
        int r = stream_step(ctx, words + i, n - i, &used, sources, sources ? n_sources : 0);
        if (r < 0) {
            ctx->status = VKVG_STATUS_INVALID_STATUS;
            break;
        }
        if (r > 0)
            break;
        i += used;
    }
    return ctx->status;
}

vkvg_status_t vkvg_amd_replay(VkvgContext ctx, const uint32_t *words, size_t n) {
    return vkvg_amd_replay_sources(ctx, words, n, NULL, 0);
}

/* ---------------------------------------------------------------------------- batch */
vkvg_status_t vkvg_amd_upload_batch(VkvgDevice dev, uint32_t n, const VkvgSurface *surfs,
                                    const uint32_t *const *streams, const size_t *n_words) {
    if (vkvg_device_status(dev))
        return VKVG_STATUS_DEVICE_ERROR;
    if (!n)
        return VKVG_STATUS_SUCCESS;
    VkvgContext *ctxs = (VkvgContext *)calloc(n, sizeof(VkvgContext));
    vkh_target *tg = (vkh_target *)calloc(n, sizeof(vkh_target));
    frame_t all;
    memset(&all, 0, sizeof(all));
    uint32_t n_draws_total = 0;
    vkvg_status_t st = VKVG_STATUS_SUCCESS;
    if (!ctxs || !tg) {
        st = VKVG_STATUS_NO_MEMORY;
        goto done;
    }
    for (uint32_t i = 0; i < n && !st; i++) {
        if (vkvg_surface_status(surfs[i]) || surfs[i]->dev != dev) {
            st = VKVG_STATUS_INVALID_SURFACE;
            break;
        }
        ctxs[i] = vkvg_create(surfs[i]);
        if (vkvg_status(ctxs[i])) {
            st = vkvg_status(ctxs[i]);
            ctxs[i] = NULL;
            break;
        }
        ctxs[i]->deferred = true;
        st = vkvg_amd_replay(ctxs[i], streams[i], n_words[i]);
        if (st)
            break;
        /* append this context's frame, rebasing indices */
        frame_t *f = &ctxs[i]->f;
        uint32_t ib = all.items.n, sb = all.sps.n, stb = all.styles.n, db = all.dashes.n, gb = all.grads.n,
                 srb = all.srcs.n;
        VEC_RESERVE(all.items, all.items.n + f->items.n);
        memcpy(all.items.p + all.items.n, f->items.p, sizeof(vkh_item) * f->items.n);
        all.items.n += f->items.n;
        for (uint32_t k = 0; k < f->sps.n; k++) {
            vkh_subpath sp = f->sps.p[k];
            sp.first_item += ib;
            VEC_PUSH(all.sps, sp);
        }
        for (uint32_t k = 0; k < f->styles.n; k++) {
            vkh_stroke_style s = f->styles.p[k];
            s.dash_first += db;
            VEC_PUSH(all.styles, s);
        }
        for (uint32_t k = 0; k < f->dashes.n; k++)
            VEC_PUSH(all.dashes, f->dashes.p[k]);
        for (uint32_t k = 0; k < f->grads.n; k++)
            VEC_PUSH(all.grads, f->grads.p[k]);
        for (uint32_t k = 0; k < f->srcs.n; k++)
            VEC_PUSH(all.srcs, f->srcs.p[k]);
        for (uint32_t k = 0; k < f->draws.n; k++) {
            vkh_draw d = f->draws.p[k];
            d.first_subpath += sb;
            if (d.kind == VKH_DRAW_STROKE && d.style >= 0)
                d.style += (int32_t)stb;
            if (d.grad >= 0)
                d.grad += (int32_t)gb;
            if (d.source >= 0)
                d.source += (int32_t)srb;
            d.target = i;
            VEC_PUSH(all.draws, d);
        }
        tg[i].rgba = surfs[i]->img;
        tg[i].width = surfs[i]->width;
        tg[i].height = surfs[i]->height;
        tg[i].sig = surfs[i]->sig;
        tg[i].stencil = ctxs[i]->stencil_use ? surfs[i]->stencil : NULL;
        tg[i].stencil_load = 0; /* a fresh context: its stencil starts from zero */
    }
    n_draws_total = all.draws.n;
    if (!st && all.draws.n) {
        vkh_frame fr;
        fr.items = all.items.p;
        fr.n_items = all.items.n;
        fr.subpaths = all.sps.p;
        fr.n_subpaths = all.sps.n;
        fr.draws = all.draws.p;
        fr.n_draws = all.draws.n;
        fr.styles = all.styles.p;
        fr.n_styles = all.styles.n;
        fr.dashes = all.dashes.p;
        fr.n_dashes = all.dashes.n;
        fr.grads = all.grads.p;
        fr.n_grads = all.grads.n;
        fr.sources = all.srcs.p;
        fr.n_sources = all.srcs.n;
        fr.targets = tg;
        fr.n_targets = n;
        if (vkvg_hip_upload(dev->eng, &fr)) {
            fprintf(stderr, "vkvg: batch upload failed: %s\n", vkvg_hip_last_error(dev->eng));
            st = VKVG_STATUS_DEVICE_ERROR;
        }
    }
done:
    for (uint32_t i = 0; ctxs && i < n; i++) {
        if (ctxs[i]) {
            ctxs[i]->f.draws.n = 0; /* already rendered as part of the batch */
            vkvg_destroy(ctxs[i]);
        }
    }
    frame_free(&all);
    free(ctxs);
    free(tg);
    if (!st && n_draws_total == 0)
        st = VKVG_STATUS_INVALID_STATUS; /* nothing to execute */
    return st;
}

vkvg_status_t vkvg_amd_render_batch(VkvgDevice dev, uint32_t n, const VkvgSurface *surfs,
                                    const uint32_t *const *streams, const size_t *n_words) {
    vkvg_status_t st = vkvg_amd_upload_batch(dev, n, surfs, streams, n_words);
    if (st == VKVG_STATUS_INVALID_STATUS)
        return VKVG_STATUS_SUCCESS; /* empty batch */
    if (st)
        return st;
    if (vkvg_hip_execute(dev->eng)) {
        fprintf(stderr, "vkvg: batch render failed: %s\n", vkvg_hip_last_error(dev->eng));
        return VKVG_STATUS_DEVICE_ERROR;
    }
    return VKVG_STATUS_SUCCESS;
}

/* ---------------------------------------------------------------------------- text
 * include/vkvg.h:1645-1745. Fonts (FreeType / HarfBuzz in the reference) are outside this backend:
 * the entry points exist so that clients link; they draw nothing and leave the status unchanged. */
void vkvg_select_font_face(VkvgContext ctx, const char *name) {
    if (!vkvg_status(ctx) && ctx->recording)
        vkvg_rec_text(ctx->recording, VKVG_CMD_SET_FONT_FACE, name);
}
void vkvg_load_font_from_path(VkvgContext ctx, const char *path, const char *name) {
    (void)name;
    if (!vkvg_status(ctx) && ctx->recording)
        vkvg_rec_text(ctx->recording, VKVG_CMD_SET_FONT_PATH, path);
}
void vkvg_load_font_from_memory(VkvgContext ctx, unsigned char *buf, long size, const char *name) {
    (void)ctx; (void)buf; (void)size; (void)name;
}
void vkvg_set_font_size(VkvgContext ctx, uint32_t size) {
    if (!vkvg_status(ctx) && ctx->recording)
        vkvg_rec_u32(ctx->recording, VKVG_CMD_SET_FONT_SIZE, size);
}
void vkvg_show_text(VkvgContext ctx, const char *utf8) {
    if (!vkvg_status(ctx) && ctx->recording)
        vkvg_rec_text(ctx->recording, VKVG_CMD_SHOW_TEXT, utf8);
}
void vkvg_text_extents(VkvgContext ctx, const char *utf8, vkvg_text_extents_t *extents) {
    (void)ctx; (void)utf8;
    if (extents)
        memset(extents, 0, sizeof(*extents));
}
void vkvg_font_extents(VkvgContext ctx, vkvg_font_extents_t *extents) {
    (void)ctx;
    if (extents)
        memset(extents, 0, sizeof(*extents));
}
VkvgText vkvg_text_run_create(VkvgContext ctx, const char *text) { (void)ctx; (void)text; return NULL; }
VkvgText vkvg_text_run_create_with_length(VkvgContext ctx, const char *text, uint32_t length) {
    (void)ctx; (void)text; (void)length;
    return NULL;
}
void vkvg_text_run_destroy(VkvgText run) { (void)run; }
void vkvg_show_text_run(VkvgContext ctx, VkvgText run) { (void)ctx; (void)run; }
void vkvg_text_run_get_extents(VkvgText run, vkvg_text_extents_t *extents) {
    (void)run;
    if (extents)
        memset(extents, 0, sizeof(*extents));
}
uint32_t vkvg_text_run_get_glyph_count(VkvgText run) { (void)run; return 0; }
void vkvg_text_run_get_glyph_position(VkvgText run, uint32_t index, vkvg_glyph_info_t *info) {
    (void)run; (void)index;
    if (info)
        memset(info, 0, sizeof(*info));
}
