/* Differentiable renderer of the vkvg HIP backend (C ABI).
 *
 * The reference has no counterpart: these entry points are new. They render closed cubic paths
 * (non-zero fills and strokes) with analytic area coverage into an FP32 premultiplied RGBA image
 * on opaque white, and give the gradient of a loss with respect to the control points, colours
 * and stroke widths, given the loss's gradient with respect to the image.
 *
 * Geometry (fixed by the scene's topology, so the function is smooth almost everywhere): cubic q
 * of a path with n cubics runs from control point 3q through 3q+1, 3q+2 to 3q+3 (mod 3n) and is
 * sampled at t = i/16, i = 0..15. A fill uses the closed polyline; a stroke of half-width w/2 uses
 * one rectangle per polyline segment plus two bevel triangles per vertex. Coverage of a pixel is
 * min(1, |integral of the winding number over the pixel|); compositing is premultiplied
 * source-over, out = c k + out (1 - c.a k) with c = (r a, g a, b a, a).
 *
 * All parameter, image and gradient pointers are DEVICE pointers (HIP), laid out as
 *   ctrl  : float[2 * n_ctrl]  (x, y) pairs, n_ctrl = 3 * sum(ncubic)
 *   rgba  : float[4 * n_paths] straight (not premultiplied) colour and alpha
 *   width : float[n_paths]     stroke widths (ignored for fills)
 *   image : float[4 * width * height] row-major premultiplied RGBA
 * Calls are synchronous with respect to the host. Errors return a negative value;
 * vkvg_diff_last_error() describes the last one.
 */
#ifndef VKVG_DIFF_H
#define VKVG_DIFF_H

#include <stdint.h>

#ifndef vkvg_public
#define vkvg_public __attribute__((visibility("default")))
#endif

#ifdef __cplusplus
extern "C" {
#endif

typedef struct vkvg_diff_scene_t *VkvgDiffScene;

/* Topology: ncubic[i] >= 1 cubics per closed path, is_stroke[i] != 0 for a stroked path (host
 * arrays). Returns NULL on invalid input or allocation failure. */
vkvg_public VkvgDiffScene vkvg_diff_create(uint32_t width, uint32_t height, uint32_t n_paths, const uint32_t *ncubic,
                               const uint32_t *is_stroke);
vkvg_public void vkvg_diff_destroy(VkvgDiffScene s);

/* Renders the image. With keep != 0 the compositing checkpoints the backward pass needs are kept
 * (coverage itself is recomputed by the backward pass). */
vkvg_public int vkvg_diff_forward(VkvgDiffScene s, const float *ctrl, const float *rgba, const float *width, float *image,
                      int keep);

/* The same forward evaluated in double precision from the same float parameters, into a double
 * image (4 * width * height): the reference for finite differences, whose perturbations are far
 * below the spacing of float coordinates. Leaves no checkpoints. */
vkvg_public int vkvg_diff_forward_f64(VkvgDiffScene s, const float *ctrl, const float *rgba, const float *width,
                                      double *image);

/* Gradients of the loss given grad_image = dL/dimage. Must follow a vkvg_diff_forward with
 * keep != 0 on the same parameters. Outputs are overwritten. */
vkvg_public int vkvg_diff_backward(VkvgDiffScene s, const float *ctrl, const float *rgba, const float *width,
                       const float *grad_image, float *grad_ctrl, float *grad_rgba, float *grad_width);

/* Device time of the last forward / backward call in milliseconds, and frame statistics:
 * out[0] edges, out[1] segments, out[2] commands, out[3] batches. */
vkvg_public float vkvg_diff_last_ms(VkvgDiffScene s, int backward);
vkvg_public void vkvg_diff_stats(VkvgDiffScene s, uint32_t out[4]);
vkvg_public const char *vkvg_diff_last_error(VkvgDiffScene s);

#ifdef __cplusplus
}
#endif

#endif
