/*
 * vkvg_amd.h — extensions of the MI355X backend that have no counterpart in the reference
 * API. They serve callers that need more than the drop-in surface:
 *   - replay of a recorded command stream through the vkvg_* entry points (the reference's
 *     recording feature, src/recording/vkvg_record_internal.c:105-249, plays the same role);
 *   - per-pixel coverage signatures, used to prove 1-sample mask parity;
 *   - batched rendering of many surfaces in one frame;
 *   - backend statistics and tile-size control.
 */
#ifndef VKVG_AMD_H
#define VKVG_AMD_H

#include "vkvg.h"
#include "vkvg_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Replays a command stream (uint32 opcodes with bit-cast float operands, see
 * vkvg_amd/stream.py for the opcode table) by calling the vkvg_* entry points in order.
 * Returns the context status afterwards. */
vkvg_public vkvg_status_t vkvg_amd_replay(VkvgContext ctx, const uint32_t *words, size_t n_words);
/* The same, with the surfaces that the stream's surface-source opcodes index (out of range:
 * VKVG_STATUS_INVALID_INDEX). */
vkvg_public vkvg_status_t vkvg_amd_replay_sources(VkvgContext ctx, const uint32_t *words, size_t n_words,
                                                  const VkvgSurface *sources, uint32_t n_sources);

/* Coverage signatures: when enabled, every draw adds k * key(n) to a per-pixel uint32, where
 * k is the number of blends the draw made at that pixel and n the draw's number on the
 * surface. Two renderers produce equal signature images exactly when all per-draw 1-sample
 * masks agree. */
vkvg_public vkvg_status_t vkvg_amd_surface_enable_signature(VkvgSurface surf);
vkvg_public vkvg_status_t vkvg_amd_surface_read_signature(VkvgSurface surf, uint32_t *out);
/* raw premultiplied RGBA8 as stored in HBM (no unpremultiply) */
vkvg_public vkvg_status_t vkvg_amd_surface_read_premultiplied(VkvgSurface surf, unsigned char *out);
/* device pointer of the surface image (width*height*4 bytes) */
vkvg_public void *vkvg_amd_surface_device_ptr(VkvgSurface surf);

/* Renders n streams onto n surfaces as ONE frame (one pass of the kernel pipeline over all
 * surfaces' draws). Equivalent to creating a context per surface, replaying its stream and
 * destroying the context, surface by surface. */
vkvg_public vkvg_status_t vkvg_amd_render_batch(VkvgDevice dev, uint32_t n, const VkvgSurface *surfaces,
                                                const uint32_t *const *streams, const size_t *n_words);
/* Same as vkvg_amd_render_batch without running the pipeline: the merged frame is left
 * resident on the device for vkvg_amd_execute. */
vkvg_public vkvg_status_t vkvg_amd_upload_batch(VkvgDevice dev, uint32_t n, const VkvgSurface *surfaces,
                                                const uint32_t *const *streams, const size_t *n_words);
/* Uploads the context's pending frame to the device without rendering it (the frame is then
 * owned by the device: vkvg_amd_execute renders it, as many times as called). */
vkvg_public vkvg_status_t vkvg_amd_upload(VkvgContext ctx);
/* Runs the pipeline on the device's resident frame; returns 0 on success. */
vkvg_public int vkvg_amd_execute(VkvgDevice dev);

/* PNG decoding behind vkvg_surface_create_from_image (no counterpart in the reference, which links
 * stb_image): 8-bit RGBA, not premultiplied, top row first, as stbi_load returns with 4 components.
 * Host only. Returns NULL on a malformed or unsupported file; free the result with vkvg_amd_png_free. */
vkvg_public uint8_t *vkvg_amd_png_decode(const uint8_t *buf, size_t n, uint32_t *width, uint32_t *height);
vkvg_public void     vkvg_amd_png_free(uint8_t *rgba);

/* Backend controls */
vkvg_public void       vkvg_amd_device_get_stats(VkvgDevice dev, vkh_stats *out);
vkvg_public int        vkvg_amd_device_set_tile_size(VkvgDevice dev, uint32_t tile);
vkvg_public vkh_engine *vkvg_amd_device_engine(VkvgDevice dev);

#ifdef __cplusplus
}
#endif
#endif
