"""GPU readback: vkvg_surface_write_to_memory and vkvg_surface_write_to_png against the oracle.

The reference reads the BGRA image back as RGBA and unpremultiplies on the host with truncation:
p = (unsigned char)((double)p / (a / 255.f)) (src/vkvg_surface.c:393-467, :283-391). Here the
device kernel k_unpremultiply does it. Surfaces are rendered WITHOUT the white background, so the
readback sees translucent pixels and pixels with alpha 0. Status codes follow
gunit_tests/surface.cpp:94-119."""
import ctypes
import io
import os

import numpy as np
import pytest

from oracle import ovg
from vkvg_amd import scenes
from parity_util import compare, render_oracle

pytestmark = pytest.mark.gpu

VKVG_STATUS_SUCCESS, VKVG_STATUS_INVALID_STATUS, VKVG_STATUS_WRITE_ERROR = 0, 6, 8  # include/vkvg.h


def _render(dev, scene):
    import vkvg_amd
    s = vkvg_amd.Surface(dev, scene.width, scene.height, signature=True)
    s.replay(scene.words())
    return s


def _check_scene(dev, scene, tmp_path):
    surf = _render(dev, scene)
    try:
        premul, sig, rgba = surf.premultiplied(), surf.signature(), surf.rgba()
        png = os.path.join(str(tmp_path), scene.name + ".png")
        surf.write_png(png)
    finally:
        surf.close()
    o_premul, o_sig, _ = render_oracle(scene)
    r = compare(premul, sig, o_premul, o_sig)
    print(scene.name, r, flush=True)
    assert r["mask_mismatch_px"] == 0 and r["rgba_max_err"] == 0, r
    alpha = o_premul[..., 3]
    n0, ntr = int(np.count_nonzero(alpha == 0)), int(np.count_nonzero((alpha > 0) & (alpha < 255)))
    assert n0 > 0 and ntr > 0, (n0, ntr)  # the readback sees empty and translucent pixels
    want = ovg.unpremultiply(o_premul)
    assert np.array_equal(rgba, want), int(np.count_nonzero(np.any(rgba != want, axis=2)))
    assert not np.any(rgba[alpha == 0])
    # the PNG holds the same unpremultiplied bytes (8-bit RGBA); PIL decodes it independently
    from PIL import Image
    with open(png, "rb") as f:
        data = f.read()
    im = Image.open(io.BytesIO(data))
    assert im.mode == "RGBA" and im.size == (scene.width, scene.height)
    assert np.array_equal(np.asarray(im), want)
    return rgba


def test_write_to_memory_and_png_translucent_512(device, tmp_path):
    _check_scene(device, scenes.config1(white=False), tmp_path)


def test_write_to_memory_and_png_translucent_4096_frame(device, tmp_path):
    # the dense 4096x4096 scene (20,000 paths, gradients) on a transparent surface
    _check_scene(device, scenes.config3(white=False), tmp_path)


def test_colour_above_alpha_and_zero_alpha_bitmap(device):
    """Bytes stored as given by create_from_bitmap, including colour > alpha (the quotient exceeds
    255 and wraps through the 8-bit conversion) and alpha 0 (reads 0)."""
    import vkvg_amd
    rng = np.random.default_rng(9)
    img = rng.integers(0, 256, (37, 53, 4), dtype=np.uint8)
    img[::5, :, 3] = 0
    img[1::7, :, 3] = 255
    img[2, :4] = [[200, 0, 0, 100], [255, 255, 255, 1], [10, 20, 30, 0], [128, 64, 32, 64]]
    s = vkvg_amd.Surface.from_bitmap(device, img)
    try:
        got = s.rgba()
        assert np.array_equal(s.premultiplied(), img)
    finally:
        s.close()
    want = ovg.unpremultiply(img)
    assert np.array_equal(got, want)
    assert list(got[2, 0]) == [254, 0, 0, 100]      # 200 / (100/255) = 510 -> low byte 254
    assert list(got[2, 2]) == [0, 0, 0, 0]


def test_write_status_codes_of_the_reference(device, tmp_path):
    """gunit_tests/surface.cpp:94-105 (SurfWrite) and :107-119 (SurfClear)."""
    import vkvg_amd
    L = vkvg_amd.lib()
    assert L.vkvg_surface_write_to_memory(None, None) == VKVG_STATUS_INVALID_STATUS
    assert L.vkvg_surface_write_to_png(None, None) == VKVG_STATUS_INVALID_STATUS
    s = vkvg_amd.Surface(device, 8, 8)
    try:
        assert L.vkvg_surface_write_to_memory(s.h, None) == VKVG_STATUS_WRITE_ERROR
        assert L.vkvg_surface_write_to_png(s.h, None) == VKVG_STATUS_WRITE_ERROR
        buf = np.ones(64, np.uint32)
        assert L.vkvg_surface_write_to_memory(s.h, buf.ctypes.data) == VKVG_STATUS_SUCCESS
        assert not buf.any()  # a new surface is transparent black
        bad = os.path.join(str(tmp_path), "no_such_dir", "x.png")
        assert L.vkvg_surface_write_to_png(s.h, bad.encode()) == VKVG_STATUS_WRITE_ERROR
        # SurfClear: paint red, destroy the context, clear the surface: every pixel reads 0x00000000
        from vkvg_amd.stream import Stream
        st = Stream()
        st.set_source_rgb(1, 0, 0)
        st.paint()
        s.replay(st.words())
        assert np.all(s.rgba()[..., 0] == 255)
        L.vkvg_surface_clear(s.h)
        assert L.vkvg_surface_write_to_memory(s.h, buf.ctypes.data) == VKVG_STATUS_SUCCESS
        assert not buf.any()
    finally:
        s.close()


def test_surface_clear_also_clears_the_clip_stencil(device):
    """vkvg_surface_clear clears colour and stencil (src/vkvg_surface.c:32-36): a context that clipped
    and flushed paints everywhere after the clear, since its clip bit was cleared with the stencil."""
    import vkvg_amd
    from vkvg_amd.stream import Stream
    L = vkvg_amd.lib()
    L.vkvg_rectangle.argtypes = [ctypes.c_void_p] + [ctypes.c_float] * 4
    L.vkvg_set_source_rgb.argtypes = [ctypes.c_void_p] + [ctypes.c_float] * 3
    for f in (L.vkvg_clip, L.vkvg_paint):
        f.argtypes = [ctypes.c_void_p]
    s = vkvg_amd.Surface(device, 32, 24)
    try:
        ctx = L.vkvg_create(s.h)
        L.vkvg_rectangle(ctx, 4, 4, 8, 8)
        L.vkvg_clip(ctx)
        L.vkvg_set_source_rgb(ctx, 0, 0, 1)
        L.vkvg_paint(ctx)
        L.vkvg_flush(ctx)
        img = s.rgba()
        assert np.count_nonzero(img[..., 3]) == 64  # clipped to the 8x8 square
        L.vkvg_surface_clear(s.h)
        L.vkvg_set_source_rgb(ctx, 0, 1, 0)
        L.vkvg_paint(ctx)
        L.vkvg_destroy(ctx)
        img = s.rgba()
        assert np.all(img == [0, 255, 0, 255])
    finally:
        s.close()
