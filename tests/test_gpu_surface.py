"""Surface sources on the GPU: bit-exact against the oracle (RGBA and coverage signatures) for
random bitmaps sampled with every extend mode and filter under offsets, pattern matrices, rotated
and scaled CTMs, opacity, clips and save / restore; a source that the GPU itself rendered; and
vkvg_surface_create_from_bitmap keeping the bytes it is given."""
import numpy as np
import pytest

from vkvg_amd import scenes
from parity_util import compare

pytestmark = pytest.mark.gpu


def _gpu(device, sc, srcs, tile=16):
    import vkvg_amd
    device.set_tile_size(tile)
    ss = [vkvg_amd.Surface.from_bitmap(device, b) for b in srcs]
    t = vkvg_amd.Surface(device, sc.width, sc.height, signature=True)
    try:
        t.replay(sc.words(), sources=ss)
        return t.premultiplied(), t.signature()
    finally:
        t.close()
        for s in ss:
            s.close()


@pytest.mark.parametrize("variant", range(4))
def test_surface_mix_matches_oracle(device, variant):
    from oracle import ovg
    sc, srcs = scenes.surface_mix(variant)
    tile = (16, 8, 32, 16)[variant]
    print(f"case {sc.name} tile {tile}", flush=True)
    g_rgba, g_sig = _gpu(device, sc, srcs, tile)
    o_rgba, o_sig, _, st = ovg.render(sc.width, sc.height, sc.words(), sources=srcs)
    assert st == 0
    r = compare(g_rgba, g_sig, o_rgba, o_sig)
    assert r["mask_mismatch_px"] == 0 and r["rgba_max_err"] == 0, r


def test_bitmap_surface_keeps_its_bytes(device):
    import vkvg_amd
    b = (np.arange(37 * 23 * 4) * 7 % 256).astype(np.uint8).reshape(23, 37, 4)
    s = vkvg_amd.Surface.from_bitmap(device, b)
    try:
        assert np.array_equal(s.premultiplied(), b)
    finally:
        s.close()


def test_source_rendered_on_the_gpu(device):
    import vkvg_amd
    from oracle import ovg
    from vkvg_amd.stream import EXTEND_REFLECT, FILTER_BILINEAR, Stream
    a = scenes.clip_mix(2, size=120, n=50, flushes=False)
    print("case rendered source", flush=True)
    device.set_tile_size(16)
    sa = vkvg_amd.Surface(device, a.width, a.height)
    t = vkvg_amd.Surface(device, 200, 160, signature=True)
    try:
        sa.replay(a.words())
        a_oracle = ovg.render(a.width, a.height, a.words())[0]
        assert np.array_equal(sa.premultiplied(), a_oracle)
        st = Stream()
        st.set_source_rgb(0.2, 0.3, 0.4)
        st.paint()
        st.rotate(0.3)
        st.set_source_pattern(0, extend=EXTEND_REFLECT, filt=FILTER_BILINEAR, matrix=(0.7, 0.1, -0.1, 0.7, 5, 9))
        st.set_opacity(0.8)
        st.arc(100, 70, 60, 0, 6.2831853)
        st.fill()
        t.replay(st.words(), sources=[sa])
        o_rgba, o_sig, _, status = ovg.render(200, 160, st.words(), sources=[a_oracle])
        assert status == 0
        r = compare(t.premultiplied(), t.signature(), o_rgba, o_sig)
        assert r["mask_mismatch_px"] == 0 and r["rgba_max_err"] == 0, r
    finally:
        t.close()
        sa.close()


def test_source_released_before_the_frame_renders(device):
    # the application drops its last reference to the source surface (and the context's pattern)
    # before the context renders: the frame still samples the source's pixels
    import ctypes
    import vkvg_amd
    from oracle import ovg
    from vkvg_amd.stream import Stream
    L = vkvg_amd.lib()
    b = (np.arange(40 * 30 * 4) * 13 % 256).astype(np.uint8).reshape(30, 40, 4)
    b[..., 3] = 255
    src = vkvg_amd.Surface.from_bitmap(device, b)
    t = vkvg_amd.Surface(device, 64, 48, signature=True)
    st = Stream()
    st.set_source_surface(0, 5, 3)
    st.paint()
    st.set_source_rgb(0, 0, 1)  # the context lets go of the surface pattern
    w = st.words()
    try:
        print("case source released before render", flush=True)
        ctx = L.vkvg_create(t.h)
        hs = (ctypes.c_void_p * 1)(src.h)
        assert L.vkvg_amd_replay_sources(ctx, w.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)), w.size, hs, 1) == 0
        src.close()  # the last reference the application holds
        other = vkvg_amd.Surface(device, 40, 30)  # a new image of the same size, painted red
        red = Stream()
        red.set_source_rgb(1, 0, 0)
        red.paint()
        other.replay(red.words())
        L.vkvg_destroy(ctx)  # renders the frame
        other.close()
        o_rgba, o_sig, _, status = ovg.render(64, 48, w, sources=[b])
        assert status == 0
        r = compare(t.premultiplied(), t.signature(), o_rgba, o_sig)
        assert r["mask_mismatch_px"] == 0 and r["rgba_max_err"] == 0, r
    finally:
        t.close()
        src.close()


@pytest.mark.parametrize("variant", range(8))
def test_grand_mix_matches_oracle(device, variant):
    # every feature interleaved at random (sources, operators, transforms, clips, save / restore,
    # dashes, clears, flushes), three tile sizes
    from oracle import ovg
    sc, srcs = scenes.grand_mix(variant)
    tile = (8, 16, 32)[variant % 3]
    print(f"case {sc.name} tile {tile}", flush=True)
    g_rgba, g_sig = _gpu(device, sc, srcs, tile)
    o_rgba, o_sig, _, st = ovg.render(sc.width, sc.height, sc.words(), sources=srcs)
    assert st == 0
    r = compare(g_rgba, g_sig, o_rgba, o_sig)
    assert r["mask_mismatch_px"] == 0 and r["rgba_max_err"] == 0, r
