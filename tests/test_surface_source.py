"""CPU pins of the oracle's surface sources (no GPU).

A surface source samples the source image at u' = ((M^-1 (f - xy)) / wh) * size (f: the pixel
centre; shaders/vkvg_main.frag:72-82) with the sampler that the pattern's extend and filter select
(src/vkvg_context_internal.c:705-772). With power-of-two source sizes the division and the
multiplication are exact, so every expected image below is integer arithmetic on texel indices:
floor(x + 0.5 - ox) shifted, scaled, wrapped (mod / mirror / clamp) or averaged. Painting with OVER
onto a transparent target leaves the sampled bytes unchanged. The last tests pin the reference's
push-constant quirks: a CTM change rebuilds the inverse matrix (dropping a pattern matrix), and
vkvg_restore multiplies a surface pattern's matrix in once more.
"""
import numpy as np

from oracle import ovg
from vkvg_amd.stream import (EXTEND_NONE, EXTEND_PAD, EXTEND_REFLECT, EXTEND_REPEAT, FILTER_BILINEAR,
                             FILTER_NEAREST, Stream)

SW, SH = 32, 16  # source size (powers of two)


def _src(seed=1, even=False):
    rng = np.random.default_rng(seed)
    s = rng.integers(0, 256, size=(SH, SW, 4), dtype=np.uint8)
    return (s & 0xFE) if even else s


def _paint(st, w=96, h=56, sources=None):
    rgba, _, _, status = ovg.render(w, h, st.words(), sources=sources)
    assert status == 0
    return rgba


def _expect(src, w, h, ix, iy, extend=EXTEND_NONE):
    """target[y, x] = src[wrap(iy[y]), wrap(ix[x])], transparent for the border."""
    out = np.zeros((h, w, 4), np.uint8)

    def wrap(i, n):
        if extend == EXTEND_REPEAT:
            return np.mod(i, n), np.ones_like(i, bool)
        if extend == EXTEND_REFLECT:
            t = np.mod(i, 2 * n)
            return np.where(t < n, t, 2 * n - 1 - t), np.ones_like(i, bool)
        if extend == EXTEND_PAD:
            return np.clip(i, 0, n - 1), np.ones_like(i, bool)
        return np.clip(i, 0, n - 1), (i >= 0) & (i < n)

    xs, okx = wrap(np.asarray(ix), SW)
    ys, oky = wrap(np.asarray(iy), SH)
    ok = oky[:, None] & okx[None, :]
    out[ok] = src[ys[:, None], xs[None, :]][ok]
    return out


def test_nearest_places_the_surface_at_its_offset_with_a_transparent_border():
    src = _src()
    for ox, oy in ((10, 7), (10.25, 6.75), (-5, 3)):
        st = Stream()
        st.set_source_surface(0, ox, oy)
        st.paint()
        w, h = 96, 56
        ix = np.floor(np.arange(w) + 0.5 - ox).astype(np.int64)
        iy = np.floor(np.arange(h) + 0.5 - oy).astype(np.int64)
        assert np.array_equal(_paint(st, w, h, [src]), _expect(src, w, h, ix, iy)), (ox, oy)


def test_extend_modes_wrap_repeat_mirror_and_clamp():
    src = _src(2)
    w, h = 96, 56
    for ext in (EXTEND_REPEAT, EXTEND_REFLECT, EXTEND_PAD, EXTEND_NONE):
        st = Stream()
        st.translate(9, 5)  # the inverse CTM moves the surface to (9, 5)
        st.set_source_pattern(0, extend=ext, filt=FILTER_NEAREST)
        st.paint()
        ix, iy = np.arange(w) - 9, np.arange(h) - 5
        assert np.array_equal(_paint(st, w, h, [src]), _expect(src, w, h, ix, iy, ext)), ext


def test_scaled_ctm_magnifies_around_the_device_offset():
    # p = f - (x, y) first, then the inverse CTM: scale 2 doubles texels, anchored at the device
    # offset (the reference does not transform (x, y) by the CTM)
    src = _src(3)
    st = Stream()
    st.scale(2, 2)
    st.set_source_surface(0, 4, 2)
    st.paint()
    w, h = 96, 56
    ix = (np.arange(w) - 4) // 2
    iy = (np.arange(h) - 2) // 2
    assert np.array_equal(_paint(st, w, h, [src]), _expect(src, w, h, ix, iy))


def test_pattern_matrix_is_applied_without_inversion():
    src = _src(4)
    st = Stream()
    st.set_source_pattern(0, extend=EXTEND_REPEAT, matrix=(0.5, 0, 0, 0.25, 0, 0))
    st.paint()
    w, h = 96, 56
    ix, iy = np.arange(w) // 2, np.arange(h) // 4
    assert np.array_equal(_paint(st, w, h, [src]), _expect(src, w, h, ix, iy, EXTEND_REPEAT))


def test_bilinear_at_texel_centres_and_half_texel_offsets():
    # u' - 0.5 integral: the weights are (1, 0): nearest. Half a texel off: the average of two
    # neighbours, exact because the texels are even.
    src = _src(5, even=True)
    w, h = 64, 32
    st = Stream()
    st.set_source_pattern(0, extend=EXTEND_PAD, filt=FILTER_BILINEAR)
    st.paint()
    assert np.array_equal(_paint(st, w, h, [src]),
                          _expect(src, w, h, np.arange(w), np.arange(h), EXTEND_PAD))
    st = Stream()
    st.translate(0.5, 0)
    st.set_source_pattern(0, extend=EXTEND_PAD, filt=FILTER_BILINEAR)
    st.paint()
    a = _expect(src, w, h, np.arange(w) - 1, np.arange(h), EXTEND_PAD).astype(np.int32)
    b = _expect(src, w, h, np.arange(w), np.arange(h), EXTEND_PAD).astype(np.int32)
    assert np.array_equal(_paint(st, w, h, [src]), ((a + b) // 2).astype(np.uint8))


def test_opacity_scales_the_sampled_texel():
    src = _src(6, even=True)
    st = Stream()
    st.set_opacity(0.5)
    st.set_source_surface(0, 0, 0)
    st.paint()
    w, h = 48, 24
    exp = _expect(src, w, h, np.arange(w), np.arange(h)) // 2
    assert np.array_equal(_paint(st, w, h, [src]), exp)


def test_a_ctm_change_rebuilds_the_inverse_and_drops_the_pattern_matrix():
    src = _src(7)
    w, h = 96, 56
    st = Stream()
    st.set_source_pattern(0, extend=EXTEND_REPEAT, matrix=(0.5, 0, 0, 0.5, 0, 0))
    st.translate(3, 0)  # matInv := inverse(CTM): the pattern matrix is gone
    st.paint()
    assert np.array_equal(_paint(st, w, h, [src]),
                          _expect(src, w, h, np.arange(w) - 3, np.arange(h), EXTEND_REPEAT))


def test_restore_multiplies_the_surface_pattern_matrix_in_again():
    # save keeps matInv = M; restore puts it back, then re-applies the pattern: M * M
    src = _src(8)
    w, h = 96, 56
    st = Stream()
    st.set_source_pattern(0, extend=EXTEND_REPEAT, matrix=(0.5, 0, 0, 0.5, 0, 0))
    st.save()
    st.set_source_rgb(1, 0, 0)
    st.restore()
    st.paint()
    assert np.array_equal(_paint(st, w, h, [src]),
                          _expect(src, w, h, np.arange(w) // 4, np.arange(h) // 4, EXTEND_REPEAT))


def test_surface_source_is_clipped_and_fills_take_it():
    src = _src(9)
    w, h = 96, 56
    st = Stream()
    st.set_source_surface(0, 0, 0)
    st.rectangle(4, 2, 20, 10)
    st.clip()
    st.rectangle(0, 0, 40, 40)
    st.fill()
    exp = np.zeros((h, w, 4), np.uint8)
    exp[2:12, 4:24] = src[2:12, 4:24]
    assert np.array_equal(_paint(st, w, h, [src]), exp)


def test_out_of_range_source_index_is_an_error():
    st = Stream()
    st.set_source_surface(1, 0, 0)
    _, _, _, status = ovg.render(8, 8, st.words(), sources=[_src()])
    assert status == 7  # VKVG_STATUS_INVALID_INDEX


def test_off_diagonal_inverse_matrix_swaps_the_axes():
    # CTM (xx yx xy yy) = (0 2 1 0), inverse (0 1 0.5 0) exactly: u = 0.5 * py from the pixel's
    # y, v = px from its x (a transposed matrix would give u = px, v = 0.5 * py)
    src = _src(10)
    w, h = 40, 48
    st = Stream()
    st.set_matrix((0, 2, 1, 0, 0, 0))
    st.set_source_pattern(0, extend=EXTEND_REPEAT)
    st.paint()
    exp = src[np.arange(w)[None, :] % SH, (np.arange(h)[:, None] // 2) % SW]
    assert np.array_equal(_paint(st, w, h, [src]), exp)


def test_pattern_matrix_applies_after_the_inverse_ctm():
    # vkvg_matrix_multiply(matInv, matInv, pattern): first the inverse CTM (x - 8), then the
    # pattern's scale: texel (x - 8) // 2, not x // 2 - 8
    src = _src(11)
    w, h = 96, 56
    st = Stream()
    st.translate(8, 0)
    st.set_source_pattern(0, extend=EXTEND_REPEAT, matrix=(0.5, 0, 0, 1, 0, 0))
    st.paint()
    assert np.array_equal(_paint(st, w, h, [src]),
                          _expect(src, w, h, (np.arange(w) - 8) // 2, np.arange(h), EXTEND_REPEAT))
