"""CPU checks of the restated reference test programs (vkvg_amd.scenes.reference_test_scenes): every
scene renders through the oracle without an error status, and the scenes whose image follows from
geometry alone match that closed form (pixel centres strictly inside axis-aligned edges, opaque
sources, so the top-left rule and the blend never decide)."""
import numpy as np

from oracle import ovg
from vkvg_amd import scenes


def _by_name(w=512, h=512):
    return {sc.name: sc for sc in scenes.reference_test_scenes(w, h)}


def _rgba(sc):
    rgba, _, _, status = ovg.render(sc.width, sc.height, sc.words())
    assert status == 0, sc.name
    return rgba


def test_every_reference_scene_renders():
    scs = scenes.reference_test_scenes()
    assert len({sc.name for sc in scs}) == len(scs)
    for sc in scs:
        _, _, stats, status = ovg.render(sc.width, sc.height, sc.words())
        assert status == 0, sc.name
        assert stats["draws"] >= 1, sc.name


def test_colinear_points_stroke_as_one_straight_band():
    # tests/colinear.c: width 10 through (100,100) (100,200) (100,300), BUTT caps: the band
    # x in [95, 105), y in [100, 300) on the grey background, nothing else
    got = _rgba(_by_name()["colinear"])
    band = np.zeros(got.shape[:2], bool)
    band[100:300, 95:105] = True
    assert (got[band] == [0, 255, 0, 255]).all()
    bg = got[~band]
    assert (bg == bg[0]).all() and bg[0][3] == 255 and bg[0][0] == bg[0][1] == bg[0][2]


def test_fill_non_zero_keeps_the_counter_wound_hole():
    # tests/fill_non_zero.c: a clockwise [100,300) x [100,250) and a counter-clockwise
    # [150,250) x [150,200): the inner winding is 0, a hole
    got = _rgba(_by_name()["fill_non_zero"])
    exp = np.zeros(got.shape[:2], bool)
    exp[100:250, 100:300] = True
    exp[150:200, 150:250] = False
    assert (got[exp] == [25, 229, 25, 255]).all()  # (0.1, 0.9, 0.1) truncated to bytes
    assert (got[~exp] == 0).all()


def test_paint_with_a_path_fills_the_path():
    # tests/simple_paint.c paint_rect: the rectangle only, after the clear
    got = _rgba(_by_name()["simple_paint_rect"])
    exp = np.zeros(got.shape[:2], bool)
    exp[100:300, 100:400] = True
    assert (got[exp] == [255, 0, 0, 255]).all() and (got[~exp] == 0).all()


def test_zero_length_dashes_with_butt_caps_draw_nothing():
    # inverse_colinear with dash {0, 60} and BUTT caps: every dash has length 0
    for name, sc in _by_name().items():
        if name.startswith("inverse_colinear_") and "c0" in name and "d0_" in name:
            assert (_rgba(sc) == 0).all(), name


def test_gradient_extend_is_not_read():
    # tests/gradient.c paint_repeat sets EXTEND_REPEAT on a linear gradient; the reference's
    # source setup reads the extend of surface patterns only, so the scene is the PAD gradient:
    # constant red left of x = 0 (no pixel), blue right of x = 300
    got = _rgba(_by_name()["gradient_paint_repeat"])
    assert (got[:, 301:] == [0, 0, 255, 255]).all()


def test_gradient_pattern_matrix_is_inverted_onto_the_control_points():
    # src/vkvg_context_internal.c:797-822: the control points go through the INVERSE of the pattern
    # matrix, then the CTM. Scale 0.5 inverts exactly to 2, so the gradient with the matrix equals
    # the gradient with doubled control points (a matrix applied uninverted would halve them).
    from vkvg_amd.stream import Stream
    stops = [[0, 1, 0, 0, 1], [0.5, 0, 1, 0, 1], [1.0, 0, 0, 1, 1]]

    def lin(geom, m):
        st = Stream()
        st.translate(30, 20)
        st.set_linear(*geom, stops, matrix=m)
        st.rectangle(0, 0, 400, 140)
        st.fill()
        return _rgba(scenes.Scene("l", 480, 200, st, 1))

    def rad(geom, m):
        st = Stream()
        st.set_radial(*geom, stops, matrix=m)
        st.paint()
        return _rgba(scenes.Scene("r", 400, 300, st, 0))

    half = (0.5, 0, 0, 0.5, 0, 0)
    assert np.array_equal(lin((0, 0, 200, 0), half), lin((0, 0, 400, 0), None))
    assert not np.array_equal(lin((0, 0, 200, 0), half), lin((0, 0, 100, 0), None))
    # radii are distances: doubled too (the centres are chosen so vkvg_pattern_edit_radial moves nothing)
    assert np.array_equal(rad((50, 50, 5, 60, 55, 40), half), rad((100, 100, 10, 120, 110, 80), None))
    # a translation moves the gradient the other way: cp - t
    assert np.array_equal(lin((0, 0, 200, 0), (1, 0, 0, 1, 50, 0)), lin((-50, 0, 150, 0), None))
    # a singular matrix is replaced by the identity
    assert np.array_equal(lin((0, 0, 200, 0), (0, 0, 0, 0, 0, 0)), lin((0, 0, 200, 0), None))


def test_every_reference_source_scene_renders():
    for sc, srcs in scenes.reference_source_scenes():
        imgs = []
        for w, h, st in srcs:
            rgba, _, _, status = ovg.render(w, h, st.words())
            assert status == 0
            imgs.append(rgba)
        rgba, _, stats, status = ovg.render(sc.width, sc.height, sc.words(), sources=imgs)
        assert status == 0 and stats["draws"] >= 1, sc.name
        assert rgba.any(), sc.name


def test_paint_surf_offsets_place_the_drawn_source():
    # tests/paint_surf.c paint_with_offset: the 256 x 256 source at (100, 100) over half green:
    # outside it the green stays, inside it the source composited over the green
    sc, srcs = [x for x in scenes.reference_source_scenes() if x[0].name == "paint_surf_paint_with_offset"][0]
    w, h, st = srcs[0]
    src = ovg.render(w, h, st.words())[0]
    got = ovg.render(sc.width, sc.height, sc.words(), sources=[src])[0]
    outside = np.ones(got.shape[:2], bool)
    outside[100:356, 100:356] = False
    assert (got[outside] == got[0, 0]).all() and got[0, 0, 1] > 0 and got[0, 0, 0] == 0
    inside = got[100:356, 100:356].astype(np.int32)
    # OVER: s + round(d (255 - sa) / 255) per channel, with d the uniform green
    d = got[0, 0].astype(np.int32)
    s = src.astype(np.int32)
    exp = np.minimum(s + np.floor((d[None, None, :] * (255 - s[..., 3:4]) + 127) / 255).astype(np.int32), 255)
    assert np.abs(inside - exp).max() <= 1
