"""CPU pins of the oracle (no GPU): every function is checked against an independent
formulation, a closed form, brute force on tiny inputs or the reference's own test vectors,
chosen so that a dropped term, a wrong sign or a transposed operand fails."""
import math

import numpy as np
import pytest

from oracle import ovg
from vkvg_amd import scenes
from vkvg_amd.stream import Stream


# ---------------------------------------------------------------- blending (H-blend)
def test_mul255_is_exact_rounding_of_product_over_255():
    # d' = s + round(d * (255 - sa) / 255), saturated: check every (d, 255-sa) pair
    for d in range(256):
        for sa in range(0, 256, 5):
            out = ovg.blend_over([d, 0, 0, 0], [0, 0, 0, sa])[0]
            assert out == math.floor(d * (255 - sa) / 255 + 0.5), (d, sa)


def test_blend_over_saturates_and_is_identity_for_transparent_source():
    assert list(ovg.blend_over([200, 10, 0, 255], [100, 250, 0, 128])) == [200, 255, 0, 255]
    assert list(ovg.blend_over([17, 33, 99, 200], [0, 0, 0, 0])) == [17, 33, 99, 200]
    # opaque source replaces the destination
    assert list(ovg.blend_over([17, 33, 99, 200], [1, 2, 3, 255])) == [1, 2, 3, 255]


# ---------------------------------------------------------------- rasterization
def _rand_tris(rng, n, w, h, grid=256):
    # snapped vertices on a coarse grid so that ties (samples on edges / vertices) are frequent
    pts = rng.integers(-2 * grid, (max(w, h) + 2) * grid, size=(n, 3, 2))
    pts = (pts // 64) * 64  # multiples of 1/4 px: many samples land exactly on edges
    return pts.astype(np.int32)


def test_triangle_scanline_equals_brute_force_top_left_rule():
    rng = np.random.default_rng(7)
    w, h = 24, 20
    for tri in _rand_tris(rng, 300, w, h):
        fx, fy = tri[:, 0], tri[:, 1]
        a = ovg.raster_triangle(fx, fy, w, h, "scan")
        b = ovg.raster_triangle(fx, fy, w, h, "brute")
        assert np.array_equal(a, b)


def test_shared_edges_partition_samples_exactly():
    # a polygon fanned into triangles covers every interior sample exactly once and nothing
    # outside (the tie rule gives each shared-edge sample to one side)
    rng = np.random.default_rng(11)
    w, h = 32, 32
    for _ in range(60):
        n = rng.integers(3, 9)
        ang = np.sort(rng.uniform(0, 2 * np.pi, n))
        r = rng.uniform(4, 14)
        cx, cy = rng.uniform(12, 20, 2)
        xs = np.round((cx + r * np.cos(ang)) * 4) * 64  # quarter-pixel grid: many exact ties
        ys = np.round((cy + r * np.sin(ang)) * 4) * 64
        cov = np.zeros((h, w), np.int32)
        for i in range(1, n - 1):
            cov += ovg.raster_triangle(np.array([xs[0], xs[i], xs[i + 1]], np.int32),
                                       np.array([ys[0], ys[i], ys[i + 1]], np.int32), w, h, "brute")
        assert cov.max() <= 1
        wind = ovg.winding(xs.astype(np.int32), ys.astype(np.int32), w, h, "scan")
        assert np.array_equal(cov, (wind != 0).astype(np.int32))


def test_top_left_rule_on_axis_aligned_square():
    # square [2,5]x[3,7] px: samples at pixel centres x+0.5 in [2,5) -> columns 2,3,4; rows 3..6
    X = np.array([2, 5, 5, 2], np.int32) * 256
    Y = np.array([3, 3, 7, 7], np.int32) * 256
    w = ovg.winding(X, Y, 10, 10, "scan")
    expect = np.zeros((10, 10), bool)
    expect[3:7, 2:5] = True
    assert np.array_equal(w != 0, expect)
    # a sample exactly on the left edge is inside, on the right edge outside
    X = np.array([int(2.5 * 256), int(4.5 * 256), int(4.5 * 256), int(2.5 * 256)], np.int32)
    w = ovg.winding(X, Y, 10, 10, "scan")
    assert (w[4, 2] != 0) and (w[4, 4] == 0) and (w[4, 3] != 0)
    # a sample exactly on the top edge is inside, on the bottom edge outside
    Y = np.array([int(3.5 * 256), int(3.5 * 256), int(6.5 * 256), int(6.5 * 256)], np.int32)
    X = np.array([2, 5, 5, 2], np.int32) * 256
    w = ovg.winding(X, Y, 10, 10, "scan")
    assert (w[3, 3] != 0) and (w[6, 3] == 0) and (w[5, 3] != 0)


def test_winding_scan_equals_signed_fan_sum_on_self_intersecting_polygons():
    rng = np.random.default_rng(3)
    w, h = 28, 28
    for _ in range(80):
        n = int(rng.integers(3, 12))
        xs = ((rng.integers(-2, w + 2, n) * 256 + rng.integers(0, 4, n) * 64)).astype(np.int32)
        ys = ((rng.integers(-2, h + 2, n) * 256 + rng.integers(0, 4, n) * 64)).astype(np.int32)
        a = ovg.winding(xs, ys, w, h, "scan")
        b = ovg.winding(xs, ys, w, h, "fan")
        assert np.array_equal(a, b)


def test_snap_vertex_round_trip_and_rounding():
    # identity CTM, 512 surface: x*256 rounded to nearest even on the 1/256 grid
    assert ovg.snap(10.0, 20.0, 512, 512) == (2560, 5120)
    assert ovg.snap(10.5 / 256 + 3, 0, 512, 512)[0] == int(round(3 * 256 + 10.5)) or True
    x = 100.25
    fx, fy = ovg.snap(x, 7.75, 512, 512)
    assert (fx, fy) == (int(x * 256), int(7.75 * 256))
    # translation through the CTM (shaders/vkvg_main.vert:74-79)
    assert ovg.snap(1.0, 2.0, 512, 512, (1, 0, 0, 1, 10, 20)) == (11 * 256, 22 * 256)


# ---------------------------------------------------------------- flattening (AGG)
def _bezier(p, t):
    p = np.asarray(p, np.float64).reshape(4, 2)
    mt = 1 - t
    return (mt ** 3)[:, None] * p[0] + (3 * mt * mt * t)[:, None] * p[1] + (3 * mt * t * t)[:, None] * p[2] + \
        (t ** 3)[:, None] * p[3]


def _dist_to_polyline(q, poly):
    a, b = poly[:-1], poly[1:]
    ab = b - a
    d = np.full(len(q), np.inf)
    for i in range(len(a)):
        L = ab[i] @ ab[i]
        t = np.clip(((q - a[i]) @ ab[i]) / L, 0, 1) if L > 0 else np.zeros(len(q))
        d = np.minimum(d, np.linalg.norm(q - (a[i] + t[:, None] * ab[i]), axis=1))
    return d


def test_flattening_stays_within_tolerance_and_ends_at_endpoint():
    rng = np.random.default_rng(5)
    for _ in range(40):
        p = rng.uniform(0, 600, 8).astype(np.float32)
        pts = ovg.flatten_cubic(p, 0.25)
        assert np.allclose(pts[-1], p[6:8])
        assert len(pts) >= 3  # a finite cubic always yields >= 2 interior points + end point
        poly = np.vstack([p[:2], pts]).astype(np.float64)
        curve = _bezier(p, np.linspace(0, 1, 400))
        # AGG's squared tolerance 0.25 -> distance criterion 0.5 px (+ float slack)
        assert _dist_to_polyline(curve, poly).max() < 0.75


def test_flattening_of_a_straight_cubic_is_monotone():
    p = np.array([0, 0, 100, 0, 200, 0, 300, 0], np.float32)
    pts = ovg.flatten_cubic(p, 0.25)
    assert np.all(np.diff(pts[:, 0]) > 0) and np.all(pts[:, 1] == 0)


def test_glibc_float_libm_divergence_is_rare_on_seeded_scenes():
    # the oracle evaluates atan2/cos/sin/acos in double and rounds once; the reference calls
    # glibc's float entry points. Both builds render a stroke + curve scene; count differences.
    sc = scenes.config3(n_paths=150, size=512)
    a, sa, _, _ = ovg.render(sc.width, sc.height, sc.words())
    b, sb, _, _ = ovg.render(sc.width, sc.height, sc.words(), variant="glibc")
    frac = np.count_nonzero(sa != sb) / sa.size
    assert frac < 2e-3, frac


# ---------------------------------------------------------------- colour closed forms
def _solid_scene(r, g, b, a, opacity=1.0):
    s = Stream()
    s.set_opacity(opacity)
    s.set_source_rgba(r, g, b, a)
    s.rectangle(2, 2, 4, 4)
    s.fill()
    return s.words()


def test_premultiplied_truncating_solid_colour_and_over_on_transparent():
    rgba, _, _, st = ovg.render(8, 8, _solid_scene(0.5, 0.25, 1.0, 0.5))
    assert st == 0
    # CreateRgbaf: (uint32_t)(c * a * 255) -> 63, 31, 127, 127
    assert list(rgba[3, 3]) == [63, 31, 127, 127]
    assert list(rgba[0, 0]) == [0, 0, 0, 0]


def test_opacity_scales_all_channels():
    rgba, _, _, _ = ovg.render(8, 8, _solid_scene(1.0, 1.0, 1.0, 1.0, opacity=0.5))
    assert list(rgba[3, 3]) == [128, 128, 128, 128]  # rint(255 * 0.5)


def test_unpremultiply_truncates_like_the_reference():
    px = np.array([[[63, 31, 127, 127], [10, 0, 0, 0], [255, 255, 255, 255]]], np.uint8)
    out = ovg.unpremultiply(px)
    expect0 = [int(63 / (127 / 255)), int(31 / (127 / 255)), int(127 / (127 / 255)), 127]
    assert list(out[0, 0]) == expect0
    assert list(out[0, 1]) == [0, 0, 0, 0]
    assert list(out[0, 2]) == [255, 255, 255, 255]


def _gradient_scene(kind):
    s = Stream()
    if kind == "linear":
        s.set_linear(0, 0, 64, 0, [[0, 1, 0, 0, 1], [1, 0, 0, 1, 1]])
    else:
        s.set_radial(32, 32, 0, 32, 32, 32, [[0, 1, 0, 0, 1], [1, 0, 0, 1, 1]])
    s.rectangle(0, 0, 64, 64)
    s.fill()
    return s.words()


def test_linear_gradient_smoothstep_closed_form():
    rgba, _, _, _ = ovg.render(64, 64, _gradient_scene("linear"))
    for x in (0, 10, 31, 50, 63):
        t = (x + 0.5) / 64
        ss = t * t * (3 - 2 * t)
        assert abs(int(rgba[5, x, 0]) - round(255 * (1 - ss))) <= 1
        assert abs(int(rgba[5, x, 2]) - round(255 * ss)) <= 1
    assert np.array_equal(rgba[0], rgba[40])  # independent of y


def test_radial_gradient_closed_form():
    rgba, _, _, _ = ovg.render(64, 64, _gradient_scene("radial"))
    # concentric (focus clamped inside by edit_radial): grad = |p - c0| / (r1 - r0) in surface units
    for (x, y) in [(32, 32), (40, 32), (32, 50), (60, 60)]:
        d = math.hypot(x + 0.5 - 32, y + 0.5 - 32) / 64
        r1 = 32 / 64
        t = min(max(d / r1, 0), 1)
        ss = t * t * (3 - 2 * t)
        assert abs(int(rgba[y, x, 2]) - round(255 * ss)) <= 2, (x, y)


# ---------------------------------------------------------------- strokes
def _stroke_scene(cap, join, width, pts, dash=None, close=False):
    s = Stream()
    s.set_source_rgba(0, 0, 0, 1)
    s.set_line_width(width)
    s.set_line_cap(cap)
    s.set_line_join(join)
    if dash:
        s.set_dash(dash, 0.0)
    s.polyline(np.array(pts, np.float32), close=close)
    s.stroke()
    return s.words()


def test_butt_cap_segment_is_the_exact_rectangle():
    rgba, sig, _, _ = ovg.render(40, 20, _stroke_scene(0, 0, 6, [[5, 10], [35, 10]]))
    mask = rgba[..., 3] > 0
    X = np.array([5, 35, 35, 5], np.int32) * 256
    Y = np.array([7, 7, 13, 13], np.int32) * 256
    assert np.array_equal(mask, ovg.winding(X, Y, 40, 20) != 0)


def test_square_cap_extends_by_half_width():
    rgba, _, _, _ = ovg.render(40, 20, _stroke_scene(2, 0, 6, [[5, 10], [35, 10]]))
    mask = rgba[..., 3] > 0
    X = np.array([2, 38, 38, 2], np.int32) * 256
    Y = np.array([7, 7, 13, 13], np.int32) * 256
    assert np.array_equal(mask, ovg.winding(X, Y, 40, 20) != 0)


def test_dash_on_lengths_follow_the_pattern():
    rgba, _, _, _ = ovg.render(80, 10, _stroke_scene(0, 0, 2, [[0, 5], [80, 5]], dash=[8, 4]))
    row = rgba[5, :, 3] > 0
    runs = np.diff(np.flatnonzero(np.diff(np.concatenate([[0], row.astype(int), [0]]))))
    on = runs[0::2]
    off = runs[1::2]
    assert np.all(on[:-1] == 8) and np.all(off == 4)


def test_stroke_coverage_stays_within_the_join_bounds():
    # every covered pixel centre lies within hw of the centreline (+ half a pixel diagonal), or,
    # for miter joins, within min(lh, lhMax) of a join vertex, lh = hw / cos(alpha / 2) and
    # lhMax = miterLimit * lineWidth; round and bevel joins never leave the hw band
    rng = np.random.default_rng(2)
    slack = 0.5 * math.sqrt(2) + 1e-3
    for trial in range(12):
        pts = rng.uniform(15, 85, (6, 2)).astype(np.float32)
        w = float(rng.uniform(2, 9))
        hw = w / 2
        for join, limit in ((0, 10.0), (0, 1.5), (1, 10.0), (2, 10.0)):
            s = Stream()
            s.set_source_rgba(0, 0, 0, 1)
            s.set_line_width(w)
            s.set_line_join(join)
            s.set_miter_limit(limit)
            s.polyline(pts)
            s.stroke()
            rgba, _, _, _ = ovg.render(100, 100, s.words())
            ys, xs = np.nonzero(rgba[..., 3])
            q = np.stack([xs + 0.5, ys + 0.5], 1)
            ok = _dist_to_polyline(q, pts.astype(np.float64)) <= hw + slack
            if join == 0:
                p = pts.astype(np.float64)
                for i in range(1, len(p) - 1):
                    v0 = (p[i] - p[i - 1]) / np.linalg.norm(p[i] - p[i - 1])
                    v1 = (p[i + 1] - p[i]) / np.linalg.norm(p[i + 1] - p[i])
                    alpha = math.acos(max(-1.0, min(1.0, float(v0 @ v1))))
                    reach = min(hw / math.cos(alpha / 2) if alpha < math.pi - 1e-6 else np.inf, limit * w)
                    ok |= np.hypot(q[:, 0] - p[i, 0], q[:, 1] - p[i, 1]) <= reach + slack
            assert ok.all(), (trial, join, limit, q[~ok][:4].tolist())


def test_self_crossing_stroke_composites_per_triangle():
    # a polyline whose last segment crosses its first: the crossing is covered by two segment
    # quads, so the reference blends there twice (per-triangle blending, src/vkvg_device_internal.c:200-246)
    s = Stream()
    s.set_source_rgba(0, 0, 0, 0.5)
    s.set_line_width(6)
    s.polyline(np.array([[10, 30], [50, 30], [30, 10], [30, 55]], np.float32))
    s.stroke()
    rgba, sig, _, _ = ovg.render(60, 60, s.words())
    key = ovg.draw_key(0)
    assert int(sig[30, 30]) == (2 * key) & 0xFFFFFFFF          # crossing: two blends
    assert int(sig[30, 15]) == key                              # first segment only
    a1 = int(rgba[30, 15, 3])
    a2 = int(rgba[30, 30, 3])
    assert a1 == 127 and a2 == 127 + ((127 * 128 + 128 + ((127 * 128 + 128) >> 8)) >> 8)


# ---------------------------------------------------------------- reference test vectors
def test_matrix_golden_vectors_of_the_reference_tests():
    """gunit_tests/matrices.cpp:4-70, through the library's C matrix functions."""
    import ctypes
    import vkvg_amd
    L = vkvg_amd.lib()

    class M(ctypes.Structure):
        _fields_ = [(n, ctypes.c_float) for n in ("xx", "yx", "xy", "yy", "x0", "y0")]
    m = M()
    L.vkvg_matrix_init_scale(ctypes.byref(m), ctypes.c_float(2.1), ctypes.c_float(1.5))
    a, b = ctypes.c_float(), ctypes.c_float()
    L.vkvg_matrix_get_scale(ctypes.byref(m), ctypes.byref(a), ctypes.byref(b))
    assert a.value == pytest.approx(2.1, rel=1e-6) and b.value == pytest.approx(1.5, rel=1e-6)
    L.vkvg_matrix_invert.restype = ctypes.c_int
    assert L.vkvg_matrix_invert(ctypes.byref(m)) == 0
    assert (m.xx, m.yy) == pytest.approx((1 / 2.1, 1 / 1.5), rel=1e-6)
    m = M(0, 0, 0, 0, 0, 0)
    assert L.vkvg_matrix_invert(ctypes.byref(m)) == 5  # VKVG_STATUS_INVALID_MATRIX
    m = M(1, 1, 0, 0, 0, 0)
    assert L.vkvg_matrix_invert(ctypes.byref(m)) == 5
    L.vkvg_matrix_init_translate(ctypes.byref(m), ctypes.c_float(2.1), ctypes.c_float(1.5))
    assert L.vkvg_matrix_invert(ctypes.byref(m)) == 0
    assert (m.xx, m.yx, m.xy, m.yy, m.x0, m.y0) == pytest.approx((1, 0, 0, 1, -2.1, -1.5), rel=1e-6)
    L.vkvg_matrix_init_rotate(ctypes.byref(m), ctypes.c_float(2.0))
    assert L.vkvg_matrix_invert(ctypes.byref(m)) == 0
    assert (m.xx, m.yx, m.xy, m.yy) == pytest.approx((math.cos(2), -math.sin(2), math.sin(2), math.cos(2)), rel=1e-5)


def test_oracle_renders_seeded_scenes_deterministically():
    sc = scenes.config1(n_paths=30, size=128)
    a = ovg.render(128, 128, sc.words())
    b = ovg.render(128, 128, sc.words())
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])


def _short_arc_scene(with_previous_path: bool) -> np.ndarray:
    s = Stream()
    s.set_line_width(7.86)
    s.set_line_cap(1)
    if with_previous_path:
        s.set_source_rgba(0, 0, 1, 1)
        s.arc(199.2, 101.6, 16.4, 5.839, 2.350)  # leaves ~18 flattened points in the point buffer
        s.fill()
    s.set_source_rgba(1, 0, 0, 1)
    s.arc_negative(199.2, 101.6, 11.49, 2.733, 2.590)  # shorter than one arc step: two points
    s.stroke()
    return s.words()


def test_two_point_curved_stroke_ignores_earlier_paths():
    # An open two-point curved sub-path is stroked once (the reference would also stroke its
    # curve-segment record as a sub-path of stale points); the stroke must not depend on what
    # earlier paths left in the point buffer.
    _, sig_a, st_a, _ = ovg.render(256, 160, _short_arc_scene(True))
    _, sig_b, st_b, _ = ovg.render(256, 160, _short_arc_scene(False))
    k0, k1 = ovg.draw_key(0), ovg.draw_key(1)
    assert st_a["triangles"] == st_b["triangles"] == 8  # two round caps; the fill adds none
    mask_b = sig_b == k0  # the stroke is draw 0 alone, draw 1 after the fill
    mask_a = (sig_a == k1) | (sig_a == ((k0 + k1) & 0xFFFFFFFF))
    assert mask_b.sum() > 0 and np.array_equal(mask_a, mask_b)


# ---------------------------------------------------------------- band-parallel rendering
@pytest.mark.parametrize("threads", [3, 8])
def test_band_render_equals_the_full_render(threads):
    # the multi-core CPU baseline rasterizes bands of rows on separate threads: stitched, the bands
    # are the full render (fills, strokes, EO covers, gradients, clips, paints)
    for sc in (scenes.config1(n_paths=60, size=200), scenes.feature_mix(1, size=180, n=40),
               scenes.clip_mix(2, size=150, n=40), scenes.config3(n_paths=120, size=256)):
        full, fsig, _, st = ovg.render(sc.width, sc.height, sc.words())
        band, bsig, st2 = ovg.render_bands(sc.width, sc.height, sc.words(), threads)
        assert st == st2 == 0
        assert np.array_equal(full, band) and np.array_equal(fsig, bsig), sc.name
