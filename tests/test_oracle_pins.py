"""Pins of the oracle's flattening, arcs, stroke joins and caps, and focal radial gradients.

Each expectation is hand-derived from the reference's stated rules or computed by an independent
closed form (union of convex polygons in double precision, SVG's centre parametrisation, the
ray-circle intersection solved geometrically), never by retyping the oracle's code. The
`check_*` functions are also run against deliberately changed builds of the oracle in
tests/test_oracle_mutants.py, each of which must fail at least one of them.

Reference: AGG subdivision src/vkvg_context_internal.c:1313-1461; arcs src/vkvg_context.c:394-503,
src/vkvg_context_internal.c:245-252, :1473-1580; joins :924-1163; caps :1165-1239; stroke driver
src/vkvg_context.c:822-948; radial gradients shaders/vkvg_main.frag:112-144, src/vkvg_pattern.c:95-119.
"""
from __future__ import annotations

import math

import numpy as np

from oracle import ovg
from vkvg_amd.stream import Stream, JOIN_MITER, JOIN_ROUND, JOIN_BEVEL, CAP_BUTT, CAP_ROUND

from pin_util import arc_step, steps_below, fan, rect, union_mask, stroke_stream, coverage_counts

PI = math.pi


# ================================================================ AGG flattening
# Level >= 1 nodes: the flatness tests apply to the given control points. Tolerance 0.25 is
# compared with squared distances. Coordinates are dyadic so every midpoint is exact.
#   d2, d3 = |cross(P2 - P4, chord)|, |cross(P3 - P4, chord)| (distance x chord length)
#   both > 1.7: if (d2+d3)^2 <= 0.25 |chord|^2, angles a12, a23, a34 from atan2;
#               da1 + da2 < 0.01 -> P1234;  da1 > 0.01 -> P2;  da2 > 0.01 -> P3
#   only d2 > 1.7 (resp. d3): da < 0.01 -> P2 and P3;  da > 0.01 -> P2 (resp. P3)
#   neither: |P1234 - chord midpoint|^2 <= 0.25 -> P1234
# A da >= pi is folded as da = 2/pi - da (M_2_PIF, where 2*pi was meant).
AGG_NODES = {
    # P1..P4 -> emitted points (the end point P4 always follows)
    "collinear": ([0, 0, 10, 0, 20, 0, 30, 0], [(15, 0)]),
    # chord 1000 long, both handles 0.125 off it: d2 = d3 = 125, (250)^2 <= 0.25e6;
    # da1 = da2 = atan(0.125/300) ~ 4.2e-4 -> sum < 0.01 -> P1234 = (500, 0.09375)
    "flat_small_angles": ([0, 0, 300, 0.125, 700, 0.125, 1000, 0], [(500, 0.09375)]),
    # da1 = atan(0.1875/12) ~ 0.0156 > 0.01 -> P2
    "first_angle": ([0, 0, 12, 0.1875, 988, 0.1875, 1000, 0], [(12, 0.1875)]),
    # da1 ~ 3.8e-4, da2 = atan(0.1875/12) ~ 0.0156 -> P3
    "second_angle": ([0, 0, 500, 0.1875, 988, 0.1875, 1000, 0], [(988, 0.1875)]),
    # heading -x: a12 = pi - 3.75e-4, a23 = -pi + 1.28e-4 -> da1 = 6.28268 >= pi, folded to
    # 2/pi - 6.28268 = -5.646; da2 = 0.01029; sum < 0.01 -> P1234 = (317, 0.1171875).
    # (Folding with 2*pi would give da1 = 5.0e-4, sum 0.0108 >= 0.01, da2 > 0.01 -> P3 = (12, 0.125).)
    "wrapped_angle": ([1000, 0, 500, 0.1875, 12, 0.125, 0, 0], [(317, 0.1171875)]),
    # d3 = 0: da = |atan2(-0.1875, 488) - atan2(0.1875, 12)| ~ 0.0160 > 0.01 -> P2
    "d2_only_sharp": ([0, 0, 12, 0.1875, 500, 0, 1000, 0], [(12, 0.1875)]),
    # d3 = 0: da ~ 1.3e-3 < 0.01 -> P2, P3
    "d2_only_flat": ([0, 0, 500, 0.1875, 700, 0, 1000, 0], [(500, 0.1875), (700, 0)]),
    # d2 = 0: da ~ 0.0160 -> P3
    "d3_only_sharp": ([0, 0, 500, 0, 988, 0.1875, 1000, 0], [(988, 0.1875)]),
    # d2 = 0: da ~ 1.3e-3 -> P2, P3
    "d3_only_flat": ([0, 0, 300, 0, 500, 0.1875, 1000, 0], [(300, 0), (500, 0.1875)]),
    # d2 = d3 = 1.0 <= 1.7; P1234 = (1, 0.375), 0.375^2 = 0.14 <= 0.25 -> P1234
    # (a distance compared with 0.25 unsquared would subdivide)
    "squared_tolerance_emit": ([0, 0, 0, 0.5, 2, 0.5, 2, 0], [(1, 0.375)]),
    # P1234 = (1, 0.5625), 0.5625^2 > 0.25 -> subdivide; both children pass the same test at
    # their own midpoints (0.3125, 0.421875) and (1.6875, 0.421875); the node midpoint is not emitted
    "squared_tolerance_split": ([0, 0, 0, 0.75, 2, 0.75, 2, 0], [(0.3125, 0.421875), (1.6875, 0.421875)]),
}


def check_agg_nodes():
    for name, (p, pts) in AGG_NODES.items():
        got = ovg.flatten_cubic_level(p, 1)
        want = np.array(list(pts) + [(p[6], p[7])], np.float32)
        assert got.shape == want.shape and np.array_equal(got, want), (name, got.tolist(), want.tolist())


def check_agg_level0_subdivides():
    # level 0 always splits: the collinear cubic's halves (0,0)(5,0)(10,0)(15,0) and
    # (15,0)(20,0)(25,0)(30,0) each emit their midpoint
    got = ovg.flatten_cubic([0, 0, 10, 0, 20, 0, 30, 0])
    assert np.array_equal(got, np.array([[7.5, 0], [22.5, 0], [30, 0]], np.float32)), got.tolist()


def _curve_path(ctm_scale):
    s = Stream()
    if ctm_scale != 1:
        s.scale(ctm_scale, ctm_scale)
    s.move_to(0, 0)
    s.curve_to(0, 1, 4, 1, 4, 0)
    s.fill()
    return ovg.geometry(s.words())["points"]


def check_curve_tolerance_follows_ctm_scale():
    # (0,0)(0,1)(4,1)(4,0): level-1 children (0,0)(0,.5)(1,.75)(2,.75) and its mirror; their
    # midpoints (0.625, 0.5625) / (3.375, 0.5625) lie sqrt(0.17578) = 0.419 px from the chord
    # midpoints: emitted with tolerance 0.25 (0.1758 <= 0.25), split under a CTM scale of 2
    # (tolerance 0.25 / 2 = 0.125)
    p1 = _curve_path(1)
    assert np.array_equal(p1, np.array([[0, 0], [0.625, 0.5625], [3.375, 0.5625], [4, 0]], np.float32)), p1.tolist()
    p2 = _curve_path(2)
    assert len(p2) > len(p1), p2.tolist()


def test_agg_nodes_hand_derived():
    check_agg_nodes()


def test_agg_level0_always_subdivides():
    check_agg_level0_subdivides()


def test_curve_tolerance_follows_the_ctm_scale():
    check_curve_tolerance_follows_ctm_scale()


# ================================================================ arcs
def _arc_points(xc, yc, r, a1, a2, negative=False, scale=1.0, prefix=None):
    s = Stream()
    if scale != 1:
        s.scale(scale, scale)
    if prefix:
        prefix(s)
    (s.arc_negative if negative else s.arc)(xc, yc, r, a1, a2)
    s.fill()
    g = ovg.geometry(s.words())
    return g["points"].astype(np.float64), g["pathes"]


# (radius, span, CTM scale, expected number of points): first point, one point per k >= 1 with
# k*step < span, last point. Steps: r=10 -> pi/10; r=100 -> pi/40; r=2 -> pi/3 (pi/2 > pi/3);
# r=20 -> pi/20; r=20 at scale 2 -> 40 px -> pi/16; r=29 -> pi/29; r=31 -> pi/12.4.
ARC_COUNTS = [
    (10.0, PI / 2 + 0.1, 1.0, 7),   # 1.6708 / 0.31416 = 5.32
    (100.0, 1.0, 1.0, 14),          # 1 / 0.07854 = 12.7
    (2.0, 2.5, 1.0, 4),             # 2.5 / 1.0472 = 2.39
    (20.0, 1.0, 1.0, 8),            # 1 / 0.15708 = 6.37
    (20.0, 1.0, 2.0, 7),            # 1 / 0.19635 = 5.09
    (29.0, 1.0, 1.0, 11),           # 1 / 0.10833 = 9.23
    (31.0, 1.0, 1.0, 5),            # 1 / 0.25335 = 3.95
]


def check_arc_counts():
    for r, span, sc, n in ARC_COUNTS:
        assert steps_below(span, arc_step(r, sc)) + 2 == n  # the table agrees with the stated rule
        pts, pathes = _arc_points(50, 40, r, 0.25, 0.25 + span, scale=sc)
        assert len(pts) == n, (r, span, sc, len(pts), n)
        neg, _ = _arc_points(50, 40, r, 0.25 + span, 0.25, negative=True, scale=sc)
        assert len(neg) == n, ("negative", r, span, sc, len(neg), n)


def check_arc_points_on_circle():
    for r, span, sc, n in ARC_COUNTS:
        a1 = 0.25
        step = arc_step(r, sc)
        for negative in (False, True):
            start, end = (a1 + span, a1) if negative else (a1, a1 + span)
            pts, pathes = _arc_points(50, 40, r, start, end, negative=negative, scale=sc)
            d = pts - [50, 40]
            assert np.abs(np.hypot(d[:, 0], d[:, 1]) - r).max() <= 2e-5 * max(r, 50), (r, negative)
            ang = np.unwrap(np.arctan2(d[:, 1], d[:, 0]))
            sgn = -1 if negative else 1
            expect = np.array([start] + [start + sgn * k * step for k in range(1, n - 1)] + [end])
            assert np.abs(ang - expect).max() < 2e-5 * n, (r, negative, (ang - expect).tolist())
            # first and last points are the exact float evaluations at the given angles
            for q, a in ((pts[0], start), (pts[-1], end)):
                fx = np.float32(np.float32(math.cos(np.float32(a))) * np.float32(r)) + np.float32(50)
                fy = np.float32(np.float32(math.sin(np.float32(a))) * np.float32(r)) + np.float32(40)
                assert (np.float32(q[0]), np.float32(q[1])) == (fx, fy), (r, negative, q, fx, fy)
            # one curve segment record holding every point, an open path
            assert pathes[0] & ovg.P_MASK == n and not pathes[0] & ovg.P_CLOSED
            assert pathes[0] & ovg.P_CURVES and pathes[1] & ovg.P_CURVES and pathes[1] & ovg.P_MASK == n


def check_full_circle_closes():
    # r = 10.3: step pi/10.3, 2*pi/step = 20.6 -> 20 stepped points after the first, no end point
    pts, pathes = _arc_points(60, 60, 10.3, 0, 2 * PI)
    assert len(pts) == 21
    assert pathes[0] & ovg.P_CLOSED and pathes[0] & ovg.P_CURVES


def check_first_arc_convex_flag():
    _, pathes = _arc_points(60, 60, 10.0, 0, 1.0)
    assert pathes[0] & ovg.P_CONVEX  # an arc that starts the path marks it simple-convex
    _, pathes = _arc_points(60, 60, 10.0, 0, 1.0, prefix=lambda s: s.move_to(0, 0))
    assert not pathes[0] & ovg.P_CONVEX


def check_arc_joins_current_point():
    # an arc after a current point adds a straight line to its start point
    pts, pathes = _arc_points(60, 60, 10.0, 0, 1.0, prefix=lambda s: s.move_to(0, 0))
    assert tuple(pts[0]) == (0, 0) and tuple(pts[1]) == (70, 60)
    assert len(pts) == 1 + steps_below(1.0, arc_step(10.0)) + 2


def _svg_center(x1, y1, x2, y2, fa, fs, rx, ry, phi):
    """Centre of an SVG elliptical arc (SVG 1.1, implementation notes F.6.5), in double."""
    c, s = math.cos(phi), math.sin(phi)
    xp = c * (x1 - x2) / 2 + s * (y1 - y2) / 2
    yp = -s * (x1 - x2) / 2 + c * (y1 - y2) / 2
    lam = xp * xp / (rx * rx) + yp * yp / (ry * ry)
    if lam > 1:
        rx, ry = rx * math.sqrt(lam), ry * math.sqrt(lam)
    num = rx * rx * ry * ry - rx * rx * yp * yp - ry * ry * xp * xp
    den = rx * rx * yp * yp + ry * ry * xp * xp
    k = math.sqrt(max(0.0, num / den)) * (-1 if fa == fs else 1)
    cxp, cyp = k * rx * yp / ry, -k * ry * xp / rx
    return c * cxp - s * cyp + (x1 + x2) / 2, s * cxp + c * cyp + (y1 + y2) / 2, rx, ry


def check_elliptic_arc_on_svg_ellipse():
    for (x1, y1, x2, y2, fa, fs, rx, ry, phi) in [
            (10, 50, 90, 30, 0, 1, 50, 30, 0.3), (10, 50, 90, 30, 1, 1, 50, 30, 0.3),
            (10, 50, 90, 30, 0, 0, 50, 30, 0.3), (10, 50, 90, 30, 1, 0, 50, 30, -0.7),
            (20, 20, 60, 60, 0, 1, 10, 10, 0.0)]:  # radii too small: scaled up (lambda > 1)
        s = Stream()
        s.move_to(x1, y1)
        s.elliptic_arc_to(x2, y2, fa, fs, rx, ry, phi)
        s.stroke()
        pts = ovg.geometry(s.words())["points"].astype(np.float64)
        cx, cy, erx, ery = _svg_center(x1, y1, x2, y2, fa, fs, rx, ry, phi)
        c, sn = math.cos(phi), math.sin(phi)
        dx, dy = pts[:, 0] - cx, pts[:, 1] - cy
        u = (c * dx + sn * dy) / erx
        v = (-sn * dx + c * dy) / ery
        assert np.abs(u * u + v * v - 1).max() < 2e-3, (x1, fa, fs, np.abs(u * u + v * v - 1).max())
        assert tuple(pts[0]) == (x1, y1) and np.hypot(*(pts[-1] - [x2, y2])) < 1e-3
        # parametric angle advances in the sweep direction with a constant step (points[1] is the
        # arc's own start point, within float rounding of the current point and so not merged)
        th = np.unwrap(np.arctan2(v, u))
        dth = np.diff(th[1:-1])
        step = max(0.001, min(PI, arc_step(min(erx, ery)) * 0.1))
        assert np.all(dth * (1 if fs else -1) > 0)
        assert np.abs(np.abs(dth[1:]) - step).max() < 1e-3, (x1, fa, fs)


def test_arc_point_counts_follow_the_arc_step():
    check_arc_counts()


def test_arc_points_lie_on_the_circle_at_stepped_angles():
    check_arc_points_on_circle()


def test_full_circle_closes_the_path():
    check_full_circle_closes()


def test_first_arc_marks_the_path_convex():
    check_first_arc_convex_flag()


def test_arc_after_a_current_point_adds_a_line():
    check_arc_joins_current_point()


def test_elliptic_arc_points_lie_on_the_svg_ellipse():
    check_elliptic_arc_on_svg_ellipse()


def test_rectangle_is_a_closed_convex_quad():
    s = Stream()
    s.rectangle(3, 4, 10, 20)
    s.fill()
    g = ovg.geometry(s.words())
    assert g["points"].tolist() == [[3, 4], [13, 4], [13, 24], [3, 24]]
    assert g["pathes"].tolist() == [4 | ovg.P_CLOSED | ovg.P_CONVEX]


# ================================================================ joins
HW = 5.2  # line width 10.4: offset lines at x.8 / x.2, away from pixel centres


def _right_angle(join, turn, width=2 * HW, miter=None, size=(72, 64)):
    """Open polyline (10,y0) -> (50,y0) -> (50,y1), butt caps. turn=+1: down (right turn on
    screen, outer corner up-right); turn=-1: up (outer corner down-right)."""
    y0 = 10 if turn > 0 else 54
    y1 = 54 if turn > 0 else 10
    return stroke_stream([[10, y0], [50, y0], [50, y1]], width, join=join, miter=miter), y0, y1


def _right_angle_polys(join, turn, hw=HW, miter_cut=None):
    y0 = 10 if turn > 0 else 54
    y1 = 54 if turn > 0 else 10
    base = [rect(10, y0 - hw, 50, y0 + hw), rect(50 - hw, min(y0, y1), 50 + hw, max(y0, y1))]
    oy = y0 - turn * hw  # outer offset line of the first segment
    if join == JOIN_MITER:
        corner = [rect(50, min(y0, oy), 50 + hw, max(y0, oy))]
        if miter_cut is not None:  # square cut by the line at distance miter_cut along the bisector
            t = miter_cut * math.sqrt(2) - hw  # cut meets the outer offset lines hw + t from p0
            corner = [[(50, y0), (50, oy), (50 + t, oy), (50 + hw, y0 - turn * t), (50 + hw, y0)]]
    elif join == JOIN_BEVEL:
        corner = [[(50, oy), (50 + hw, y0), (50, y0)]]
    else:
        a0 = -PI / 2 if turn > 0 else PI / 2
        corner = [fan((50, y0), hw, a0, 0.0, arc_step(hw))]
    return base + corner


def check_right_angle_joins():
    for join in (JOIN_MITER, JOIN_ROUND, JOIN_BEVEL):
        for turn in (1, -1):
            words, _, _ = _right_angle(join, turn)
            k, stats = coverage_counts(words, 72, 64)
            want = union_mask(_right_angle_polys(join, turn), 72, 64)
            assert np.array_equal(k > 0, want), (join, turn, np.argwhere((k > 0) != want)[:8].tolist())
            assert k.max() == 1, (join, turn)  # joins and segment quads do not overlap


def check_round_join_vertex_count():
    # round join fan: one triangle per stepped arc vertex plus one; total = 2 quads + fan
    for hw, n_arc in ((HW, 2), (37.5, 7)):  # steps pi/5.2 (1.5708/0.6042 = 2.6), pi/15 (7.5)
        assert steps_below(PI / 2, arc_step(hw)) == n_arc
        words = stroke_stream([[100, 100], [300, 100], [300, 300]], 2 * hw, join=JOIN_ROUND)
        g = ovg.geometry(words, 400, 400)
        assert len(g["tris"]) == 2 + 2 + n_arc + 1, (hw, len(g["tris"]))
        # the stepped vertices lie on the circle of radius hw around the corner
        v = g["verts"].astype(np.float64)
        r = np.hypot(v[:, 0] - 300, v[:, 1] - 100)
        assert np.sum(np.abs(r - hw) < 1e-4) >= n_arc + 2


def check_large_round_join_mask():
    hw, x, y = 37.5, 150.25, 50.25  # offset lines at .75 / .25, away from pixel centres
    words = stroke_stream([[20, y], [x, y], [x, 180]], 2 * hw, join=JOIN_ROUND)
    k, _ = coverage_counts(words, 200, 220)
    polys = [rect(20, y - hw, x, y + hw), rect(x - hw, y, x + hw, 180),
             fan((x, y), hw, -PI / 2, 0.0, arc_step(hw))]
    want = union_mask(polys, 200, 220, margin=0.005)
    assert np.array_equal(k > 0, want), np.argwhere((k > 0) != want)[:8].tolist()


def check_miter_limit_switch():
    # right angle: lh = hw / cos(pi/4) = 0.7071 * width; the miter is cut when lh > limit * width
    # (lhMax = miterLimit * lineWidth): 2 triangles for a full miter, 3 for the cut one
    def tris(limit):
        words, _, _ = _right_angle(JOIN_MITER, 1, miter=limit)
        return len(ovg.geometry(words)["tris"])
    assert tris(0.72) == 4 and tris(10.0) == 4
    assert tris(0.70) == 5 and tris(0.5) == 5
    # the cut shape: lw 40.4, limit 0.6. The cut is centred lhMax = 0.6*40.4 = 24.24 px from the
    # corner along the outward bisector o = (1, -1)/sqrt(2); its two vertices lie
    # (lh - lhMax) * cos(alpha/2) = (28.567 - 24.24) / sqrt(2) = 3.06 px either side of that centre
    # along the bisector of the directions b = (1, 1)/sqrt(2) -- short of the offset lines (the
    # reference's cut is not flush with them). The segment quads run from the caps to those
    # vertices and to the inner corner p0 - lh * o = (79.8, 60.2).
    hw = 20.2
    words = stroke_stream([[10, 40], [100, 40], [100, 140]], 2 * hw, join=JOIN_MITER, miter=0.6)
    k, _ = coverage_counts(words, 150, 150)
    r2 = math.sqrt(0.5)
    lh, lhmax = hw / r2, 0.6 * 2 * hw
    cx, cy = 100 + lhmax * r2, 40 - lhmax * r2
    x = (lh - lhmax) * r2
    e1, e2 = (cx - x * r2, cy - x * r2), (cx + x * r2, cy + x * r2)
    inner = (100 - hw, 40 + hw)
    polys = [[(10, 40 - hw), e1, inner, (10, 40 + hw)], [e1, e2, inner], [e2, (100 + hw, 140), (100 - hw, 140), inner]]
    want = union_mask(polys, 150, 150, margin=2e-4, snap=True)
    assert np.array_equal(k > 0, want), np.argwhere((k > 0) != want)[:8].tolist()
    assert k.max() == 1


def _arc_stroke(join, r, width, a1=0.3, a2=2.9):
    def ops(s):
        s.arc(100, 100, r, a1, a2)
    return stroke_stream(None, width, join=join, curve_ops=ops)


def check_curve_internal_joins():
    # inside a curve the join is ROUND where dot < 0.8 and MITER otherwise, whatever the user's
    # join: the arc's stroke equals the stroke of the same points as straight lines with that join
    for r, width, forced in ((3.0, 20.0, JOIN_ROUND),    # step pi/3: dot 0.5; round fans of pi/10 steps
                             (100.0, 8.0, JOIN_MITER)):  # step pi/40: dot 0.997
        pts = ovg.geometry(_arc_stroke(JOIN_MITER, r, width))["points"]
        poly = {j: ovg.geometry(stroke_stream(pts, width, join=j)) for j in (0, 1, 2)}
        assert len(poly[forced]["tris"]) != len(poly[JOIN_BEVEL if forced == JOIN_ROUND else JOIN_ROUND]["tris"])
        for user in (JOIN_MITER, JOIN_ROUND, JOIN_BEVEL):
            g = ovg.geometry(_arc_stroke(user, r, width))
            assert np.array_equal(g["tris"], poly[forced]["tris"]), (r, user)
            assert np.array_equal(g["verts"], poly[forced]["verts"]), (r, user)


def test_right_angle_joins_are_the_closed_form_regions():
    check_right_angle_joins()


def test_round_join_vertex_count_follows_the_arc_step():
    check_round_join_vertex_count()


def test_large_round_join_region():
    check_large_round_join_mask()


def test_miter_limit_switches_at_limit_times_width():
    check_miter_limit_switch()


def test_curve_internal_joins_ignore_the_user_join():
    check_curve_internal_joins()


# ================================================================ caps
def check_round_caps():
    # segment (x0,y0)->(x1,y1), width 2*hw: rectangle plus two half-disc fans behind the ends
    # rightwards, down-right, up-left (the n.y < 0 branch of the start angle), down-left
    for (x0, y0, x1, y1), hw in (((40.25, 50.25, 160.25, 50.25), 20.2), ((30.3, 20.1, 110.2, 80.4), 12.35),
                                 ((120.1, 90.3, 20.2, 30.1), 7.7), ((150.3, 20.2, 60.1, 95.3), 10.6)):
        words = stroke_stream([[x0, y0], [x1, y1]], 2 * hw, cap=CAP_ROUND)
        k, _ = coverage_counts(words, 200, 110)
        d = math.atan2(y1 - y0, x1 - x0)
        nx, ny = math.cos(d), math.sin(d)
        px, py = -ny * hw, nx * hw
        body = [(x0 + px, y0 + py), (x1 + px, y1 + py), (x1 - px, y1 - py), (x0 - px, y0 - py)]
        st = arc_step(hw)
        start = fan((x0, y0), hw, d + PI / 2, d + 3 * PI / 2, st)
        end = fan((x1, y1), hw, d + PI / 2, d - PI / 2, st)
        want = union_mask([body, start, end], 200, 110, margin=1e-3, snap=True)
        assert np.array_equal(k > 0, want), ((x0, y0), np.argwhere((k > 0) != want)[:8].tolist())
        assert k.max() == 1


def test_round_caps_are_half_disc_fans_behind_the_end_points():
    check_round_caps()


# ================================================================ focal radial gradients
def _radial_fp64(W, H, c0, r0, c1, r1, stops):
    """The radial gradient as the fragment shader defines it, in double: coordinates normalised by
    the surface size (radii by the width); the gradient length along the ray from the focus c0
    through p is where that ray leaves circle 1, found as proj + sqrt(r1^2 - perp^2) with proj and
    perp the components of c1 - c0 along and across the ray; t = (|p - c0| - r0) / (length - r0);
    colours chained with smoothstep between consecutive stops. Circles are first made nested as
    vkvg_pattern_edit_radial requires (r0 <= r1 - 1, c0 moved to within r1 - r0 - 1 of c1)."""
    c0 = np.array(c0, float)
    c1 = np.array(c1, float)
    r0 = min(r0, r1 - 1)
    u = c0 - c1
    L = math.hypot(*u)
    if L + r0 + 1 >= r1:
        c0 = c1 + u / L * (r1 - r0 - 1)
    Y, X = np.mgrid[0:H, 0:W] + 0.5
    n = np.array([W, H], float)
    C0, C1, R0, R1 = c0 / n, c1 / n, r0 / W, r1 / W
    dx, dy = X / W - C0[0], Y / H - C0[1]
    dist = np.hypot(dx, dy)
    ux, uy = dx / dist, dy / dist
    e = C1 - C0
    proj = ux * e[0] + uy * e[1]
    length = proj + np.sqrt(R1 * R1 - ((e @ e) - proj * proj))
    t = (dist - R0) / (length - R0)
    st = np.asarray(stops, float)
    cols = np.concatenate([st[:, 1:4] * st[:, 4:5], st[:, 4:5]], 1)
    col = np.broadcast_to(cols[0], (H, W, 4)).copy()
    for i in range(1, len(st)):
        s = np.clip((t - st[i - 1, 0]) / (st[i, 0] - st[i - 1, 0]), 0, 1)
        s = s * s * (3 - 2 * s)
        col = col * (1 - s[..., None]) + cols[i] * s[..., None]
    return np.rint(np.clip(col, 0, 1) * 255).astype(np.int64)


RADIAL_CASES = [
    # (W, H), c0, r0, c1, r1, stops
    ((96, 64), (30, 20), 4, (48, 32), 40, [[0, 1, 0, 0, 1], [1, 0, 0, 1, 1]]),           # focal, W != H
    ((96, 64), (30, 20), 4, (48, 32), 40, [[0, 1, 0, 0, 1], [0.4, 0, 1, 0, 0.6], [1, 0, 0, 1, 1]]),
    ((128, 128), (10, 10), 5, (64, 64), 30, [[0, 1, 0, 0, 1], [1, 0, 0, 1, 1]]),         # c0 outside: pulled in
    ((80, 120), (40, 60), 50, (44, 60), 30, [[0, 1, 0.5, 0, 1], [1, 0, 0, 1, 0.5]]),     # r0 > r1 - 1: clamped
    ((64, 64), (20, 44), 0, (32, 32), 30, [[0.2, 0, 0, 1, 1], [0.5, 1, 1, 0, 1], [0.9, 0, 1, 0, 0.7]]),
]


def _render_radial(W, H, c0, r0, c1, r1, stops):
    s = Stream()
    s.set_radial(c0[0], c0[1], r0, c1[0], c1[1], r1, stops)
    s.rectangle(0, 0, W, H)
    s.fill()
    rgba, _, _, st = ovg.render(W, H, s.words())
    assert st == 0
    return rgba.astype(np.int64)


def check_focal_radial():
    for case in RADIAL_CASES:
        (W, H) = case[0]
        got = _render_radial(W, H, *case[1:])
        want = _radial_fp64(W, H, *case[1:])
        err = np.abs(got - want)
        # float32 evaluation against double: at most 1 LSB, on a small fraction of pixels
        assert err.max() <= 1 and np.count_nonzero(err) <= 0.001 * err.size, (case, err.max(), np.count_nonzero(err))


def test_focal_radial_gradient_against_fp64_geometry():
    check_focal_radial()


# ================================================================ dashes
# src/vkvg_context.c:853-867, :898-916 and src/vkvg_context_internal.c:1240-1264: the dash state
# (on/off, index, offset within the pattern) runs along the whole sub-path; the offset is taken
# modulo the pattern's total; a cap closes every dash.
def _dash_runs(pts, dashes, offset, width=2.0, size=(120, 70), axis_row=None, axis_col=None):
    s = Stream()
    s.set_source_rgba(0, 0, 0, 1)
    s.set_line_width(width)
    s.set_dash(dashes, offset)
    s.polyline(np.asarray(pts, np.float32))
    s.stroke()
    rgba, _, _, st = ovg.render(size[0], size[1], s.words())
    assert st == 0
    line = rgba[axis_row, :, 3] > 0 if axis_row is not None else rgba[:, axis_col, 3] > 0
    edges = np.flatnonzero(np.diff(np.concatenate([[0], line.astype(int), [0]])))
    return [(int(a), int(b)) for a, b in zip(edges[0::2], edges[1::2])]  # [start, end) pixel runs


def check_dash_offset_and_wrap():
    # The walk starts "on" with the offset as its distance to the next toggle, and every toggle
    # alternates start and end caps: the first dash starts at the offset (the offset delays the
    # pattern; cairo's offset advances into it instead). Horizontal line x 0.5 .. 100.5 at y 5.5,
    # pattern [8, 4], offset 3: dashes [3.5, 11.5), [15.5, 23.5), ... cover pixel columns 3..10,
    # 15..22, ...
    runs = _dash_runs([[0.5, 5.5], [100.5, 5.5]], [8, 4], 3.0, axis_row=5)
    assert runs[:4] == [(3, 11), (15, 23), (27, 35), (39, 47)], runs
    # an offset beyond the pattern is taken modulo its total (15 = 12 + 3)
    assert _dash_runs([[0.5, 5.5], [100.5, 5.5]], [8, 4], 15.0, axis_row=5) == runs
    # an odd number of entries: toggles at 0, 5, 7, 10, 15, 17, 20, ... alternate start / end, so
    # the roles of the entries swap on every repetition: dashes [0, 5), [7, 10), [15, 17), [20, 25)
    runs3 = _dash_runs([[0.5, 5.5], [100.5, 5.5]], [5, 2, 3], 0.0, axis_row=5)
    assert runs3[:4] == [(0, 5), (7, 10), (15, 17), (20, 25)], runs3


def check_dash_continues_around_a_corner():
    # (0.5, 5.5) -> (30.5, 5.5) -> (30.5, 60.5), pattern [8, 4]: the horizontal leg holds on
    # [0, 8), [12, 20), [24, 30); the dash started at 24 runs 2 more px down the vertical leg
    # (to y 7.5), then off 4, on 8 from y 11.5
    runs = _dash_runs([[0.5, 5.5], [30.5, 5.5], [30.5, 60.5]], [8, 4], 0.0, axis_col=30, size=(60, 70))
    assert runs[1:3] == [(11, 19), (23, 31)], runs


def check_invalid_dashes():
    for d in ([0, 0], [-1, 4]):
        s = Stream()
        s.set_dash(d, 0.0)
        s.polyline(np.array([[0, 0], [10, 10]], np.float32))
        s.stroke()
        assert ovg.render(16, 16, s.words())[3] == 13  # VKVG_STATUS_INVALID_DASH


def test_dash_offset_and_pattern_wrap():
    check_dash_offset_and_wrap()


def test_dash_continues_around_a_corner():
    check_dash_continues_around_a_corner()


def test_invalid_dash_patterns():
    check_invalid_dashes()
