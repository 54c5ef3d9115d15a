"""Seeded synthetic scenes for the five benchmark configurations.

Random numbers come from the PCG64 bit generator (O'Neill 2014) through its raw 64-bit
output stream, which is stable across numpy versions; every conversion to floats is done
here. Scene k uses seed 0x5EED0000 + 16*k + variant.

Frozen choices not fixed by the configuration text are marked (*) in the docstrings.
Every scene starts with set_source_rgb(1,1,1); paint() onto a transparent surface.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

from .stream import (CAP_BUTT, CAP_ROUND, CAP_SQUARE, FILL_EVEN_ODD, FILL_NON_ZERO, JOIN_BEVEL, JOIN_MITER,
                     JOIN_ROUND, Stream)


class Rng:
    """Uniform / log-uniform / normal draws from the PCG64 raw stream."""

    def __init__(self, seed: int) -> None:
        self._bg = np.random.PCG64(seed)

    def u(self, n: int | None = None) -> np.ndarray | float:
        k = 1 if n is None else n
        raw = self._bg.random_raw(k).astype(np.uint64)
        v = (raw >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)  # [0,1) with 53 bits
        return float(v[0]) if n is None else v

    def uniform(self, lo, hi, n=None):
        return lo + (hi - lo) * self.u(n)

    def loguniform(self, lo, hi, n=None):
        return np.exp(math.log(lo) + (math.log(hi) - math.log(lo)) * self.u(n)) if n is not None else \
            math.exp(math.log(lo) + (math.log(hi) - math.log(lo)) * self.u())

    def randint(self, lo, hi, n=None):
        """integer uniform in [lo, hi] inclusive"""
        v = np.floor(lo + (hi - lo + 1) * self.u(n)).astype(np.int64) if n is not None else \
            int(math.floor(lo + (hi - lo + 1) * self.u()))
        return np.minimum(v, hi) if n is not None else min(v, hi)

    def normal(self, n: int) -> np.ndarray:
        u1 = self.u(n)
        u2 = self.u(n)
        return np.sqrt(-2.0 * np.log1p(-u1)) * np.cos(2.0 * math.pi * u2)


@dataclass
class Scene:
    name: str
    width: int
    height: int
    stream: Stream
    n_paths: int
    ctrl_points: int = 0
    gradients: int = 0
    gradient_stops: int = 0
    meta: dict = field(default_factory=dict)

    def words(self) -> np.ndarray:
        return self.stream.words()

    def algorithmic_bytes(self, fp32_out: bool = False) -> int:
        """8*(control points) + 32*paths + (24+20*stops)*gradients + 4*W*H (16*W*H for FP32 out)."""
        px = self.width * self.height
        return (8 * self.ctrl_points + 32 * self.n_paths + 24 * self.gradients + 20 * self.gradient_stops
                + (16 if fp32_out else 4) * px)


def _white(s: Stream) -> None:
    s.set_source_rgb(1.0, 1.0, 1.0)
    s.paint()


def config1(variant: int = 0, n_paths: int = 200, size: int = 512, white: bool = True) -> Scene:
    """512x512: closed paths of n~U{4..12} cubics, control points ~U[0,size)^2, non-zero fill,
    rgb~U[0,1], alpha~U[0.5,1]. `white=False` leaves the surface transparent underneath."""
    rng = Rng(0x5EED0000 + 16 * 1 + variant)
    s = Stream()
    if white:
        _white(s)
    cps = 0
    for _ in range(n_paths):
        n = rng.randint(4, 12)
        pts = rng.uniform(0, size, 2 + 6 * n).astype(np.float32)
        col = rng.u(3)
        a = rng.uniform(0.5, 1.0)
        s.set_source_rgba(col[0], col[1], col[2], a)
        s.curves(pts[:2], pts[2:].reshape(n, 6), close=True)
        s.fill()
        cps += 1 + 3 * n
    return Scene("c1_512_nz_cubics", size, size, s, n_paths, ctrl_points=cps)


def _stroke_style(rng: Rng, s: Stream) -> None:
    s.set_line_width(float(rng.loguniform(0.5, 20.0)))
    s.set_line_join([JOIN_MITER, JOIN_ROUND, JOIN_BEVEL][rng.randint(0, 2)])
    s.set_line_cap([CAP_BUTT, CAP_ROUND, CAP_SQUARE][rng.randint(0, 2)])
    if rng.u() < 0.25:
        s.set_dash([8.0, 4.0], 0.0)
    else:
        s.set_dash([], 0.0)


def _c2_content(rng: Rng, s: Stream, W: int, H: int, n_lines: int, n_stars: int) -> tuple[int, int]:
    """(a) polylines of m~U{2..64} points ~U over the surface, width log-U[0.5,20], join/cap
    uniform, miter limit 10, dash [8,4] on 25%, alpha~U[0.5,1]; (b) even-odd stars with
    n~U{5..31}, r log-U[10,200] (*), centre ~U, vertex k at angle 2*pi*k*s/n, s=floor((n-1)/2) (*)."""
    cps = 0
    s.set_miter_limit(10.0)
    s.set_fill_rule(FILL_NON_ZERO)
    for _ in range(n_lines):
        m = rng.randint(2, 64)
        xy = np.stack([rng.uniform(0, W, m), rng.uniform(0, H, m)], 1).astype(np.float32)
        col = rng.u(3)
        s.set_source_rgba(col[0], col[1], col[2], rng.uniform(0.5, 1.0))
        _stroke_style(rng, s)
        s.polyline(xy)
        s.stroke()
        cps += m
    s.set_dash([], 0.0)
    s.set_fill_rule(FILL_EVEN_ODD)
    for _ in range(n_stars):
        n = rng.randint(5, 31)
        r = rng.loguniform(10.0, 200.0)
        cx, cy = rng.uniform(0, W), rng.uniform(0, H)
        step = (n - 1) // 2
        k = np.arange(n)
        ang = 2.0 * math.pi * k * step / n
        xy = np.stack([cx + r * np.cos(ang), cy + r * np.sin(ang)], 1).astype(np.float32)
        col = rng.u(3)
        s.set_source_rgba(col[0], col[1], col[2], rng.uniform(0.5, 1.0))
        s.polyline(xy, close=True)
        s.fill()
        cps += n
    s.set_fill_rule(FILL_NON_ZERO)
    return cps, n_lines + n_stars


def config2(variant: int = 0, n_lines: int = 5000, n_stars: int = 2000, width: int = 1920,
            height: int = 1080) -> Scene:
    rng = Rng(0x5EED0000 + 16 * 2 + variant)
    s = Stream()
    _white(s)
    cps, n = _c2_content(rng, s, width, height, n_lines, n_stars)
    return Scene("c2_1080p_stroke_eo", width, height, s, n, ctrl_points=cps)


def config3(variant: int = 0, n_paths: int = 20000, size: int = 4096, blobs: int = 16,
            sigma: float = 300.0, scale_lo: float = 8.0, scale_hi: float = 256.0, white: bool = True) -> Scene:
    """4096x4096 dense scene: paths of 1..40 cubics around centres from a 16-blob Gaussian mixture
    (means ~U, sigma 300, clamped to the surface (*)); control points = centre + N(0, s^2),
    s log-U[8,256] (*). 60% non-zero fill (closed), 40% open stroke with width U[0.5,8] and
    join/cap uniform (*). 30% of sources are gradients, half linear / half radial (*), with
    2..4 stops, offsets sorted U with ends at 0 and 1 (*). alpha~U[0.2,1]."""
    rng = Rng(0x5EED0000 + 16 * 3 + variant)
    k = size / 4096.0
    means = rng.uniform(0, size, 2 * blobs).reshape(blobs, 2)
    s = Stream()
    if white:
        _white(s)
    s.set_miter_limit(10.0)
    cps = grads = stops_total = 0
    for _ in range(n_paths):
        b = rng.randint(0, blobs - 1)
        c = means[b] + sigma * k * rng.normal(2)
        c = np.clip(c, 0, size - 1)
        sc = rng.loguniform(scale_lo * k, scale_hi * k)
        n = rng.randint(1, 40)
        pts = (np.repeat(c[None, :], 1 + 3 * n, 0).reshape(-1) + sc * rng.normal(2 + 6 * n)).astype(np.float32)
        alpha = rng.uniform(0.2, 1.0)
        if rng.u() < 0.3:
            ns = rng.randint(2, 4)
            offs = np.sort(rng.u(ns))
            offs[0], offs[-1] = 0.0, 1.0
            st = np.zeros((ns, 5), np.float32)
            st[:, 0] = offs
            st[:, 1:4] = rng.u(3 * ns).reshape(ns, 3)
            st[:, 4] = alpha
            if rng.u() < 0.5:
                a0 = c + sc * rng.normal(2)
                a1 = c + sc * rng.normal(2)
                s.set_linear(a0[0], a0[1], a1[0], a1[1], st)
            else:
                r1 = sc * rng.uniform(1.0, 3.0)
                r0 = r1 * rng.uniform(0.0, 0.5)
                f = c + 0.25 * sc * rng.normal(2)
                s.set_radial(f[0], f[1], r0, c[0], c[1], r1, st)
            grads += 1
            stops_total += ns
        else:
            col = rng.u(3)
            s.set_source_rgba(col[0], col[1], col[2], alpha)
        if rng.u() < 0.6:
            s.curves(pts[:2], pts[2:].reshape(n, 6), close=True)
            s.fill()
        else:
            s.set_line_width(float(rng.uniform(0.5, 8.0)))
            s.set_line_join([JOIN_MITER, JOIN_ROUND, JOIN_BEVEL][rng.randint(0, 2)])
            s.set_line_cap([CAP_BUTT, CAP_ROUND, CAP_SQUARE][rng.randint(0, 2)])
            s.curves(pts[:2], pts[2:].reshape(n, 6), close=False)
            s.stroke()
        cps += 1 + 3 * n
    return Scene("c3_4k_dense", size, size, s, n_paths, ctrl_points=cps, gradients=grads,
                 gradient_stops=stops_total)


def config4(variant: int = 0, n_surfaces: int = 256, side_lo: int = 64, side_hi: int = 2048,
            paths_lo: int = 10, paths_hi: int = 5000) -> list[Scene]:
    """256 surfaces with side log-U[64,2048] (square (*)), each holding log-U{10..5000} paths
    drawn as in config 2, polylines:stars = 5:2 (*)."""
    rng = Rng(0x5EED0000 + 16 * 4 + variant)
    out = []
    for i in range(n_surfaces):
        side = int(round(rng.loguniform(side_lo, side_hi)))
        npaths = int(round(rng.loguniform(paths_lo, paths_hi)))
        n_lines = max(1, (npaths * 5) // 7)
        n_stars = max(0, npaths - n_lines)
        s = Stream()
        _white(s)
        cps, n = _c2_content(rng, s, side, side, n_lines, n_stars)
        out.append(Scene(f"c4_{i}", side, side, s, n, ctrl_points=cps))
    return out


def feature_mix(variant: int = 0, size: int = 384, n: int = 60) -> Scene:
    """Exercises the API surface beyond the benchmark configs: closed and open strokes with every
    join / cap, dash offsets, arcs, ellipses, rounded rectangles, quadratic and relative curves,
    transforms, opacity, the CLEAR and DIFFERENCE operators, fill_preserve + stroke, paint with a
    pending path, clear."""
    rng = Rng(0x5EED0000 + 16 * 9 + variant)
    s = Stream()
    _white(s)
    cps = 0
    for i in range(n):
        kind = i % 12
        col = rng.u(3)
        s.set_source_rgba(col[0], col[1], col[2], rng.uniform(0.3, 1.0))
        s.set_line_width(float(rng.loguniform(0.7, 14.0)))
        s.set_line_join(int(rng.randint(0, 2)))
        s.set_line_cap(int(rng.randint(0, 2)))
        s.set_miter_limit(float(rng.uniform(1.5, 12.0)))
        cx, cy = rng.uniform(0, size), rng.uniform(0, size)
        r = rng.loguniform(5, size / 3)
        if kind == 0:  # closed polygon stroke (closing join + index patch)
            m = rng.randint(3, 9)
            ang = np.sort(rng.uniform(0, 2 * math.pi, m))
            s.polyline(np.stack([cx + r * np.cos(ang), cy + r * np.sin(ang)], 1), close=True)
            s.stroke()
            cps += m
        elif kind == 1:  # dashed closed / open with offset
            s.set_dash([float(rng.uniform(2, 12)), float(rng.uniform(1, 8))], float(rng.uniform(0, 10)))
            m = rng.randint(2, 7)
            s.polyline(np.stack([rng.uniform(0, size, m), rng.uniform(0, size, m)], 1), close=bool(i % 2))
            s.stroke()
            s.set_dash([], 0.0)
            cps += m
        elif kind == 2:  # arc fill + arc_negative stroke
            s.arc(cx, cy, r, float(rng.uniform(0, 6)), float(rng.uniform(0, 6)))
            s.fill()
            s.arc_negative(cx, cy, r * 0.7, float(rng.uniform(0, 6)), float(rng.uniform(0, 6)))
            s.stroke()
        elif kind == 3:  # ellipse fill_preserve + stroke
            s.ellipse(r, r * rng.uniform(0.3, 1.0), cx, cy, float(rng.uniform(0, 3)))
            s.fill_preserve()
            s.set_source_rgba(0, 0, 0, 0.8)
            s.stroke()
        elif kind == 4:  # rounded rectangle, even-odd
            s.set_fill_rule(FILL_EVEN_ODD)
            s.rounded_rectangle(cx - r, cy - r / 2, 2 * r, r, r * 0.3)
            s.rectangle(cx - r / 2, cy - r / 4, r, r / 2)
            s.fill()
            s.rounded_rectangle2(cx - r / 3, cy - r / 3, r, r * 0.8, r * 0.2, r * 0.1)
            s.stroke()
            s.set_fill_rule(FILL_NON_ZERO)
        elif kind == 5:  # quadratic + relative curves, open stroke
            s.move_to(cx, cy)
            s.quadratic_to(cx + r, cy - r, cx + 2 * r, cy)
            s.rel_curve_to(r, r, -r, r, 0, 2 * r)
            s.rel_line_to(-r, 0)
            s.stroke()
            cps += 8
        elif kind == 6:  # transformed gradient fill
            s.translate(cx, cy)
            s.rotate(float(rng.uniform(0, 3)))
            s.scale(float(rng.uniform(0.5, 2)), float(rng.uniform(0.5, 2)))
            s.set_linear(-r, 0, r, 0, [[0, 1, 0, 0, 0.9], [0.5, 0, 1, 0, 0.6], [1, 0, 0, 1, 0.9]])
            s.rectangle(-r / 2, -r / 2, r, r)
            s.fill()
            s.identity_matrix()
        elif kind == 7:  # radial gradient stroke with opacity
            s.set_opacity(float(rng.uniform(0.3, 0.9)))
            s.set_radial(cx, cy, 0, cx, cy, r, [[0, 1, 1, 0, 1], [1, 0.2, 0, 0.6, 0.5]])
            s.arc(cx, cy, r * 0.8, 0, 6.2831853)
            s.stroke()
            s.set_opacity(1.0)
        elif kind == 8:  # DIFFERENCE then CLEAR operator on small shapes
            s.set_operator(3)
            s.rectangle(cx, cy, r / 2, r / 3)
            s.fill()
            s.set_operator(0)
            s.rectangle(cx + 2, cy + 2, r / 5, r / 5)
            s.fill()
            s.set_operator(2)
        elif kind == 9:  # self-intersecting cubic loop, non-zero
            pts = rng.uniform(0, size, 2 + 6 * 3).astype(np.float32)
            s.curves(pts[:2], pts[2:].reshape(3, 6), close=True)
            s.fill()
            cps += 10
        elif kind == 10:  # elliptic arcs with curve-forced joins, stroked closed
            s.move_to(cx, cy)
            s.elliptic_arc_to(cx + r, cy + r / 2, bool(i % 2), bool(i % 3), r, r / 2, float(rng.uniform(0, 1)))
            s.line_to(cx, cy + r)
            s.close_path()
            s.stroke()
        else:  # paint with a pending path = fill; new sub paths
            s.move_to(cx, cy)
            s.line_to(cx + r, cy)
            s.line_to(cx + r / 2, cy + r)
            s.new_sub_path()
            s.move_to(cx + r / 3, cy + r / 4)
            s.line_to(cx + r, cy + r)
            s.line_to(cx, cy + r / 2)
            s.paint()
    return Scene("feature_mix", size, size, s, n, ctrl_points=cps)


def clear_scene(size: int = 96) -> Scene:
    s = Stream()
    _white(s)
    s.set_source_rgba(1, 0, 0, 1)
    s.rectangle(10, 10, 30, 30)
    s.fill()
    s.clear()
    s.set_source_rgba(0, 0, 1, 0.5)
    s.rectangle(20, 20, 40, 40)
    s.fill()
    return Scene("clear", size, size, s, 2)


def long_dash_scene(size: int = 640, n: int = 48) -> Scene:
    """Dashed strokes whose pieces (segments between two points) carry from a few to several hundred
    dash boundaries, around the counts where their emission moves from one thread to a wavefront
    and where a wavefront takes a second and third round of 64 caps: per stroke a dash pattern of
    1, 2, 3 or 5 entries (one of them possibly 0), a dash offset, a cap and a join; segment lengths
    chosen as the pattern's mean boundary spacing times U{40..260}; open and closed polylines of
    2-4 points (a closed one ends its last piece in a gap half the time)."""
    rng = Rng(0x5EED0000 + 16 * 12)
    s = Stream()
    _white(s)
    s.set_miter_limit(10.0)
    patterns = ([3.0], [2.0, 1.0], [1.5, 0.0, 2.5], [2.0, 1.0, 0.5, 1.5, 1.0])
    for i in range(n):
        pat = patterns[i % len(patterns)]
        spacing = sum(pat) / len(pat)
        s.set_dash(pat, float(rng.uniform(0.0, 2.0 * sum(pat))))
        s.set_line_width(float(rng.uniform(0.5, 5.0)))
        s.set_line_cap([CAP_BUTT, CAP_ROUND, CAP_SQUARE][i % 3])
        s.set_line_join([JOIN_MITER, JOIN_ROUND, JOIN_BEVEL][(i // 3) % 3])
        s.set_source_rgba(*rng.u(3), float(rng.uniform(0.5, 1.0)))
        m = 2 + i % 3
        x, y = rng.uniform(20, size - 20, 2)
        pts = [(x, y)]
        for _ in range(m - 1):
            length = spacing * rng.randint(40, 260)
            ang = rng.uniform(0.0, 2.0 * math.pi)
            x = min(max(x + length * math.cos(ang), 5.0), size - 5.0)
            y = min(max(y + length * math.sin(ang), 5.0), size - 5.0)
            pts.append((x, y))
        s.polyline(np.array(pts, np.float32), close=(i % 4 == 3))
        s.stroke()
    s.set_dash([], 0.0)
    return Scene("long_dash", size, size, s, n)


def runs_scene(size: int = 256, runs: int = 4, run_len: int = 150) -> Scene:
    """Long runs of draws of one kind between single draws of the other: run_len short strokes
    (no fill edges) then one filled triangle, and run_len small fills (no stroke triangles) then one
    stroke, repeated. A 64-wide run of consecutive fill edges (or stroke triangles) then spans more
    than 64 draws, the case the edge kernels' per-wavefront draw lookup hands to a plain search."""
    rng = Rng(0x5EED0000 + 16 * 7)
    s = Stream()
    _white(s)
    n = 0
    for _ in range(runs):
        s.set_line_width(1.5)
        for _ in range(run_len):
            x, y = rng.uniform(4, size - 4, 2)
            s.set_source_rgba(*rng.u(3), 0.8)
            s.move_to(x, y)
            s.line_to(x + rng.uniform(-6, 6), y + rng.uniform(-6, 6))
            s.stroke()
            n += 1
        x, y = rng.uniform(10, size - 10, 2)
        s.set_source_rgba(*rng.u(3), 0.6)
        s.move_to(x, y)
        s.line_to(x + 9, y + 2)
        s.line_to(x + 3, y + 8)
        s.fill()
        n += 1
        for _ in range(run_len):
            x, y = rng.uniform(4, size - 8, 2)
            s.set_source_rgba(*rng.u(3), 0.7)
            s.rectangle(x, y, rng.uniform(0.5, 4), rng.uniform(0.5, 4))
            s.fill()
            n += 1
        x, y = rng.uniform(10, size - 10, 2)
        s.set_source_rgba(*rng.u(3), 0.9)
        s.move_to(x, y)
        s.line_to(x + 12, y + 5)
        s.stroke()
        n += 1
    return Scene("runs", size, size, s, n)


def clip_reference_scenes(width: int = 1024, height: int = 800) -> list[Scene]:
    """The call sequences of the reference's visual clip tests (tests/clip.c:3-155: clipped paints
    with both fill rules, reset_clip, transformed clips, a clip_preserve + CLEAR-operator fill, nested
    clips inside save / restore) and of tests/save_restore.c:3-18 (20 nested saves with transformed
    strokes), each on a transparent surface. The tests themselves only write images."""
    out = []

    def clipped_paint(rule, second):
        st = Stream()
        st.set_fill_rule(rule)
        st.set_source_rgb(1, 0, 0)
        st.paint()
        st.set_source_rgb(0, 1, 0)
        st.rectangle(100, 100, 300, 200)
        st.clip()
        st.paint()
        if second:
            st.reset_clip()
            st.rectangle(200, 200, 300, 200)
            st.set_source_rgb(0, 0, 1)
            st.paint()
        return st

    for rule, tag in ((FILL_NON_ZERO, "ec"), (FILL_EVEN_ODD, "eo")):
        out.append(Scene(f"clipped_paint_{tag}", width, height, clipped_paint(rule, False), 0))
        out.append(Scene(f"clipped_paint2_{tag}", width, height, clipped_paint(rule, True), 1))
        st = Stream()
        st.set_fill_rule(rule)
        st.set_source_rgb(1, 0, 0)
        st.paint()
        st.rotate(0.2)
        st.set_source_rgb(0, 1, 0)
        st.rectangle(100, 100, 300, 200)
        st.clip()
        st.paint()
        out.append(Scene(f"clipped_transformed_{tag}", width, height, st, 0))
    st = Stream()
    st.set_fill_rule(FILL_NON_ZERO)
    st.set_source_rgb(1, 0, 0)
    st.paint()
    st.rotate(0.1)
    st.set_source_rgb(0, 1, 0)
    st.rectangle(100, 100, 300, 200)
    st.clip()
    st.paint()
    st.translate(200, 100)
    st.set_source_rgb(0, 0, 1)
    st.paint()
    out.append(Scene("clip_transformed_ec", width, height, st, 0))
    st = Stream()  # test_clip
    st.clear()
    st.set_fill_rule(FILL_EVEN_ODD)
    st.polyline(np.array([[10, 10], [400, 150], [900, 10], [700, 450], [900, 750], [500, 650], [100, 800],
                          [150, 400]], np.float32))
    st.clip_preserve()
    st.set_operator(0)
    st.fill_preserve()
    st.clip()
    st.set_operator(2)
    st.set_source_rgb(1, 0, 0)
    st.paint()
    out.append(Scene("test_clip", width, height, st, 1))
    st = Stream()  # test_clip2
    st.clear()
    st.set_fill_rule(FILL_NON_ZERO)
    st.rectangle(50, 50, 500, 500)
    st.clip()
    st.set_source_rgb(1, 0, 0)
    st.paint()
    st.save()
    st.rectangle(100, 100, 350, 350)
    st.clip()
    st.save()
    st.set_source_rgb(1, 1, 0)
    st.paint()
    st.rectangle(200, 200, 200, 200)
    st.clip()
    st.set_source_rgb(0, 1, 0)
    st.paint()
    st.restore()
    st.rectangle(350, 350, 420, 420)
    st.set_source_rgb(0, 0, 1)
    st.fill()
    st.restore()
    out.append(Scene("test_clip2", width, height, st, 1))
    st = Stream()  # tests/save_restore.c recurse_draw, depth 20
    for depth in range(1, 21):
        st.save()
        st.translate(5, 5)
        st.rectangle(float(depth), float(depth), 200, 200)
        st.set_source_rgb(0, 1, 0)
        st.stroke()
    for _ in range(20):
        st.restore()
    out.append(Scene("save_restore", width, height, st, 20))
    return out


def clip_mix(variant: int = 0, size: int = 320, n: int = 80, flushes: bool = True) -> Scene:
    """Random fills, strokes and paints interleaved with clip / clip_preserve (both fill rules,
    transformed, curved), reset_clip, clear, save / restore with state changes in between (nesting
    deep enough to pass the six stencil levels), gradients saved and restored, and flushes (frame
    boundaries that the stencil must survive)."""
    rng = Rng(0x5EED0000 + 16 * 11 + variant)
    s = Stream()
    _white(s)
    depth = 0
    draws = 0
    for i in range(n):
        r = rng.u()
        cx, cy = rng.uniform(0, size), rng.uniform(0, size)
        rad = rng.loguniform(8, size / 2)
        if r < 0.16:  # clip
            s.set_fill_rule(FILL_EVEN_ODD if rng.u() < 0.4 else FILL_NON_ZERO)
            kind = rng.randint(0, 3)
            if kind == 0:
                s.rectangle(cx - rad, cy - rad * 0.7, 2 * rad, 1.4 * rad)
            elif kind == 1:
                s.arc(cx, cy, rad, 0, 6.2831853)
            elif kind == 2:  # star polygon: self-intersecting, fill rules differ
                m = 5
                ang = np.arange(m) * (4 * math.pi / m) + rng.uniform(0, 1)
                s.polyline(np.stack([cx + rad * np.cos(ang), cy + rad * np.sin(ang)], 1), close=True)
            else:
                pts = rng.uniform(-size * 0.1, size * 1.1, 2 + 6 * 2).astype(np.float32)
                s.curves(pts[:2], pts[2:].reshape(2, 6), close=True)
            if rng.u() < 0.3:
                s.clip_preserve()
                s.set_source_rgba(*rng.u(3), 0.5)
                s.stroke()
                draws += 1
            else:
                s.clip()
            draws += 1
        elif r < 0.30 and depth < 16:
            s.save()
            depth += 1
            s.translate(float(rng.uniform(-8, 8)), float(rng.uniform(-8, 8)))
            s.set_line_width(float(rng.loguniform(0.8, 10)))
            s.set_opacity(float(rng.uniform(0.5, 1.0)))
            if rng.u() < 0.3:
                s.set_linear(0, 0, size, size, [[0, *rng.u(3), 1], [1, *rng.u(3), 0.7]])
        elif r < 0.42 and depth > 0:
            s.restore()
            depth -= 1
        elif r < 0.45:
            s.reset_clip()
        elif r < 0.47:
            s.set_source_rgba(*rng.u(3), rng.uniform(0.3, 1))
            s.paint()
            draws += 1
        elif r < 0.48:
            s.clear()
            draws += 1
        elif r < 0.51 and flushes:
            s.flush()
        else:
            s.set_source_rgba(*rng.u(3), rng.uniform(0.3, 1.0))
            if rng.u() < 0.5:
                s.set_fill_rule(FILL_EVEN_ODD if rng.u() < 0.3 else FILL_NON_ZERO)
                s.rectangle(cx - rad / 2, cy - rad / 3, rad, rad * 0.6)
                s.arc(cx + rad / 3, cy, rad / 3, 0, 6.2831853)
                s.fill()
            else:
                s.set_line_join(int(rng.randint(0, 2)))
                m = rng.randint(2, 6)
                s.polyline(np.stack([rng.uniform(0, size, m), rng.uniform(0, size, m)], 1), close=bool(i % 2))
                s.stroke()
            draws += 1
    while depth > 0:
        s.restore()
        depth -= 1
        s.set_source_rgba(*rng.u(3), 0.4)
        s.paint()
        draws += 1
    return Scene(f"clip_mix{variant}", size, size, s, draws)


def deep_clip_scene(size: int = 256, depth: int = 14) -> Scene:
    """Nested save + clip to shrinking concentric squares, one paint per level, then restores with a
    paint after each: crosses the six-level stencil byte twice (two whole-stencil copies)."""
    s = Stream()
    s.set_source_rgb(1, 1, 1)
    s.paint()
    for d in range(1, depth + 1):
        s.save()
        s.rectangle(float(8 * d), float(8 * d), float(size - 16 * d), float(size - 16 * d))
        s.clip()
        s.set_source_rgb((d * 37 % 256) / 255.0, (d * 71 % 256) / 255.0, (d * 113 % 256) / 255.0)
        s.paint()
    for d in range(depth, 0, -1):
        s.restore()
        s.set_source_rgba((d * 53 % 256) / 255.0, (d * 29 % 256) / 255.0, (d * 97 % 256) / 255.0, 0.5)
        s.paint()
    return Scene("deep_clip", size, size, s, 2 * depth)


def _f32(v: float) -> float:
    return float(np.float32(v))


PIF = _f32(math.pi)  # M_PIF of the reference (tests/common/test.h:15)

SHAPE_RECTANGLE, SHAPE_ROUNDED_RECTANGLE, SHAPE_CIRCLE, SHAPE_STAR, SHAPE_RANDOM = 1, 2, 3, 5, 6
_STAR = ((0, 85), (75, 75), (100, 10), (125, 75), (200, 85), (150, 125), (160, 190), (100, 150), (40, 190),
         (50, 125), (0, 85))


def randoms(shape: int = SHAPE_RANDOM, mode: str = "fill", n: int = 500, size: int = 512, variant: int = 0,
            size_fact: float = 0.2) -> Scene:
    """The reference's random-shape benchmark (tests/perfs/randoms.c:17-79): a context set up as
    tests/common/test.c:686-697 does (line width 2, miter joins, butt caps, no dash, non-zero rule,
    vkvg_clear), then n shapes from draw_random_shape (tests/common/test.c:702-793) -- rectangles,
    rounded rectangles (four arcs), full circles (one arc), stars (a closed 11-point polyline), or a
    random pick among them -- each with a random translucent colour and filled, stroked, or filled
    (preserve) then stroked. The reference draws its uniform numbers from a fixed table
    (tests/common/rnd.c); here they come from the PCG64 stream (*), in the same order of use."""
    rng = Rng(0x5EED0000 + 16 * 6 + 4 * variant + {"fill": 0, "stroke": 1, "fill_stroke": 2}[mode])
    w = h = float(size)
    s = Stream()
    s.set_line_width(2.0)
    s.set_line_join(JOIN_MITER)
    s.set_line_cap(CAP_BUTT)
    s.set_dash([], 0.0)
    s.set_fill_rule(FILL_NON_ZERO)
    s.clear()
    f = _f32
    pi, pi_2 = f(math.pi), f(math.pi / 2)

    def rnd():
        return f(rng.u())

    def shape_calls(kind):
        s.set_source_rgba(rnd(), rnd(), rnd(), rnd())
        if kind == SHAPE_RECTANGLE:
            z = math.trunc(f(f(f(size_fact * w) * rnd()) + 1.0))
            v = math.trunc(f(f(f(size_fact * h) * rnd()) + 1.0))
            x = math.trunc(f((w - z) * rnd()))
            y = math.trunc(f((h - v) * rnd()))
            s.rectangle(x + 1, y + 1, z, v)
        elif kind == SHAPE_ROUNDED_RECTANGLE:
            z = math.trunc(f(f(f(size_fact * w) * rnd()) + 1.0))
            v = math.trunc(f(f(f(size_fact * h) * rnd()) + 1.0))
            x = math.trunc(f((w - z) * rnd()))
            y = math.trunc(f((h - v) * rnd()))
            r = math.trunc(f(f(f(0.2 * z) * rnd()) + 1.0))
            if r > v / 2 or r > z / 2:
                r = min(v / 2, z / 2)
            s.move_to(x, y + r)
            s.arc(x + r, y + r, r, pi, -pi_2)
            s.line_to(x + z - r, y)
            s.arc(x + z - r, y + r, r, -pi_2, 0)
            s.line_to(x + z, y + v - r)
            s.arc(x + z - r, y + v - r, r, 0, pi_2)
            s.line_to(x + r, y + v)
            s.arc(x + r, y + v - r, r, pi_2, pi)
            s.line_to(x, y + r)
            s.close_path()
        elif kind == SHAPE_CIRCLE:
            x = f(rnd() * w)
            y = f(rnd() * h)
            r = math.trunc(f(f(f(size_fact * min(w, h)) * rnd()) + 1.0))
            s.arc(x, y, r, 0, f(pi * 2.0))
        elif kind in (4, SHAPE_STAR):
            x = f(rnd() * w)
            y = f(rnd() * h)
            z = f(f(rnd() * size_fact) + 0.15)
            pts = np.array([[f(x + f(px * z)), f(y + f(py * z))] for px, py in _STAR], np.float32)
            s.polyline(pts, close=True)
        else:  # SHAPE_RANDOM: one of rectangle .. star (1 + rndf() * 4), with a second colour draw
            shape_calls(int(1 + rnd() * 4))

    for _ in range(n):
        shape_calls(shape)
        if mode == "fill":
            s.fill()
        elif mode == "stroke":
            s.stroke()
        else:
            s.fill_preserve()
            s.stroke()
    names = {1: "rectangles", 2: "rounded_rects", 3: "circles", 5: "stars", 6: "random"}
    return Scene(f"randoms_{names[shape]}_{mode}", size, size, s, n)


def reference_test_scenes(width: int = 512, height: int = 512, random_size: int = 60) -> list[Scene]:
    """The call sequences of the reference's visual test programs (tests/*.c), one scene per test
    function, on a transparent surface of the tests' default size (512 x 512,
    tests/common/test.c:38-39). `_initCtx` (tests/common/test.c:686-697) sets width 2, MITER, BUTT,
    no dash, NON_ZERO and clears. Float constants are rounded as the C code rounds them (M_PIF, the
    float products, (float)-M_PI_2). Tests that draw random shapes use this module's generator and
    `random_size` items (the reference draws `test_size` = 500 with its own generator) (*). The
    interactive tests/inverse_colinear.c is swept over its key toggles (join, cap, closed, arcs at
    either end, dash on / off and the dash's first entry, point count). The tests only write
    images: these scenes are parity cases, with no expected image of their own."""
    out: list[Scene] = []

    def add(name, st, n):
        out.append(Scene(name, width, height, st, n))

    def init_ctx(st):  # tests/common/test.c:686-697 with the default globals (:677-684)
        st.set_line_width(2)
        st.set_line_join(JOIN_MITER)
        st.set_line_cap(CAP_BUTT)
        st.set_dash([], 0)
        st.set_fill_rule(FILL_NON_ZERO)
        st.clear()

    # tests/line_caps.c:3-52
    st = Stream()
    st.set_source_rgba(0.9, 0.9, 0.9, 1)
    st.paint()
    x, y, dx, dy = 20.0, 20.0, 40.0, 60.0
    st.set_line_width(30)
    st.set_source_rgba(0, 0, 0, 1)
    st.move_to(x, y)
    st.rel_line_to(0, dy)
    st.stroke()
    st.set_line_cap(CAP_SQUARE)
    st.move_to(x + dx, y)
    st.rel_line_to(0, dy)
    st.stroke()
    st.set_line_cap(CAP_ROUND)
    st.move_to(x + 2 * dx, y)
    for mx, my, lx, ly in ((None, None, 0, dy), (dx, -dy, dx, dy), (dx, -dy / 2, dx, 0), (dx, dy / 2, dx, -dy),
                           (dx, dy, 0, -dy), (dx * 2, dy, -dx, -dy), (dx * 3, dy / 2, -dx, 0)):
        if mx is not None:
            st.rel_move_to(mx, my)
        st.rel_line_to(lx, ly)
    st.stroke()
    st.set_line_cap(CAP_BUTT)
    st.set_line_width(1)
    st.set_source_rgba(1, 0, 0, 1)
    st.move_to(x, y)
    st.rel_line_to(0, dy)
    for _ in range(2):
        st.rel_move_to(dx, -dy)
        st.rel_line_to(0, dy)
    st.stroke()
    add("line_caps", st, 4)

    # tests/colinear.c:3-19
    st = Stream()
    st.set_source_rgba(0.7, 0.7, 0.7, 1)
    st.paint()
    st.set_source_rgba(0, 1, 0, 1)
    st.set_line_width(10)
    st.polyline(np.array([[100, 100], [100, 200], [100, 300]], np.float32))
    st.stroke()
    add("colinear", st, 1)

    # tests/fill_non_zero.c:3-24
    st = Stream()
    st.save()
    st.set_line_width(30)
    st.set_fill_rule(FILL_NON_ZERO)
    st.set_source_rgba(0.1, 0.9, 0.1, 1.0)
    st.move_to(100, 100)
    for lx, ly in ((200, 0), (0, 150), (-200, 0)):
        st.rel_line_to(lx, ly)
    st.close_path()
    st.move_to(150, 150)
    for lx, ly in ((0, 50), (100, 0), (0, -50)):
        st.rel_line_to(lx, ly)
    st.close_path()
    st.fill()
    add("fill_non_zero", st, 1)

    # tests/radial_gradient.c:3-15
    st = Stream()
    st.set_radial(250, 250, 0, 300, 300, 260, [[0, 1, 0, 0, 1], [0.5, 0, 1, 0, 1], [1, 0, 0, 0, 1]])
    st.paint()
    add("radial_gradient", st, 0)

    # tests/line_join.c:3-72 (test)
    st = Stream()
    x, y = 250.0, 150.0
    st.set_line_width(100)
    st.set_source_rgba(0, 1, 0, 1)
    st.set_line_join(JOIN_ROUND)
    st.move_to(x, y)
    for lx, ly in ((-50, 30), (0, 60), (50, 30)):
        st.rel_line_to(lx, ly)
    st.close_path()
    st.stroke()
    st.set_source_rgba(0, 0, 1, 1)
    st.move_to(x + 250, y)
    for lx, ly in ((50, -30), (50, 0), (50, 30), (0, 60), (-50, 70), (-50, 0), (-50, -70)):
        st.rel_line_to(lx, ly)
    st.close_path()
    st.stroke()
    add("line_join", st, 2)

    # tests/line_join.c:74-122 (test2) and :124-172 (test3)
    for name, legs, starts, shift in (
            ("line_join2", ((70, 50), (-70, 50)), (((100, 100), 1, 1), ((200, 600), 1, -1), ((400, 100), -1, 1),
                                                  ((500, 600), -1, -1)), 500),
            ("line_join3", ((50, 70), (50, -70)), (((100, 100), 1, 1), ((500, 200), -1, 1), ((100, 400), 1, -1),
                                                  ((500, 500), -1, -1)), 450)):
        st = Stream()
        st.set_fill_rule(FILL_EVEN_ODD)
        st.set_line_width(30)
        st.set_line_join(JOIN_ROUND)
        st.translate(-50, -50)
        st.set_source_rgb(0.5, 0, 0)
        reps = 5 if name == "line_join2" else 4
        for _ in range(2):
            for (px, py), sx, sy in starts:
                st.move_to(px, py)
                for _ in range(reps):
                    for lx, ly in legs:
                        st.rel_line_to(sx * lx, sy * ly)
                st.stroke()
            st.set_line_join(JOIN_BEVEL)
            st.translate(shift, 0)
        add(name, st, 8)

    # tests/dashes.c:3-57, the dash offset after 1, 5 and 50 runs (dash_offset += 0.2f per run)
    off = np.float32(0)
    for run in range(1, 51):
        off = np.float32(off + np.float32(0.2))
        if run not in (1, 5, 50):
            continue
        st = Stream()
        st.clear()
        st.set_source_rgb(1, 1, 1)
        st.paint()
        st.set_line_cap(CAP_ROUND)
        st.set_line_join(JOIN_ROUND)
        st.set_dash([100, 20, 20, 20], float(off))
        st.set_line_width(4)
        st.set_source_rgb(0, 0, 1)
        st.move_to(50, 50)
        for lx, ly in ((500, 0), (0, 200), (200, 0), (0, 500), (-700, 0)):
            st.rel_line_to(lx, ly)
        st.close_path()
        st.stroke()
        st.set_dash([0, 20], float(off))
        st.set_source_rgb(0, 1, 0)
        st.move_to(100, 100)
        for lx, ly in ((400, 0), (0, 200), (200, 0), (0, 400), (-600, 0)):
            st.rel_line_to(lx, ly)
        st.close_path()
        st.stroke()
        st.set_line_width(10)
        st.set_source_rgb(1, 0, 0)
        st.set_dash([80, 20], float(off))
        st.rectangle(200, 300, 200, 200)
        st.stroke()
        add(f"dashes_run{run}", st, 3)

    # tests/dashes.c:59-93 (_long_curve, _long_path) with main's globals (:99-103): dash {0, 10},
    # width 4, ROUND caps
    w, h = float(width), float(height)
    for name in ("dashes_path", "dashes_curve"):
        rng = Rng(0x5EED0100 + (name == "dashes_curve"))
        st = Stream()
        st.set_line_width(4)
        st.set_line_join(JOIN_MITER)
        st.set_line_cap(CAP_ROUND)
        st.set_dash([0, 10], 0)
        st.set_fill_rule(FILL_NON_ZERO)
        st.clear()
        st.set_source_rgba(*[_f32(v) for v in rng.u(4)])
        if name == "dashes_path":
            ww, hh = w - 10, h - 10
            pts = np.stack([ww * rng.u(random_size + 1), hh * rng.u(random_size + 1)], 1)
            st.polyline(pts.astype(np.float32))
        else:
            st.move_to(_f32(w * rng.u()), _f32(h * rng.u()))
            for _ in range(random_size):
                cx1, cy1, cx2, cy2, x2, y2 = (w * rng.u(), h * rng.u(), w * rng.u(), h * rng.u(), w * rng.u(),
                                              h * rng.u())
                st.curve_to(cx1, cy1, cx2, cy2, x2, y2)
        st.stroke()
        add(name, st, 1)

    # tests/arcs.c:15-26 (scaled_up), :28-38 (sizes), :40-86 (test)
    def growing_arcs(st, y, count, end):  # tests/arcs.c:3-13, tests/circles.c:3-10
        x = np.float32(2)
        for i in range(1, count):
            x = np.float32(x + np.float32(0.5) * np.float32(i))
            st.arc(float(np.float32(x + 2)), y, float(np.float32(0.5) * np.float32(i)), 0, end)
            x = np.float32(x + np.float32(np.float32(0.5) * np.float32(i) + np.float32(5)))

    st = Stream()
    st.set_source_rgb(1, 1, 1)
    st.paint()
    st.set_source_rgb(0, 0, 0)
    st.scale(10, 10)
    st.arc(20, 20, 2.0, 0, _f32(np.float32(PIF) / np.float32(2)))
    st.stroke()
    add("arcs_scaled_up", st, 1)
    st = Stream()
    st.set_source_rgb(1, 1, 1)
    st.paint()
    st.set_source_rgb(0, 0, 0)
    growing_arcs(st, 100, 40, _f32(np.float32(PIF) * np.float32(1.5)))
    st.stroke()
    add("arcs_sizes", st, 1)
    st = Stream()
    st.set_source_rgb(1, 1, 1)
    st.paint()
    st.set_source_rgb(0, 0, 0)
    half, full = _f32(np.float32(PIF) / np.float32(2)), _f32(np.float32(PIF) * np.float32(2))
    st.set_source_rgb(1, 0, 1)
    st.set_line_width(5.0)
    st.arc(100, 100, 20, 0, half)
    st.stroke()
    st.set_source_rgb(0, 1, 1)
    st.arc_negative(100, 100, 20, 0, half)
    st.stroke()
    st.set_source_rgb(1, 0, 1)
    st.arc(100, 200, 20, half, 0)
    st.stroke()
    st.set_source_rgb(0, 1, 1)
    st.arc_negative(100, 200, 20, half, 0)
    st.stroke()
    st.set_source_rgb(0, 0, 1)
    st.set_line_width(10.0)
    st.arc(350, 100, 40, 0, full)
    st.stroke()
    st.set_source_rgb(0, 1, 0)
    st.set_line_width(1.0)
    for rgb in ((0, 1, 0), (1, 0, 0)):
        if rgb == (1, 0, 0):
            st.set_source_rgb(*rgb)
            st.scale(3, 3)
        st.arc(150, 100, 3.5, 0, full)
        st.stroke()
        st.arc(200, 200, 10, 0, full)
        st.fill()
    add("arcs_test", st, 10)

    # tests/stroke.c:3-52
    st = Stream()
    st.set_line_width(1)
    st.set_source_rgba(1, 0, 0, 1)
    st.polyline(np.array([[200.5, 200.5], [400.5, 200.5], [400.5, 400.5], [200.5, 400.5]], np.float32), close=True)
    st.stroke()
    for rgb, o in (((0, 1, 0), 300.5), ((0.5, 1, 0), 320.5)):
        st.set_source_rgba(*rgb, 1)
        st.polyline(np.array([[o, o], [o + 200, o], [o + 200, o + 200], [o, o + 200]], np.float32))
        st.stroke()
    st.set_line_width(40)
    st.set_source_rgba(0.5, 0.6, 1, 1.0)
    st.move_to(700, 475)
    st.line_to(400, 475)
    st.stroke()
    st.set_source_rgba(0, 0.5, 0.5, 0.5)
    st.move_to(300, 200)
    st.arc(200, 200, 100, 0, PIF)
    st.stroke()
    st.set_line_width(20)
    st.set_source_rgba(0.1, 0.1, 0.1, 0.5)
    st.move_to(100, 60)
    st.line_to(400, 600)
    st.stroke()
    st.set_source_rgba(1, 1, 1, 1)
    st.set_line_width(1)
    st.rectangle(600.5, 200.5, 100, 60)
    st.stroke()
    add("stroke", st, 7)

    # tests/curve.c:4-14 (test3)
    st = Stream()
    st.clear()
    st.set_line_width(20)
    st.set_source_rgb(1, 0, 0)
    st.move_to(80, 170)
    st.curve_to(100, 170, 160, 171, 180, 170)
    st.close_path()
    st.stroke()
    add("curve_test3", st, 1)
    # tests/curve.c:16-69 (test)
    st = Stream()
    st.set_line_width(20)
    st.scale(2, 2)
    st.set_line_join(JOIN_BEVEL)
    st.set_source_rgb(0.5, 0, 0)
    st.translate(200, 30)
    st.arc(200, 200, 20, 0, full)
    st.set_source_rgba(0.5, 0.0, 1.0, 0.5)
    st.move_to(100, 100)
    st.line_to(200, 100)
    st.curve_to(250, 100, 300, 150, 300, 200)
    st.line_to(300, 300)
    st.curve_to(300, 350, 250, 400, 200, 400)
    st.line_to(100, 400)
    st.curve_to(50, 400, 10, 350, 10, 300)
    st.line_to(10, 200)
    st.curve_to(10, 150, 50, 100, 100, 100)
    st.fill_preserve()
    st.set_source_rgba(0.1, 0.3, 0.7, 0.5)
    st.stroke()
    add("curve_test", st, 2)
    # tests/curve.c:71-100 (curved_rect)
    st = Stream()
    x, y, cw, ch, r = 50.0, 50.0, 150.0, 140.0, 30.0
    mpi2 = _f32(math.pi / 2)
    st.scale(2, 2)
    st.set_line_width(15)
    st.set_source_rgba(0, 0.5, 0.4, 1)
    st.move_to(x, y + r)
    st.arc(x + r, y + r, r, PIF, -mpi2)
    st.line_to(x + cw - r, y)
    st.arc(x + cw - r, y + r, r, -mpi2, 0)
    st.line_to(x + cw, y + ch - r)
    st.arc(x + cw - r, y + ch - r, r, 0, mpi2)
    st.line_to(x + r, y + ch)
    st.arc(x + r, y + ch - r, r, mpi2, PIF)
    st.line_to(x, y + r)
    st.close_path()
    st.fill_preserve()
    st.set_source_rgba(0.5, 0, 0, 0.5)
    st.stroke()
    add("curve_curved_rect", st, 2)
    # tests/curve.c:102-121 (test2)
    st = Stream()
    st.move_to(100, 400)
    st.curve_to(100, 100, 600, 700, 600, 400)
    st.curve_to(1000, 100, 100, 800, 1000, 800)
    st.curve_to(1000, 500, 700, 500, 700, 100)
    st.close_path()
    st.set_source_rgba(1, 0, 0, 1)
    st.set_line_width(40)
    st.stroke()
    add("curve_test2", st, 1)
    # tests/curve.c:123-140 (random_curves_stroke), :142-185 (_long_curv and its fill-rule variants)
    rng = Rng(0x5EED0102)
    st = Stream()
    init_ctx(st)
    for _ in range(random_size):
        st.set_source_rgba(*[_f32(v) for v in rng.u(4)])
        st.move_to(_f32(rng.u() * w), _f32(rng.u() * h))
        st.curve_to(*[(w if k % 2 == 0 else h) * rng.u() for k in range(6)])
        st.stroke()
    add("curve_random_curves_stroke", st, random_size)
    for k, (rule, stroke_too) in enumerate(((FILL_NON_ZERO, False), (FILL_EVEN_ODD, False),
                                            (FILL_NON_ZERO, True), (FILL_EVEN_ODD, True))):
        rng = Rng(0x5EED0110 + k)
        st = Stream()
        init_ctx(st)
        st.set_fill_rule(rule)
        st.set_source_rgba(*[_f32(v) for v in rng.u(4)])
        st.move_to(_f32(rng.u() * w), _f32(rng.u() * h))
        for _ in range(random_size):
            st.curve_to(*[(w if j % 2 == 0 else h) * rng.u() for j in range(6)])
        if stroke_too:
            st.fill_preserve()
            st.set_source_rgba(*[_f32(v) for v in rng.u(4)])
            st.stroke()
        else:
            st.fill()
        add(f"curve_long_curv_{'fill_stroke' if stroke_too else 'fill'}_{'nz' if rule else 'eo'}", st, 1 + stroke_too)

    # tests/gradient.c:3-9 (create_grad), :11-65 (paint), :67-76 (paint_repeat: the extend of a gradient
    # is not read by the reference, src/vkvg_context_internal.c:705-772, so it paints the PAD gradient)
    def grad3(x0):
        return (x0, 0, 300, 0), [[0, 1, 0, 0, 1], [0.5, 0, 1, 0, 1], [1, 0, 0, 1, 1]]

    st = Stream()
    init_ctx(st)
    st.rectangle(0, 0, 400, 460)
    st.set_linear(*grad3(0)[0], grad3(0)[1])
    st.fill()
    gx = 100.0
    st.rectangle(gx, 200, 50, 50)
    st.set_linear(*grad3(gx)[0], grad3(gx)[1])
    st.fill()
    gx += 100
    st.rectangle(10, 10, 300, 50)
    st.set_linear(10, 0, 300, 0, [[0, 0, 0, 1, 0], [1, 1, 0, 0, 1]])
    st.fill()
    for step in (200, 100):
        gx += step
        st.rectangle(gx, 200, 50, 50)
        st.set_linear(*grad3(gx)[0], grad3(gx)[1])
        st.fill()
    add("gradient_paint", st, 5)
    st = Stream()
    init_ctx(st)
    st.set_linear(*grad3(0)[0], grad3(0)[1])
    st.paint()
    add("gradient_paint_repeat", st, 0)
    # tests/gradient.c:78-89 (test), :91-108 (test2), :110-132 (gradient_transform), :134-148 (gradient_alpha)
    st = Stream()
    init_ctx(st)
    st.set_linear(*grad3(0)[0], grad3(0)[1])
    st.rectangle(100, 100, 200, 200)
    st.set_line_width(20)
    st.stroke()
    add("gradient_test", st, 1)
    st = Stream()
    init_ctx(st)
    st.set_source_rgb(1, 0, 0)
    st.paint()
    st.set_line_width(20)
    st.set_linear(100, 0, 300, 0, [[0, 1, 1, 1, 1], [1, 1, 1, 0, 0]])
    st.rectangle(100, 100, 200, 200)
    st.fill()
    add("gradient_test2", st, 1)
    st = Stream()
    init_ctx(st)
    st.translate(200, 100)
    st.rotate(0.5)
    st.set_line_width(20)
    st.set_linear(0, 0, 400, 0, [[0, 1, 0, 0, 1], [0.5, 0, 1, 0, 1], [1, 0, 0, 1, 1]])
    st.rectangle(0, 0, 400, 200)
    st.stroke()
    add("gradient_transform", st, 1)
    st = Stream()
    init_ctx(st)
    st.set_source_rgb(0, 1, 0)
    st.paint()
    st.set_linear(10, 0, 300, 0, [[1, 1, 0, 0, 0.5], [1, 1, 0, 0, 0.5]])
    st.rectangle(10, 10, 300, 200)
    st.fill()
    add("gradient_alpha", st, 1)

    # tests/fill_and_stroke.c:3-21
    st = Stream()
    for px in (100, 300):
        st.move_to(px, 100)
        st.rel_line_to(50, -80)
        st.rel_line_to(50, 80)
        st.close_path()
    st.set_line_width(10.0)
    st.set_source_rgb(0, 0, 1)
    st.fill_preserve()
    st.set_source_rgb(1, 0, 0)
    st.stroke()
    add("fill_and_stroke", st, 2)

    # tests/simple_paint.c:3-45: paint, with a rotation or a scale, with and without a path
    for name, xf, rect, alpha in (("paint", None, False, 0.5), ("paint_with_rotation", "rot", False, 1),
                                  ("paint_with_scale", "scale", False, 1), ("paint_rect", None, True, 1),
                                  ("paint_rect_with_rotation", "rot", True, 1),
                                  ("paint_rect_with_scale", "scale", True, 1)):
        st = Stream()
        init_ctx(st)
        if xf == "rot":
            st.rotate(45)
        elif xf == "scale":
            st.scale(0.2, 0.2)
        st.set_source_rgba(1, 0, 0, alpha)
        if rect:
            st.rectangle(100, 100, 300, 200)
        st.paint()
        add(f"simple_{name}", st, 0)

    # tests/compositing.c:3-15 (compositing), :17-30 (opacity)
    for name in ("compositing", "opacity"):
        st = Stream()
        st.clear()
        st.set_source_rgba(1, 0, 0, 0.5 if name == "compositing" else 1.0)
        st.rectangle(100, 100, 200, 200)
        st.fill()
        if name == "opacity":
            st.set_opacity(0.5)
        st.set_source_rgba(0, 0, 1, 0.5 if name == "compositing" else 1.0)
        st.rectangle(200, 200, 200, 200)
        st.fill()
        add(name, st, 2)

    # tests/circles.c:12-22 (scaled_up), :24-34 (fill_and_stroke), :36-52 (sizes)
    st = Stream()
    st.clear()
    st.set_source_rgb(1, 1, 1)
    st.scale(100, 100)
    st.arc(2, 2, 0.5, 0, full)
    st.stroke()
    add("circles_scaled_up", st, 1)
    st = Stream()
    st.clear()
    st.set_source_rgba(0, 0.1, 0.8, 0.5)
    st.set_line_width(10)
    st.arc(300, 300, 150.0, 0, full)
    st.fill_preserve()
    st.stroke()
    add("circles_fill_and_stroke", st, 2)
    st = Stream()
    st.clear()
    st.set_source_rgb(1, 1, 1)
    growing_arcs(st, 100, 40, full)
    st.fill()
    growing_arcs(st, 200, 40, full)
    st.stroke()
    st.set_source_rgba(0, 0, 1, 0.4)
    growing_arcs(st, 300, 40, full)
    st.fill_preserve()
    st.set_line_width(5)
    st.stroke()
    add("circles_sizes", st, 4)

    # tests/lines.c: horizontal, vertical, horzAndVert, multilines, multi_segments (random lines)
    for k, name in enumerate(("horizontal", "vertical", "horz_and_vert", "multilines", "multi_segments")):
        rng = Rng(0x5EED0120 + k)
        st = Stream()
        init_ctx(st)
        n_draws = 1
        if name in ("horizontal", "vertical", "horz_and_vert"):
            n_draws = random_size
            for _ in range(random_size):
                st.set_source_rgba(*[_f32(v) for v in rng.u(4)])
                x1, y1 = _f32(w * rng.u()), _f32(h * rng.u())
                st.move_to(x1, y1)
                if name == "horz_and_vert":
                    st.line_to(_f32(w * rng.u()) + 1, _f32(h * rng.u()) + 1)
                else:
                    v = _f32(500.0 * rng.u())
                    st.line_to(x1 + v, y1) if name == "horizontal" else st.line_to(x1, y1 + v)
                st.stroke()
        elif name == "multilines":
            st.set_source_rgba(*[_f32(v) for v in rng.u(4)])
            for _ in range(random_size):
                st.move_to(_f32(w * rng.u()), _f32(h * rng.u()))
                st.line_to(_f32(w * rng.u()) + 1, _f32(h * rng.u()) + 1)
            st.stroke()
        else:
            st.set_source_rgba(*[_f32(v) for v in rng.u(4)])
            pts = np.stack([(w - 10) * rng.u(random_size + 1), (h - 10) * rng.u(random_size + 1)], 1)
            st.polyline(pts.astype(np.float32))
            st.stroke()
        add(f"lines_{name}", st, n_draws)

    # tests/transform.c cairo_tests (rotation += 0.002f per run; runs 1 and 300) with
    # _test_rounded_rect (the branch of a 200 x 200 rectangle with radius 102.4: four corner curves)
    rot = np.float32(0)
    for run in range(1, 301):
        rot = np.float32(rot + np.float32(0.002))
        if run not in (1, 300):
            continue
        st = Stream()
        st.set_fill_rule(FILL_EVEN_ODD)
        st.set_source_rgba(1.0, 1.0, 1.0, 1)
        st.paint()
        st.set_line_width(10)
        st.set_source_rgb(0, 0, 1)
        st.translate(float(width // 2), float(height // 2))
        st.rotate(float(rot))
        st.arc(0, 0, 150, 0, _f32(math.pi * 1.5))
        st.stroke()
        x0, y0, x1, y1 = -100.0, -100.0, 100.0, 100.0
        st.move_to(x0, (y0 + y1) / 2)
        st.curve_to(x0, y0, x0, y0, (x0 + x1) / 2, y0)
        st.curve_to(x1, y0, x1, y0, x1, (y0 + y1) / 2)
        st.curve_to(x1, y1, x1, y1, (x1 + x0) / 2, y1)
        st.curve_to(x0, y1, x0, y1, x0, (y0 + y1) / 2)
        st.close_path()
        st.set_source_rgb(0.5, 0.5, 1)
        st.fill_preserve()
        st.set_source_rgba(0.5, 0, 0, 0.5)
        st.set_line_width(10.0)
        st.stroke()
        add(f"transform_cairo_tests_run{run}", st, 3)

    # tests/rect_fill.c test / test_evenodd, tests/fill.c test
    for rule, tag in ((FILL_NON_ZERO, "nz"), (FILL_EVEN_ODD, "eo")):
        st = Stream()
        st.clear()
        st.set_fill_rule(rule)
        st.set_source_rgba(0, 0, 1, 0.5)
        st.rectangle(100, 100, 200, 200)
        st.fill()
        st.rectangle(200, 200, 200, 200)
        st.set_source_rgba(1, 0, 0, 0.5)
        st.fill()
        add(f"rect_fill_{tag}", st, 2)
    st = Stream()
    init_ctx(st)
    st.set_source_rgba(0.1, 0.9, 0.1, 1.0)
    st.move_to(100, 100)
    for lx, ly in ((50, 200), (150, -100), (100, 200), (-100, 100), (-10, -100), (-190, -50)):
        st.rel_line_to(lx, ly)
    st.close_path()
    st.fill()
    add("fill", st, 1)

    # tests/gradient2.c draw (interactive): linear and radial, 1 to 4 handle points, a hovered one
    g2 = np.array([[150, 150], [250, 150], [125, 125], [145, 125]], np.float32)

    def vlen(a, b):
        d = (b - a).astype(np.float32)
        return float(np.sqrt(np.float32(d[0] * d[0] + d[1] * d[1]), dtype=np.float32))

    for radial in (False, True):
        for n_pts in range(1, 5):
            hover = (n_pts + radial) % 3 - 1  # -1, 0 or 1
            st = Stream()
            st.clear()
            g2stops = [[0.0, 1, 0, 0, 0.5], [0.3, 0, 1, 0, 0.5], [0.6, 0, 0, 1, 0.5], [1.0, 0, 0, 0, 0.5]]
            if radial:
                st.set_radial(float(g2[2, 0]), float(g2[2, 1]), vlen(g2[2], g2[3]), float(g2[0, 0]), float(g2[0, 1]),
                              vlen(g2[0], g2[1]), g2stops)
            else:
                st.set_linear(float(g2[0, 0]), float(g2[0, 1]), float(g2[1, 0]), float(g2[1, 1]), g2stops)
            st.paint()
            st.set_dash([], 0)
            st.set_line_width(1)
            for i in range(n_pts):
                st.set_source_rgba(*((0.5, 0.5, 1.0, 0.7) if hover == i else (1.0, 0.5, 0.5, 0.9)))
                st.arc(float(g2[i, 0]), float(g2[i, 1]), 7.0, 0, full)
                st.fill_preserve()
                st.stroke()
            add(f"gradient2_{'radial' if radial else 'linear'}_{n_pts}pts", st, 2 * n_pts)

    # tests/bezier.c draw (interactive): one cubic, swept over closed / filled / join / cap and the
    # hovered handle (-1: none); the handles are drawn after a vkvg_flush
    bz = np.array([[150, 150], [500, 340], [470, 150], [160, 350]], np.float32)
    k = 0
    for closed in (False, True):
        for filled in (False, True):
            for join in (JOIN_MITER, JOIN_ROUND, JOIN_BEVEL):
                for cap in (CAP_BUTT, CAP_ROUND, CAP_SQUARE):
                    hover = k % 5 - 1
                    k += 1
                    st = Stream()
                    st.clear()
                    st.set_line_width(10.0)
                    st.set_line_join(join)
                    st.set_line_cap(cap)
                    st.move_to(float(bz[0, 0]), float(bz[0, 1]))
                    st.curve_to(*[float(v) for v in bz[1:].reshape(-1)])
                    if closed:
                        st.close_path()
                    if filled:
                        st.set_source_rgba(0.4, 0.6, 0.4, 1)
                        st.fill_preserve()
                    st.set_source_rgba(1, 0, 0, 0.3)
                    st.stroke()
                    st.flush()
                    st.set_line_width(1)
                    st.set_source_rgba(0.5, 0.5, 0.5, 0.6)
                    for i in range(4):
                        if i == hover:
                            st.arc(float(bz[i, 0]), float(bz[i, 1]), 7.0, 0, full)
                            st.set_source_rgba(0.5, 0.5, 1.0, 0.6)
                            st.fill_preserve()
                            st.set_source_rgba(0.5, 0.5, 0.5, 0.6)
                        else:
                            st.arc(float(bz[i, 0]), float(bz[i, 1]), 5.0, 0, full)
                            st.fill_preserve()
                        st.stroke()
                    if hover >= 0:
                        st.stroke_preserve()
                        st.set_dash([], 0)
                        st.set_line_width(2)
                        st.set_source_rgba(0, 0, 0, 1)
                        st.stroke()
                        st.set_source_rgba(0.5, 0.5, 1.0, 0.7)
                        st.arc(float(bz[hover, 0]), float(bz[hover, 1]), 5.0, 0, full)
                        st.fill_preserve()
                        st.stroke()
                    else:
                        st.stroke()
                    add(f"bezier_{'z' if closed else ''}{'f' if filled else ''}j{join}c{cap}_h{hover}", st, 12)

    # tests/inverse_colinear.c:31-72 (draw), swept over the toggles of its key handler (:74-128)
    ic_pts = np.array([[150, 150], [200, 300], [250, 150], [280, 350], [300, 100]], np.float32)
    a15 = _f32(np.float32(PIF) * np.float32(1.5))
    k = 0
    for join in (JOIN_MITER, JOIN_ROUND, JOIN_BEVEL):
        for cap in (CAP_BUTT, CAP_ROUND, CAP_SQUARE):
            for closed in (False, True):
                for dash in (None, (0.0, 60.0), (80.0, 60.0)):
                    start_arc, end_arc = bool(k & 1), bool(k & 2)
                    n_pts = 2 + k % 4
                    limit = (10.0, 2.5, 40.0)[k % 3]
                    k += 1
                    st = Stream()
                    st.clear()
                    if dash is not None:
                        st.set_dash(list(dash), 0)
                    st.set_source_rgba(1, 0, 0, 0.3)
                    st.set_line_width(50.0)
                    st.set_line_join(join)
                    st.set_line_cap(cap)
                    st.set_miter_limit(limit)
                    if start_arc:
                        st.arc_negative(float(ic_pts[0, 0]), float(ic_pts[0, 1]), 200, a15, PIF)
                    else:
                        st.move_to(float(ic_pts[0, 0]), float(ic_pts[0, 1]))
                    for i in range(1, n_pts - 1):
                        st.line_to(float(ic_pts[i, 0]), float(ic_pts[i, 1]))
                    last = ic_pts[n_pts - 1]
                    if end_arc:
                        st.arc_negative(float(last[0]), float(last[1]), 200, a15, PIF)
                    else:
                        st.line_to(float(last[0]), float(last[1]))
                    if closed:
                        st.close_path()
                    st.stroke()
                    add(f"inverse_colinear_j{join}c{cap}{'z' if closed else ''}"
                        f"{'' if dash is None else 'd%d' % int(dash[0])}_{k - 1}", st, 1)
    return out


def reference_source_scenes(width: int = 512, height: int = 512):
    """The reference's visual tests that take a source surface or a pattern matrix, as a list of
    (scene, sources): each source is (width, height, stream), an image that vkvg itself draws and
    that the scene's stream indexes. tests/paint_surf.c:3-88 draws its 256 x 256 source with
    createSurf (:3-18: paint red at 0.5, stroke a blue rectangle at 0.5, width 10) and paints it
    with offsets, repeats, a rotation, a scale and a path; tests/pattern_transforms.c:34-105 sets
    pattern matrices on a surface pattern (img_scale, icon_sheet: the images of data/ are not
    decoded here; createSurf's image of 300 x 200 and 64 x 48 stands in) and on linear gradients
    (pat_scale; pat_rotate after 1, 100 and 1000 runs, angle += 0.001f per run). Rotation
    matrices are rounded to float here; the reference computes them with sinf / cosf."""
    out = []

    def create_surf(w, h):  # tests/paint_surf.c:3-18
        st = Stream()
        st.set_line_width(10.0)
        st.set_source_rgba(1, 0, 0, 0.5)
        st.paint()
        st.set_source_rgba(0, 0, 1, 0.5)
        st.rectangle(5.0, 5.0, float(w) - 10.0, float(h) - 10.0)
        st.stroke()
        return (w, h, st)

    def init_ctx(st):  # tests/common/test.c:686-697
        st.set_line_width(2)
        st.set_line_join(JOIN_MITER)
        st.set_line_cap(CAP_BUTT)
        st.set_dash([], 0)
        st.set_fill_rule(FILL_NON_ZERO)
        st.clear()

    src = create_surf(256, 256)
    for name in ("paint", "paint_with_offset", "paint_multiple", "paint_with_rotation", "paint_with_scale",
                 "paint_rect"):
        st = Stream()
        init_ctx(st)
        if name == "paint_with_offset":
            st.set_source_rgba(0, 1, 0, 0.5)
            st.paint()
            st.set_source_surface(0, 100, 100)
            st.paint()
        elif name == "paint_multiple":
            for i in range(10):
                st.set_source_surface(0, i * 20, i * 20)
                st.paint()
        else:
            st.set_source_surface(0, 0, 0)
            if name == "paint_with_rotation":
                st.rotate(45)
            elif name == "paint_with_scale":
                st.scale(0.2, 0.2)
            elif name == "paint_rect":
                st.rectangle(100, 100, 300, 200)
            st.paint()
        out.append((Scene(f"paint_surf_{name}", width, height, st, 0), [src]))

    # tests/pattern_transforms.c:34-49 (img_scale), :50-67 (icon_sheet)
    st = Stream()
    init_ctx(st)
    st.translate(100, 100)
    st.set_source_pattern(0, matrix=(2.0, 0, 0, 2.0, 0, 0))
    st.rectangle(0, 0, 200, 140)
    st.fill()
    out.append((Scene("pattern_img_scale", width, height, st, 1), [create_surf(300, 200)]))
    scale = np.float32(0.5)
    c, s_ = np.float32(math.cos(0.2)), np.float32(math.sin(0.2))
    # init_translate(22, 27), then scale and rotate: each premultiplies the current matrix
    m = np.array([[1, 0], [0, 1]], np.float32), np.array([22, 27], np.float32)
    lin = np.array([[scale, 0], [0, scale]], np.float32) @ m[0]
    rot = np.array([[c, s_], [-s_, c]], np.float32)
    lin = (rot @ lin).astype(np.float32)
    st = Stream()
    init_ctx(st)
    st.translate(100, 100)
    st.set_source_pattern(0, matrix=(float(lin[0, 0]), float(lin[0, 1]), float(lin[1, 0]), float(lin[1, 1]),
                                     float(m[1][0]), float(m[1][1])))
    st.rectangle(0, 0, float(np.float32(24) / scale), float(np.float32(24) / scale))
    st.fill()
    out.append((Scene("pattern_icon_sheet", width, height, st, 1), [create_surf(64, 48)]))

    # tests/pattern_transforms.c:6-12 (create_grad), :69-84 (pat_scale), :86-105 (pat_rotate)
    stops = [[0, 1, 0, 0, 1], [0.5, 0, 1, 0, 1], [1.0, 0, 0, 1, 1]]
    st = Stream()
    init_ctx(st)
    st.translate(100, 100)
    st.set_linear(0, 0, 200, 0, stops, matrix=(0.5, 0, 0, 0.5, 0, 0))
    st.rectangle(0, 0, 200, 140)
    st.fill()
    out.append((Scene("pattern_pat_scale", width, height, st, 1), []))
    angle = np.float32(0)
    for run in range(1, 1001):
        angle = np.float32(angle + np.float32(0.001))
        if run not in (1, 100, 1000):
            continue
        ca, sa = float(np.float32(math.cos(float(angle)))), float(np.float32(math.sin(float(angle))))
        st = Stream()
        init_ctx(st)
        st.set_linear(0, 0, 200, 0, stops, matrix=(ca, sa, -sa, ca, 100, 70))
        st.rectangle(0, 0, 200, 140)
        st.fill()
        out.append((Scene(f"pattern_pat_rotate_run{run}", width, height, st, 1), []))
    return out


def surface_mix(variant: int = 0, size: int = 256, n: int = 40, n_sources: int = 3):
    """Surface sources: random bitmaps of uneven sizes painted and used to fill and stroke, placed
    with set_source_surface offsets or as patterns with every extend mode, nearest and bilinear
    filters, pattern matrices, rotated / scaled CTMs, opacity, clips and save / restore. Returns the
    scene and the source bitmaps (H x W x 4 uint8) that its surface-source opcodes index."""
    rng = Rng(0x5EED0000 + 16 * 12 + variant)
    sources = []
    for k in range(n_sources):
        w, h = int(rng.randint(5, 70)), int(rng.randint(5, 70))
        px = (rng.u(w * h * 4) * 256).astype(np.int64).clip(0, 255).astype(np.uint8).reshape(h, w, 4)
        if k % 2 == 0:  # premultiplied-valid texels for half of the sources
            a = px[..., 3:4].astype(np.uint16)
            px[..., :3] = (px[..., :3].astype(np.uint16) * a // 255).astype(np.uint8)
        sources.append(px)
    s = Stream()
    _white(s)
    depth = 0
    for i in range(n):
        r = rng.u()
        k = int(rng.randint(0, n_sources - 1))
        if r < 0.35:
            s.set_source_surface(k, float(rng.uniform(-20, size)), float(rng.uniform(-20, size)))
        elif r < 0.7:
            mat = None
            if rng.u() < 0.5:
                a = float(rng.uniform(0, 6.3))
                sc = float(rng.loguniform(0.3, 3))
                mat = (sc * math.cos(a), sc * math.sin(a), -sc * math.sin(a), sc * math.cos(a),
                       float(rng.uniform(-30, 30)), float(rng.uniform(-30, 30)))
            s.set_source_pattern(k, extend=int(rng.randint(0, 3)), filt=int(rng.randint(0, 5)), matrix=mat)
        elif r < 0.78:
            s.rotate(float(rng.uniform(-0.5, 0.5)))
        elif r < 0.84:
            s.scale(float(rng.uniform(0.6, 1.6)), float(rng.uniform(0.6, 1.6)))
        elif r < 0.88:
            s.identity_matrix()
        elif r < 0.92 and depth < 8:
            s.save()
            depth += 1
            cx, cy = rng.uniform(0, size, 2)
            s.arc(float(cx), float(cy), float(rng.uniform(20, size / 2)), 0, 6.2831853)
            s.clip()
        elif depth > 0 and r < 0.96:
            s.restore()
            depth -= 1
        s.set_opacity(float(rng.uniform(0.4, 1.0)) if rng.u() < 0.3 else 1.0)
        what = rng.u()
        if what < 0.3:
            s.paint()
        elif what < 0.7:
            cx, cy = rng.uniform(0, size, 2)
            s.rectangle(float(cx - 40), float(cy - 30), float(rng.uniform(10, 120)), float(rng.uniform(10, 90)))
            s.arc(float(cx), float(cy), float(rng.uniform(5, 50)), 0, 6.2831853)
            s.fill()
        else:
            s.set_line_width(float(rng.uniform(2, 14)))
            m = int(rng.randint(2, 5))
            s.polyline(np.stack([rng.uniform(0, size, m), rng.uniform(0, size, m)], 1))
            s.stroke()
    while depth > 0:
        s.restore()
        depth -= 1
    return Scene(f"surface_mix{variant}", size, size, s, n), sources


def grand_mix(variant: int = 0, size: int = 288, n: int = 120):
    """Everything interleaved at random: fills (both rules) and strokes (joins, caps, dashes) of lines,
    curves and arcs, solid / gradient / surface sources, opacity, operators, transforms, clips,
    reset_clip, save / restore, clear, paint and flushes. Returns the scene and its source bitmaps."""
    rng = Rng(0x5EED0000 + 16 * 13 + variant)
    sources = []
    for _ in range(2):
        w, h = int(rng.randint(8, 48)), int(rng.randint(8, 48))
        px = (rng.u(w * h * 4) * 256).astype(np.int64).clip(0, 255).astype(np.uint8).reshape(h, w, 4)
        a = px[..., 3:4].astype(np.uint16)
        px[..., :3] = (px[..., :3].astype(np.uint16) * a // 255).astype(np.uint8)
        sources.append(px)
    s = Stream()
    _white(s)
    depth = 0
    for i in range(n):
        r = rng.u()
        cx, cy = rng.uniform(0, size), rng.uniform(0, size)
        rad = float(rng.loguniform(6, size / 2.5))
        if r < 0.08:
            s.set_source_rgba(*rng.u(3), float(rng.uniform(0.2, 1)))
        elif r < 0.13:
            s.set_linear(cx - rad, cy, cx + rad, cy + rad / 2, [[0, *rng.u(3), 1], [0.5, *rng.u(3), 0.7],
                                                               [1, *rng.u(3), 0.9]])
        elif r < 0.17:
            s.set_radial(cx, cy, rad * 0.1, cx + 3, cy, rad, [[0, *rng.u(3), 0.8], [1, *rng.u(3), 1]])
        elif r < 0.21:
            s.set_source_surface(int(rng.randint(0, 1)), float(cx - 20), float(cy - 20))
        elif r < 0.24:
            s.set_source_pattern(int(rng.randint(0, 1)), extend=int(rng.randint(0, 3)),
                                 filt=int(rng.randint(0, 5)), matrix=(0.8, 0.2, -0.2, 0.8, 3, 1))
        elif r < 0.28:
            s.set_opacity(float(rng.uniform(0.3, 1.0)))
        elif r < 0.30:
            s.set_operator(int(rng.randint(0, 3)))
        elif r < 0.33:
            s.rotate(float(rng.uniform(-0.3, 0.3)))
        elif r < 0.35:
            s.identity_matrix()
        elif r < 0.41:
            s.set_fill_rule(FILL_EVEN_ODD if rng.u() < 0.5 else FILL_NON_ZERO)
            if rng.u() < 0.5:
                s.arc(cx, cy, rad, 0, 6.2831853)
            else:
                m = 5
                ang = np.arange(m) * (4 * math.pi / m)
                s.polyline(np.stack([cx + rad * np.cos(ang), cy + rad * np.sin(ang)], 1), close=True)
            if rng.u() < 0.3:
                s.clip_preserve()
                s.stroke()
            else:
                s.clip()
        elif r < 0.44:
            s.reset_clip()
        elif r < 0.50 and depth < 14:
            s.save()
            depth += 1
            s.set_line_width(float(rng.uniform(1, 9)))
        elif r < 0.56 and depth > 0:
            s.restore()
            depth -= 1
        elif r < 0.57:
            s.clear()
        elif r < 0.59:
            s.paint()
        elif r < 0.61:
            s.flush()
        elif r < 0.80:
            s.set_fill_rule(FILL_EVEN_ODD if rng.u() < 0.3 else FILL_NON_ZERO)
            pts = rng.uniform(-size * 0.1, size * 1.1, 2 + 6 * 3).astype(np.float32)
            s.curves(pts[:2], pts[2:].reshape(3, 6), close=True)
            s.rectangle(float(cx), float(cy), float(rad), float(rad * 0.5))
            s.fill()
        else:
            s.set_line_width(float(rng.uniform(0.8, 12)))
            s.set_line_join(int(rng.randint(0, 2)))
            s.set_line_cap(int(rng.randint(0, 2)))
            if rng.u() < 0.3:
                s.set_dash([float(rng.uniform(2, 10)), float(rng.uniform(1, 6))], float(rng.uniform(0, 5)))
            m = int(rng.randint(2, 6))
            s.polyline(np.stack([rng.uniform(0, size, m), rng.uniform(0, size, m)], 1), close=bool(i % 3 == 0))
            if rng.u() < 0.3:
                s.arc(cx, cy, rad / 2, 0, float(rng.uniform(1, 6)))
            s.stroke()
            s.set_dash([], 0.0)
    while depth > 0:
        s.restore()
        depth -= 1
    s.set_operator(2)
    s.set_source_rgba(0.1, 0.2, 0.3, 0.5)
    s.paint()
    return Scene(f"grand_mix{variant}", size, size, s, n), sources


@dataclass
class DiffScene:
    """Differentiable-render scene: topology plus FP32 parameters and the target's control points."""
    name: str
    width: int
    height: int
    ncubic: np.ndarray      # cubics per closed path
    stroke: np.ndarray      # bool per path
    ctrl: np.ndarray        # (n_ctrl, 2) float32, 3 per cubic
    rgba: np.ndarray        # (n_paths, 4) float32, straight colour and alpha
    lwidth: np.ndarray      # (n_paths,) float32 stroke widths (unused for fills)
    target_ctrl: np.ndarray  # control points of the target scene

    @property
    def n_paths(self) -> int:
        return len(self.ncubic)


def config5(variant: int = 0, n_paths: int = 1024, size: int = 512, smin: float = 8.0,
            smax: float = 64.0) -> DiffScene:
    """512x512 FP32: closed paths of n~U{3..8} cubics, centre ~U over the surface, control points
    = centre + N(0, s^2) with s log-U[8,64] (*); half of the paths (p = 0.5 (*)) stroked with
    width U[1,6], the others non-zero filled; rgb~U[0,1], alpha~U[0.5,1] (*). The target moves
    every control point by N(0, 4^2) px."""
    rng = Rng(0x5EED0000 + 16 * 5 + variant)
    ncubic, stroke, ctrl, rgba, lw = [], [], [], [], []
    for _ in range(n_paths):
        n = rng.randint(3, 8)
        c = rng.uniform(0, size, 2)
        s = rng.loguniform(smin, smax)
        ctrl.append(c[None, :] + s * rng.normal(6 * n).reshape(3 * n, 2))
        st = rng.u() < 0.5
        stroke.append(st)
        lw.append(rng.uniform(1.0, 6.0) if st else 0.0)
        rgba.append(np.concatenate([rng.u(3), [rng.uniform(0.5, 1.0)]]))
        ncubic.append(n)
    ctrl = np.concatenate(ctrl).astype(np.float32)
    target = (ctrl + 4.0 * rng.normal(ctrl.size).reshape(ctrl.shape)).astype(np.float32)
    return DiffScene("c5_512_diff", size, size, np.array(ncubic, np.int64), np.array(stroke, bool), ctrl,
                     np.array(rgba, np.float32), np.array(lw, np.float32), target)
