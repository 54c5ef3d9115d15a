"""Closed-form geometry for the oracle pins: pixel-centre masks of unions of convex polygons,
computed in double precision, and the arc step of the reference's `_get_arc_step`
(src/vkvg_context_internal.c:245-252) as the rule it states."""
from __future__ import annotations

import math

import numpy as np

from oracle import ovg
from vkvg_amd.stream import Stream


def arc_step(radius: float, scale: float = 1.0) -> float:
    """min(pi/3, pi/r) below 30 px of (scaled) radius, min(pi/3, pi/(0.4 r)) from 30 px on."""
    r = radius * abs(scale)
    return min(math.pi / 3, math.pi / r) if r < 30 else min(math.pi / 3, math.pi / (0.4 * r))


def steps_below(span: float, step: float) -> int:
    """Number of k >= 1 with k * step < span (the span is chosen away from multiples of step)."""
    n = math.floor(span / step)
    assert abs(span / step - round(span / step)) > 1e-3, "span too close to a multiple of the step"
    return n if n * step < span else n - 1


def fan(center, radius, a_start, a_end, step):
    """Convex polygon: the centre, the point at a_start, the points at a_start + k*step strictly
    between a_start and a_end (either direction), and the point at a_end."""
    cx, cy = center
    sgn = 1.0 if a_end >= a_start else -1.0
    angs = [a_start]
    k = 1
    while sgn * (a_start + sgn * k * step) < sgn * a_end:
        angs.append(a_start + sgn * k * step)
        k += 1
    angs.append(a_end)
    return [(cx, cy)] + [(cx + radius * math.cos(a), cy + radius * math.sin(a)) for a in angs]


def rect(x0, y0, x1, y1):
    return [(x0, y0), (x1, y0), (x1, y1), (x0, y1)]


def _inside_convex(poly, X, Y):
    P = np.asarray(poly, np.float64)
    n = len(P)
    area = 0.0
    for i in range(n):
        a, b = P[i], P[(i + 1) % n]
        area += a[0] * b[1] - b[0] * a[1]
    s = 1.0 if area > 0 else -1.0
    inside = np.ones(X.shape, bool)
    for i in range(n):
        a, b = P[i], P[(i + 1) % n]
        cr = (b[0] - a[0]) * (Y - a[1]) - (b[1] - a[1]) * (X - a[0])
        inside &= s * cr > 0
    return inside


def _edge_distance(poly, X, Y, shared=frozenset()):
    P = np.asarray(poly, np.float64)
    d = np.full(X.shape, np.inf)
    for i in range(len(P)):
        a, b = P[i], P[(i + 1) % len(P)]
        if (tuple(a), tuple(b)) in shared:
            continue
        ab = b - a
        L = ab @ ab
        if L == 0:
            continue
        t = np.clip(((X - a[0]) * ab[0] + (Y - a[1]) * ab[1]) / L, 0, 1)
        d = np.minimum(d, np.hypot(X - (a[0] + t * ab[0]), Y - (a[1] + t * ab[1])))
    return d


def union_mask(polys, w, h, margin=0.02, snap=False):
    """Pixel-centre mask of the union of convex polygons; asserts that no pixel centre lies within
    `margin` px of a polygon edge, so the 1/256 px vertex snap cannot change the mask. With `snap`
    the vertices are first rounded to the 1/256 px grid (the rasterizer's sub-pixel precision) and
    the margin only guards against float-rounding differences of the vertices. Edges that two
    polygons share (the same two vertices) are interior and not checked."""
    if snap:
        polys = [[(round(x * 256) / 256, round(y * 256) / 256) for x, y in p] for p in polys]
    Y, X = np.mgrid[0:h, 0:w].astype(np.float64) + 0.5
    m = np.zeros((h, w), bool)
    dmin = np.full((h, w), np.inf)
    edges = {}
    for p in polys:
        P = [tuple(map(float, q)) for q in p]
        for i in range(len(P)):
            key = frozenset((P[i], P[(i + 1) % len(P)]))
            edges[key] = edges.get(key, 0) + 1
    for p in polys:
        P = [tuple(map(float, q)) for q in p]
        shared = {(P[i], P[(i + 1) % len(P)]) for i in range(len(P))
                  if edges[frozenset((P[i], P[(i + 1) % len(P)]))] > 1}
        m |= _inside_convex(p, X, Y)
        dmin = np.minimum(dmin, _edge_distance(p, X, Y, shared))
    assert dmin.min() > margin, f"a pixel centre lies {dmin.min():.4f} px from an edge"
    return m


def stroke_stream(points, width, join=0, cap=0, miter=None, close=False, curve_ops=None):
    s = Stream()
    s.set_source_rgba(0, 0, 0, 1)
    s.set_line_width(width)
    s.set_line_join(join)
    s.set_line_cap(cap)
    if miter is not None:
        s.set_miter_limit(miter)
    if curve_ops is not None:
        curve_ops(s)
    else:
        s.polyline(np.asarray(points, np.float32), close=close)
    s.stroke()
    return s.words()


def coverage_counts(words, w, h):
    """Per-pixel number of blends of the stream's single draw (from the signature image)."""
    rgba, sig, stats, st = ovg.render(w, h, words)
    assert st == 0
    key = ovg.draw_key(0)
    inv = pow(key, -1, 1 << 32)  # keys are odd: k = sig * key^-1 mod 2^32
    k = (sig.astype(np.uint64) * inv) & 0xFFFFFFFF
    return k.astype(np.int64), stats
