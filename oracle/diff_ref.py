"""FP64 CPU reference of the differentiable renderer. TEST INFRASTRUCTURE ONLY: imported by
tests/ and nothing else; the product path (vkvg_amd/csrc/vkh_diff.hip) never calls it.

The reference has no differentiable renderer; this module defines the one built here and is the
checker for its HIP kernels:

* geometry: every closed path is a loop of n cubics; cubic q runs from control point 3q through
  3q+1, 3q+2 to 3q+3 (mod 3n) and is sampled at t = i/16, i = 0..15 (Bernstein form), giving a
  closed polyline of 16n points. A fill uses the polyline's edges. A stroke of half-width
  h = width/2 uses, per polyline segment p_j -> p_j+1 with unit left normal n_j = (-d.y, d.x)/|d|,
  the rectangle (p_j + h n_j, p_j+1 + h n_j, p_j+1 - h n_j, p_j - h n_j) and, at every vertex,
  the two bevel triangles (p_j, p_j + s h n_j-1, p_j + s h n_j), s = +-1, reordered so that all
  stroke primitives have the rectangles' orientation.
* coverage of a draw at a pixel: A = integral over the pixel of the winding number w(x, y) =
  sum over edges of sgn(dy) [y in the edge's y-range] [x > x_edge(y)], k = min(1, |A|).
* compositing: premultiplied source-over in floating point onto opaque white,
  out = c k + out (1 - c.a k), c = (r a, g a, b a, a).
* gradients: d/dtheta of a pixel's A is the boundary integral of the normal velocity, per edge
  piece inside the pixel: dA/da = -(dy, -dx) * int (1-s) ds, dA/db = -(dy, -dx) * int s ds
  (edge a -> b, d = b - a, s in [0, 1] along the edge); chained through the stroke normals and
  half-widths and the Bernstein weights to control points, colours and widths.

The gradient is pinned by central differences of the float64 forward (tests/test_diff_cpu.py).
"""
from __future__ import annotations

import math

import numpy as np

NSEG = 16  # polyline segments per cubic


def bernstein(nseg: int = NSEG) -> np.ndarray:
    t = np.arange(nseg, dtype=np.float64) / nseg
    u = 1.0 - t
    return np.stack([u * u * u, 3 * u * u * t, 3 * u * t * t, t * t * t], 1)


class Scene:
    """Topology: ncubic[i] cubics per closed path, stroke[i] True for a stroked path."""

    def __init__(self, width: int, height: int, ncubic, stroke):
        self.width, self.height = int(width), int(height)
        self.ncubic = np.asarray(ncubic, np.int64)
        self.stroke = np.asarray(stroke, bool)
        self.first_ctrl = np.concatenate([[0], np.cumsum(3 * self.ncubic)])

    @property
    def n_paths(self) -> int:
        return len(self.ncubic)

    @property
    def n_ctrl(self) -> int:
        return int(self.first_ctrl[-1])


def flatten(sc: Scene, ctrl: np.ndarray, p: int) -> np.ndarray:
    B = bernstein()
    c = ctrl[sc.first_ctrl[p]:sc.first_ctrl[p + 1]]
    n = int(sc.ncubic[p])
    out = []
    for q in range(n):
        P = np.stack([c[3 * q], c[3 * q + 1], c[3 * q + 2], c[(3 * q + 3) % (3 * n)]])
        out.append(B @ P)
    return np.concatenate(out)


def _normals(pts: np.ndarray) -> np.ndarray:
    d = np.roll(pts, -1, 0) - pts
    L = np.sqrt((d * d).sum(1))
    with np.errstate(invalid="ignore", divide="ignore"):
        n = np.stack([-d[:, 1] / L, d[:, 0] / L], 1)
    n[L == 0] = 0.0
    return n


def edges_of(sc: Scene, pts: np.ndarray, p: int, h: float):
    """Returns (E, 4) edges and, per edge endpoint, (point, normal index, sigma)."""
    m = len(pts)
    if not sc.stroke[p]:
        E = np.concatenate([pts, np.roll(pts, -1, 0)], 1)
        src = [((j, 0, 0), ((j + 1) % m, 0, 0)) for j in range(m)]
        return E, src
    n = _normals(pts)
    V = lambda i, k, s: pts[i] + s * h * n[k]  # noqa: E731
    edges, src = [], []

    def poly(vs):
        for a in range(len(vs)):
            b = (a + 1) % len(vs)
            edges.append(np.concatenate([V(*vs[a]), V(*vs[b])]))
            src.append((vs[a], vs[b]))

    for j in range(m):
        j1 = (j + 1) % m
        poly([(j, j, 1), (j1, j, 1), (j1, j, -1), (j, j, -1)])
        jp = (j - 1) % m
        n0, n1 = n[jp], n[j]
        cr = n0[0] * n1[1] - n0[1] * n1[0]
        for s in (1, -1):
            tri = [(j, 0, 0), (j, jp, s), (j, j, s)]
            if cr > 0:
                tri = [tri[0], tri[2], tri[1]]
            poly(tri)
    return np.array(edges), src


def pieces(e, W: int, H: int):
    """Pieces of edge e = (ax, ay, bx, by) per pixel row: yields (row, col, dy, fm, s0, s1) with
    col = -1 for the part left of the image; dy signed (= (s1-s0)*(by-ay)), fm the piece's mean
    x inside the cell, s0 < s1 the piece's parameter range on the edge."""
    ax, ay, bx, by = (float(v) for v in e)
    if ay == by:
        return
    y_lo, y_hi = min(ay, by), max(ay, by)
    for r in range(max(0, int(math.floor(y_lo))), min(H - 1, int(math.floor(y_hi))) + 1):
        ylo, yhi = max(y_lo, r), min(y_hi, r + 1)
        if ylo >= yhi:
            continue
        sa, sb = (ylo - ay) / (by - ay), (yhi - ay) / (by - ay)
        s0, s1 = min(sa, sb), max(sa, sb)
        x0, x1 = ax + s0 * (bx - ax), ax + s1 * (bx - ax)
        xl, xr = min(x0, x1), max(x0, x1)

        sL, sR = (s0, s1) if x0 <= x1 else (s1, s0)

        def s_of(x):  # interior split points only, clamped to the row's range
            return min(max((x - ax) / (bx - ax), s0), s1)

        dy_row = (yhi - ylo) if by > ay else (ylo - yhi)

        def emit(col, xa, xb):
            whole = True
            if bx == ax:
                ps0, ps1 = s0, s1
            else:
                u = sL if xa <= xl else s_of(xa)
                v = sR if xb >= xr else s_of(xb)
                ps0, ps1 = min(u, v), max(u, v)
                whole = xa <= xl and xb >= xr
            dy = dy_row if whole else (ps1 - ps0) * (by - ay)
            return (r, col, dy, 0.5 * (xa + xb) - col if col >= 0 else 0.0, ps0, ps1)

        if xl < 0:
            yield emit(-1, xl, min(xr, 0.0))
        c_lo = max(int(math.floor(xl)), 0)
        c_hi = min(int(math.floor(xr)) if xr > xl else int(math.floor(xl)), W - 1)
        for c in range(c_lo, c_hi + 1):
            xa, xb = max(xl, float(c)), min(xr, float(c + 1))
            if xb < xa or (xb == xa and xr > xl):
                continue
            yield emit(c, xa, xb)


def coverage_area(E: np.ndarray, W: int, H: int) -> np.ndarray:
    acc = np.zeros((H, W + 1))
    for e in E:
        for r, c, dy, fm, _, _ in pieces(e, W, H):
            if c < 0:
                acc[r, 0] += dy
            else:
                acc[r, c] += dy * (1.0 - fm)
                acc[r, c + 1] += dy * fm
    return np.cumsum(acc, 1)[:, :W]


def premul(rgba: np.ndarray) -> np.ndarray:
    r, g, b, a = rgba
    return np.array([r * a, g * a, b * a, a])


def forward(sc: Scene, ctrl, rgba, width, keep: bool = False):
    ctrl = np.asarray(ctrl, np.float64).reshape(-1, 2)
    rgba = np.asarray(rgba, np.float64).reshape(-1, 4)
    width = np.asarray(width, np.float64)
    W, H = sc.width, sc.height
    out = np.ones((H, W, 4))
    tape = []
    for p in range(sc.n_paths):
        pts = flatten(sc, ctrl, p)
        E, src = edges_of(sc, pts, p, 0.5 * width[p])
        A = coverage_area(E, W, H)
        k = np.minimum(1.0, np.abs(A))
        c = premul(rgba[p])
        if keep:
            tape.append((pts, E, src, A, out.copy()))
        out = c * k[..., None] + out * (1.0 - c[3] * k)[..., None]
    return (out, tape) if keep else out


def backward(sc: Scene, ctrl, rgba, width, grad_out):
    """Returns (d ctrl (n_ctrl, 2), d rgba (n_paths, 4), d width (n_paths,))."""
    ctrl = np.asarray(ctrl, np.float64).reshape(-1, 2)
    rgba = np.asarray(rgba, np.float64).reshape(-1, 4)
    width = np.asarray(width, np.float64)
    W, H = sc.width, sc.height
    _, tape = forward(sc, ctrl, rgba, width, keep=True)
    g = np.array(grad_out, np.float64).reshape(H, W, 4)
    gctrl = np.zeros_like(ctrl)
    grgba = np.zeros_like(rgba)
    gwidth = np.zeros_like(width)
    B = bernstein()
    for p in range(sc.n_paths - 1, -1, -1):
        pts, E, src, A, prev = tape[p]
        k = np.minimum(1.0, np.abs(A))
        c = premul(rgba[p])
        dk = (g * (c - c[3] * prev)).sum(2)
        gc = (g * k[..., None]).sum((0, 1))
        gc[3] += -((g * prev).sum(2) * k).sum()
        r_, g_, b_, a_ = rgba[p]
        grgba[p] += [a_ * gc[0], a_ * gc[1], a_ * gc[2], r_ * gc[0] + g_ * gc[1] + b_ * gc[2] + gc[3]]
        fp = np.where(np.abs(A) < 1.0, np.sign(A), 0.0)
        G = dk * fp
        g = g * (1.0 - c[3] * k)[..., None]
        # edge gradients by the boundary formula
        gE = np.zeros((len(E), 4))
        for ei, e in enumerate(E):
            dx, dy = e[2] - e[0], e[3] - e[1]
            for r, col, _, _, s0, s1 in pieces(e, W, H):
                if col < 0:
                    continue
                I1 = 0.5 * (s1 * s1 - s0 * s0)
                I0 = (s1 - s0) - I1
                Gv = G[r, col]
                gE[ei] += Gv * np.array([-dy * I0, dx * I0, -dy * I1, dx * I1])
        # chain to points, normals, half-width
        m = len(pts)
        gp = np.zeros((m, 2))
        gn = np.zeros((m, 2))
        gh = 0.0
        h = 0.5 * width[p]
        n = _normals(pts) if sc.stroke[p] else None
        for ei, (va, vb) in enumerate(src):
            for (i, kk, s), gv in ((va, gE[ei, :2]), (vb, gE[ei, 2:])):
                gp[i] += gv
                if s:
                    gn[kk] += s * h * gv
                    gh += s * float(n[kk] @ gv)
        if sc.stroke[p]:
            d = np.roll(pts, -1, 0) - pts
            L = np.sqrt((d * d).sum(1))
            for j in range(m):
                if L[j] == 0:
                    continue
                gnj = gn[j]
                gd = (np.array([gnj[1], -gnj[0]]) - d[j] * float(n[j] @ gnj) / L[j]) / L[j]
                gp[j] -= gd
                gp[(j + 1) % m] += gd
            gwidth[p] += 0.5 * gh
        # Bernstein weights to control points
        nc = int(sc.ncubic[p])
        f = sc.first_ctrl[p]
        for q in range(nc):
            gq = gp[q * NSEG:(q + 1) * NSEG]
            idx = [f + 3 * q, f + 3 * q + 1, f + 3 * q + 2, f + (3 * q + 3) % (3 * nc)]
            for mm in range(4):
                gctrl[idx[mm]] += (B[:, mm:mm + 1] * gq).sum(0)
    return gctrl, grgba, gwidth


def loss(img: np.ndarray, target: np.ndarray) -> float:
    d = np.asarray(img, np.float64) - np.asarray(target, np.float64)
    return float((d * d).sum())
