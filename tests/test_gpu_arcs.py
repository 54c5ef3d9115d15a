"""Arcs evaluated on the GPU (arc items expanded by the flattening kernels) against the oracle,
which evaluates every point on the CPU: the reference's random-shape benchmark scenes
(tests/perfs/randoms.c:17-79 through draw_random_shape, tests/common/test.c:702-793) and an arc
stress scene (large radii, both directions, spans across 2 pi, CTM scales, arcs after a current
point, close_path removing an arc's last point, elliptic arcs and rounded rectangles)."""
import math

import numpy as np
import pytest

from vkvg_amd import scenes
from vkvg_amd.scenes import Rng, Scene
from vkvg_amd.stream import Stream
from parity_util import compare, render_gpu, render_oracle

pytestmark = pytest.mark.gpu


def _check(dev, scene, tile=16):
    g_rgba, g_sig, st = render_gpu(dev, scene, tile)
    o_rgba, o_sig, _ = render_oracle(scene)
    r = compare(g_rgba, g_sig, o_rgba, o_sig)
    print(scene.name, tile, r, {k: st[k] for k in ("points", "edges")}, flush=True)
    assert r["mask_mismatch_px"] == 0 and r["rgba_max_err"] == 0, r


@pytest.mark.parametrize("shape", [scenes.SHAPE_RECTANGLE, scenes.SHAPE_ROUNDED_RECTANGLE, scenes.SHAPE_CIRCLE,
                                   scenes.SHAPE_STAR, scenes.SHAPE_RANDOM])
@pytest.mark.parametrize("mode", ["fill", "stroke", "fill_stroke"])
def test_reference_random_shapes(device, shape, mode):
    _check(device, scenes.randoms(shape, mode))


def _last_circle_point(cx, cy, r):
    """The last stepped point of vkvg_arc(cx, cy, r, 0, 2 pi) for r < 30, in the reference's float
    arithmetic: step = min(pi/3, pi/r), angles step, 2 step, ... below 2 pi."""
    f32 = np.float32
    pif = f32(math.pi)
    step = min(f32(pif / f32(3)), f32(pif / f32(r)))
    a, last, a2 = f32(step), f32(step), f32(f32(2) * pif)
    while a < a2:
        last = a
        a = f32(a + step)
    cx, cy = f32(cx), f32(cy)
    return (float(f32(f32(f32(math.cos(float(last))) * f32(r)) + cx)),
            float(f32(f32(f32(math.sin(float(last))) * f32(r)) + cy)))


def arc_stress(variant=0, size=640, n=70):
    rng = Rng(0x5EED7000 + variant)
    s = Stream()
    s.set_source_rgb(1, 1, 1)
    s.paint()
    for i in range(n):
        s.set_source_rgba(*rng.u(3), 0.4 + 0.6 * rng.u())
        cx, cy = rng.uniform(0, size, 2)
        r = float(rng.loguniform(0.5, 900.0))
        a1 = float(rng.uniform(-8, 8))
        a2 = a1 + float(rng.uniform(-9, 9))
        kind = i % 7
        if kind in (1, 4):
            s.save()
            s.scale(float(rng.uniform(0.3, 3.0)), float(rng.uniform(0.3, 3.0)))
        if kind == 0:
            s.arc(cx, cy, r, a1, a2)
        elif kind == 1:
            s.arc_negative(cx, cy, r, a1, a2)
        elif kind == 2:  # arc after a current point, then a line and a close
            s.move_to(cx - r, cy)
            s.arc(cx, cy, r * 0.5, a1, a1 + float(rng.uniform(0.2, 6.0)))
            s.line_to(cx, cy)
            s.close_path()
        elif kind == 3:  # full circles, one closed by the arc itself, one closed again
            s.arc(cx, cy, r, 0, 2 * math.pi)
            s.close_path()
            s.arc_negative(cx + 3, cy, r * 0.7, 2 * math.pi, 0)
        elif kind == 4:
            s.rounded_rectangle(cx - r / 2, cy - r / 3, r, r * 0.7, float(rng.uniform(0, r / 2)))
        elif kind == 5:  # elliptic arcs, both flags
            s.move_to(cx, cy)
            s.elliptic_arc_to(cx + r, cy + r / 3, bool(i % 2), bool((i // 2) % 2), r * 0.8, r * 0.4,
                              float(rng.uniform(0, 3)))
            s.rounded_rectangle2(cx - r / 4, cy - r / 4, r / 2, r / 3, r / 10, r / 14)
        else:  # a full circle whose last stepped point is the path's first point: the circle's
            # own close_path removes that point from the arc item
            r = float(np.float32(rng.uniform(2, 25)))
            px, py = _last_circle_point(cx, cy, r)
            s.move_to(px, py)
            s.arc(cx, cy, r, 0, 2 * math.pi)
        if kind in (1, 4):
            s.restore()
        if i % 3 == 0:
            s.set_line_width(float(rng.uniform(0.5, 12)))
            s.set_line_join(int(i // 3) % 3)
            s.set_line_cap(int(i // 5) % 3)
            s.stroke()
        elif i % 3 == 1:
            s.set_fill_rule(int(i // 3) % 2)
            s.fill()
        else:
            s.fill_preserve()
            s.set_source_rgba(0, 0, 0, 0.5)
            s.set_line_width(2)
            s.stroke()
    return Scene(f"arc_stress_{variant}", size, size, s, n)


@pytest.mark.parametrize("variant", [0, 1, 2])
def test_arc_stress(device, variant):
    _check(device, arc_stress(variant))


@pytest.mark.parametrize("tile", [8, 32])
def test_arc_stress_tiles(device, tile):
    _check(device, arc_stress(3, size=400, n=40), tile=tile)
