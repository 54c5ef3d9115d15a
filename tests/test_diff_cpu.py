"""The FP64 reference of the differentiable renderer, pinned on the CPU: closed-form pixel areas
and the analytic gradient against float64 central differences of its own forward."""
import numpy as np

from oracle import diff_ref as D
from vkvg_amd import scenes


def _poly_edges(xy):
    xy = np.asarray(xy, np.float64)
    return np.concatenate([xy, np.roll(xy, -1, 0)], 1)


def test_rectangle_coverage_is_the_exact_overlap():
    x0, y0, x1, y1 = 1.3, 0.6, 4.75, 3.2
    E = _poly_edges([[x0, y0], [x0, y1], [x1, y1], [x1, y0]])
    A = D.coverage_area(E, 6, 5)
    ox = np.clip(np.minimum(np.arange(6) + 1, x1) - np.maximum(np.arange(6), x0), 0, None)
    oy = np.clip(np.minimum(np.arange(5) + 1, y1) - np.maximum(np.arange(5), y0), 0, None)
    assert np.allclose(np.abs(A), oy[:, None] * ox[None, :], atol=1e-12)


def test_triangle_coverage_sums_to_its_area_and_flips_sign_with_orientation():
    tri = np.array([[1.2, 0.7], [7.9, 2.3], [3.1, 6.6]])
    area = 0.5 * abs((tri[1, 0] - tri[0, 0]) * (tri[2, 1] - tri[0, 1]) - (tri[1, 1] - tri[0, 1]) * (tri[2, 0] - tri[0, 0]))
    A = D.coverage_area(_poly_edges(tri), 10, 8)
    B = D.coverage_area(_poly_edges(tri[::-1]), 10, 8)
    assert abs(abs(A.sum()) - area) < 1e-12
    assert np.allclose(A, -B, atol=1e-12)
    assert np.all(np.abs(A) <= 1 + 1e-12)


def test_shape_left_of_the_image_contributes_nothing():
    sq = _poly_edges([[-5.5, 1.0], [-5.5, 3.0], [-1.5, 3.0], [-1.5, 1.0]])
    assert np.abs(D.coverage_area(sq, 6, 5)).max() < 1e-12


def _tiny_scene():
    sc = D.Scene(24, 20, [3, 3, 4], [False, True, False])
    rng = np.random.default_rng(7)
    ctrl = np.concatenate([c + s * rng.standard_normal((3 * n, 2))
                           for c, s, n in (([8, 9], 4, 3), ([14, 10], 5, 3), ([12, 8], 6, 4))])
    rgba = np.array([[0.9, 0.2, 0.1, 0.8], [0.1, 0.5, 0.9, 0.7], [0.3, 0.8, 0.2, 0.6]])
    width = np.array([0.0, 2.5, 0.0])
    target = D.forward(sc, ctrl + 1.5 * rng.standard_normal(ctrl.shape), rgba, width)
    return sc, ctrl, rgba, width, target


def test_reference_gradient_matches_float64_central_differences():
    sc, ctrl, rgba, width, target = _tiny_scene()
    img = D.forward(sc, ctrl, rgba, width)
    gc, gr, gw = D.backward(sc, ctrl, rgba, width, 2.0 * (img - target))
    rng = np.random.default_rng(3)
    h = 1e-6

    def L(c, r, w):
        return D.loss(D.forward(sc, c, r, w), target)

    checks = []
    for i in rng.choice(ctrl.size, 14, replace=False):
        cp, cm = ctrl.copy().reshape(-1), ctrl.copy().reshape(-1)
        cp[i] += h
        cm[i] -= h
        checks.append((gc.reshape(-1)[i], (L(cp.reshape(-1, 2), rgba, width) - L(cm.reshape(-1, 2), rgba, width)) / (2 * h)))
    for i in rng.choice(rgba.size, 5, replace=False):
        rp, rm = rgba.copy().reshape(-1), rgba.copy().reshape(-1)
        rp[i] += h
        rm[i] -= h
        checks.append((gr.reshape(-1)[i], (L(ctrl, rp.reshape(-1, 4), width) - L(ctrl, rm.reshape(-1, 4), width)) / (2 * h)))
    wp, wm = width.copy(), width.copy()
    wp[1] += h
    wm[1] -= h
    checks.append((gw[1], (L(ctrl, rgba, wp) - L(ctrl, rgba, wm)) / (2 * h)))
    an = np.array([c[0] for c in checks])
    fd = np.array([c[1] for c in checks])
    scale = np.abs(fd).max()
    assert scale > 1e-3
    # relative error per parameter, floored at 1e-3 of the largest gradient
    err = np.abs(an - fd) / np.maximum(np.abs(fd), 1e-3 * scale)
    assert err.max() < 1e-4, (an, fd, err)


def test_config5_scene_shapes():
    sc = scenes.config5(n_paths=16, size=128)
    assert sc.ctrl.shape == (3 * int(sc.ncubic.sum()), 2) and sc.ctrl.dtype == np.float32
    assert sc.rgba.shape == (16, 4) and sc.target_ctrl.shape == sc.ctrl.shape
    assert np.all((sc.lwidth[sc.stroke] >= 1) & (sc.lwidth[sc.stroke] <= 6))
