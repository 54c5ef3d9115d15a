"""Differentiable renderer on the GPU: forward against the FP64 reference, gradients against the
FP64 reference gradient and against central differences of the GPU forward, and the
torch.autograd.Function harness."""
import numpy as np
import pytest

from oracle import diff_ref as D
from vkvg_amd import scenes

pytestmark = pytest.mark.gpu


def _setup(sc):
    import torch
    from vkvg_amd import diff
    R = diff.DiffRenderer(sc.width, sc.height, sc.ncubic, sc.stroke)
    ctrl, rgba, width, tctrl = diff.scene_tensors(sc)
    return torch, diff, R, ctrl, rgba, width, tctrl


def _ref_scene(sc):
    return D.Scene(sc.width, sc.height, sc.ncubic, sc.stroke)


def _rel_err(a, b, floor_frac=1e-3):
    a, b = np.asarray(a, np.float64).ravel(), np.asarray(b, np.float64).ravel()
    floor = floor_frac * np.abs(b).max()
    return np.abs(a - b) / np.maximum(np.abs(b), floor)


def test_forward_matches_fp64_reference():
    sc = scenes.config5(n_paths=20, size=80, smin=4.0, smax=16.0)
    torch, diff, R, ctrl, rgba, width, _ = _setup(sc)
    img = R.forward(ctrl, rgba, width).cpu().numpy()
    ref = D.forward(_ref_scene(sc), sc.ctrl, sc.rgba, sc.lwidth)
    err = np.abs(img - ref)
    print("forward max err", err.max(), "mean", err.mean(), R.stats(), flush=True)
    assert err.max() < 2e-4
    R.close()


def test_gradients_match_fp64_reference():
    sc = scenes.config5(n_paths=20, size=80, smin=4.0, smax=16.0)
    torch, diff, R, ctrl, rgba, width, tctrl = _setup(sc)
    target = R.forward(tctrl, rgba, width, keep=False)
    img = R.forward(ctrl, rgba, width, keep=True)
    g_img = 2.0 * (img - target)
    gc, gr, gw = R.backward(ctrl, rgba, width, g_img)
    rc, rr, rw = D.backward(_ref_scene(sc), sc.ctrl, sc.rgba, sc.lwidth, g_img.cpu().numpy().astype(np.float64))
    for name, a, b in (("ctrl", gc, rc), ("rgba", gr, rr), ("width", gw, rw)):
        e = _rel_err(a.cpu().numpy(), b)
        print(name, "max rel err", e.max(), flush=True)
        assert e.max() < 1e-3, name
    R.close()


def test_gradients_match_central_differences_on_256_parameters():
    # The full 512x512, 1024-path scene. Central differences with h = 1e-3 (px for control points
    # and widths, absolute for colours) of the same forward evaluated in double precision: near
    # x = 500 px neighbouring float coordinates are 3e-5 px apart, 3% of h, so differences of the
    # float forward would measure its rounding. The denominator is the actual (float) step.
    # Tolerance 1e-3 relative, floored at 1e-3 of the largest gradient checked.
    sc = scenes.config5()
    torch, diff, R, ctrl, rgba, width, tctrl = _setup(sc)
    target = R.forward64(tctrl, rgba, width)
    img = R.forward(ctrl, rgba, width, keep=True)
    gc, gr, gw = R.backward(ctrl, rgba, width, (2.0 * (img.double() - target)).float())
    rng = np.random.default_rng(11)
    flat = [("ctrl", ctrl, gc), ("rgba", rgba, gr), ("width", width, gw)]
    picks = []
    for _ in range(256):
        kind = rng.choice(3, p=[0.8, 0.15, 0.05])
        name, t, g = flat[kind]
        if name == "width":
            i = int(rng.choice(np.flatnonzero(sc.stroke)))
        else:
            i = int(rng.integers(t.numel()))
        picks.append((name, t, g, i))
    h = 1e-3

    def loss():
        return float(((R.forward64(ctrl, rgba, width) - target) ** 2).sum())

    an, fd = [], []
    for name, t, g, i in picks:
        v = t.view(-1)
        old = float(v[i])
        v[i] = old + h
        xp = float(v[i])
        lp = loss()
        v[i] = old - h
        xm = float(v[i])
        lm = loss()
        v[i] = old
        fd.append((lp - lm) / (xp - xm))
        an.append(float(g.view(-1)[i]))
    an, fd = np.array(an), np.array(fd)
    e = _rel_err(an, fd, floor_frac=1e-3)
    print("FD: max rel err", e.max(), "median", np.median(e), "largest |g|", np.abs(fd).max(), flush=True)
    assert np.isfinite(an).all()
    assert e.max() < 1e-3
    R.close()


def test_fp64_forward_agrees_with_fp32_forward():
    sc = scenes.config5(n_paths=64, size=192)
    torch, diff, R, ctrl, rgba, width, _ = _setup(sc)
    a = R.forward(ctrl, rgba, width).double()
    b = R.forward64(ctrl, rgba, width)
    assert float((a - b).abs().max()) < 1e-4
    R.close()


def test_autograd_function_gives_the_same_gradients():
    sc = scenes.config5(n_paths=40, size=128)
    torch, diff, R, ctrl, rgba, width, tctrl = _setup(sc)
    target = R.forward(tctrl, rgba, width).detach()
    c, r, w = (x.clone().requires_grad_(True) for x in (ctrl, rgba, width))
    img = diff.render(R, c, r, w)
    loss = ((img - target) ** 2).sum()
    loss.backward()
    img2 = R.forward(ctrl, rgba, width, keep=True)
    gc, gr, gw = R.backward(ctrl, rgba, width, 2.0 * (img2 - target))
    for a, b in ((c.grad, gc), (r.grad, gr), (w.grad, gw)):
        assert torch.allclose(a, b, rtol=1e-4, atol=1e-5 * float(b.abs().max()))
    R.close()
