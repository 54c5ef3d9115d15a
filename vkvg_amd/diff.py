"""Differentiable renderer: torch.autograd.Function over the HIP kernels of
vkvg_amd/csrc/vkh_diff.hip (C ABI of include/vkvg_diff.h).

    scene = DiffRenderer(width, height, ncubic, stroke)          # topology, built once
    img = render(scene, ctrl, rgba, width)                       # (H, W, 4) float32 on cuda
    loss = ((img.double() - target.double()) ** 2).sum(); loss.backward()

ctrl (n_ctrl, 2), rgba (n_paths, 4) and width (n_paths,) are float32 CUDA tensors; gradients
flow to all three. There is no CPU fallback: the HIP library must be built.
"""
from __future__ import annotations

import ctypes
import time

import numpy as np

from . import VkvgError, lib


def _bind():
    L = lib()
    if getattr(L, "_diff_bound", False):
        return L
    vp, u32, f32, i32 = ctypes.c_void_p, ctypes.c_uint32, ctypes.c_float, ctypes.c_int
    L.vkvg_diff_create.argtypes = [u32, u32, u32, vp, vp]
    L.vkvg_diff_create.restype = vp
    L.vkvg_diff_destroy.argtypes = [vp]
    L.vkvg_diff_forward.argtypes = [vp, vp, vp, vp, vp, i32]
    L.vkvg_diff_forward.restype = i32
    L.vkvg_diff_forward_f64.argtypes = [vp, vp, vp, vp, vp]
    L.vkvg_diff_forward_f64.restype = i32
    L.vkvg_diff_backward.argtypes = [vp, vp, vp, vp, vp, vp, vp, vp]
    L.vkvg_diff_backward.restype = i32
    L.vkvg_diff_last_ms.argtypes = [vp, i32]
    L.vkvg_diff_last_ms.restype = f32
    L.vkvg_diff_stats.argtypes = [vp, vp]
    L.vkvg_diff_last_error.argtypes = [vp]
    L.vkvg_diff_last_error.restype = ctypes.c_char_p
    L._diff_bound = True
    return L


class DiffRenderer:
    """Owns the device-side topology and work buffers of one scene shape."""

    def __init__(self, width: int, height: int, ncubic, stroke):
        self._L = _bind()
        nc = np.ascontiguousarray(np.asarray(ncubic, np.uint32))
        st = np.ascontiguousarray(np.asarray(stroke, np.uint32))
        if nc.shape != st.shape or nc.ndim != 1 or len(nc) == 0:
            raise ValueError("ncubic and stroke must be equal-length 1-D arrays")
        self.width, self.height, self.n_paths = int(width), int(height), len(nc)
        self.n_ctrl = int(3 * nc.sum())
        self._h = self._L.vkvg_diff_create(self.width, self.height, self.n_paths, nc.ctypes.data, st.ctypes.data)
        if not self._h:
            raise VkvgError("vkvg_diff_create failed")

    def close(self):
        if self._h:
            self._L.vkvg_diff_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc: int, what: str):
        if rc != 0:
            raise VkvgError(f"{what}: {self._L.vkvg_diff_last_error(self._h).decode()}")

    def _args(self, ctrl, rgba, width):
        import torch
        for name, t, shape in (("ctrl", ctrl, (self.n_ctrl, 2)), ("rgba", rgba, (self.n_paths, 4)),
                               ("width", width, (self.n_paths,))):
            if not (t.is_cuda and t.dtype == torch.float32 and tuple(t.shape) == shape and t.is_contiguous()):
                raise ValueError(f"{name}: expected contiguous float32 cuda tensor of shape {shape}")
        return ctrl.data_ptr(), rgba.data_ptr(), width.data_ptr()

    def forward(self, ctrl, rgba, width, keep: bool = False):
        import torch
        img = torch.empty((self.height, self.width, 4), dtype=torch.float32, device=ctrl.device)
        c, r, w = self._args(ctrl, rgba, width)
        torch.cuda.synchronize()
        self._check(self._L.vkvg_diff_forward(self._h, c, r, w, img.data_ptr(), 1 if keep else 0), "forward")
        return img

    def forward64(self, ctrl, rgba, width):
        """The same forward in double precision (float parameters in, float64 image out): the
        function finite differences are taken of."""
        import torch
        img = torch.empty((self.height, self.width, 4), dtype=torch.float64, device=ctrl.device)
        c, r, w = self._args(ctrl, rgba, width)
        torch.cuda.synchronize()
        self._check(self._L.vkvg_diff_forward_f64(self._h, c, r, w, img.data_ptr()), "forward_f64")
        return img

    def backward(self, ctrl, rgba, width, grad_img):
        import torch
        g = grad_img.contiguous().to(torch.float32)
        gc, gr, gw = torch.empty_like(ctrl), torch.empty_like(rgba), torch.empty_like(width)
        c, r, w = self._args(ctrl, rgba, width)
        torch.cuda.synchronize()
        self._check(self._L.vkvg_diff_backward(self._h, c, r, w, g.data_ptr(), gc.data_ptr(), gr.data_ptr(),
                                               gw.data_ptr()), "backward")
        return gc, gr, gw

    def last_ms(self, backward: bool = False) -> float:
        return float(self._L.vkvg_diff_last_ms(self._h, 1 if backward else 0))

    def stats(self) -> dict:
        out = (ctypes.c_uint32 * 4)()
        self._L.vkvg_diff_stats(self._h, out)
        return {"edges": out[0], "segments": out[1], "commands": out[2], "batches": out[3]}


def _function():
    import torch

    class _Render(torch.autograd.Function):
        @staticmethod
        def forward(ctx, scene, ctrl, rgba, width):
            ctx.scene = scene
            ctx.save_for_backward(ctrl, rgba, width)
            return scene.forward(ctrl, rgba, width, keep=True)

        @staticmethod
        def backward(ctx, grad_img):
            ctrl, rgba, width = ctx.saved_tensors
            gc, gr, gw = ctx.scene.backward(ctrl, rgba, width, grad_img)
            return None, gc, gr, gw

    return _Render


_RENDER = None


def render(scene: DiffRenderer, ctrl, rgba, width):
    """Differentiable render; keeps the backward checkpoints when any input requires grad.
    Calling backward on an older graph after a newer render of the same scene is not supported."""
    global _RENDER
    import torch
    if not (ctrl.requires_grad or rgba.requires_grad or width.requires_grad) or not torch.is_grad_enabled():
        return scene.forward(ctrl, rgba, width, keep=False)
    if _RENDER is None:
        _RENDER = _function()
    return _RENDER.apply(scene, ctrl, rgba, width)


def scene_tensors(sc, device="cuda"):
    """Torch tensors (ctrl, rgba, width, target_ctrl) of a scenes.DiffScene."""
    import torch
    f = lambda a: torch.tensor(np.ascontiguousarray(a, np.float32), device=device)  # noqa: E731
    return f(sc.ctrl), f(sc.rgba), f(sc.lwidth), f(sc.target_ctrl)


def bench(steps: int = 10, warmup: int = 2, variant: int = 0) -> dict:
    """Plain forward vs forward + backward (L2 loss against the perturbed target) on the
    512x512, 1024-path differentiable scene. Times include the whole pipeline of each call."""
    import torch
    from . import scenes
    sc = scenes.config5(variant)
    R = DiffRenderer(sc.width, sc.height, sc.ncubic, sc.stroke)
    ctrl, rgba, width, tctrl = scene_tensors(sc)
    target = R.forward(tctrl, rgba, width)

    def fwd():
        return R.forward(ctrl, rgba, width, keep=False)

    def fwd_bwd():
        img = R.forward(ctrl, rgba, width, keep=True)
        g = 2.0 * (img - target)
        return R.backward(ctrl, rgba, width, g)

    for _ in range(warmup):
        fwd()
        fwd_bwd()
    torch.cuda.synchronize()
    t = time.perf_counter()
    dev_f = []
    for _ in range(steps):
        fwd()
        dev_f.append(R.last_ms())
    torch.cuda.synchronize()
    t_f = (time.perf_counter() - t) / steps * 1e3
    t = time.perf_counter()
    dev_fb = []
    for _ in range(steps):
        fwd_bwd()
        dev_fb.append(R.last_ms() + R.last_ms(True))
    torch.cuda.synchronize()
    t_fb = (time.perf_counter() - t) / steps * 1e3
    st = R.stats()
    R.close()
    return {"forward_ms": t_f, "fwd_bwd_ms": t_fb, "ratio": t_fb / t_f,
            "device_forward_ms": float(np.mean(dev_f)), "device_fwd_bwd_ms": float(np.mean(dev_fb)),
            "device_ratio": float(np.mean(dev_fb) / np.mean(dev_f)), "paths": sc.n_paths,
            "surface": f"{sc.width}x{sc.height} FP32 RGBA", **st}


if __name__ == "__main__":
    import json
    import sys
    try:  # torch brings up its device before the library does
        import torch
        torch.cuda.init()
    except Exception:
        pass
    steps = int(sys.argv[1]) if len(sys.argv) > 1 else 10
    print(json.dumps(bench(steps=steps, warmup=1)))
