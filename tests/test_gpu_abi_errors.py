"""The backend ABI rejects malformed frames with an error and a message, before any kernel runs:
stencil draws on a target without a stencil plane, unknown draw kinds, stencil indices out of range,
surface sources out of range (also for a paint). Frames are built directly as vkh_frame structs."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

u32, i32, f32, vp = ctypes.c_uint32, ctypes.c_int32, ctypes.c_float, ctypes.c_void_p


class Draw(ctypes.Structure):  # vkh_draw, include/vkvg_hip.h
    _fields_ = [("kind", u32), ("op", u32), ("color", u32), ("grad", i32), ("opacity", f32), ("mat", f32 * 6),
                ("first_subpath", u32), ("n_subpaths", u32), ("style", i32), ("target", u32),
                ("key_index", u32), ("source", i32)]


class Target(ctypes.Structure):  # vkh_target
    _fields_ = [("rgba", vp), ("width", u32), ("height", u32), ("sig", vp), ("stencil", vp),
                ("stencil_load", u32), ("reserved", u32)]


class Source(ctypes.Structure):  # vkh_surface_source
    _fields_ = [("rgba", vp), ("width", u32), ("height", u32), ("extend", u32), ("filter", u32),
                ("rect", f32 * 4), ("mat_inv", f32 * 6)]


class Frame(ctypes.Structure):  # vkh_frame
    _fields_ = [("items", vp), ("n_items", u32), ("subpaths", vp), ("n_subpaths", u32), ("draws", vp),
                ("n_draws", u32), ("styles", vp), ("n_styles", u32), ("dashes", vp), ("n_dashes", u32),
                ("grads", vp), ("n_grads", u32), ("targets", vp), ("n_targets", u32), ("sources", vp),
                ("n_sources", u32)]


def _draw(kind, style=-1, source=-1):
    d = Draw()
    d.kind, d.op, d.color, d.grad, d.opacity = kind, 2, 0xFF0000FF, -1, 1.0
    d.mat[0] = d.mat[3] = 1.0
    d.style, d.source = style, source
    return d


def _upload(eng, draws, target, sources=()):
    import vkvg_amd
    L = vkvg_amd.lib()
    L.vkvg_hip_upload.argtypes = [vp, ctypes.POINTER(Frame)]
    L.vkvg_hip_upload.restype = ctypes.c_int
    ds = (Draw * len(draws))(*draws)
    ss = (Source * max(1, len(sources)))(*sources)
    f = Frame()
    f.draws, f.n_draws = ctypes.addressof(ds), len(draws)
    f.targets, f.n_targets = ctypes.addressof(target), 1
    f.sources, f.n_sources = (ctypes.addressof(ss), len(sources)) if sources else (None, 0)
    r = L.vkvg_hip_upload(eng, ctypes.byref(f))
    return r, L.vkvg_hip_last_error(eng).decode()


def test_malformed_frames_are_rejected(device):
    import vkvg_amd
    eng = device._engine()
    s = vkvg_amd.Surface(device, 32, 32)
    try:
        t = Target()
        t.rgba, t.width, t.height = s.device_ptr(), 32, 32
        print("case abi errors", flush=True)
        r, msg = _upload(eng, [_draw(5)], t)  # clip draw without a stencil plane
        assert r == -2 and "stencil" in msg, (r, msg)
        r, msg = _upload(eng, [_draw(9)], t)  # reset without a stencil plane
        assert r == -2 and "stencil" in msg, (r, msg)
        r, msg = _upload(eng, [_draw(12)], t)
        assert r == -2 and "kind" in msg, (r, msg)
        r, msg = _upload(eng, [_draw(3, source=0)], t)  # a paint with a source, none given
        assert r == -2 and "source" in msg, (r, msg)
        src = Source()
        src.rgba, src.width, src.height = None, 8, 8
        r, msg = _upload(eng, [_draw(3, source=0)], t, [src])
        assert r == -2 and "surface source" in msg, (r, msg)
        plane = vkvg_amd.Surface(device, 32, 32)  # any 32x32x4-byte device buffer serves as a plane
        try:
            t.stencil = plane.device_ptr()
            r, msg = _upload(eng, [_draw(7, style=6)], t)  # saved bit 6 does not exist
            assert r == -2 and "index" in msg, (r, msg)
            r, msg = _upload(eng, [_draw(10, style=3)], t)  # copy 3 does not exist
            assert r == -2 and "index" in msg, (r, msg)
            r, msg = _upload(eng, [_draw(7, style=5)], t)  # a valid frame passes
            assert r == 0, (r, msg)
        finally:
            plane.close()
    finally:
        s.close()
