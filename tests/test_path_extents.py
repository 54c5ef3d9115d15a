"""vkvg_path_extents (src/vkvg_context.c:684-696, src/vkvg_context_internal.c:1879-1917).

The host part, vkvg_hip_item_extents, runs the flattening kernels' subdivision code compiled for the
host; on the CPU it is compared with the extents of the oracle's flattening of the same cubic
(the oracle's AGG restatement, pinned by tests/test_oracle.py), bit for bit. On the GPU the public
entry point is checked on whole paths: lines and curves, the empty path, a path consumed by a fill
or kept by fill_preserve, and the reference's FLT_MIN start of the maxima."""
import ctypes

import numpy as np
import pytest

FLT_MAX = float(np.finfo(np.float32).max)
FLT_MIN = float(np.finfo(np.float32).tiny)
ITEM_CUBIC, ITEM_NO_END = 1, 4


class _Item(ctypes.Structure):
    _fields_ = [("p", ctypes.c_float * 8), ("flags", ctypes.c_uint32), ("tol", ctypes.c_float)]


def _lib():
    import vkvg_amd
    L = vkvg_amd.lib()
    L.vkvg_hip_item_extents.argtypes = [ctypes.POINTER(_Item), ctypes.POINTER(ctypes.c_float)]
    L.vkvg_hip_item_extents.restype = None
    return L


def _item_extents(L, p, flags, tol, start=(FLT_MAX, FLT_MAX, FLT_MIN, FLT_MIN)):
    it = _Item((ctypes.c_float * 8)(*[float(v) for v in p]), flags, tol)
    ext = (ctypes.c_float * 4)(*start)
    L.vkvg_hip_item_extents(ctypes.byref(it), ext)
    return np.array(ext[:], np.float32)


def _expected(pts, start=(FLT_MAX, FLT_MAX, FLT_MIN, FLT_MIN)):
    x0, y0, x1, y1 = (np.float32(v) for v in start)
    for x, y in pts:
        x0, x1 = (x if x < x0 else x0), (x if x > x1 else x1)
        y0, y1 = (y if y < y0 else y0), (y if y > y1 else y1)
    return np.array([x0, y0, x1, y1], np.float32)


def test_cubic_item_extents_equal_the_oracle_flattening():
    from oracle import ovg
    L = _lib()
    rng = np.random.default_rng(7)
    for k in range(400):
        scale = float(10.0 ** rng.uniform(-1, 3.5))
        p = (rng.uniform(-1, 1, 8) * scale + rng.uniform(-500, 500, 2).repeat(4).reshape(2, 4).T.reshape(-1)) \
            .astype(np.float32)
        if k % 10 == 0:
            p[2:4] = p[0:2]  # first control point on the start
        if k % 15 == 0:
            p[4:6] = p[6:8]  # second control point on the end
        tol = float(np.float32((0.25, 0.0625, 2.5, 0.01)[k % 4]))
        pts = ovg.flatten_cubic(p, tol)  # subdivision points, then the end point
        got = _item_extents(L, p, ITEM_CUBIC, tol)
        assert np.array_equal(got, _expected(pts)), (k, p, tol)
        got = _item_extents(L, p, ITEM_CUBIC | ITEM_NO_END, tol)
        assert np.array_equal(got, _expected(pts[:-1])), (k, "no end")


def test_point_item_and_the_flt_min_start():
    L = _lib()
    assert np.array_equal(_item_extents(L, [3, -4, 0, 0, 0, 0, 0, 0], 0, 0),
                          np.array([3, -4, 3, FLT_MIN], np.float32))
    got = _item_extents(L, [-3, -4, 0, 0, 0, 0, 0, 0], 0, 0, start=(0, 1, 2, 3))
    assert np.array_equal(got, np.array([-3, -4, 2, 3], np.float32))


@pytest.mark.gpu
def test_path_extents_through_the_api(device):
    from oracle import ovg
    import vkvg_amd
    L = _lib()
    vp, f32 = ctypes.c_void_p, ctypes.c_float
    L.vkvg_path_extents.argtypes = [vp] + [ctypes.POINTER(f32)] * 4
    for name, args in (("vkvg_move_to", [f32] * 2), ("vkvg_line_to", [f32] * 2), ("vkvg_curve_to", [f32] * 6),
                       ("vkvg_close_path", []), ("vkvg_fill", []), ("vkvg_fill_preserve", []),
                       ("vkvg_scale", [f32] * 2)):
        getattr(L, name).argtypes = [vp] + args
    s = vkvg_amd.Surface(device, 64, 64)
    ctx = L.vkvg_create(s.h)
    try:
        def ext():
            v = [f32() for _ in range(4)]
            L.vkvg_path_extents(ctx, *[ctypes.byref(x) for x in v])
            return np.array([x.value for x in v], np.float32)

        assert np.array_equal(ext(), np.zeros(4, np.float32))  # no path
        L.vkvg_move_to(ctx, 10, 20)
        assert np.array_equal(ext(), np.zeros(4, np.float32))  # a lone point is discarded
        L.vkvg_move_to(ctx, 10, 20)
        L.vkvg_line_to(ctx, 30, 5)
        L.vkvg_line_to(ctx, 12, 40)
        assert np.array_equal(ext(), np.array([10, 5, 30, 40], np.float32))
        # vkvg_path_extents ran _finish_path: the sub-path is committed and the next one is empty, so
        # the curve starts a sub-path at its first control point (src/vkvg_context.c:541-566)
        c = np.array([90, 80, 90, 80, -30, 70, 50, 10], np.float32)
        L.vkvg_curve_to(ctx, *[float(v) for v in c[2:]])
        exp = _expected(np.concatenate([np.array([[10, 20], [30, 5], [12, 40], [90, 80]], np.float32),
                                        ovg.flatten_cubic(c, 0.25)]))
        assert np.array_equal(ext(), exp)
        L.vkvg_fill_preserve(ctx)
        assert np.array_equal(ext(), exp)  # the path is kept
        L.vkvg_fill(ctx)
        assert np.array_equal(ext(), np.zeros(4, np.float32))  # the path is consumed
        L.vkvg_move_to(ctx, -10, -20)
        L.vkvg_line_to(ctx, -30, -5)
        assert np.array_equal(ext(), np.array([-30, -20, FLT_MIN, FLT_MIN], np.float32))
        L.vkvg_fill(ctx)
        L.vkvg_scale(ctx, 4, 4)  # the tolerance follows the scale: 0.25 / 4
        L.vkvg_move_to(ctx, 0, 0)
        c = np.array([0, 0, 5, 9, 9, -3, 14, 6], np.float32)
        L.vkvg_curve_to(ctx, *[float(v) for v in c[2:]])
        assert np.array_equal(ext(), _expected(np.concatenate([np.zeros((1, 2), np.float32),
                                                               ovg.flatten_cubic(c, 0.0625)])))
    finally:
        L.vkvg_destroy(ctx)
        s.close()
