"""The flattening kernel's AGG traversal (the stack form: ancestors kept per lane, one split per
iteration, right children popped by counting the path's leading ones) and the stackless form of the
overflow kernel, both compiled for the host from the kernels' own code (vkvg_hip_flatten_points),
emit exactly the oracle's subdivision points (src/vkvg_context_internal.c:1313-1461, restated in
oracle/ovg_geom.c) for cubics of every scale and tolerance, including ones whose subdivision goes
deeper than the levels the kernel keeps on its stack."""
import ctypes

import numpy as np
import pytest

ITEM_CUBIC = 1


class _Item(ctypes.Structure):
    _fields_ = [("p", ctypes.c_float * 8), ("flags", ctypes.c_uint32), ("tol", ctypes.c_float)]


def _points(L, p, tol, stack):
    it = _Item((ctypes.c_float * 8)(*[float(v) for v in p]), ITEM_CUBIC, tol)
    cap = 1 << 16
    buf = np.zeros(2 * cap, np.float32)
    n = L.vkvg_hip_flatten_points(ctypes.byref(it), int(stack), buf.ctypes.data_as(ctypes.POINTER(ctypes.c_float)),
                                  cap)
    assert n <= cap
    return buf[:2 * n].reshape(-1, 2)


@pytest.fixture(scope="module")
def L():
    import vkvg_amd
    lib = vkvg_amd.lib()
    lib.vkvg_hip_flatten_points.argtypes = [ctypes.POINTER(_Item), ctypes.c_int, ctypes.POINTER(ctypes.c_float),
                                            ctypes.c_uint32]
    lib.vkvg_hip_flatten_points.restype = ctypes.c_uint32
    return lib


def _cases():
    rng = np.random.default_rng(11)
    out = []
    for k in range(600):
        scale = float(10.0 ** rng.uniform(-1, 4.5))  # up to ~30 000 px: subdivision past the stacked levels
        p = (rng.uniform(-1, 1, 8) * scale + rng.uniform(-500, 500, 2).repeat(4).reshape(2, 4).T.reshape(-1)) \
            .astype(np.float32)
        if k % 10 == 0:
            p[2:4] = p[0:2]  # first control point on the start
        if k % 15 == 0:
            p[4:6] = p[6:8]  # second control point on the end
        if k % 17 == 0:  # cusp: control points crossed over
            p[2:4], p[4:6] = p[6:8] + (p[6:8] - p[0:2]), p[0:2] - (p[6:8] - p[0:2])
        if k % 23 == 0:
            p[:] = p[0]  # a point
        tol = float(np.float32((0.25, 0.0625, 2.5, 0.01)[k % 4]))
        out.append((p, tol))
    return out


def test_stack_and_stackless_traversals_emit_the_oracle_points(L):
    from oracle import ovg
    deepest = 0
    for k, (p, tol) in enumerate(_cases()):
        want = ovg.flatten_cubic(p, tol)[:-1]  # the oracle appends the end point
        got_stack = _points(L, p, tol, 1)
        got_plain = _points(L, p, tol, 0)
        assert np.array_equal(got_stack.view(np.uint32), want.view(np.uint32)), (k, p, tol)
        assert np.array_equal(got_plain.view(np.uint32), want.view(np.uint32)), (k, p, tol)
        deepest = max(deepest, len(want))
    # some cubic had more points than a tree of 8 levels (more than the kernel stacks) has leaves:
    # its nodes below the stacked levels were recomputed from the deepest stacked ancestor
    assert deepest > 256
