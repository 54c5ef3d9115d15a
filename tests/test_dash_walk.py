"""The dash walk of the stroke kernels (whole dash periods taken in one step where the float
additions are exact), compiled for the host from the kernels' own code (vkvg_hip_dash_walk), ends
every segment in the state the reference's loop reaches one dash at a time: the loop of
_draw_dashed_segment, src/vkvg_context_internal.c:1246-1259 (flip dashOn, curDashOffset +=
dashes[curDash] in float, next dash, while curDashOffset < segment length), restated here with
numpy float32 arithmetic over many segments at once."""
import ctypes

import numpy as np
import pytest


@pytest.fixture(scope="module")
def L():
    import vkvg_amd
    lib = vkvg_amd.lib()
    f = lib.vkvg_hip_dash_walk
    f.argtypes = [ctypes.POINTER(ctypes.c_float), ctypes.c_uint32, ctypes.c_float, ctypes.POINTER(ctypes.c_int),
                  ctypes.POINTER(ctypes.c_uint32), ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_int)]
    f.restype = ctypes.c_uint32
    return f


def _loop(pats, off, cur, on, seg):
    """The reference's loop, one dash at a time, for every case in lock-step."""
    n = np.array([len(p) for p in pats])
    D = np.zeros((len(pats), n.max()), np.float32)
    for i, p in enumerate(pats):
        D[i, :len(p)] = p
    off = off.astype(np.float32).copy()
    cur = cur.copy()
    on = on.copy()
    caps = np.zeros(len(pats), np.int64)
    rows = np.arange(len(pats))
    while True:
        a = off < seg
        if not a.any():
            return off, cur, on, caps
        on[a] = ~on[a]
        off[a] = (off[a] + D[rows[a], cur[a]]).astype(np.float32)
        cur[a] += 1
        cur[a & (cur == n)] = 0
        caps[a] += 1


def _cases():
    rng = np.random.default_rng(5)
    fixed = [[8.0, 4.0], [3.3, 1.7, 0.1], [0.1], [7.5], [1e-3, 5.0], [0.0, 4.0], [5.0, 0.0, 2.0],
             [1000.5, 3.0], [2.0 ** -20, 1.0], [1.0, 2.0, 3.0, 4.0, 5.0, 6.0, 7.0]]
    pats, offs, curs, ons, segs = [], [], [], [], []
    for k in range(1500):
        if k % 3:
            p = fixed[k % len(fixed)]
        else:  # random dashes of random scale and precision
            m = int(rng.integers(1, 9))
            p = list((rng.uniform(0.05, 1, m) * 10.0 ** rng.uniform(-1, 2, m)).astype(np.float32))
            if k % 7 == 0:
                p[int(rng.integers(0, m))] = 0.0
            if sum(p) == 0:
                p[0] = 1.0
        p = [float(np.float32(v)) for v in p]
        tot = np.float32(sum(np.float32(v) for v in p))
        r = k % 5
        off = np.float32(rng.uniform(0, 1)) * tot  # an offset with a fine lowest bit
        if r == 1:
            off = np.float32(0.0)
        elif r == 2:
            off = np.float32(np.floor(off))  # a coarse one
        seg = np.float32(10.0 ** rng.uniform(-2, 4.2))
        if k % 11 == 0:
            seg = np.float32(tot * int(rng.integers(1, 400)))  # ends on a period boundary
        pats.append(p)
        offs.append(off)
        curs.append(int(rng.integers(0, len(p))))
        ons.append(bool(k & 1))
        segs.append(seg)
    return pats, np.array(offs, np.float32), np.array(curs), np.array(ons), np.array(segs, np.float32)


def test_dash_walk_matches_the_loop(L):
    pats, off, cur, on, seg = _cases()
    w_off, w_cur, w_on, w_caps = _loop(pats, off, cur, on, seg)
    jumped = 0
    for i, p in enumerate(pats):
        d = (ctypes.c_float * len(p))(*p)
        o = ctypes.c_float(off[i])
        c = ctypes.c_uint32(int(cur[i]))
        b = ctypes.c_int(int(on[i]))
        over = ctypes.c_int(7)
        caps = L(d, len(p), float(seg[i]), ctypes.byref(b), ctypes.byref(c), ctypes.byref(o), ctypes.byref(over))
        assert over.value == 0
        got = np.array([o.value], np.float32).view(np.uint32)[0]
        want = w_off[i:i + 1].view(np.uint32)[0]
        assert (caps, c.value, bool(b.value), got) == (w_caps[i], w_cur[i], bool(w_on[i]), want), (i, p, off[i], seg[i])
        jumped += caps > 4 * len(p)
    assert jumped > 500  # most segments span several periods, so the whole-period step is exercised


def test_dash_walk_stops_on_a_segment_its_dashes_cannot_cross(L):
    """Past 2^24 dash lengths the offset's float additions round the dashes away and the reference's
    loop never ends: the walk stops at its cap bound and flags the segment (the frame is refused)."""
    for p, seg in (([8.0, 4.0], 1e9), ([8.0, 4.0], float("inf")), ([0.5], 3e7), ([3.0, 0.0, 1.0], 1e12)):
        d = (ctypes.c_float * len(p))(*p)
        o, c, b, over = ctypes.c_float(0.0), ctypes.c_uint32(0), ctypes.c_int(1), ctypes.c_int(0)
        caps = L(d, len(p), seg, ctypes.byref(b), ctypes.byref(c), ctypes.byref(o), ctypes.byref(over))
        assert over.value == 1 and caps >= 1 << 24 and c.value < len(p), (p, seg, caps)
    # a long segment the dashes still cross ends normally
    d = (ctypes.c_float * 2)(8.0, 4.0)
    o, c, b, over = ctypes.c_float(0.0), ctypes.c_uint32(0), ctypes.c_int(1), ctypes.c_int(0)
    caps = L(d, 2, 6.0e6, ctypes.byref(b), ctypes.byref(c), ctypes.byref(o), ctypes.byref(over))
    assert over.value == 0 and caps == 1000000
