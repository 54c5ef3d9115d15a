"""ctypes binding of the CPU oracle (test infrastructure only).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import this module.
The shipped library never loads the oracle.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIBS: dict[str, ctypes.CDLL] = {}
_OVERRIDE: list[str] = []  # a library path that replaces libovg.so (mutation tests)


def build(quiet: bool = True) -> None:
    subprocess.run(["make", "-C", _HERE, "-s"], check=True,
                   stdout=subprocess.DEVNULL if quiet else None)


class using:
    """Context manager: every call inside it goes to the oracle library at `path` (a build of
    the oracle's sources with a deliberate change, for the mutation tests)."""

    def __init__(self, path: str) -> None:
        self.path = path

    def __enter__(self):
        _OVERRIDE.append(self.path)
        return self

    def __exit__(self, *exc):
        _OVERRIDE.pop()
        return False


def _lib(variant: str = "") -> ctypes.CDLL:
    if _OVERRIDE:
        name = path = _OVERRIDE[-1]
    else:
        name = "libovg_glibc.so" if variant == "glibc" else "libovg.so"
        path = os.path.join(_HERE, name)
    if name in _LIBS:
        return _LIBS[name]
    if not os.path.exists(path):
        build()
    lib = ctypes.CDLL(path)
    u32p = ctypes.POINTER(ctypes.c_uint32)
    i32p = ctypes.POINTER(ctypes.c_int32)
    lib.ovg_render_stream.argtypes = [ctypes.c_uint32, ctypes.c_uint32, u32p, ctypes.c_size_t,
                                      ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
    lib.ovg_render_stream.restype = ctypes.c_int
    lib.ovg_render_stream_src.argtypes = [ctypes.c_uint32, ctypes.c_uint32, u32p, ctypes.c_size_t, ctypes.c_void_p,
                                          ctypes.c_uint32, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
    lib.ovg_render_stream_src.restype = ctypes.c_int
    lib.ovg_render_stream_rows.argtypes = [ctypes.c_uint32, ctypes.c_uint32, u32p, ctypes.c_size_t, ctypes.c_uint32,
                                           ctypes.c_uint32, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
    lib.ovg_render_stream_rows.restype = ctypes.c_int
    lib.ovg_unpremultiply.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t]
    lib.ovg_draw_key.argtypes = [ctypes.c_uint32]
    lib.ovg_draw_key.restype = ctypes.c_uint32
    lib.ovg_snap_vertex.argtypes = [ctypes.c_float, ctypes.c_float, ctypes.c_uint32, ctypes.c_uint32,
                                    ctypes.POINTER(ctypes.c_float), i32p, i32p]
    lib.ovg_sample_in_triangle.argtypes = [i32p, i32p, ctypes.c_int32, ctypes.c_int32]
    lib.ovg_sample_in_triangle.restype = ctypes.c_int
    for fn in ("ovg_raster_triangle_scan", "ovg_raster_triangle_brute"):
        getattr(lib, fn).argtypes = [i32p, i32p, ctypes.c_uint32, ctypes.c_uint32, i32p]
    for fn in ("ovg_winding_scan", "ovg_winding_fan"):
        getattr(lib, fn).argtypes = [i32p, i32p, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, i32p]
    lib.ovg_flatten_cubic.argtypes = [ctypes.c_float] * 9 + [ctypes.POINTER(ctypes.c_float), ctypes.c_uint32]
    lib.ovg_flatten_cubic.restype = ctypes.c_uint32
    lib.ovg_flatten_cubic_level.argtypes = [ctypes.c_float] * 9 + [ctypes.c_uint32, ctypes.POINTER(ctypes.c_float),
                                                                    ctypes.c_uint32]
    lib.ovg_flatten_cubic_level.restype = ctypes.c_uint32
    lib.ovg_geometry_dump.argtypes = [ctypes.c_uint32, ctypes.c_uint32, u32p, ctypes.c_size_t,
                                      ctypes.POINTER(_Geometry)]
    lib.ovg_geometry_dump.restype = ctypes.c_int
    lib.ovg_blend_over.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
    _LIBS[name] = lib
    return lib


class _Geometry(ctypes.Structure):
    _fields_ = [("points", ctypes.c_void_p), ("cap_points", ctypes.c_uint32), ("n_points", ctypes.c_uint32),
                ("pathes", ctypes.c_void_p), ("cap_pathes", ctypes.c_uint32), ("n_pathes", ctypes.c_uint32),
                ("verts", ctypes.c_void_p), ("cap_verts", ctypes.c_uint32), ("n_verts", ctypes.c_uint32),
                ("inds", ctypes.c_void_p), ("cap_inds", ctypes.c_uint32), ("n_inds", ctypes.c_uint32)]


def _p(a: np.ndarray, t=ctypes.c_uint32):
    return a.ctypes.data_as(ctypes.POINTER(t))


class _Source(ctypes.Structure):
    _fields_ = [("rgba", ctypes.c_void_p), ("width", ctypes.c_uint32), ("height", ctypes.c_uint32)]


def render(width: int, height: int, words: np.ndarray, variant: str = "", with_sig: bool = True,
           sources: list | None = None):
    """Returns (premultiplied RGBA8 HxWx4, signature HxW uint32 or None, stats dict, status).
    `sources`: HxWx4 uint8 images the stream's surface-source opcodes index."""
    lib = _lib(variant)
    words = np.ascontiguousarray(words, np.uint32)
    rgba = np.zeros((height, width, 4), np.uint8)
    sig = np.zeros((height, width), np.uint32) if with_sig else None
    stats = np.zeros(3, np.uint64)
    srcs = [np.ascontiguousarray(s, np.uint8) for s in (sources or [])]
    tab = (_Source * max(1, len(srcs)))(*[_Source(s.ctypes.data, s.shape[1], s.shape[0]) for s in srcs])
    st = lib.ovg_render_stream_src(width, height, _p(words), words.size, ctypes.addressof(tab), len(srcs),
                                   rgba.ctypes.data, sig.ctypes.data if sig is not None else None,
                                   stats.ctypes.data)
    return rgba, sig, {"draws": int(stats[0]), "triangles": int(stats[1]), "fragments": int(stats[2])}, st


def render_bands(width: int, height: int, words: np.ndarray, threads: int, with_sig: bool = True):
    """The same image as render(), rasterized as `threads` bands of rows on as many threads (the
    C calls release the GIL); every band repeats the geometry. Returns (rgba, sig or None, status)."""
    from concurrent.futures import ThreadPoolExecutor
    lib = _lib()
    words = np.ascontiguousarray(words, np.uint32)
    rgba = np.zeros((height, width, 4), np.uint8)
    sig = np.zeros((height, width), np.uint32) if with_sig else None
    edges = [height * k // threads for k in range(threads + 1)]

    def band(k):
        r0, r1 = edges[k], edges[k + 1]
        if r0 == r1:
            return 0
        b_rgba = np.zeros((height, width, 4), np.uint8)
        b_sig = np.zeros((height, width), np.uint32) if with_sig else None
        st = lib.ovg_render_stream_rows(width, height, _p(words), words.size, r0, r1, b_rgba.ctypes.data,
                                        b_sig.ctypes.data if b_sig is not None else None, None)
        rgba[r0:r1] = b_rgba[r0:r1]
        if with_sig:
            sig[r0:r1] = b_sig[r0:r1]
        return st

    with ThreadPoolExecutor(max_workers=threads) as ex:
        sts = list(ex.map(band, range(threads)))
    return rgba, sig, max(sts) if sts else 0


def unpremultiply(rgba: np.ndarray) -> np.ndarray:
    lib = _lib()
    src = np.ascontiguousarray(rgba, np.uint8)
    out = np.empty_like(src)
    lib.ovg_unpremultiply(src.ctypes.data, out.ctypes.data, src.size // 4)
    return out


def draw_key(n: int) -> int:
    return int(_lib().ovg_draw_key(n))


def snap(x: float, y: float, w: int, h: int, mat6=(1, 0, 0, 1, 0, 0)) -> tuple[int, int]:
    lib = _lib()
    m = (ctypes.c_float * 6)(*mat6)
    fx, fy = ctypes.c_int32(), ctypes.c_int32()
    lib.ovg_snap_vertex(x, y, w, h, m, ctypes.byref(fx), ctypes.byref(fy))
    return fx.value, fy.value


def raster_triangle(fx, fy, w, h, method: str = "scan") -> np.ndarray:
    lib = _lib()
    X = np.ascontiguousarray(fx, np.int32)
    Y = np.ascontiguousarray(fy, np.int32)
    cov = np.zeros((h, w), np.int32)
    fn = lib.ovg_raster_triangle_scan if method == "scan" else lib.ovg_raster_triangle_brute
    fn(_p(X, ctypes.c_int32), _p(Y, ctypes.c_int32), w, h, _p(cov, ctypes.c_int32))
    return cov


def sample_in_triangle(fx, fy, sx, sy) -> bool:
    X = np.ascontiguousarray(fx, np.int32)
    Y = np.ascontiguousarray(fy, np.int32)
    return bool(_lib().ovg_sample_in_triangle(_p(X, ctypes.c_int32), _p(Y, ctypes.c_int32), sx, sy))


def winding(px, py, w, h, method: str = "scan") -> np.ndarray:
    lib = _lib()
    X = np.ascontiguousarray(px, np.int32)
    Y = np.ascontiguousarray(py, np.int32)
    out = np.zeros((h, w), np.int32)
    fn = lib.ovg_winding_scan if method == "scan" else lib.ovg_winding_fan
    fn(_p(X, ctypes.c_int32), _p(Y, ctypes.c_int32), X.size, w, h, _p(out, ctypes.c_int32))
    return out


def flatten_cubic(p, tol: float = 0.25, cap: int = 1 << 16, variant: str = "") -> np.ndarray:
    lib = _lib(variant)
    out = np.zeros(2 * cap, np.float32)
    n = lib.ovg_flatten_cubic(*[float(v) for v in p], tol, _p(out, ctypes.c_float), cap)
    return out[:2 * min(n, cap)].reshape(-1, 2)


def flatten_cubic_level(p, level: int, tol: float = 0.25, cap: int = 1 << 16) -> np.ndarray:
    """AGG subdivision started at `level` (>= 1: the flatness tests apply to `p` itself)."""
    out = np.zeros(2 * cap, np.float32)
    n = _lib().ovg_flatten_cubic_level(*[float(v) for v in p], tol, level, _p(out, ctypes.c_float), cap)
    return out[:2 * min(n, cap)].reshape(-1, 2)


P_CLOSED, P_CURVES, P_CONVEX, P_MASK = 0x80000000, 0x40000000, 0x20000000, 0x1FFFFFFF


def geometry(words: np.ndarray, width: int = 256, height: int = 256, cap: int = 1 << 16):
    """Geometry of the stream's first fill / stroke before rasterization: dict with the path's
    points (n x 2 float32), path records (uint32, CLOSED / CURVES / CONVEX bits), and for a
    stroke the emitted vertices (n x 2 float32) and triangles (n x 3 uint32)."""
    words = np.ascontiguousarray(words, np.uint32)
    pts = np.zeros(2 * cap, np.float32)
    pth = np.zeros(cap, np.uint32)
    vts = np.zeros(2 * cap, np.float32)
    ind = np.zeros(3 * cap, np.uint32)
    g = _Geometry(pts.ctypes.data, cap, 0, pth.ctypes.data, cap, 0, vts.ctypes.data, cap, 0, ind.ctypes.data, 3 * cap, 0)
    st = _lib().ovg_geometry_dump(width, height, _p(words), words.size, ctypes.byref(g))
    assert g.n_points <= cap and g.n_verts <= cap and g.n_inds <= 3 * cap
    return {"status": st, "points": pts[:2 * g.n_points].reshape(-1, 2), "pathes": pth[:g.n_pathes].copy(),
            "verts": vts[:2 * g.n_verts].reshape(-1, 2), "tris": ind[:g.n_inds].reshape(-1, 3)}


def blend_over(d, s) -> np.ndarray:
    dd = np.array(d, np.uint8)
    ss = np.array(s, np.uint8)
    _lib().ovg_blend_over(dd.ctypes.data, ss.ctypes.data)
    return dd
