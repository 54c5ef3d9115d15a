"""vkvg_amd — MI355X-native backend of the vkvg 2-D vector graphics C API.

The product is the shared library ``libvkvg.so`` (C API of include/vkvg.h, extensions of
include/vkvg_amd.h, HIP backend of include/vkvg_hip.h). This module is a thin ctypes layer
used by the tests, the benchmark and the differentiable-rendering harness. There is no
fallback: without the built library every entry point raises.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libvkvg.so")
_lib = None


class VkvgError(RuntimeError):
    pass


class Stats(ctypes.Structure):
    _fields_ = [("points", ctypes.c_uint64), ("stroke_vertices", ctypes.c_uint64),
                ("stroke_triangles", ctypes.c_uint64), ("edges", ctypes.c_uint64),
                ("path_tiles", ctypes.c_uint64), ("segments", ctypes.c_uint64),
                ("commands", ctypes.c_uint64), ("tiles", ctypes.c_uint64), ("tile_size", ctypes.c_uint32),
                ("ms_total", ctypes.c_float), ("ms_fine", ctypes.c_float), ("ms_device", ctypes.c_float),
                ("ms_stage", ctypes.c_float * 6), ("graph", ctypes.c_uint32), ("fine_words", ctypes.c_uint32)]

    STAGES = ("flatten", "geometry", "tile_setup", "binning", "coarse", "fine")

    def as_dict(self) -> dict:
        d = {k: getattr(self, k) for k, _ in self._fields_ if k != "ms_stage"}
        d["ms_stage"] = {n: float(v) for n, v in zip(self.STAGES, self.ms_stage)}
        return d


EXPORTED = [
    # include/vkvg.h
    "vkvg_matrix_init_identity", "vkvg_matrix_init", "vkvg_matrix_init_translate", "vkvg_matrix_init_scale",
    "vkvg_matrix_init_rotate", "vkvg_matrix_translate", "vkvg_matrix_scale", "vkvg_matrix_rotate",
    "vkvg_matrix_multiply", "vkvg_matrix_transform_distance", "vkvg_matrix_transform_point", "vkvg_matrix_invert",
    "vkvg_matrix_get_scale", "vkvg_device_create", "vkvg_device_destroy", "vkvg_device_status",
    "vkvg_device_reference", "vkvg_device_get_reference_count", "vkvg_device_set_dpy", "vkvg_device_get_dpy",
    "vkvg_surface_create", "vkvg_surface_create_from_image", "vkvg_surface_create_from_bitmap", "vkvg_surface_status", "vkvg_surface_reference",
    "vkvg_surface_get_reference_count", "vkvg_surface_destroy", "vkvg_surface_clear", "vkvg_surface_get_vk_image",
    "vkvg_surface_get_vk_format", "vkvg_surface_get_width", "vkvg_surface_get_height", "vkvg_surface_write_to_png",
    "vkvg_surface_write_to_memory", "vkvg_surface_resolve", "vkvg_create", "vkvg_destroy", "vkvg_status",
    "vkvg_status_to_string", "vkvg_reference", "vkvg_get_reference_count", "vkvg_flush", "vkvg_new_path",
    "vkvg_close_path", "vkvg_new_sub_path", "vkvg_get_current_point", "vkvg_path_extents", "vkvg_line_to", "vkvg_rel_line_to",
    "vkvg_move_to", "vkvg_rel_move_to", "vkvg_arc", "vkvg_arc_negative", "vkvg_curve_to", "vkvg_rel_curve_to",
    "vkvg_quadratic_to", "vkvg_rel_quadratic_to", "vkvg_rectangle", "vkvg_rounded_rectangle", "vkvg_ellipse",
    "vkvg_elliptic_arc_to", "vkvg_rel_elliptic_arc_to", "vkvg_rounded_rectangle2",
    "vkvg_get_required_instance_extensions", "vkvg_get_required_device_extensions", "vkvg_get_device_requirements", "vkvg_has_current_point", "vkvg_stroke",
    "vkvg_stroke_preserve", "vkvg_fill", "vkvg_fill_preserve", "vkvg_paint", "vkvg_clear", "vkvg_reset_clip",
    "vkvg_clip", "vkvg_clip_preserve", "vkvg_save", "vkvg_restore", "vkvg_set_opacity",
    "vkvg_get_opacity", "vkvg_set_source_color", "vkvg_set_source_rgba", "vkvg_set_source_rgb", "vkvg_set_source",
    "vkvg_set_source_surface", "vkvg_pattern_create_for_surface",
    "vkvg_set_line_width", "vkvg_set_miter_limit", "vkvg_get_miter_limit", "vkvg_set_line_cap",
    "vkvg_set_line_join", "vkvg_set_operator", "vkvg_set_fill_rule", "vkvg_set_dash", "vkvg_get_dash",
    "vkvg_get_line_width", "vkvg_get_line_cap", "vkvg_get_line_join", "vkvg_get_operator", "vkvg_get_fill_rule",
    "vkvg_get_source", "vkvg_get_target", "vkvg_translate", "vkvg_scale", "vkvg_rotate", "vkvg_transform",
    "vkvg_set_matrix", "vkvg_get_matrix", "vkvg_identity_matrix", "vkvg_pattern_status", "vkvg_pattern_reference",
    "vkvg_pattern_get_reference_count", "vkvg_pattern_create_linear", "vkvg_pattern_edit_linear",
    "vkvg_pattern_get_linear_points", "vkvg_pattern_create_radial", "vkvg_pattern_edit_radial",
    "vkvg_pattern_get_color_stop_count", "vkvg_pattern_get_color_stop_rgba", "vkvg_pattern_destroy",
    "vkvg_pattern_add_color_stop", "vkvg_pattern_set_extend", "vkvg_pattern_set_filter", "vkvg_pattern_get_extend",
    "vkvg_pattern_get_filter", "vkvg_pattern_get_type", "vkvg_pattern_set_matrix", "vkvg_pattern_get_matrix",
    "vkvg_device_set_context_cache_size", "vkvg_select_font_face", "vkvg_load_font_from_path",
    "vkvg_load_font_from_memory", "vkvg_set_font_size", "vkvg_show_text", "vkvg_text_extents", "vkvg_font_extents",
    "vkvg_text_run_create", "vkvg_text_run_create_with_length", "vkvg_text_run_destroy", "vkvg_show_text_run",
    "vkvg_text_run_get_extents", "vkvg_text_run_get_glyph_count", "vkvg_text_run_get_glyph_position",
    "vkvg_start_recording", "vkvg_stop_recording", "vkvg_replay", "vkvg_replay_command", "vkvg_recording_get_command",
    "vkvg_recording_get_count", "vkvg_recording_get_data", "vkvg_recording_destroy",
    # include/vkvg_amd.h
    "vkvg_amd_replay", "vkvg_amd_replay_sources", "vkvg_amd_surface_enable_signature", "vkvg_amd_surface_read_signature",
    "vkvg_amd_surface_read_premultiplied", "vkvg_amd_surface_device_ptr", "vkvg_amd_render_batch",
    "vkvg_amd_device_get_stats", "vkvg_amd_device_set_tile_size", "vkvg_amd_device_engine",
    "vkvg_amd_upload", "vkvg_amd_execute", "vkvg_amd_upload_batch", "vkvg_amd_png_decode", "vkvg_amd_png_free",
    # include/vkvg_hip.h
    "vkvg_hip_engine_create", "vkvg_hip_engine_destroy", "vkvg_hip_set_tile_size", "vkvg_hip_set_stream",
    "vkvg_hip_image_alloc", "vkvg_hip_image_free", "vkvg_hip_image_clear", "vkvg_hip_stencil_clear", "vkvg_hip_image_download",
    "vkvg_hip_image_upload", "vkvg_hip_image_read_unpremultiplied", "vkvg_hip_image_read_unpremultiplied_noalloc", "vkvg_hip_render", "vkvg_hip_upload",
    "vkvg_hip_execute", "vkvg_hip_get_stats",
    "vkvg_hip_last_error", "vkvg_hip_item_extents",
    # engine tuning (include/vkvg_hip.h) and the differentiable renderer (include/vkvg_diff.h)
    "vkvg_hip_autotune", "vkvg_hip_retile", "vkvg_hip_set_fine_words", "vkvg_hip_set_graph",
    "vkvg_diff_create", "vkvg_diff_destroy", "vkvg_diff_forward", "vkvg_diff_forward_f64", "vkvg_diff_backward", "vkvg_diff_last_ms", "vkvg_diff_stats", "vkvg_diff_last_error",
]


def lib() -> ctypes.CDLL:
    """Loads libvkvg.so; raises when it has not been built.

    torch ships its own copy of the HIP and HSA runtimes (torch/lib/libamdhip64.so, SONAME
    libamdhip64.so.7). When torch is loaded first, libvkvg.so's libamdhip64.so.7 dependency
    resolves to that copy and the process holds one HIP runtime. When libvkvg.so is loaded first,
    it maps ROCm's libamdhip64.so.7 and a later torch import maps its own copy as well: two HIP and
    two HSA runtimes in one process, after which device creation can fail. So torch, when it is
    installed, is imported before the library is loaded."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise VkvgError(f"{LIB_PATH} is missing: build it with `python -m vkvg_amd.build` "
                        "(the HIP backend has no CPU fallback)")
    try:
        import torch  # noqa: F401  (one HIP runtime per process: torch's, see above)
    except ImportError:
        pass
    L = ctypes.CDLL(LIB_PATH)
    vp, u32, f32, sz = ctypes.c_void_p, ctypes.c_uint32, ctypes.c_float, ctypes.c_size_t
    L.vkvg_device_create.argtypes = [vp]
    L.vkvg_device_create.restype = vp
    L.vkvg_device_status.argtypes = [vp]
    L.vkvg_device_status.restype = ctypes.c_int
    L.vkvg_device_destroy.argtypes = [vp]
    L.vkvg_surface_create.argtypes = [vp, u32, u32]
    L.vkvg_surface_create.restype = vp
    L.vkvg_surface_status.argtypes = [vp]
    L.vkvg_surface_status.restype = ctypes.c_int
    L.vkvg_surface_destroy.argtypes = [vp]
    L.vkvg_surface_write_to_memory.argtypes = [vp, vp]
    L.vkvg_surface_write_to_memory.restype = ctypes.c_int
    L.vkvg_surface_write_to_png.argtypes = [vp, ctypes.c_char_p]
    L.vkvg_surface_write_to_png.restype = ctypes.c_int
    L.vkvg_surface_clear.argtypes = [vp]
    L.vkvg_create.argtypes = [vp]
    L.vkvg_create.restype = vp
    L.vkvg_destroy.argtypes = [vp]
    L.vkvg_flush.argtypes = [vp]
    L.vkvg_status.argtypes = [vp]
    L.vkvg_status.restype = ctypes.c_int
    L.vkvg_amd_replay.argtypes = [vp, ctypes.POINTER(u32), sz]
    L.vkvg_amd_replay.restype = ctypes.c_int
    L.vkvg_amd_replay_sources.argtypes = [vp, ctypes.POINTER(u32), sz, ctypes.POINTER(vp), u32]
    L.vkvg_amd_replay_sources.restype = ctypes.c_int
    L.vkvg_surface_create_from_bitmap.argtypes = [vp, vp, u32, u32]
    L.vkvg_surface_create_from_bitmap.restype = vp
    L.vkvg_amd_surface_enable_signature.argtypes = [vp]
    L.vkvg_amd_surface_enable_signature.restype = ctypes.c_int
    L.vkvg_amd_surface_read_signature.argtypes = [vp, vp]
    L.vkvg_amd_surface_read_signature.restype = ctypes.c_int
    L.vkvg_amd_surface_read_premultiplied.argtypes = [vp, vp]
    L.vkvg_amd_surface_read_premultiplied.restype = ctypes.c_int
    L.vkvg_amd_surface_device_ptr.argtypes = [vp]
    L.vkvg_amd_surface_device_ptr.restype = vp
    L.vkvg_amd_render_batch.argtypes = [vp, u32, ctypes.POINTER(vp), ctypes.POINTER(ctypes.POINTER(u32)),
                                        ctypes.POINTER(sz)]
    L.vkvg_amd_render_batch.restype = ctypes.c_int
    L.vkvg_amd_device_get_stats.argtypes = [vp, ctypes.POINTER(Stats)]
    L.vkvg_amd_device_set_tile_size.argtypes = [vp, u32]
    L.vkvg_amd_device_set_tile_size.restype = ctypes.c_int
    L.vkvg_amd_device_engine.argtypes = [vp]
    L.vkvg_amd_device_engine.restype = vp
    L.vkvg_amd_upload.argtypes = [vp]
    L.vkvg_amd_upload.restype = ctypes.c_int
    L.vkvg_amd_execute.argtypes = [vp]
    L.vkvg_amd_execute.restype = ctypes.c_int
    L.vkvg_amd_upload_batch.argtypes = L.vkvg_amd_render_batch.argtypes
    L.vkvg_amd_upload_batch.restype = ctypes.c_int
    L.vkvg_hip_set_stream.argtypes = [vp, vp]
    L.vkvg_hip_last_error.argtypes = [vp]
    L.vkvg_hip_last_error.restype = ctypes.c_char_p
    _lib = L
    return L


def _u32p(a: np.ndarray):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32))


class Device:
    def __init__(self) -> None:
        L = lib()
        self._info = (ctypes.c_uint8 * 48)()  # vkvg_device_create_info_t, zeroed: 1 sample, no Vulkan handles
        self.h = L.vkvg_device_create(self._info)
        st = L.vkvg_device_status(self.h)
        if st:
            raise VkvgError(f"vkvg_device_create failed (status {st})")

    def close(self) -> None:
        if self.h:
            lib().vkvg_device_destroy(self.h)
            self.h = None

    def set_tile_size(self, t: int) -> None:
        if lib().vkvg_amd_device_set_tile_size(self.h, t):
            raise VkvgError(f"bad tile size {t}")

    def stats(self) -> dict:
        s = Stats()
        lib().vkvg_amd_device_get_stats(self.h, ctypes.byref(s))
        return s.as_dict()

    def _engine(self):
        L = lib()
        L.vkvg_amd_device_engine.restype = ctypes.c_void_p
        L.vkvg_amd_device_engine.argtypes = [ctypes.c_void_p]
        e = L.vkvg_amd_device_engine(self.h)
        if not e:
            raise VkvgError("device has no engine")
        return e

    def set_graph(self, on: bool) -> None:
        """Replay the resident frame as one captured hipGraph after its first execution."""
        L = lib()
        L.vkvg_hip_set_graph.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.vkvg_hip_set_graph(self._engine(), 1 if on else 0)

    def set_fine_words(self, words: int) -> None:
        L = lib()
        L.vkvg_hip_set_fine_words.argtypes = [ctypes.c_void_p, ctypes.c_uint32]
        if L.vkvg_hip_set_fine_words(self._engine(), words):
            raise VkvgError(f"bad fine capacity {words}")

    def autotune(self, reps: int = 3) -> dict:
        """Times tile 8/16/32 x fine capacity 2048/4096/8192 words on the resident frame and keeps
        the fastest; returns the table and the choice."""
        class Entry(ctypes.Structure):
            _fields_ = [("tile", ctypes.c_uint32), ("fine_words", ctypes.c_uint32), ("ms", ctypes.c_float)]

        class Tune(ctypes.Structure):
            _fields_ = [("n", ctypes.c_uint32), ("best", ctypes.c_uint32), ("e", Entry * 16)]
        L = lib()
        L.vkvg_hip_autotune.argtypes = [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_void_p]
        L.vkvg_hip_autotune.restype = ctypes.c_int
        t = Tune()
        if L.vkvg_hip_autotune(self._engine(), reps, ctypes.byref(t)):
            L.vkvg_hip_last_error.restype = ctypes.c_char_p
            L.vkvg_hip_last_error.argtypes = [ctypes.c_void_p]
            raise VkvgError(f"autotune failed: {L.vkvg_hip_last_error(self._engine()).decode()}")
        table = [{"tile": t.e[k].tile, "fine_words": t.e[k].fine_words, "ms": float(t.e[k].ms)} for k in range(t.n)]
        return {"best": table[t.best], "table": table}

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


class Surface:
    def __init__(self, dev: Device, width: int, height: int, signature: bool = False) -> None:
        L = lib()
        self.dev = dev
        self.width, self.height = width, height
        self.h = L.vkvg_surface_create(dev.h, width, height)
        st = L.vkvg_surface_status(self.h)
        if st:
            raise VkvgError(f"vkvg_surface_create failed (status {st})")
        if signature and L.vkvg_amd_surface_enable_signature(self.h):
            raise VkvgError("cannot enable coverage signatures")

    @classmethod
    def from_bitmap(cls, dev: "Device", rgba: np.ndarray, signature: bool = False) -> "Surface":
        """vkvg_surface_create_from_bitmap: the bytes (H x W x 4) become the surface's pixels."""
        L = lib()
        a = np.ascontiguousarray(rgba, np.uint8)
        h, w = a.shape[:2]
        self = cls.__new__(cls)
        self.dev, self.width, self.height = dev, w, h
        self.h = L.vkvg_surface_create_from_bitmap(dev.h, a.ctypes.data, w, h)
        st = L.vkvg_surface_status(self.h)
        if st:
            raise VkvgError(f"vkvg_surface_create_from_bitmap failed (status {st})")
        if signature and L.vkvg_amd_surface_enable_signature(self.h):
            raise VkvgError("cannot enable coverage signatures")
        return self

    @classmethod
    def from_image(cls, dev: "Device", path: str) -> "Surface":
        """vkvg_surface_create_from_image: a PNG file's pixels (8-bit RGBA as decoded)."""
        L = lib()
        L.vkvg_surface_create_from_image.argtypes = [ctypes.c_void_p, ctypes.c_char_p]
        L.vkvg_surface_create_from_image.restype = ctypes.c_void_p
        for f in (L.vkvg_surface_get_width, L.vkvg_surface_get_height):
            f.argtypes = [ctypes.c_void_p]
            f.restype = ctypes.c_uint32
        self = cls.__new__(cls)
        self.dev = dev
        self.h = L.vkvg_surface_create_from_image(dev.h, os.fsencode(path))
        st = L.vkvg_surface_status(self.h)
        if st:
            raise VkvgError(f"vkvg_surface_create_from_image failed (status {st})")
        self.width, self.height = L.vkvg_surface_get_width(self.h), L.vkvg_surface_get_height(self.h)
        return self

    def replay(self, words: np.ndarray, sources: list | None = None) -> None:
        """vkvg_create + replay of the stream + vkvg_destroy (which renders the frame); `sources` are
        the surfaces the stream's surface-source opcodes index."""
        L = lib()
        w = np.ascontiguousarray(words, np.uint32)
        ctx = L.vkvg_create(self.h)
        st = L.vkvg_status(ctx)
        if st:
            raise VkvgError(f"vkvg_create failed (status {st})")
        srcs = sources or []
        hs = (ctypes.c_void_p * max(1, len(srcs)))(*[s.h for s in srcs])
        st = L.vkvg_amd_replay_sources(ctx, _u32p(w), w.size, hs, len(srcs))
        L.vkvg_destroy(ctx)
        if st:
            raise VkvgError(f"replay failed (status {st}): "
                            f"{L.vkvg_hip_last_error(L.vkvg_amd_device_engine(self.dev.h)).decode()}")

    def upload(self, words: np.ndarray) -> None:
        """Replays the stream and leaves its frame resident on the device (not rendered)."""
        L = lib()
        w = np.ascontiguousarray(words, np.uint32)
        ctx = L.vkvg_create(self.h)
        st = L.vkvg_status(ctx)
        if st:
            raise VkvgError(f"vkvg_create failed (status {st})")
        st = L.vkvg_amd_replay(ctx, _u32p(w), w.size)
        if not st:
            st = L.vkvg_amd_upload(ctx)
        L.vkvg_destroy(ctx)
        if st:
            raise VkvgError(f"upload failed (status {st})")

    def premultiplied(self) -> np.ndarray:
        out = np.empty((self.height, self.width, 4), np.uint8)
        if lib().vkvg_amd_surface_read_premultiplied(self.h, out.ctypes.data):
            raise VkvgError("readback failed")
        return out

    def rgba(self) -> np.ndarray:
        """vkvg_surface_write_to_memory: RGBA with the reference's truncating unpremultiply."""
        out = np.empty((self.height, self.width, 4), np.uint8)
        if lib().vkvg_surface_write_to_memory(self.h, out.ctypes.data):
            raise VkvgError("readback failed")
        return out

    def signature(self) -> np.ndarray:
        out = np.empty((self.height, self.width), np.uint32)
        if lib().vkvg_amd_surface_read_signature(self.h, out.ctypes.data):
            raise VkvgError("signature readback failed (not enabled?)")
        return out

    def write_png(self, path: str) -> None:
        if lib().vkvg_surface_write_to_png(self.h, path.encode()):
            raise VkvgError("write_to_png failed")

    def device_ptr(self) -> int:
        return int(lib().vkvg_amd_surface_device_ptr(self.h))

    def close(self) -> None:
        if self.h:
            lib().vkvg_surface_destroy(self.h)
            self.h = None


def execute(dev: Device) -> None:
    """Runs the pipeline on the device's resident frame."""
    L = lib()
    if L.vkvg_amd_execute(dev.h):
        raise VkvgError(L.vkvg_hip_last_error(L.vkvg_amd_device_engine(dev.h)).decode())


def upload_batch(dev: Device, surfaces: list, streams: list) -> None:
    L = lib()
    n = len(surfaces)
    ws = [np.ascontiguousarray(s, np.uint32) for s in streams]
    hs = (ctypes.c_void_p * n)(*[s.h for s in surfaces])
    ps = (ctypes.POINTER(ctypes.c_uint32) * n)(*[_u32p(w) for w in ws])
    ls = (ctypes.c_size_t * n)(*[w.size for w in ws])
    st = L.vkvg_amd_upload_batch(dev.h, n, hs, ps, ls)
    if st:
        raise VkvgError(f"upload_batch failed (status {st})")


def render_batch(dev: Device, surfaces: list[Surface], streams: list[np.ndarray]) -> None:
    L = lib()
    n = len(surfaces)
    ws = [np.ascontiguousarray(s, np.uint32) for s in streams]
    hs = (ctypes.c_void_p * n)(*[s.h for s in surfaces])
    ps = (ctypes.POINTER(ctypes.c_uint32) * n)(*[_u32p(w) for w in ws])
    ls = (ctypes.c_size_t * n)(*[w.size for w in ws])
    st = L.vkvg_amd_render_batch(dev.h, n, hs, ps, ls)
    if st:
        raise VkvgError(f"render_batch failed (status {st})")
