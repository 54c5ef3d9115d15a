"""Scene command stream: a flat uint32 array of opcodes with fixed operands.

One stream drives both the HIP library (``vkvg_amd_replay`` calls the public vkvg_* entry
points in order) and the CPU oracle, so both see exactly the same API call sequence.
Floats are stored bit-cast (IEEE binary32). Opcode values are part of the stream format.
"""
from __future__ import annotations

import numpy as np

OP_END = 0
OP_MOVE_TO = 1
OP_LINE_TO = 2
OP_CURVE_TO = 3
OP_CLOSE_PATH = 4
OP_ARC = 5
OP_ARC_NEG = 6
OP_RECTANGLE = 7
OP_FILL = 8
OP_STROKE = 9
OP_FILL_PRESERVE = 10
OP_STROKE_PRESERVE = 11
OP_PAINT = 12
OP_SET_SOURCE_RGBA = 13
OP_SET_SOURCE_RGB = 14
OP_SET_LINE_WIDTH = 15
OP_SET_LINE_JOIN = 16
OP_SET_LINE_CAP = 17
OP_SET_MITER_LIMIT = 18
OP_SET_FILL_RULE = 19
OP_SET_DASH = 20
OP_SET_OPACITY = 21
OP_SET_OPERATOR = 22
OP_LINEAR = 23
OP_RADIAL = 24
OP_NEW_PATH = 25
OP_NEW_SUB_PATH = 26
OP_REL_LINE_TO = 27
OP_REL_MOVE_TO = 28
OP_QUADRATIC_TO = 29
OP_TRANSLATE = 30
OP_SCALE = 31
OP_ROTATE = 32
OP_IDENTITY_MATRIX = 33
OP_CLEAR = 34
OP_ELLIPSE = 35
OP_REL_CURVE_TO = 36
OP_SET_MATRIX = 37
OP_ELLIPTIC_ARC_TO = 38
OP_ROUNDED_RECT = 39
OP_SET_SOURCE_COLOR = 40
OP_CLIP = 41
OP_CLIP_PRESERVE = 42
OP_RESET_CLIP = 43
OP_SAVE = 44
OP_RESTORE = 45
OP_FLUSH = 46
OP_SET_SOURCE_SURFACE = 47
OP_SOURCE_PATTERN = 48
OP_ROUNDED_RECT2 = 49
OP_LINEAR_MATRIX = 50  # pattern matrix (6 floats), then the operands of OP_LINEAR
OP_RADIAL_MATRIX = 51  # pattern matrix (6 floats), then the operands of OP_RADIAL

# enum values of include/vkvg.h
JOIN_MITER, JOIN_ROUND, JOIN_BEVEL = 0, 1, 2
CAP_BUTT, CAP_ROUND, CAP_SQUARE = 0, 1, 2
FILL_EVEN_ODD, FILL_NON_ZERO = 0, 1
OPERATOR_CLEAR, OPERATOR_SOURCE, OPERATOR_OVER, OPERATOR_DIFFERENCE = 0, 1, 2, 3
EXTEND_NONE, EXTEND_REPEAT, EXTEND_REFLECT, EXTEND_PAD = 0, 1, 2, 3
FILTER_FAST, FILTER_GOOD, FILTER_BEST, FILTER_NEAREST, FILTER_BILINEAR, FILTER_GAUSSIAN = 0, 1, 2, 3, 4, 5


def _f(*vals) -> np.ndarray:
    return np.asarray(vals, dtype=np.float32).view(np.uint32)


class Stream:
    """Builder mirroring the vkvg path/paint/state calls."""

    def __init__(self) -> None:
        self._chunks: list[np.ndarray] = []
        self.n_paths = 0  # fill/stroke calls recorded (for paths/s metrics)

    # -- low level
    def _op(self, op: int, *floats: float) -> None:
        self._chunks.append(np.concatenate([np.array([op], np.uint32), _f(*floats)]) if floats
                            else np.array([op], np.uint32))

    def _opu(self, op: int, *u32: int) -> None:
        self._chunks.append(np.array((op,) + tuple(int(v) for v in u32), np.uint32))

    def words(self) -> np.ndarray:
        if not self._chunks:
            return np.zeros(1, np.uint32)
        return np.ascontiguousarray(np.concatenate(self._chunks + [np.zeros(1, np.uint32)]))

    # -- path
    def move_to(self, x, y): self._op(OP_MOVE_TO, x, y)
    def line_to(self, x, y): self._op(OP_LINE_TO, x, y)
    def rel_line_to(self, x, y): self._op(OP_REL_LINE_TO, x, y)
    def rel_move_to(self, x, y): self._op(OP_REL_MOVE_TO, x, y)
    def curve_to(self, x1, y1, x2, y2, x3, y3): self._op(OP_CURVE_TO, x1, y1, x2, y2, x3, y3)
    def rel_curve_to(self, x1, y1, x2, y2, x3, y3): self._op(OP_REL_CURVE_TO, x1, y1, x2, y2, x3, y3)
    def quadratic_to(self, x1, y1, x2, y2): self._op(OP_QUADRATIC_TO, x1, y1, x2, y2)
    def close_path(self): self._op(OP_CLOSE_PATH)
    def new_path(self): self._op(OP_NEW_PATH)
    def new_sub_path(self): self._op(OP_NEW_SUB_PATH)
    def arc(self, xc, yc, r, a1, a2): self._op(OP_ARC, xc, yc, r, a1, a2)
    def arc_negative(self, xc, yc, r, a1, a2): self._op(OP_ARC_NEG, xc, yc, r, a1, a2)
    def rectangle(self, x, y, w, h): self._op(OP_RECTANGLE, x, y, w, h)
    def rounded_rectangle(self, x, y, w, h, r): self._op(OP_ROUNDED_RECT, x, y, w, h, r)
    def ellipse(self, rx, ry, x, y, rot): self._op(OP_ELLIPSE, rx, ry, x, y, rot)
    def rounded_rectangle2(self, x, y, w, h, rx, ry): self._op(OP_ROUNDED_RECT2, x, y, w, h, rx, ry)

    def elliptic_arc_to(self, x, y, large, sweep, rx, ry, phi):
        self._chunks.append(np.concatenate([np.array([OP_ELLIPTIC_ARC_TO], np.uint32), _f(x, y),
                                            np.array([int(large), int(sweep)], np.uint32), _f(rx, ry, phi)]))

    def polyline(self, xy: np.ndarray, close: bool = False) -> None:
        """move_to(xy[0]) then line_to for the rest (vectorised)."""
        xy = np.asarray(xy, np.float32).reshape(-1, 2)
        n = xy.shape[0]
        w = np.empty((n, 3), np.uint32)
        w[:, 0] = OP_LINE_TO
        w[0, 0] = OP_MOVE_TO
        w[:, 1:] = xy.view(np.uint32)
        self._chunks.append(w.reshape(-1))
        if close:
            self.close_path()

    def curves(self, p0, ctrl: np.ndarray, close: bool = False) -> None:
        """move_to(p0) then one curve_to per row of ctrl (n x 6)."""
        ctrl = np.asarray(ctrl, np.float32).reshape(-1, 6)
        n = ctrl.shape[0]
        w = np.empty((n, 7), np.uint32)
        w[:, 0] = OP_CURVE_TO
        w[:, 1:] = ctrl.view(np.uint32)
        self.move_to(float(p0[0]), float(p0[1]))
        self._chunks.append(w.reshape(-1))
        if close:
            self.close_path()

    # -- paint
    def fill(self): self._op(OP_FILL); self.n_paths += 1
    def stroke(self): self._op(OP_STROKE); self.n_paths += 1
    def fill_preserve(self): self._op(OP_FILL_PRESERVE); self.n_paths += 1
    def stroke_preserve(self): self._op(OP_STROKE_PRESERVE); self.n_paths += 1
    def paint(self): self._op(OP_PAINT)
    def clear(self): self._op(OP_CLEAR)

    # -- clip and graphic-state stack
    def clip(self): self._op(OP_CLIP)
    def clip_preserve(self): self._op(OP_CLIP_PRESERVE)
    def reset_clip(self): self._op(OP_RESET_CLIP)
    def save(self): self._op(OP_SAVE)
    def restore(self): self._op(OP_RESTORE)
    # -- surface sources (index into the surfaces passed with the stream)
    def set_source_surface(self, index, x, y):
        self._chunks.append(np.concatenate([np.array([OP_SET_SOURCE_SURFACE, index], np.uint32), _f(x, y)]))

    def set_source_pattern(self, index, extend=EXTEND_NONE, filt=FILTER_FAST, matrix=None):
        """vkvg_pattern_create_for_surface + set_extend / set_filter (/ set_matrix), set_source, destroy."""
        m = (1, 0, 0, 1, 0, 0) if matrix is None else tuple(matrix)
        self._chunks.append(np.concatenate([np.array([OP_SOURCE_PATTERN, index, extend, filt, matrix is not None],
                                                     np.uint32), _f(*m)]))

    def flush(self):
        """vkvg_flush: renders what is recorded so far (the next draws go into a new frame)."""
        self._op(OP_FLUSH)

    # -- state
    def set_source_rgba(self, r, g, b, a): self._op(OP_SET_SOURCE_RGBA, r, g, b, a)
    def set_source_rgb(self, r, g, b): self._op(OP_SET_SOURCE_RGB, r, g, b)
    def set_source_color(self, c): self._opu(OP_SET_SOURCE_COLOR, c)
    def set_line_width(self, w): self._op(OP_SET_LINE_WIDTH, w)
    def set_line_join(self, j): self._opu(OP_SET_LINE_JOIN, j)
    def set_line_cap(self, c): self._opu(OP_SET_LINE_CAP, c)
    def set_miter_limit(self, m): self._op(OP_SET_MITER_LIMIT, m)
    def set_fill_rule(self, r): self._opu(OP_SET_FILL_RULE, r)
    def set_opacity(self, o): self._op(OP_SET_OPACITY, o)
    def set_operator(self, o): self._opu(OP_SET_OPERATOR, o)

    def set_dash(self, dashes, offset):
        d = list(dashes)
        self._chunks.append(np.concatenate([np.array([OP_SET_DASH, len(d)], np.uint32), _f(offset),
                                            _f(*d) if d else np.zeros(0, np.uint32)]))

    def _gradient(self, op, geom, stops, matrix=None):
        st = np.asarray(stops, np.float32).reshape(-1, 5)
        head = [np.array([op], np.uint32)] if matrix is None else \
            [np.array([op + (OP_LINEAR_MATRIX - OP_LINEAR)], np.uint32), _f(*matrix)]
        self._chunks.append(np.concatenate(head + [_f(*geom), np.array([st.shape[0]], np.uint32),
                                                   st.reshape(-1).view(np.uint32)]))

    def set_linear(self, x0, y0, x1, y1, stops, matrix=None):
        """create a linear pattern, add (offset, r, g, b, a) stops, set its matrix (xx, yx, xy, yy,
        x0, y0) when one is given, set it as source, destroy it."""
        self._gradient(OP_LINEAR, (x0, y0, x1, y1), stops, matrix)

    def set_radial(self, cx0, cy0, r0, cx1, cy1, r1, stops, matrix=None):
        self._gradient(OP_RADIAL, (cx0, cy0, r0, cx1, cy1, r1), stops, matrix)

    def translate(self, x, y): self._op(OP_TRANSLATE, x, y)
    def scale(self, x, y): self._op(OP_SCALE, x, y)
    def rotate(self, a): self._op(OP_ROTATE, a)
    def identity_matrix(self): self._op(OP_IDENTITY_MATRIX)
    def set_matrix(self, m6): self._op(OP_SET_MATRIX, *m6)
