"""Strokes whose reference loops would not end: a dashed segment too long for its dash pattern is
refused with an error (past some 2^24 dash lengths the float additions of the dash loop,
src/vkvg_context_internal.c:1246-1259, round every dash away; the device's dash walk stops at its
cap bound and the frame is rejected before any geometry is emitted), and the next frame renders
normally; a round-capped stroke so wide that the reference's arc step rounds away
(src/vkvg_context_internal.c:245-252, :1165-1239) renders with the arc step raised to
VKH_MIN_ARC_STEP."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _dashed_line(x1):
    from vkvg_amd.stream import Stream
    s = Stream()
    s.set_source_rgba(0.0, 0.0, 0.0, 1.0)
    s.set_line_width(2.0)
    s.set_dash([8.0, 4.0], 0.0)
    s.polyline(np.array([[0.0, 10.0], [x1, 10.0]], np.float32))
    s.stroke()
    return s.words()


def test_endless_dash_walk_is_refused(device):
    import vkvg_amd
    surf = vkvg_amd.Surface(device, 64, 64)
    try:
        print("case dashed segment of 1e9 px", flush=True)
        with pytest.raises(vkvg_amd.VkvgError):
            surf.upload(_dashed_line(1.0e9))
            vkvg_amd.execute(device)
        L = vkvg_amd.lib()
        msg = L.vkvg_hip_last_error(L.vkvg_amd_device_engine(device.h)).decode()
        assert "dash" in msg, msg
        print("case dashed segment of 60 px after it", flush=True)
        surf.replay(_dashed_line(60.0))
        img = surf.premultiplied()
        assert img[10, 2, 3] == 255 and img[10, 10, 3] == 0  # in the first dash, in the first gap
    finally:
        surf.close()


def test_round_caps_of_a_huge_width_render(device):
    import vkvg_amd
    from vkvg_amd.stream import CAP_ROUND, Stream
    s = Stream()
    s.set_source_rgba(0.0, 0.0, 1.0, 1.0)
    s.set_line_width(1.0e9)
    s.set_line_cap(CAP_ROUND)
    s.polyline(np.array([[10.0, 30.0], [50.0, 30.0]], np.float32))
    s.stroke()
    surf = vkvg_amd.Surface(device, 64, 64)
    try:
        print("case round caps of width 1e9", flush=True)
        surf.replay(s.words())
        img = surf.premultiplied()
        assert img[30, 30, 3] == 255 and img[0, 0, 3] == 255 and img[63, 63, 3] == 255
    finally:
        surf.close()
