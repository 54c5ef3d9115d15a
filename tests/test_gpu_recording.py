"""Recording API (include/vkvg.h:1960-1970) on the GPU: a scene's calls are recorded on one context
(which draws nothing meanwhile), replayed on another, and the result compared with the oracle's
rendering of the same calls (1-sample masks through the coverage signature, RGBA8)."""
import ctypes

import numpy as np
import pytest

from vkvg_amd import scenes
from parity_util import compare, render_oracle

pytestmark = pytest.mark.gpu


def _bind(L):
    vp = ctypes.c_void_p
    L.vkvg_start_recording.argtypes = [vp]
    L.vkvg_stop_recording.argtypes = [vp]
    L.vkvg_stop_recording.restype = vp
    L.vkvg_replay.argtypes = [vp, vp]
    L.vkvg_recording_get_count.argtypes = [vp]
    L.vkvg_recording_get_count.restype = ctypes.c_uint32
    L.vkvg_recording_destroy.argtypes = [vp]


@pytest.mark.parametrize("scene", [scenes.feature_mix(0), scenes.feature_mix(2), scenes.clip_mix(1, flushes=False),
                                   scenes.config2(n_lines=300, n_stars=100, width=500, height=300)],
                         ids=["feature_mix0", "feature_mix2", "clip_mix1", "c2_small"])
def test_recorded_scene_replays_like_the_calls(device, scene):
    import vkvg_amd
    L = vkvg_amd.lib()
    _bind(L)
    w = np.ascontiguousarray(scene.words(), np.uint32)
    rec_surf = vkvg_amd.Surface(device, scene.width, scene.height)
    out = vkvg_amd.Surface(device, scene.width, scene.height, signature=True)
    try:
        ctx = L.vkvg_create(rec_surf.h)
        L.vkvg_start_recording(ctx)
        assert L.vkvg_amd_replay(ctx, w.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)), w.size) == 0
        rec = L.vkvg_stop_recording(ctx)
        L.vkvg_destroy(ctx)
        assert rec and L.vkvg_recording_get_count(rec) > 0
        assert not rec_surf.premultiplied().any()  # nothing was drawn while recording
        ctx = L.vkvg_create(out.h)
        L.vkvg_replay(ctx, rec)
        assert L.vkvg_status(ctx) == 0
        L.vkvg_destroy(ctx)
        L.vkvg_recording_destroy(rec)
        g_rgba, g_sig = out.premultiplied(), out.signature()
    finally:
        rec_surf.close()
        out.close()
    o_rgba, o_sig, _ = render_oracle(scene)
    r = compare(g_rgba, g_sig, o_rgba, o_sig)
    print(scene.name, r, flush=True)
    assert r["mask_mismatch_px"] == 0 and r["rgba_max_err"] == 0, r
