"""GPU parity on the reference's own visual tests (tests/*.c call sequences restated as scenes by
vkvg_amd.scenes.reference_test_scenes): every scene through the HIP library, one by one and all
in one batch, bit-exact against the oracle (coverage signature and RGBA)."""
import pytest

from vkvg_amd import scenes
from parity_util import compare, render_gpu, render_oracle

pytestmark = pytest.mark.gpu


def _exact(name, r):
    assert r["mask_mismatch_px"] == 0 and r["rgba_max_err"] == 0, (name, r)


def test_reference_test_scenes_one_by_one(device):
    for sc in scenes.reference_test_scenes():
        print("case", sc.name, flush=True)
        g_rgba, g_sig, _ = render_gpu(device, sc)
        o_rgba, o_sig, _ = render_oracle(sc)
        _exact(sc.name, compare(g_rgba, g_sig, o_rgba, o_sig))


@pytest.mark.parametrize("tile", [8, 32])
def test_reference_test_scenes_on_a_wider_surface(device, tile):
    # 1024 x 800: the parts the tests draw beyond 512 x 512 (curve test2, the dash rectangles)
    for sc in scenes.reference_test_scenes(1024, 800):
        if not sc.name.startswith(("curve_", "dashes_", "stroke", "line_join")):
            continue
        print("case", sc.name, tile, flush=True)
        g_rgba, g_sig, _ = render_gpu(device, sc, tile)
        o_rgba, o_sig, _ = render_oracle(sc)
        _exact(sc.name, compare(g_rgba, g_sig, o_rgba, o_sig))


def test_reference_test_scenes_in_one_batch(device):
    import vkvg_amd
    subs = scenes.reference_test_scenes()
    surfs = [vkvg_amd.Surface(device, s.width, s.height, signature=True) for s in subs]
    try:
        device.set_tile_size(16)
        print("case reference scenes batch", len(subs), flush=True)
        vkvg_amd.render_batch(device, surfs, [s.words() for s in subs])
        for sc, s in zip(subs, surfs):
            o_rgba, o_sig, _ = render_oracle(sc)
            _exact(sc.name, compare(s.premultiplied(), s.signature(), o_rgba, o_sig))
    finally:
        for s in surfs:
            s.close()


def test_reference_source_scenes(device):
    # sources drawn by the library (as the reference's tests draw theirs with vkvg), the oracle's
    # sources drawn by the oracle; pattern matrices on surface patterns and on gradients
    import numpy as np
    import vkvg_amd
    from oracle import ovg
    device.set_tile_size(16)
    for sc, specs in scenes.reference_source_scenes():
        print("case", sc.name, flush=True)
        srcs, o_srcs = [], []
        t = vkvg_amd.Surface(device, sc.width, sc.height, signature=True)
        try:
            for w, h, st in specs:
                s = vkvg_amd.Surface(device, w, h)
                srcs.append(s)
                s.replay(st.words())
                o = ovg.render(w, h, st.words())[0]
                assert np.array_equal(s.premultiplied(), o), sc.name
                o_srcs.append(o)
            t.replay(sc.words(), sources=srcs or None)
            o_rgba, o_sig, _, status = ovg.render(sc.width, sc.height, sc.words(), sources=o_srcs or None)
            assert status == 0
            _exact(sc.name, compare(t.premultiplied(), t.signature(), o_rgba, o_sig))
        finally:
            t.close()
            for s in srcs:
                s.close()
