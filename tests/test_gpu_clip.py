"""Clip and save/restore on the GPU: every scene bit-exact against the oracle (RGBA and coverage
signatures, which fold each clip's 1-sample mask): the call sequences of the reference's clip
tests, random mixes of clips, nested saves beyond the six-level stencil byte, reset_clip, clear,
flushes that the stencil plane must survive, batches mixing surfaces with and without a stencil,
and the status of an unbalanced restore."""
import numpy as np
import pytest

from vkvg_amd import scenes
from parity_util import compare, render_gpu, render_oracle

pytestmark = pytest.mark.gpu


def _check(device, sc, tile=16):
    print(f"case {sc.name} {sc.width}x{sc.height} tile {tile}", flush=True)
    g_rgba, g_sig, _ = render_gpu(device, sc, tile)
    o_rgba, o_sig, _ = render_oracle(sc)
    r = compare(g_rgba, g_sig, o_rgba, o_sig)
    assert r["mask_mismatch_px"] == 0 and r["rgba_max_err"] == 0, (sc.name, tile, r)


def test_reference_clip_tests_match_oracle(device):
    for sc in scenes.clip_reference_scenes(1024, 800):
        _check(device, sc)


@pytest.mark.parametrize("variant", range(6))
def test_clip_mix_matches_oracle(device, variant):
    sc = scenes.clip_mix(variant, size=320, n=90)
    _check(device, sc, tile=(8, 16, 32)[variant % 3])


def test_deep_clip_nesting_matches_oracle(device):
    for tile in (8, 16, 32):
        _check(device, scenes.deep_clip_scene(256, 14), tile)
    _check(device, scenes.deep_clip_scene(400, 20), 16)  # three whole-stencil copies


def test_batch_mixes_clipped_and_plain_surfaces(device):
    import vkvg_amd
    device.set_tile_size(16)
    scs = [scenes.clip_mix(v, size=96 + 40 * v, n=70, flushes=False) for v in range(4)]
    scs.insert(2, scenes.feature_mix(1, size=150, n=30))
    surfs = [vkvg_amd.Surface(device, sc.width, sc.height, signature=True) for sc in scs]
    try:
        print("case batch of", [sc.name for sc in scs], flush=True)
        vkvg_amd.render_batch(device, surfs, [sc.words() for sc in scs])
        for sc, s in zip(scs, surfs):
            o_rgba, o_sig, _ = render_oracle(sc)
            r = compare(s.premultiplied(), s.signature(), o_rgba, o_sig)
            assert r["mask_mismatch_px"] == 0 and r["rgba_max_err"] == 0, (sc.name, r)
    finally:
        for s in surfs:
            s.close()


def test_unbalanced_restore_reports_invalid_restore(device):
    import vkvg_amd
    from vkvg_amd.stream import Stream
    st = Stream()
    st.save()
    st.restore()
    st.restore()
    s = vkvg_amd.Surface(device, 16, 16)
    try:
        with pytest.raises(vkvg_amd.VkvgError, match="status 3"):
            s.replay(st.words())
    finally:
        s.close()


def test_clipped_batch_replayed_as_graph(device):
    # clip frames start from a zero stencil at every execution: eager, graph sizing run and replay
    # each give the oracle's image, and the signatures accumulate once per execution
    import vkvg_amd
    from parity_util import render_oracle
    scs = [scenes.clip_mix(v, size=100 + 30 * v, n=60, flushes=False) for v in range(3)]
    surfs = [vkvg_amd.Surface(device, sc.width, sc.height, signature=True) for sc in scs]
    oracle = [render_oracle(sc) for sc in scs]
    try:
        print("case clipped batch, graph replay", flush=True)
        device.set_tile_size(16)
        device.set_graph(False)
        vkvg_amd.upload_batch(device, surfs, [sc.words() for sc in scs])
        vkvg_amd.execute(device)
        device.set_graph(True)
        vkvg_amd.execute(device)
        vkvg_amd.execute(device)
        assert device.stats()["graph"] == 1
        for sc, s, (o_rgba, o_sig, _) in zip(scs, surfs, oracle):
            assert np.array_equal(s.premultiplied(), o_rgba), sc.name
            want = (3 * o_sig.astype(np.uint64)) & 0xFFFFFFFF
            assert np.array_equal(s.signature().astype(np.uint64), want), sc.name
    finally:
        device.set_graph(False)
        for s in surfs:
            s.close()
