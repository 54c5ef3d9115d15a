"""Graph replay and autotuning of a resident frame: every replay equals the eager execution, and
the tuned configuration keeps bit-exact parity with the oracle."""
import numpy as np
import pytest

from vkvg_amd import scenes
from parity_util import compare, render_oracle

pytestmark = pytest.mark.gpu


def test_graph_replays_match_eager_execution(device):
    import torch
    import vkvg_amd
    sc = scenes.config3(n_paths=300, size=640)
    device.set_tile_size(16)
    device.set_graph(False)
    s = vkvg_amd.Surface(device, sc.width, sc.height, signature=True)
    try:
        s.upload(sc.words())
        vkvg_amd.execute(device)
        rgba1, sig1 = s.premultiplied(), s.signature().astype(np.uint64)
        device.set_graph(True)
        vkvg_amd.execute(device)  # eager sizing run, then capture
        n = 2
        for rep in range(3):
            x = torch.randn(1 << 20, device="cuda")  # unrelated eager work between replays
            torch.cuda.synchronize()
            vkvg_amd.execute(device)
            n += 1
            st = device.stats()
            assert st["graph"] == 1, st
            assert np.array_equal(s.premultiplied(), rgba1), rep
            sig = s.signature().astype(np.uint64)
            assert np.array_equal(sig, (n * sig1) & 0xFFFFFFFF), rep
            del x
    finally:
        device.set_graph(False)
        s.close()


def test_autotune_picks_the_fastest_and_keeps_parity(device):
    import vkvg_amd
    sc = scenes.config2(n_lines=300, n_stars=100, width=700, height=400, variant=2)
    s = vkvg_amd.Surface(device, sc.width, sc.height, signature=False)
    try:
        s.upload(sc.words())
        res = device.autotune(reps=2)
        table = res["table"]
        assert len(table) == 9
        assert res["best"]["ms"] == min(e["ms"] for e in table)
        print(res["best"], flush=True)
    finally:
        s.close()
    # the tuned tile and capacity stay selected for the next frames: parity under them
    s2 = vkvg_amd.Surface(device, sc.width, sc.height, signature=True)
    try:
        s2.replay(sc.words())
        st = device.stats()
        assert st["tile_size"] == res["best"]["tile"] and st["fine_words"] == res["best"]["fine_words"]
        o_rgba, o_sig, _ = render_oracle(sc)
        r = compare(s2.premultiplied(), s2.signature(), o_rgba, o_sig)
        assert r["mask_mismatch_px"] == 0 and r["rgba_max_err"] == 0, r
    finally:
        s2.close()
        device.set_tile_size(16)
        device.set_fine_words(4096)
