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


def test_batched_surfaces_match_oracle_eager_and_replayed(device):
    # uneven surfaces of the batched workload's kind in one frame (one upload, one pipeline), then
    # the same frame replayed as a captured graph: every surface bit-exact against its own oracle
    import vkvg_amd
    subs = scenes.config4(n_surfaces=16, side_lo=48, side_hi=320, paths_lo=5, paths_hi=400)
    surfs = [vkvg_amd.Surface(device, s.width, s.height, signature=True) for s in subs]
    oracle = [render_oracle(sc) for sc in subs]
    try:
        device.set_tile_size(16)
        device.set_graph(False)
        vkvg_amd.upload_batch(device, surfs, [s.words() for s in subs])
        vkvg_amd.execute(device)
        for sc, s, (o_rgba, o_sig, _) in zip(subs, surfs, oracle):
            r = compare(s.premultiplied(), s.signature(), o_rgba, o_sig)
            assert r["mask_mismatch_px"] == 0 and r["rgba_max_err"] == 0, (sc.name, r)
        device.set_graph(True)
        vkvg_amd.execute(device)  # sizing run, then capture
        vkvg_amd.execute(device)  # replay
        assert device.stats()["graph"] == 1
        for sc, s, (o_rgba, o_sig, _) in zip(subs, surfs, oracle):
            assert np.array_equal(s.premultiplied(), o_rgba), sc.name
            # the signatures accumulate once per execution: eager, sizing, replay
            want = (3 * o_sig.astype(np.uint64)) & 0xFFFFFFFF
            assert np.array_equal(s.signature().astype(np.uint64), want), sc.name
    finally:
        device.set_graph(False)
        for s in surfs:
            s.close()


def test_full_batch_autotuned_and_replayed_sampled_surfaces(device):
    # the benchmark's batched frame exactly as it is timed: all 256 surfaces resident in one frame,
    # autotuned, then replayed as captured graphs; every 16th surface compared with its own oracle
    # render (RGBA8 bit-exact; masks through the signature, which grows by one oracle signature per
    # execution)
    import vkvg_amd
    subs = scenes.config4()
    picked = list(range(0, len(subs), 16))
    surfs = [vkvg_amd.Surface(device, s.width, s.height, signature=(i in picked)) for i, s in enumerate(subs)]
    try:
        device.set_graph(False)
        vkvg_amd.upload_batch(device, surfs, [s.words() for s in subs])
        tune = device.autotune(reps=1)
        print("case c4 tuned", tune["best"], flush=True)
        sig0 = {i: surfs[i].signature().astype(np.uint64) for i in picked}
        device.set_graph(True)
        for _ in range(3):  # sizing run + capture, then two replays
            vkvg_amd.execute(device)
        assert device.stats()["graph"] == 1
        for i in picked:
            sc, s = subs[i], surfs[i]
            print(f"case c4 replayed surface {i} {sc.width}x{sc.height} {sc.n_paths} paths", flush=True)
            o_rgba, o_sig, _ = render_oracle(sc)
            assert np.array_equal(s.premultiplied(), o_rgba), i
            want = (sig0[i] + 3 * o_sig.astype(np.uint64)) & 0xFFFFFFFF
            assert np.array_equal(s.signature().astype(np.uint64), want), i
    finally:
        device.set_graph(False)
        device.set_tile_size(16)
        device.set_fine_words(4096)
        for s in surfs:
            s.close()
