"""GPU parity: HIP library vs CPU oracle on seeded scenes (1-sample masks bit-exact via
coverage signatures; RGBA8 within 1 LSB on >= 99.9% of pixels, max and mean reported)."""
import pytest

from vkvg_amd import scenes
from parity_util import compare, render_gpu, render_oracle

pytestmark = pytest.mark.gpu


def _check(dev, scene, tile=16):
    g_rgba, g_sig, st = render_gpu(dev, scene, tile)
    o_rgba, o_sig, _ = render_oracle(scene)
    r = compare(g_rgba, g_sig, o_rgba, o_sig)
    print(scene.name, tile, r, {k: st[k] for k in ("points", "edges", "commands")}, flush=True)
    assert r["mask_mismatch_px"] == 0, r
    assert r["frac_within_1lsb"] >= 0.999, r
    return r


def test_c1_nonzero_cubics(device):
    _check(device, scenes.config1())


def test_c2_strokes_and_eo_stars_small(device):
    _check(device, scenes.config2(n_lines=400, n_stars=150, width=640, height=360))


def test_c3_dense_small(device):
    _check(device, scenes.config3(n_paths=400, size=768))


def test_c3_full_resolution_first_1000_paths(device):
    # the real 4096x4096 dense scene, truncated to its first 1000 paths so the oracle stays fast
    _check(device, scenes.config3(n_paths=1000))


@pytest.mark.parametrize("tile", [8, 32])
def test_tile_sizes(device, tile):
    _check(device, scenes.config1(n_paths=120), tile=tile)
    _check(device, scenes.config2(n_lines=150, n_stars=60, width=500, height=300, variant=1), tile=tile)


@pytest.mark.parametrize("variant", [0, 1, 2])
def test_feature_mix(device, variant):
    # joins/caps on closed strokes, dash offsets, arcs, ellipses, EO, quadratics, relative
    # curves, transforms, gradients, opacity, DIFFERENCE and CLEAR operators, paint with a path
    _check(device, scenes.feature_mix(variant))


def test_clear(device):
    _check(device, scenes.clear_scene())


@pytest.mark.parametrize("tile", [8, 32])
def test_feature_mix_tiles(device, tile):
    _check(device, scenes.feature_mix(3, size=300, n=48), tile=tile)


def test_long_runs_of_one_draw_kind(device):
    # fill-edge and stroke-triangle runs spanning more than 64 draws per wavefront
    _check(device, scenes.runs_scene())


def test_long_dashed_pieces(device):
    # pieces of a few to several hundred dash boundaries: emitted by one thread below 64 caps, by a
    # wavefront in one, two or more rounds of 64 caps above (odd and zero-entry patterns, offsets,
    # every cap and join, closed paths ending in a gap)
    _check(device, scenes.long_dash_scene())


def test_c2_full_scene(device):
    # the whole 1920x1080 configuration: 5000 polylines (25 % dashed) and 2000 even-odd stars
    _check(device, scenes.config2())


def test_c3_full_scene(device):
    # the benchmark frame itself: 20000 paths on 4096x4096, every pixel and every draw's mask
    _check(device, scenes.config3())


def test_c4_full_batch_sampled_surfaces(device):
    # the benchmark's batched frame (all 256 surfaces in one pipeline); every 16th surface, at its
    # full size and path count, compared with its own oracle render
    import vkvg_amd
    subs = scenes.config4()
    picked = set(range(0, len(subs), 16))
    surfs = [vkvg_amd.Surface(device, s.width, s.height, signature=(i in picked)) for i, s in enumerate(subs)]
    try:
        device.set_tile_size(16)
        print("case c4 full batch", flush=True)
        vkvg_amd.render_batch(device, surfs, [s.words() for s in subs])
        for i in sorted(picked):
            sc, s = subs[i], surfs[i]
            print(f"case c4 surface {i} {sc.width}x{sc.height} {sc.n_paths} paths", flush=True)
            o_rgba, o_sig, _ = render_oracle(sc)
            r = compare(s.premultiplied(), s.signature(), o_rgba, o_sig)
            assert r["mask_mismatch_px"] == 0 and r["rgba_max_err"] == 0, (i, r)
    finally:
        for s in surfs:
            s.close()
