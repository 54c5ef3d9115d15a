"""Shared helpers of the parity tests: render one scene through the HIP library and the oracle."""
from __future__ import annotations

import numpy as np


def render_gpu(dev, scene, tile: int = 16):
    import vkvg_amd
    dev.set_tile_size(tile)
    s = vkvg_amd.Surface(dev, scene.width, scene.height, signature=True)
    try:
        s.replay(scene.words())
        return s.premultiplied(), s.signature(), dev.stats()
    finally:
        s.close()


def render_oracle(scene, variant: str = ""):
    from oracle import ovg
    rgba, sig, stats, status = ovg.render(scene.width, scene.height, scene.words(), variant=variant)
    assert status == 0, f"oracle status {status}"
    return rgba, sig, stats


def compare(gpu_rgba, gpu_sig, ora_rgba, ora_sig):
    """Returns a dict: mask mismatches (signature) and RGBA error statistics."""
    mask_bad = int(np.count_nonzero(gpu_sig != ora_sig))
    d = np.abs(gpu_rgba.astype(np.int16) - ora_rgba.astype(np.int16)).max(axis=2)
    return {
        "mask_mismatch_px": mask_bad,
        "rgba_max_err": int(d.max()),
        "rgba_mean_err": float(d.mean()),
        "frac_within_1lsb": float(np.count_nonzero(d <= 1) / d.size),
        "rgba_exact_px": int(np.count_nonzero(d == 0)),
        "pixels": int(d.size),
    }
