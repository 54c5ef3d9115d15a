"""The allocation-free readback (vkvg_hip_image_read_unpremultiplied_noalloc) against the existing
readback entry point, through the C ABI: two calls in a row, the second on a smaller surface, and
the size limit of its staging array."""
import ctypes

import numpy as np
import pytest

from vkvg_amd import scenes

pytestmark = pytest.mark.gpu


def _bind(L):
    vp, u32 = ctypes.c_void_p, ctypes.c_uint32
    for f in (L.vkvg_hip_image_read_unpremultiplied, L.vkvg_hip_image_read_unpremultiplied_noalloc):
        f.argtypes = [vp, vp, u32, u32, vp]
        f.restype = ctypes.c_int


def _read(L, eng, surf, fn):
    out = np.full((surf.height, surf.width, 4), 77, np.uint8)
    assert fn(eng, surf.device_ptr(), surf.width, surf.height, out.ctypes.data) == 0
    return out


def test_noalloc_readback_equals_the_readback(device):
    import vkvg_amd
    L = vkvg_amd.lib()
    _bind(L)
    eng = L.vkvg_amd_device_engine(device.h)
    for sc in (scenes.config1(white=False, size=512, n_paths=120), scenes.config1(variant=1, white=False, size=96,
                                                                                 n_paths=30)):
        s = vkvg_amd.Surface(device, sc.width, sc.height)
        try:
            s.replay(sc.words())
            want = _read(L, eng, s, L.vkvg_hip_image_read_unpremultiplied)
            got = _read(L, eng, s, L.vkvg_hip_image_read_unpremultiplied_noalloc)
            assert np.array_equal(got, want), sc.name
            assert np.count_nonzero(want[..., 3] == 0) > 0 and np.count_nonzero(want[..., 3] == 255) > 0
        finally:
            s.close()


def test_noalloc_readback_rejects_images_beyond_its_array(device):
    import vkvg_amd
    L = vkvg_amd.lib()
    _bind(L)
    eng = L.vkvg_amd_device_engine(device.h)
    s = vkvg_amd.Surface(device, 4097, 4096)
    try:
        out = np.zeros(4, np.uint8)  # never written: the size check comes first
        assert L.vkvg_hip_image_read_unpremultiplied_noalloc(eng, s.device_ptr(), 4097, 4096, out.ctypes.data) == -1
        assert b"staging array" in L.vkvg_hip_last_error(eng)
    finally:
        s.close()
