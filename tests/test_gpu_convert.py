"""The fine kernel packs gradient colours to RGBA8 with the hardware conversion
(v_cvt_pk_u8_f32 of x * 255); it must equal the reference rounding rintf(clamp01(x) * 255) for
every float x, in every byte position. The library's diagnostic kernel checks all 2^32 bit
patterns on the device. Its blend uses byte permutes that only the device compile takes; a second
diagnostic kernel checks that form against the per-channel blend over every input."""
import ctypes

import pytest

pytestmark = pytest.mark.gpu


def test_hardware_unorm8_conversion_equals_reference_rounding(device):
    from vkvg_amd import lib
    L = lib()
    L.vkh_check_u8_conversion.restype = ctypes.c_longlong
    assert L.vkh_check_u8_conversion() == 0


def test_device_packed_blend_equals_per_channel_blend(device):
    from vkvg_amd import lib
    L = lib()
    L.vkh_check_blend.restype = ctypes.c_longlong
    assert L.vkh_check_blend() == 0
