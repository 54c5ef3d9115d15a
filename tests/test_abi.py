"""The C-ABI library loads without a GPU and exports every entry point include/*.h declares."""
import ctypes
import os
import re
import subprocess

import vkvg_amd

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    names = set()
    for h in ("vkvg.h", "vkvg_amd.h", "vkvg_hip.h", "vkvg_diff.h"):
        src = open(os.path.join(ROOT, "include", h)).read()
        src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
        for m in re.finditer(r"vkvg_public\s+[^;]*?\b(vkvg_\w+)\s*\(", src):
            names.add(m.group(1))
    return names


def test_library_exports_every_declared_symbol():
    L = vkvg_amd.lib()
    out = subprocess.run(["nm", "-D", "--defined-only", vkvg_amd.LIB_PATH], capture_output=True, text=True,
                         check=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if line.strip()}
    declared = _declared()
    assert len(declared) > 120
    missing = sorted(declared - exported)
    assert not missing, missing
    for n in declared:
        assert hasattr(L, n)


def test_python_export_table_matches_headers():
    assert set(vkvg_amd.EXPORTED) == _declared()


def test_no_device_means_a_status_not_a_crash():
    # without a HIP device the constructor returns the static error status, never NULL
    import torch
    if torch.cuda.is_available():
        return
    L = vkvg_amd.lib()
    d = L.vkvg_device_create(None)
    assert d and L.vkvg_device_status(d) == 17  # VKVG_STATUS_INVALID_DEVICE_CREATE_INFO
    info = (ctypes.c_uint8 * 48)()
    d = L.vkvg_device_create(info)
    assert d and L.vkvg_device_status(d) == 16  # VKVG_STATUS_DEVICE_ERROR
    s = L.vkvg_surface_create(d, 16, 16)
    assert L.vkvg_surface_status(s) == 16
    c = L.vkvg_create(s)
    assert L.vkvg_status(c) != 0
    L.vkvg_move_to.argtypes = [ctypes.c_void_p, ctypes.c_float, ctypes.c_float]
    L.vkvg_move_to(c, 1.0, 2.0)  # no-op on an object in error
    L.vkvg_destroy(c)
