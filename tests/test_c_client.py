"""A C program written against include/vkvg.h, compiled with gcc and linked with libvkvg.so: struct
layouts, enum values and the reference's unit-test expectations (gunit_tests/device.cpp,
context.cpp, surface.cpp, matrices.cpp) through the C ABI itself, not through ctypes."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "tests", "c", "abi_client.c")
LIBDIR = os.path.join(ROOT, "vkvg_amd")


@pytest.fixture(scope="module")
def client(tmp_path_factory):
    exe = str(tmp_path_factory.mktemp("c_client") / "abi_client")
    r = subprocess.run(["gcc", "-std=c11", "-O1", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"), SRC,
                        "-L", LIBDIR, "-lvkvg", "-Wl,-rpath," + LIBDIR, "-lm", "-o", exe],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr
    return exe


def _run(exe, *args):
    r = subprocess.run([exe, *args], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    return r


def test_c_client_without_a_device(client):
    r = _run(client)
    assert r.returncode == 0, r.stdout + r.stderr


@pytest.mark.gpu
def test_c_client_with_a_device(client):
    r = _run(client, "device")
    assert r.returncode == 0, r.stdout + r.stderr
