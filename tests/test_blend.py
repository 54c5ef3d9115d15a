"""The packed UNORM8 blends used by the fine kernel (two channels per 16-bit lane, in 32-bit
words and in 16-bit vector lanes) equal the per-channel form for every (destination, source,
operator): compiled from the same header with gcc and with the ROCm clang (the device compiler,
which takes the clang builtins of the 16-bit form), checked exhaustively over source alpha x
destination channel x source channel."""
import os
import subprocess
import tempfile

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

PROG = r"""
#include <stdio.h>
#include "vkh_blend.h"
int main(void) {
    unsigned long long bad = 0;
    for (uint32_t op = 0; op <= 3; op++)
        for (uint32_t sa = 0; sa < 256; sa++)
            for (uint32_t dc = 0; dc < 256; dc++)
                for (uint32_t sc = 0; sc < 256; sc++) {
                    uint32_t d = dc | (((dc * 7 + 3) & 255) << 8) | ((255 - dc) << 16) | (((dc * 13) & 255) << 24);
                    uint32_t s = sc | (((sc * 5 + 1) & 255) << 8) | (((sc ^ 0x5A) & 255) << 16) | (sa << 24);
                    bad += blend_px(op, d, s) != blend_px_channels(op, d, s);
                    bad += blend_px(op, s, d) != blend_px_channels(op, s, d);
                    bad += blend_px16(d, vkh_src16_of(op, s)) != blend_px_channels(op, d, s);
                    bad += blend_px16(s, vkh_src16_of(op, d)) != blend_px_channels(op, s, d);
                }
    printf("%llu\n", bad);
    return 0;
}
"""


def _compilers():
    out = ["gcc"]
    if os.path.exists("/opt/rocm/llvm/bin/clang"):
        out.append("/opt/rocm/llvm/bin/clang")
    return out


@pytest.mark.parametrize("cc", _compilers(), ids=os.path.basename)
def test_packed_blend_equals_per_channel_blend(cc):
    with tempfile.TemporaryDirectory() as t:
        src = os.path.join(t, "b.c")
        exe = os.path.join(t, "b")
        open(src, "w").write(PROG)
        subprocess.run([cc, "-O2", "-std=c11", "-I", os.path.join(ROOT, "vkvg_amd", "csrc"), "-o", exe, src],
                       check=True)
        out = subprocess.run([exe], capture_output=True, text=True, check=True).stdout.strip()
    assert out == "0"
