"""The fine kernel's division by a precomputed reciprocal (vkvg_amd/csrc/vkh_fdiv.h) equals IEEE
float division bit for bit. The header is compiled for the host with gcc and with the ROCm clang
(the compiler of the device code) and run over random operands across the fast path's exponent
range, consecutive dividends against divisors with hard significands (all ones, 1 + ulp, powers of
two, odd integers), the stop spans and pixel coordinates the gradients divide, and the fallback
cases (zero, subnormal, huge, infinite and NaN operands)."""
import os
import shutil
import subprocess
import tempfile

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

PROG = r"""
#include <stdio.h>
#include <stdint.h>
#include <string.h>
#include "vkh_fdiv.h"

static uint64_t st = 0x9E3779B97F4A7C15ull;
static uint32_t rnd(void) { st ^= st << 13; st ^= st >> 7; st ^= st << 17; return (uint32_t)(st >> 16); }
static float fbits(uint32_t u) { float f; memcpy(&f, &u, 4); return f; }
static uint32_t ubits(float f) { uint32_t u; memcpy(&u, &f, 4); return u; }
/* random float with unbiased exponent in [lo, hi], random significand and sign */
static float rf(int lo, int hi) {
    uint32_t e = (uint32_t)(lo + (int)(rnd() % (uint32_t)(hi - lo + 1)) + 127);
    return fbits((rnd() & 0x80000000u) | (e << 23) | (rnd() & 0x7FFFFFu));
}
static unsigned long long bad = 0, n = 0;
static void check(float a, float b) {
    volatile float vb = b;
    float want = a / vb;
    float got = vkh_div_rcp(a, b, vkh_rcp_for_div(b));
    n++;
    if (ubits(got) != ubits(want) && !(want != want && got != got)) {
        if (bad < 5) printf("MISMATCH a=%a b=%a got=%a want=%a\n", a, b, got, want);
        bad++;
    }
}
int main(void) {
    for (int i = 0; i < 20000000; i++) check(rf(-60, 100), rf(-52, 52));
    for (int i = 0; i < 4000000; i++) check(rf(-3, 3), rf(-3, 3));
    const uint32_t hard[] = {0x3F7FFFFFu, 0x3F800001u, 0x3F800000u, 0x3FFFFFFFu, 0x40400000u, 0x40A00000u,
                             0x3EAAAAABu, 0x3F7FFFFEu, 0x3F000001u, 0x40E00000u, 0x41100000u, 0x3DCCCCCDu,
                             0x3F4CCCCDu, 0x3F19999Au, 0x4B7FFFFFu, 0x2C800001u, 0x53FFFFFFu, 0x3F3504F3u};
    for (unsigned h = 0; h < sizeof(hard) / sizeof(hard[0]); h++)
        for (uint32_t m = 0; m < (1u << 20); m++) {
            float b = fbits(hard[h]);
            check(fbits(0x3F800000u | (m << 3) | (rnd() & 7u)), b);
            check(fbits(0x3F800000u | (m << 3) | (rnd() & 7u)), -b);
            check(fbits(0x3F800000u | ((m * 8u) & 0x7FFFFFu)), b);
        }
    /* gradient shapes: stop-span divisions (x - e0) / (e1 - e0) and pixel / size */
    for (int i = 0; i < 4000000; i++) {
        float e0 = (float)(rnd() & 0xFFFFFF) / 16777216.0f, e1 = (float)(rnd() & 0xFFFFFF) / 16777216.0f;
        float x = (float)((int)(rnd() % 20000u) - 10000) / 4096.0f + fbits(rnd() & 0x33FFFFFFu);
        check(x - e0, e1 - e0);
        check(((float)(rnd() % 8192u) + 0.5f), (float)(1u + rnd() % 8192u));
    }
    const float sp[] = {0.0f, -0.0f, 1e-45f, 1e-40f, 0x1p-126f, 0x1p-60f, 0x1p-51f, 0x1p-50f, 0x1p50f, 0x1p51f,
                        1e30f, 3e38f, 1.0f / 0.0f, -1.0f / 0.0f, 0.0f / 0.0f, 1.0f, -2.5f, 0.3f};
    for (unsigned i = 0; i < sizeof(sp) / sizeof(sp[0]); i++)
        for (unsigned j = 0; j < sizeof(sp) / sizeof(sp[0]); j++) {
            check(sp[i], sp[j]);
            check(sp[i] * 0.75f, sp[j] * 1.25f);
        }
    printf("%llu %llu\n", bad, n);
    return 0;
}
"""


def _compilers():
    out = [["gcc"]]
    clang = "/opt/rocm/llvm/bin/clang"
    if os.path.exists(clang):
        out.append([clang])
    return out


@pytest.mark.parametrize("cc", _compilers(), ids=lambda c: os.path.basename(c[0]))
def test_division_by_reciprocal_is_ieee_division(cc):
    if shutil.which(cc[0]) is None and not os.path.exists(cc[0]):
        pytest.skip(f"{cc[0]} not available")
    with tempfile.TemporaryDirectory() as t:
        src = os.path.join(t, "d.c")
        exe = os.path.join(t, "d")
        open(src, "w").write(PROG)
        subprocess.run(cc + ["-O2", "-std=c11", "-ffp-contract=off", "-mfma", "-I",
                             os.path.join(ROOT, "vkvg_amd", "csrc"), "-o", exe, src, "-lm"], check=True)
        out = subprocess.run([exe], capture_output=True, text=True, check=True, timeout=600).stdout.strip()
    bad, n = out.splitlines()[-1].split()
    assert int(n) > 40_000_000
    assert bad == "0", out
