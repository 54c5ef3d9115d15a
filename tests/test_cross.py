"""The fine kernel's crossing arithmetic in doubles (vkvg_amd/csrc/vkh_cross.h) equals exact
integer arithmetic: the first sample row below an edge on a sample column (against a row-by-row
scan of the perturbed above-edge predicate in 128-bit integers) and the crossing column of every
sample row an edge spans, stepped row by row and computed directly (against ceil(N / D) in 128-bit
integers). Random edges cover tile-sized,
surface-sized and extreme (+-2^23 fixed point, the clamp of edge coordinates) extents, horizontal,
vertical and near-vertical slopes, and samples on the edge (the tie-break)."""
import os
import subprocess
import tempfile

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

PROG = r"""
#include <stdio.h>
#include <stdint.h>
#include "vkh_cross.h"

typedef __int128 i128;
static uint64_t st = 0x2545F4914F6CDD1Dull;
static uint32_t rnd(void) { st ^= st << 13; st ^= st >> 7; st ^= st << 17; return (uint32_t)(st >> 11); }
static int32_t rin(int32_t lo, int32_t hi) { return lo + (int32_t)(rnd() % (uint32_t)(hi - lo + 1)); }
static i128 ceil_div(i128 n, i128 d) { i128 q = n / d; if (q * d != n && ((n > 0) == (d > 0))) q++; return q; }
static int above(int32_t lx, int32_t ly, int32_t rx, int32_t ry, int32_t px, int32_t py) {
    i128 e = (i128)(rx - lx) * (py - ly) - (i128)(ry - ly) * (px - lx);
    return e < 0 || (e == 0 && ry > ly);
}
static unsigned long long bad = 0, n_rows = 0, n_cols = 0;

static void check(int32_t x0, int32_t y0, int32_t x1, int32_t y1, int32_t X0, int32_t Y0, int T) {
    const int32_t xs = 256 * X0 + 128, sy0 = 256 * Y0 + 128;
    if ((x0 < x1 ? x0 : x1) <= xs && xs < (x0 > x1 ? x0 : x1)) {
        int32_t lx = x0, ly = y0, rx = x1, ry = y1;
        if (x1 < x0) { lx = x1; ly = y1; rx = x0; ry = y0; }
        /* brute force: rows -2T .. 2T, the first that is not above (monotone), clamped to [-1, T] */
        int32_t want = 2 * T + 1;
        for (int32_t r = -2 * T; r <= 2 * T; r++)
            if (!above(lx, ly, rx, ry, xs, sy0 + 256 * r)) { want = r; break; }
        for (int32_t r = -2 * T; r <= 2 * T; r++) /* monotonicity of the brute force itself */
            if (above(lx, ly, rx, ry, xs, sy0 + 256 * r) != (r < want)) bad++;
        if (want < -1) want = -1;
        if (want > T) want = T;
        int32_t got = cross_first_row(lx, ly, rx, ry, xs, sy0, -1, T);
        n_rows++;
        if (got != want) { if (bad < 5) printf("ROW %d %d %d %d X0=%d Y0=%d got %d want %d\n", x0, y0, x1, y1, X0, Y0, got, want); bad++; }
    }
    if (y0 != y1) {
        const int down = y0 < y1;
        const int32_t xt = down ? x0 : x1, yt = down ? y0 : y1, xb = down ? x1 : x0, yb = down ? y1 : y0;
        int32_t r_lo = ((yt - 128 + 255) >> 8) - Y0, r_hi = ((yb - 128 + 255) >> 8) - 1 - Y0;
        if (r_lo < 0) r_lo = 0;
        if (r_hi > T - 1) r_hi = T - 1;
        if (r_lo <= r_hi) {
            vkh_dda s = cross_dda_start(xt, yt, xb, yb, sy0 + 256 * r_lo);
            for (int32_t r = r_lo; r <= r_hi; r++) {
                const int32_t y = sy0 + 256 * r;
                i128 N = (i128)(yb - yt) * (xt - 128) + (i128)(xb - xt) * (y - yt);
                i128 q = ceil_div(N, (i128)256 * (yb - yt));
                n_cols++;
                if ((i128)s.q != q) { if (bad < 5) printf("COL %d %d %d %d r=%d got %d\n", x0, y0, x1, y1, r, s.q); bad++; }
                if ((i128)cross_column(xt, yt, xb, yb, y) != q) bad++;
                if (r < r_hi) cross_dda_step(&s);
            }
        }
    }
}

int main(void) {
    const int32_t LIM = 1 << 23;
    for (int i = 0; i < 600000; i++) {
        int T = 8 << (rnd() % 3);
        int32_t X0 = rin(-64, 64) * T, Y0 = rin(-64, 64) * T;
        int32_t cx = 256 * X0, cy = 256 * Y0, ext;
        switch (rnd() % 4) {
        case 0: ext = 256 * T * 2; break;          /* tile-sized */
        case 1: ext = 256 * 4096; break;           /* surface-sized */
        case 2: ext = LIM; break;                  /* extreme */
        default: ext = 512; break;                 /* sub-pixel */
        }
        int32_t x0 = cx + rin(-ext, ext), y0 = cy + rin(-ext, ext), x1 = cx + rin(-ext, ext), y1 = cy + rin(-ext, ext);
        switch (rnd() % 6) {
        case 0: y1 = y0; break;                     /* horizontal */
        case 1: x1 = x0 + (int32_t)(rnd() % 3) - 1; break; /* (near-)vertical */
        case 2: x0 = 256 * X0 + 128 + (rnd() % 2 ? 0 : 256 * (int32_t)(rnd() % T)); break; /* end on a sample column */
        case 3: y0 = 256 * Y0 + 128 + 256 * (int32_t)(rnd() % T); x0 = 256 * X0 + 128 + 256 * (int32_t)(rnd() % T); break;
        default: break;
        }
        if (rnd() % 8 == 0) { /* endpoints at the clamp */
            x0 = (rnd() & 1) ? LIM : -LIM; y1 = (rnd() & 1) ? LIM : -LIM;
        }
        x0 = x0 > LIM ? LIM : x0 < -LIM ? -LIM : x0; x1 = x1 > LIM ? LIM : x1 < -LIM ? -LIM : x1;
        y0 = y0 > LIM ? LIM : y0 < -LIM ? -LIM : y0; y1 = y1 > LIM ? LIM : y1 < -LIM ? -LIM : y1;
        check(x0, y0, x1, y1, X0, Y0, T);
        check(x1, y1, x0, y0, X0, Y0, T);
    }
    printf("%llu %llu %llu\n", bad, n_rows, n_cols);
    return 0;
}
"""


def _compilers():
    out = ["gcc"]
    if os.path.exists("/opt/rocm/llvm/bin/clang"):
        out.append("/opt/rocm/llvm/bin/clang")
    return out


@pytest.mark.parametrize("cc", _compilers(), ids=os.path.basename)
def test_double_crossings_equal_integer_crossings(cc):
    with tempfile.TemporaryDirectory() as t:
        src = os.path.join(t, "c.c")
        exe = os.path.join(t, "c")
        open(src, "w").write(PROG)
        subprocess.run([cc, "-O2", "-std=gnu11", "-ffp-contract=off", "-I", os.path.join(ROOT, "vkvg_amd", "csrc"),
                        "-o", exe, src, "-lm"], check=True)
        out = subprocess.run([exe], capture_output=True, text=True, check=True, timeout=600).stdout.strip()
    bad, rows, cols = out.splitlines()[-1].split()
    assert int(rows) > 100_000 and int(cols) > 1_000_000
    assert bad == "0", out
