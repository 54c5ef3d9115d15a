"""Mutation tests of the oracle pins: the oracle's C sources are copied, one deliberate mistake is
made in the copy, the copy is compiled, and the pins of tests/test_oracle_pins.py run against it.
Each mistake must fail at least one pin; otherwise that pin would not constrain the function.

The mistakes are plausible ones for each restated function: the reference's M_2_PIF quirk
"fixed", a swapped flattening branch, an unsquared tolerance, a different arc step or threshold,
miter and bevel swapped, cairo's miter-limit semantics, curve-internal joins not forced, a round
cap on the wrong side, the other ray-circle root, and so on."""
from __future__ import annotations

import os
import shutil
import subprocess
import tempfile
from concurrent.futures import ThreadPoolExecutor

import pytest

from oracle import ovg

import test_oracle_pins as pins

ORACLE = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle")
SOURCES = ["ovg_geom.c", "ovg_raster.c", "ovg_replay.c", "ovg_state.c"]
HEADERS = ["ovg.h", "ovg_internal.h"]

# name: (file, original text, replacement, pins of which at least one must fail)
MUTANTS = {
    "two_pi_fold": ("ovg_geom.c", "OVG_2_PIF - ", "2.f * OVG_PIF - ", [pins.check_agg_nodes]),
    "cusp_branch_emits_p2": ("ovg_geom.c", "if (da1 + da2 < (float)0.01) {\n                    emit(u, x1234, y1234);",
                             "if (da1 + da2 < (float)0.01) {\n                    emit(u, x2, y2);",
                             [pins.check_agg_nodes]),
    "d2_branch_emits_p3": ("ovg_geom.c", "if ((double)da1 > 0.01) {\n                        emit(u, x2, y2);",
                           "if ((double)da1 > 0.01) {\n                        emit(u, x3, y3);", [pins.check_agg_nodes]),
    "unsquared_tolerance": ("ovg_geom.c", "if (dx * dx + dy * dy <= tol) {", "if (sqrtf(dx * dx + dy * dy) <= tol) {",
                            [pins.check_agg_nodes]),
    "level0_not_forced": ("ovg_geom.c", "if (level > 0) {", "if (level >= 0) {", [pins.check_agg_level0_subdivides]),
    "tolerance_ignores_ctm": ("ovg_geom.c", "float tol = (float)fabs(0.25f / fmaxf(sx, sy));", "float tol = 0.25f;",
                              [pins.check_curve_tolerance_follows_ctm_scale]),
    "arc_step_half_radius": ("ovg_geom.c", "OVG_PIF / (r * 0.4f)", "OVG_PIF / (r * 0.5f)", [pins.check_arc_counts]),
    "arc_step_threshold_20": ("ovg_geom.c", "if (r < 30.0f)", "if (r < 20.0f)", [pins.check_arc_counts]),
    "arc_step_ignores_scale": ("ovg_geom.c", "float r = radius * fabsf(fmaxf(sx, sy));", "float r = radius;",
                               [pins.check_arc_counts]),
    "full_circle_open": ("ovg_geom.c", "if (EQUF(a2 - a1, OVG_PIF * 2.f)) {", "if (0) {", [pins.check_full_circle_closes]),
    "arc_not_convex": ("ovg_geom.c", "c->simpleConvex = !c->pathPtr;", "c->simpleConvex = false;",
                       [pins.check_first_arc_convex_flag]),
    "elliptic_centre_flipped": ("ovg_geom.c", "if (largeArc == ccw) {", "if (largeArc != ccw) {",
                                [pins.check_elliptic_arc_on_svg_ellipse]),
    "miter_bevel_swapped": ("ovg_geom.c", "int join = c->lineJoin;",
                            "int join = c->lineJoin == 0 ? 2 : c->lineJoin == 2 ? 0 : c->lineJoin;",
                            [pins.check_right_angle_joins]),
    "cairo_miter_limit": ("ovg_geom.c", "s.lhMax = c->miterLimit * c->lineWidth;", "s.lhMax = c->miterLimit * s.hw;",
                          [pins.check_miter_limit_switch]),
    "flush_miter_cut": ("ovg_geom.c", "float x = (lh - s->lhMax) * cosHalfAlpha;",
                        "float x = (lh - s->lhMax) / cosHalfAlpha;", [pins.check_miter_limit_switch]),
    "curve_joins_not_forced": ("ovg_geom.c", "if (isCurve)\n        join =", "if (0)\n        join =",
                               [pins.check_curve_internal_joins]),
    "curve_join_threshold": ("ovg_geom.c", "(dot < 0.8f) ? 1", "(dot < 0.5f) ? 1", [pins.check_curve_internal_joins]),
    "round_cap_wrong_side": ("ovg_geom.c", "float a = L_ACOSF(n.x) + OVG_PIF_2;", "float a = L_ACOSF(n.x) - OVG_PIF_2;",
                             [pins.check_round_caps]),
    "round_step_of_width": ("ovg_geom.c", "s->arcStep = arc_step(c, s->hw);", "s->arcStep = arc_step(c, c->lineWidth);",
                            [pins.check_round_join_vertex_count, pins.check_large_round_join_mask, pins.check_round_caps]),
    "radial_other_root": ("ovg_raster.c", "float t = (-b + sqrtf(fabsf(disc)))", "float t = (-b - sqrtf(fabsf(disc)))",
                          [pins.check_focal_radial]),
    "radial_radii_by_height": ("ovg_raster.c", "float r0 = g->cp[0][2] / W, r1 = g->cp[1][2] / W;",
                               "float r0 = g->cp[0][2] / H, r1 = g->cp[1][2] / H;", [pins.check_focal_radial]),
    "radial_length_without_r0": ("ovg_raster.c", "gradLength = sqrtf(ddx * ddx + ddy * ddy) - r0;",
                                 "gradLength = sqrtf(ddx * ddx + ddy * ddy);", [pins.check_focal_radial]),
    "radial_focus_not_pulled_in": ("ovg_replay.c", "if (l + radius0 + 1.0f >= radius1) {", "if (0) {",
                                   [pins.check_focal_radial]),
    "dash_offset_not_wrapped": ("ovg_geom.c", "dc.curDashOffset = fmodf(c->dashOffset, dc.totDashLength);",
                                "dc.curDashOffset = c->dashOffset;", [pins.check_dash_offset_and_wrap]),
    "dash_starts_off": ("ovg_geom.c", "dc.dashOn = true;", "dc.dashOn = false;", [pins.check_dash_offset_and_wrap]),
    "dash_not_carried": ("ovg_geom.c", "dc->curDashOffset -= segLen;", "dc->curDashOffset = 0;",
                         [pins.check_dash_continues_around_a_corner]),
}


def _build(tmp: str, name: str) -> str:
    fname, old, new, _ = MUTANTS[name]
    d = os.path.join(tmp, name)
    os.makedirs(d)
    for f in SOURCES + HEADERS:
        shutil.copy(os.path.join(ORACLE, f), d)
    path = os.path.join(d, fname)
    with open(path) as fh:
        text = fh.read()
    assert old in text, f"{name}: the text to change is no longer in {fname}"
    with open(path, "w") as fh:
        fh.write(text.replace(old, new))
    lib = os.path.join(d, "libovg_mutant.so")
    subprocess.run(["gcc", "-O1", "-std=c11", "-fPIC", "-ffp-contract=off", "-fno-fast-math", "-shared", "-o", lib,
                    *[os.path.join(d, s) for s in SOURCES], "-lm"], check=True, capture_output=True)
    return lib


@pytest.fixture(scope="module")
def mutant_libs():
    tmp = tempfile.mkdtemp(prefix="ovg_mutants_")
    try:
        with ThreadPoolExecutor(max_workers=8) as ex:
            libs = dict(zip(MUTANTS, ex.map(lambda n: _build(tmp, n), MUTANTS)))
        yield libs
    finally:
        shutil.rmtree(tmp, ignore_errors=True)


def test_unchanged_oracle_passes_every_pin():
    for _, _, _, checks in MUTANTS.values():
        for check in checks:
            check()


@pytest.mark.parametrize("name", list(MUTANTS))
def test_mutant_fails_a_pin(mutant_libs, name):
    failed = []
    with ovg.using(mutant_libs[name]):
        for check in MUTANTS[name][3]:
            try:
                check()
            except AssertionError:
                failed.append(check.__name__)
    assert failed, f"mutant {name} passes {[c.__name__ for c in MUTANTS[name][3]]}"
