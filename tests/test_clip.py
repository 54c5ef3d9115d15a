"""CPU pins of the oracle's clip and save/restore (no GPU).

The expected images are written without the stencil: where a clip is an axis-aligned rectangle
on whole pixels, "paint under the clip" is the same image as "fill that rectangle" (both cover
the pixel centres inside it), so each clip scene is compared with a clip-free stream whose
regions were worked out by hand from src/vkvg_context.c:698-795 and :1251-1512 (the comments
give the derivation). Fills and the blend are pinned by test_oracle.py. A plausible slip in the
state machine (a wrong save bit, a restore that runs when the reference skips it, a missing
whole-stencil copy, a clip that grows instead of shrinking) changes one of these images.
"""
import numpy as np

from oracle import ovg
from vkvg_amd import scenes
from vkvg_amd.stream import FILL_EVEN_ODD, FILL_NON_ZERO, JOIN_MITER, JOIN_ROUND, Stream


def _rgba(w, h, st):
    rgba, _, _, status = ovg.render(w, h, st.words())
    assert status == 0
    return rgba


def _rect_mask(w, h, *rects):
    m = np.zeros((h, w), bool)
    for x0, y0, x1, y1 in rects:
        m[max(y0, 0):min(y1, h), max(x0, 0):min(x1, w)] = True
    return m


RED = [255, 0, 0, 255]
GREEN = [0, 255, 0, 255]
BLUE = [0, 0, 255, 255]


def test_clipped_paint_is_the_clip_rectangle_for_both_fill_rules():
    w, h = 600, 500
    for rule in (FILL_NON_ZERO, FILL_EVEN_ODD):
        sc = [x for x in scenes.clip_reference_scenes(w, h) if x.name.startswith("clipped_paint2_")]
        got = _rgba(w, h, sc[0 if rule == FILL_NON_ZERO else 1].stream)
        inner = _rect_mask(w, h, (100, 100, 400, 300))
        second = _rect_mask(w, h, (200, 200, 500, 400))  # after reset_clip: paint with a path = fill
        exp = np.zeros((h, w, 4), np.uint8)
        exp[:] = RED
        exp[inner] = GREEN
        exp[second] = BLUE
        assert np.array_equal(got, exp)


def test_clip_intersects_and_never_grows():
    w, h = 300, 200
    st = Stream()
    st.set_source_rgb(1, 0, 0)
    st.paint()
    st.rectangle(20, 30, 200, 100)
    st.clip()
    st.rectangle(120, 10, 150, 180)  # partly outside the first clip
    st.clip()
    st.set_source_rgb(0, 1, 0)
    st.paint()
    exp = np.zeros((h, w, 4), np.uint8)
    exp[:] = RED
    exp[_rect_mask(w, h, (120, 30, 220, 130))] = GREEN
    assert np.array_equal(_rgba(w, h, st), exp)


def test_even_odd_clip_keeps_the_symmetric_difference():
    w, h = 200, 200
    for rule, region in ((FILL_EVEN_ODD, "xor"), (FILL_NON_ZERO, "union")):
        st = Stream()
        st.set_source_rgb(1, 0, 0)
        st.paint()
        st.set_fill_rule(rule)
        st.rectangle(20, 20, 100, 100)
        st.rectangle(60, 60, 100, 100)
        st.clip()
        st.set_source_rgb(0, 0, 1)
        st.paint()
        a, b = _rect_mask(w, h, (20, 20, 120, 120)), _rect_mask(w, h, (60, 60, 160, 160))
        m = (a ^ b) if region == "xor" else (a | b)
        exp = np.zeros((h, w, 4), np.uint8)
        exp[:] = RED
        exp[m] = BLUE
        assert np.array_equal(_rgba(w, h, st), exp), rule


def test_clip_mask_equals_fill_mask_under_a_rotation():
    # a rotated rectangle is not pixel aligned: the clip's coverage must equal the fill's
    w, h = 500, 420
    for rule in (FILL_NON_ZERO, FILL_EVEN_ODD):
        a, b = Stream(), Stream()
        for st in (a, b):
            st.set_fill_rule(rule)
            st.set_source_rgb(1, 0, 0)
            st.paint()
            st.rotate(0.2)
            st.set_source_rgb(0, 1, 0)
            st.rectangle(100, 100, 300, 200)
        a.clip()
        a.paint()
        b.fill()
        assert np.array_equal(_rgba(w, h, a), _rgba(w, h, b))


def test_nested_clips_in_save_restore_follow_the_reference_state_machine():
    """tests/clip.c test_clip2, plus a final paint. A = [50,550)^2 clip; save#1 stores CLIP in
    bit 2; B = [100,450)^2 clip (inside A); save#2 stores it in bit 3; C = [200,400)^2 clip;
    restore#2 (a clip happened since save#2) copies bit 3 back: clip B; the fill of
    [350,770)^2 lands on [350,450)^2; restore#1 runs no stencil pass because no clip happened
    since restore#2 (curClipState is none), so the final paint is still clipped to B."""
    w, h = 800, 800
    sc = [x for x in scenes.clip_reference_scenes(w, h) if x.name == "test_clip2"][0]
    st = sc.stream
    st.set_source_rgb(1, 0, 1)
    st.set_operator(3)  # DIFFERENCE: visible on every colour below
    st.paint()
    got = _rgba(w, h, st)
    ref = Stream()
    ref.clear()
    for (x0, x1), col in (((50, 550), (1, 0, 0)), ((100, 450), (1, 1, 0)), ((200, 400), (0, 1, 0)),
                          ((350, 450), (0, 0, 1))):
        ref.set_source_rgb(*col)
        ref.rectangle(x0, x0, x1 - x0, x1 - x0)
        ref.fill()
    ref.set_source_rgb(1, 0, 1)
    ref.set_operator(3)
    ref.rectangle(100, 100, 350, 350)
    ref.fill()
    assert np.array_equal(got, _rgba(w, h, ref))


def test_deep_nesting_crosses_the_stencil_byte_with_whole_stencil_copies():
    """deep_clip_scene(256, 14): level d clips to R_d = [8d, 256-8d)^2 and paints. save_d (d >= 2)
    stores CLIP(R_{d-1}) in bit (d-2)%6+2; save_8 and save_14 first copy the whole stencil. On the
    way back only restore_14 runs its stencil pass (CLIP := R_13), and its matching copy-back
    (level 12) restores the byte saved at save_14, whose CLIP is R_13 as well; restores 13..9 run
    no pass (no clip since the last restore); restore_8's copy-back (level 6) brings the byte
    saved at save_8 whose CLIP is R_7; restores 7..1 change nothing. So the paint after
    restore_d is clipped to R_13 for d >= 9 and to R_7 for d <= 8."""
    size, depth = 256, 14
    got = _rgba(size, size, scenes.deep_clip_scene(size, depth).stream)
    ref = Stream()
    ref.set_source_rgb(1, 1, 1)
    ref.paint()

    def square(k):
        ref.rectangle(float(8 * k), float(8 * k), float(size - 16 * k), float(size - 16 * k))

    for d in range(1, depth + 1):
        ref.set_source_rgb((d * 37 % 256) / 255.0, (d * 71 % 256) / 255.0, (d * 113 % 256) / 255.0)
        square(d)
        ref.fill()
    for d in range(depth, 0, -1):
        ref.set_source_rgba((d * 53 % 256) / 255.0, (d * 29 % 256) / 255.0, (d * 97 % 256) / 255.0, 0.5)
        square(13 if d >= 9 else 7)
        ref.fill()
    assert np.array_equal(got, _rgba(size, size, ref))


def test_reset_clip_and_clear_wipe_the_saved_clip_bits():
    # clip A; save (A kept in bit 2); reset_clip clears the whole stencil, bit 2 included; the
    # restore copies that zero back: nothing is clipped afterwards. Same for clear.
    w, h = 160, 120
    for wipe in ("reset", "clear"):
        st = Stream()
        st.rectangle(10, 10, 50, 50)
        st.clip()
        st.save()
        if wipe == "reset":
            st.reset_clip()
        else:
            st.clear()
        st.restore()
        st.set_source_rgb(0, 0, 1)
        st.paint()
        exp = np.zeros((h, w, 4), np.uint8)
        exp[:] = BLUE
        assert np.array_equal(_rgba(w, h, st), exp), wipe


def test_restore_brings_back_width_colour_matrix_opacity_dash_and_resets_the_join():
    w, h = 200, 160
    pts = np.array([[20, 20], [150, 40], [30, 120], [170, 140]], np.float32)
    a = Stream()
    a.set_line_width(6)
    a.set_line_join(JOIN_ROUND)
    a.set_source_rgb(1, 0, 0)
    a.save()
    a.set_line_width(19)
    a.translate(30, 30)
    a.set_opacity(0.3)
    a.set_source_rgb(0, 0, 1)
    a.set_dash([4, 2], 1)
    a.set_line_join(JOIN_ROUND)
    a.restore()
    a.polyline(pts)
    a.stroke()
    b = Stream()  # restore reads the line join from a save field that vkvg_save never writes: MITER
    b.set_line_width(6)
    b.set_line_join(JOIN_MITER)
    b.set_source_rgb(1, 0, 0)
    b.polyline(pts)
    b.stroke()
    c = Stream()
    c.set_line_width(6)
    c.set_line_join(JOIN_ROUND)
    c.set_source_rgb(1, 0, 0)
    c.polyline(pts)
    c.stroke()
    ga, gb, gc = _rgba(w, h, a), _rgba(w, h, b), _rgba(w, h, c)
    assert np.array_equal(ga, gb)
    assert not np.array_equal(gb, gc)


def test_restore_reapplies_a_saved_gradient_under_the_restored_matrix():
    # the gradient was set under the identity, then the CTM moved; the restore finds another source
    # current and re-runs _update_cur_pattern with the saved (translated) CTM
    w, h = 240, 120
    stops = [[0, 1, 0, 0, 1], [1, 0, 0, 1, 1]]
    a = Stream()
    a.set_linear(0, 0, 120, 0, stops)
    a.translate(60, 0)
    a.save()
    a.set_source_rgb(0, 1, 0)
    a.restore()
    a.rectangle(-60, 0, 240, 120)
    a.fill()
    b = Stream()
    b.translate(60, 0)
    b.set_linear(0, 0, 120, 0, stops)
    b.rectangle(-60, 0, 240, 120)
    b.fill()
    c = Stream()  # the same gradient never re-applied: differs
    c.set_linear(0, 0, 120, 0, stops)
    c.translate(60, 0)
    c.rectangle(-60, 0, 240, 120)
    c.fill()
    ga, gb, gc = _rgba(w, h, a), _rgba(w, h, b), _rgba(w, h, c)
    assert np.array_equal(ga, gb)
    assert not np.array_equal(gb, gc)


def test_unbalanced_restore_sets_invalid_restore():
    st = Stream()
    st.save()
    st.restore()
    st.restore()
    _, _, _, status = ovg.render(16, 16, st.words())
    assert status == 3  # VKVG_STATUS_INVALID_RESTORE


def test_clip_signature_folds_the_clip_mask():
    # the clip draw adds its key where CLIP stays 0: the signature holds the clip's 1-sample mask
    w, h = 64, 48
    st = Stream()
    st.rectangle(10, 5, 30, 20)
    st.clip()
    _, sig, stats, status = ovg.render(w, h, st.words())
    assert status == 0 and stats["draws"] == 1
    exp = np.where(_rect_mask(w, h, (10, 5, 40, 25)), np.uint32(ovg.draw_key(0)), np.uint32(0))
    assert np.array_equal(sig, exp)
