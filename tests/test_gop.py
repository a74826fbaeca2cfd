"""The frame loop's coding schedule (celv2_gop_plan) against the oracle's restatement (oracle/gop.py),
hand-derived schedules and the reference's own GOP test structures. CPU only: the planner is host
code in libcelv2hip.so and makes no HIP call."""
import os
import random
import sys

import pytest

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import gop as ogop  # noqa: E402  (oracle/gop.py)

from celv2_amd import gop  # noqa: E402


def as_dicts(entries):
    return [e.as_dict() for e in entries]


def c1():
    return gop.defaults(7, 130, 1920, 1080, long_gop=True)


def test_defaults_follow_the_command_line_rules():
    # src/main.cpp:739-740: one reference, no B frames; 5 or more at qindex >= 100, 7 (15 long).
    assert gop.defaults(1, 130, 1920, 1080)["bframes"] == 0
    assert gop.defaults(4, 130, 1920, 1080)["bframes"] == 3
    assert gop.defaults(7, 90, 1920, 1080)["bframes"] == 3
    assert gop.defaults(7, 130, 1920, 1080)["bframes"] == 7
    assert c1()["bframes"] == 15
    assert gop.defaults(7, 130, 3840, 2160)["parallel_b"] == 1
    assert gop.defaults(7, 130, 7680, 4320)["parallel_b"] == 0
    for refs, q, w, h, lg in [(1, 130, 64, 48, 0), (4, 120, 352, 288, 0), (7, 130, 1920, 1080, 1), (7, 99, 7680, 4320, 1)]:
        assert gop.defaults(refs, q, w, h, bool(lg)) == ogop.defaults(refs, q, w, h, bool(lg))


def test_c1_first_minigop_by_hand():
    """32 frames, 7 references, 16-frame mini-GOPs: the schedule derived by hand from the rules."""
    p = gop.plan(c1(), 32)
    first = [(e.type, e.display, e.coded, e.qindex, e.refs, e.refresh) for e in p[:23]]
    assert first[:7] == [
        ("K", 0, True, 130, [], 0xFF),
        ("A", 16, True, 114, [0], 0x02),               # the anchor: turn 1, from the keyframe only
        ("M", 8, True, 130, [0, 16], 0x04),            # level 1 at mq; equally distant: the earlier first
        ("M", 4, True, 138, [0, 8, 16], 0x08),         # level 2: +8
        ("M", 2, True, 146, [0, 4, 8, 16], 0x10),      # level 3: +16, depth first
        ("M", 6, True, 146, [4, 8, 2, 0, 16], 0x20),
        ("M", 12, True, 138, [8, 16, 6, 4, 2, 0], 0x40),
    ]
    order = [(e.type, e.display, e.coded) for e in p[7:23]]
    assert order == [("B", 1, True), ("M", 2, False), ("B", 3, True), ("M", 4, False), ("B", 5, True),
                     ("M", 6, False), ("B", 7, True), ("M", 8, False), ("B", 9, True), ("B", 10, True),
                     ("B", 11, True), ("M", 12, False), ("B", 13, True), ("B", 14, True), ("B", 15, True),
                     ("A", 16, False)]
    leaves = [e for e in p[7:23] if e.coded]
    assert all(e.qindex == 146 and sorted(e.refs) == [0, 2, 4, 6, 8, 12, 16] for e in leaves)
    assert leaves[0].refs == [0, 2, 4, 6, 8, 12, 16]
    assert [e for e in leaves if e.display == 10][0].refs == [8, 12, 6, 4, 16, 2, 0]
    assert leaves[-1].refs == [16, 12, 8, 6, 4, 2, 0]
    # Parallel B frames: only the last is kept, in the slots above the turns.
    assert [e.refresh for e in leaves] == [0] * 9 + [0x80]


def test_c1_second_minigop_by_hand():
    p = gop.plan(c1(), 32)
    second = [(e.type, e.display, e.refs, e.refresh) for e in p[23:29]]
    assert second == [
        ("A", 31, [16, 15, 12, 8, 6, 4, 2], 0x01),   # 8 distinct frames held: the nearest 7; turn 0 (the keyframe)
        ("M", 23, [16, 15, 31, 12, 8, 6, 4], 0x10),  # the earliest turn: slot 4 (frame 2)
        ("M", 19, [16, 15, 23, 12, 8, 31, 6], 0x08),
        ("M", 21, [19, 23, 16, 15, 12, 31, 8], 0x20),
        ("M", 27, [23, 31, 21, 19, 16, 15, 12], 0x04),
        ("M", 25, [23, 27, 21, 19, 31, 16, 15], 0x40),
    ]
    leaves = [e for e in p[29:] if e.coded]
    assert [e.display for e in leaves] == [17, 18, 20, 22, 24, 26, 28, 29, 30]
    assert leaves[0].refs == [16, 15, 19, 21, 23, 25, 27]


def test_hidden_reference_and_two_b_frames_rank_the_nearest_past_first():
    """tests/test_random_decisions.cpp:205-236: 4 references, a hidden reference frame and 2 shown B
    frames per group over 160 frames; every frame's first reference is its nearest past one or a
    later frame."""
    cfg = dict(gop.defaults(4, 120, 32, 32), bframes=2, parallel_b=0)
    p = gop.plan(cfg, 160)
    coded = gop.coded(p)
    assert [(e.type, e.display) for e in coded[:4]] == [("K", 0), ("A", 3), ("B", 1), ("B", 2)]
    for d in range(3, 160, 3):
        assert [(e.type, e.display, e.coded) for e in p if d - 2 <= e.display <= d] == [
            ("A", d, True), ("B", d - 2, True), ("B", d - 1, True), ("A", d, False)]
    for e in coded[1:]:
        assert all(h <= e.refs[0] or h > e.display for h in e.refs), (e.display, e.refs)


def test_anchors_of_eight_frame_minigops():
    """tests/test_cli.cpp:1046-1090: 20 frames with --bframes 7 and no cut: no anchor within 9..15
    (here 8 and 16, then 19 ending the 3-frame rest)."""
    cfg = dict(gop.defaults(4, 120, 128, 96), bframes=7)
    anchors = [e.display for e in gop.plan(cfg, 20) if e.type == "A" and e.coded]
    assert anchors == [8, 16, 19]
    assert all(d < 9 or d > 15 for d in anchors)


def test_one_reference_codes_lone_frames():
    p = gop.plan(gop.defaults(1, 130, 64, 48), 6)
    assert [(e.type, e.refs) for e in p] == [("K", [])] + [("P", [d - 1]) for d in range(1, 6)]


def test_two_references_have_no_pyramid():
    cfg = gop.defaults(2, 130, 64, 48)  # bframes 3
    p = gop.plan(cfg, 9)
    assert not any(e.type == "M" for e in p)
    assert [(e.type, e.display) for e in gop.coded(p)] == [("K", 0), ("A", 4), ("B", 1), ("B", 2), ("B", 3),
                                                          ("A", 8), ("B", 5), ("B", 6), ("B", 7)]


def test_keyframes_split_groups_and_short_shots_take_part_of_the_key_delta():
    cfg = dict(c1(), kq=-8, level_step16=2)  # --tune animation's settings (src/main.cpp:605-608)
    key = [i in (0, 16, 32) for i in range(48)]
    p = gop.plan(cfg, 48, key)
    assert [e.display for e in p if e.type == "K"] == [0, 16, 32]
    assert all(e.qindex_delta == -8 for e in p if e.type == "K")  # 8 frames ahead or more: the whole delta
    # Frames 1..15 (the cut at 16 ends the group): mids at positions 6, 2, 4, 10, 8, steps of 2 over 8 frames.
    mids = [e for e in p if e.type == "M" and e.coded and e.display < 16]
    assert [(e.display, e.level, e.qindex_delta) for e in mids] == [(7, 1, 0), (3, 2, 2), (5, 3, 4), (11, 2, 2),
                                                                    (9, 3, 4)]
    assert [e.display for e in p if e.type == "A" and e.coded] == [15, 31, 47]
    # Keyframes leading 2 and 0 frames: lround(-8 * 3 / 9) = -3, lround(-8 / 9) = -1.
    short = gop.plan(cfg, 4, [1, 0, 0, 1])
    assert [e.qindex_delta for e in short if e.type == "K"] == [-3, -1]


def test_invalid_schedules_raise():
    with pytest.raises(ValueError, match="keyframe"):
        gop.plan(c1(), 4, [0, 1, 0, 0])
    with pytest.raises(ValueError, match="refs"):
        gop.plan(dict(c1(), refs=8), 4)
    with pytest.raises(ValueError, match="bframes"):
        gop.plan(dict(c1(), bframes=16), 4)
    assert gop.plan(c1(), 0) == []


def test_short_buffer_reports_the_count():
    import ctypes
    L = gop._lib()
    c = gop._Config(**c1())
    out = (gop._Frame * 4)()
    n = gop.I32()
    assert L.celv2_gop_plan(ctypes.byref(c), 32, None, out, 4, ctypes.byref(n)) == 2
    assert n.value == len(gop.plan(c1(), 32))
    assert out[0].type == ord("K") and out[1].display == 16


def check_invariants(cfg, frames, key, p):
    shown = [e.display for e in p if e.shown]
    assert shown == list(range(frames))
    coded_before = set()
    refs_held = set()
    for e in p:
        if not e.coded:
            assert e.display in coded_before and not e.refs
            continue
        assert e.display not in coded_before
        assert len(e.refs) == len(set(e.refs)) <= cfg["refs"]
        assert set(e.refs) <= refs_held
        assert e.refs == sorted(e.refs, key=lambda h: (abs(e.display - h), h))
        if e.type == "K":
            assert key[e.display] and not e.refs and e.refresh == 0xFF
        else:
            assert e.refs, "an inter frame predicts from something"
        coded_before.add(e.display)
        if e.refresh:
            refs_held.add(e.display)
    # Parallel B frames of one mini-GOP depend on none of each other: all but the last refresh no slot,
    # so they see the same slots (each takes the nearest of them to itself).
    group = []
    for e in p + [None]:
        if e is not None and e.type == "B" and e.coded:
            group.append(e)
            continue
        if e is not None and e.type == "M" and not e.coded:
            continue
        if group:
            assert group[-1].refresh == (0xFF & ~((1 << cfg["refs"]) - 1))
            if cfg["parallel_b"]:
                assert all(g.refresh == 0 and not g.keep_shown for g in group[:-1])
        group = []


def test_planner_matches_the_oracle_over_a_sweep():
    rng = random.Random(7)
    cases = 0
    for refs in range(1, 8):
        for bframes in (0, 1, 2, 3, 5, 7, 8, 9, 12, 15):
            for _ in range(6):
                cfg = dict(refs=refs, bframes=bframes, pyramid=rng.randint(0, 1), parallel_b=rng.randint(0, 1),
                           bq=rng.randint(-20, 30), aq=rng.randint(-40, 10), mq=rng.randint(-5, 5),
                           kq=rng.randint(-64, 8), level_step=rng.randint(0, 12),
                           level_step16=rng.choice([-1, 0, 2, 5]), qindex=rng.randint(1, 255))
                frames = rng.randint(1, 70)
                key = [i == 0 or rng.random() < 0.06 for i in range(frames)]
                p = gop.plan(cfg, frames, key)
                assert [e.as_dict() for e in p] == ogop.plan(cfg, frames, key), (cfg, frames, key)
                check_invariants(cfg, frames, key, p)
                cases += 1
    assert cases == 7 * 10 * 6
