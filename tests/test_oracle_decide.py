"""Pins the oracle's per-block choice and partition decision (kernels_ref.cpp:41-198 of CelV2, which
cannot be built here: it includes AVM's quant_common.h) by brute force over every partition tree of
small squares, forced ties, invariants, and the known answers of CelV2's own engine tests."""
import numpy as np
import pytest

from celv2_amd import synth
from celv2_amd.records import BLOCK_DTYPE, SB_DTYPE, leaf_rects
import oracle_lib as ol

LAM = 799  # lambda_q8 at qindex 120 of the stated step model


def ptr(a):
    return a.ctypes.data


def frames(workload, seed, n, w, h):
    return synth.clip_frames(synth.params(workload, seed, w, h), n)[1]


def square_costs(O, y8, refs, lam, x, y, size):
    """decide_square on one square; returns (cost, tree, leaves, per-block choice costs)."""
    ph, pw = y8.shape
    P = [ol.phases(r) for r in refs]
    F = [ol.fields(O, "orc_", y8, r)[0] for r in refs]
    rows, cols = (F[0][0].shape if F else (1, 1))
    mvx = [np.ascontiguousarray(f[0].ravel()) for f in F]
    mvy = [np.ascontiguousarray(f[1].ravel()) for f in F]
    tree = np.zeros(256, np.uint8)
    blocks = np.zeros(256, BLOCK_DTYPE)
    n = np.zeros(2, np.int32)
    cands = np.zeros(1, SB_DTYPE)
    ox, oy = x // 64 * 64, y // 64 * 64
    c = O.orc_decide_square(pw, ph, ptr(y8), ol.ptr_array(P), ol.ptr_array(mvx), ol.ptr_array(mvy), cols, rows, len(refs),
                            lam, x, y, size, ox, oy, ptr(tree), ptr(blocks), ptr(n), ptr(cands))
    cc = cands[0]["cands"]
    choice = np.minimum(cc["inter_cost"], cc["intra_cost"])
    return c, list(tree[: n[0]]), blocks[: n[1]], choice


def enumerate_trees(O, x, y, size, cost_of, part, allow_leaf):
    """Every tree of the square (x, y, size) (SB-relative), as (cost, tree list)."""
    out = []
    half = size // 2
    if allow_leaf(size):
        b = lambda xx, yy, ww, hh: O.orc_block_index(xx, yy, ww, hh)
        out.append((cost_of[b(x, y, size, size)], [0]))
        out.append((part + cost_of[b(x, y, size, half)] + cost_of[b(x, y + half, size, half)], [1, 0, 0]))
        out.append((part + cost_of[b(x, y, half, size)] + cost_of[b(x + half, y, half, size)], [2, 0, 0]))
    if size > 8:
        subs = [enumerate_trees(O, x + (q & 1) * half, y + (q >> 1) * half, half, cost_of, part, allow_leaf) for q in range(4)]
        for a in subs[0]:
            for b_ in subs[1]:
                for c in subs[2]:
                    for d in subs[3]:
                        out.append((part + a[0] + b_[0] + c[0] + d[0], [3] + a[1] + b_[1] + c[1] + d[1]))
    return out


@pytest.mark.parametrize("nrefs", [0, 1, 2])
def test_dp_is_the_cheapest_tree(nrefs):
    O = ol.oracle()
    y8 = frames(0, 41, 3, 96, 96)
    refs = y8[1:1 + nrefs]
    for (x, y, size) in [(0, 0, 16), (16, 32, 16), (40, 8, 8), (48, 48, 16)]:
        cost, tree, leaves, choice = square_costs(O, y8[0], refs, LAM, x, y, size)
        trees = enumerate_trees(O, x % 64, y % 64, size, choice, 2 * LAM, lambda s: s <= 32)
        best = min(t[0] for t in trees)
        assert cost == best
        # The chosen tree costs what the DP returned, and it is one of the cheapest.
        assert any(t[1] == tree and t[0] == cost for t in trees)
        # The first cheapest in NONE, HORZ, VERT, SPLIT order at the root.
        root = [t for t in trees if t[0] == best][0]
        assert tree[0] == root[1][0]


def test_leaves_follow_candidates_and_tile_superblocks():
    y8 = frames(1, 4, 3, 200, 136)
    for refs in ([], [y8[1]], [y8[1], y8[2]]):
        m = ol.decide_frame(y8[0], refs, LAM, 200, 136)
        for i, sb in enumerate(m):
            rects = leaf_rects(sb)
            assert len(rects) == sb["nblocks"]
            area = np.zeros((64, 64), np.int32)
            for k, (x, y, w, h) in enumerate(rects):
                area[y:y + h, x:x + w] += 1
                blk = sb["blocks"][k]
                assert (blk["w4"], blk["h4"]) == (w // 4, h // 4)
                c = sb["cands"][ol.oracle().orc_block_index(x, y, w, h)]
                inter = c["inter_cost"] < c["intra_cost"]
                assert bool(blk["inter"]) == inter
                if inter:
                    assert (blk["mv_col"], blk["mv_row"], blk["ref"]) == (c["inter"]["mv_col"], c["inter"]["mv_row"], c["inter"]["ref"])
                else:
                    assert blk["intra"] == c["intra"] and max(w, h) <= 32  # no intra block above 32
            assert (area == 1).all()
            if not refs:
                assert sb["tree"][0] == 3  # key frames split the 64x64


def test_flat_frame_ties_take_the_first_choice():
    flat = np.full((128, 128), 128, np.uint8)  # the value unavailable edges predict
    key = ol.decide_frame(flat, [], LAM)
    for sb in key:  # 64x64 split; 32x32 NONE; DC (1 bit) beats the other modes' equal SATD
        assert list(sb["tree"][:5]) == [3, 0, 0, 0, 0]
        assert all(b["intra"] == 0 for b in sb["blocks"][:4])
    inter = ol.decide_frame(flat, [flat, flat], LAM)
    for sb in inter:  # one inter 64x64 from the nearest reference, MV 0 (every candidate ties)
        assert list(sb["tree"][:1]) == [0] and sb["nblocks"] == 1
        b = sb["blocks"][0]
        assert b["inter"] == 1 and b["ref"] == 0 and b["ref2"] == 255 and b["mv_col"] == 0 and b["mv_row"] == 0


def test_threaded_engine_equals_single_thread():
    y8 = frames(1, 8, 4, 264, 200)
    want = ol.decide_frame(y8[0], [y8[1], y8[2], y8[3]], LAM, 264, 200)
    for threads in (1, 3, 8):
        e = ol.CpuEngine(264, 200, 64, threads)
        e.launch(y8[0], [y8[1], y8[2], y8[3]], LAM)
        got, ready = e.collect()
        assert ready == e.height // 64 + (1 if e.height % 64 else 0)
        assert got.tobytes() == want.tobytes()


def test_row_hook_failure_reaches_collect():
    y8 = frames(0, 2, 2, 256, 256)
    e = ol.CpuEngine(256, 256, 64, 2, fail_row=2)
    e.launch(y8[0], [y8[1]], LAM)
    with pytest.raises(RuntimeError) as ex:
        e.collect()
    assert "row hook" in str(ex.value.args[0]) and ex.value.args[1] <= 2


def test_decision_threads_values():
    O = ol.oracle()
    assert O.orc_decision_threads(48, 2) == 46
    assert O.orc_decision_threads(4, 4) == 1
    assert O.orc_decision_threads(1, 1) == 1
    assert O.orc_decision_threads(0, 1) == 1


# ---- known answers of CelV2's engine tests (tests/test_gpu_me.cpp), on this build's own content ----

def test_two_centres_are_searched():
    """A block whose best centre (zero: the content plus 3) is a dead end while the runner-up (the
    coarser level's 26, 0) is 6 px from the exact match at (20, 0) (test_gpu_me.cpp:312-332)."""
    O = ol.oracle()
    w, h, bx, by = 128, 64, 4, 2
    v = [7]

    def rnd():
        v[0] = (v[0] * 1664525 + 1013904223) & 0xFFFFFFFF
        return v[0] >> 24

    cur = np.array([rnd() for _ in range(w * h)], np.uint8).reshape(h, w)
    ref = np.array([rnd() for _ in range(w * h)], np.uint8).reshape(h, w)
    for y in range(by * 8, by * 8 + 8):
        for x in range(bx * 8, bx * 8 + 8):
            ref[y, x + 20] = cur[y, x]
            ref[y, x] = (int(cur[y, x]) + 3) & 255
    pmx = np.full(8 * 4, 13, np.int32)
    pmy = np.zeros(8 * 4, np.int32)
    mv = np.zeros(2, np.int32)
    cost = O.orc_search_cell(ptr(cur), ptr(ref), None, w, h, bx, by, ptr(pmx), ptr(pmy), 8, 4, 0, ptr(mv))
    assert list(mv) == [20, 0]
    assert cost == 4 * (26 - 20)  # an exact match 6 pels from the one centre (2 x 13, 0)


def shifted_texture(w, h, sx, sy, seed=1):
    rng = np.random.default_rng(seed)
    big = rng.integers(40, 216, (h + 64, w + 64)).astype(np.uint8)
    big = ((big.astype(np.int32) + np.roll(big, 1, 1) + np.roll(big, 1, 0)) // 3).astype(np.uint8)
    return np.ascontiguousarray(big[32 + sy:32 + sy + h, 32 + sx:32 + sx + w])


def test_large_motion_found():
    """Motion of (96, 24) px, beyond a 3-level pyramid: most of the level-0 field finds it
    (test_gpu_me.cpp:293-308)."""
    O = ol.oracle()
    rng = np.random.default_rng(4)
    base = rng.integers(0, 256, (360 // 8 + 40, 640 // 8 + 40))
    big = np.kron(base, np.ones((8, 8))).astype(np.float64)
    ys, xs = np.mgrid[0:big.shape[0], 0:big.shape[1]]
    big = np.clip(big * 0.6 + 50 + 30 * np.sin(xs / 7.0) * np.cos(ys / 5.0), 0, 255).astype(np.uint8)
    prev = np.ascontiguousarray(big[100:460, 100:740])
    cur = np.ascontiguousarray(big[100 - 24:460 - 24, 100 - 96:740 - 96])  # cur(x, y) = prev(x - 96, y - 24)
    lv = ol.fields(O, "orc_", cur, prev)
    mx, my = lv[0]
    sub = (slice(4, None), slice(13, None))
    hit = ((mx[sub] == -96 * 4) & (my[sub] == -24 * 4)).mean()
    assert hit >= 0.75, hit


def test_integer_shift_found_in_interior():
    """Content moved by (5, -3): interior leaves take MV (40, -24)/8 over >= 90% of their area
    (test_gpu_me.cpp:71-112)."""
    w, h = 640, 360
    prev = shifted_texture(w, h, 0, 0)
    cur = shifted_texture(w, h, 5, -3)
    m = ol.decide_frame(cur, [prev], LAM)
    interior = hit = 0
    cols = (w + 63) // 64
    for i, sb in enumerate(m):
        for k, (x, y, bw, bh) in enumerate(leaf_rects(sb, 64 * (i % cols), 64 * (i // cols))):
            if x < 16 or y < 16 or x + bw > w - 16 or y + bh > h - 16:
                continue
            interior += bw * bh
            b = sb["blocks"][k]
            hit += bw * bh * (b["inter"] and b["mv_col"] == 40 and b["mv_row"] == -24)
    assert interior > 0 and hit >= 0.9 * interior, (hit, interior)
