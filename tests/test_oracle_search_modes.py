"""Pins the oracle's pruned-form restrictions (CPU): the narrow full search (me_search at range 2,
src/gpu/inter_common.h:106-119, against the reference's own function: golden vectors from
tests/golden/make_golden.py and, where oracle/_ref is built, live), the narrow search cell as
me_cell's centre ranking (inter_common.h:156-193) followed by the two narrow searches (a numpy
restatement here), the ring rule that decides which blocks are narrow (against an area formulation
of the same rule), and the reference limits (each superblock decided as with its nearest k
references)."""
import ctypes
import os

import numpy as np
import pytest

import oracle_lib as ol
from celv2_amd import synth

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "narrow_search.npz")
P, I = ctypes.c_void_p, ctypes.c_int


def ptr(a):
    return a.ctypes.data


@pytest.fixture(scope="module")
def O():
    lib = ol.oracle()
    lib.orc_full_search.argtypes = [P, P] + [I] * 7 + [P]
    lib.orc_search_cell_range.argtypes = [P, P, P, I, I, I, I, P, P, I, I, I, I, P]
    lib.orc_narrow_map.argtypes = [I] * 5 + [P, P, I, P]
    return lib


def test_full_search_golden(O):
    g = np.load(GOLDEN, allow_pickle=False)
    cur, ref = np.ascontiguousarray(g["cur"]), np.ascontiguousarray(g["ref"])
    for bx, by, px, py, rng_, mvx, mvy, cost in g["rows"]:
        mv = np.zeros(2, np.int32)
        got = O.orc_full_search(ptr(cur), ptr(ref), 64, 64, int(bx), int(by), int(px), int(py), int(rng_), ptr(mv))
        assert (got, int(mv[0]), int(mv[1])) == (cost, mvx, mvy), (bx, by, px, py, rng_)


@pytest.mark.skipif(not os.path.exists(ol.REF_SO), reason="oracle/_ref not built (needs the reference tree)")
def test_full_search_live(O):
    R = ol.ref()
    R.ref_me_search.argtypes = [P, P] + [I] * 7 + [P]
    rng = np.random.default_rng(41)
    for _ in range(200):
        h, w = int(rng.integers(16, 48)), int(rng.integers(16, 48))
        cur = rng.integers(0, 256, (h, w), dtype=np.uint8)
        ref = np.clip(np.roll(cur, (1, -2), axis=(0, 1)).astype(int) + rng.integers(-9, 10, (h, w)), 0, 255).astype(np.uint8)
        bx, by = int(rng.integers(0, w // 8)), int(rng.integers(0, h // 8))
        px, py, r = int(rng.integers(-12, 13)), int(rng.integers(-12, 13)), int(rng.integers(0, 5))
        a, b = np.zeros(2, np.int32), np.zeros(2, np.int32)
        ca = O.orc_full_search(ptr(cur), ptr(ref), w, h, bx, by, px, py, r, ptr(a))
        cb = R.ref_me_search(ptr(cur), ptr(ref), w, h, bx, by, px, py, r, ptr(b))
        assert (ca, list(a)) == (cb, list(b))


def sad(cur, ref, bx, by, mx, my):
    h, w = ref.shape
    ys = np.clip(8 * by + np.arange(8) + my, 0, h - 1)
    xs = np.clip(8 * bx + np.arange(8) + mx, 0, w - 1)
    return int(np.abs(cur[8 * by:8 * by + 8, 8 * bx:8 * bx + 8].astype(int) - ref[np.ix_(ys, xs)].astype(int)).sum())


def search(cur, ref, bx, by, px, py, r):
    """(cost, mvx, mvy): the first least SAD + 4 (|dx| + |dy|) in dy-major order."""
    best = None
    for dy in range(-r, r + 1):
        for dx in range(-r, r + 1):
            c = sad(cur, ref, bx, by, px + dx, py + dy) + 4 * (abs(dx) + abs(dy))
            if best is None or c < best[0]:
                best = (c, px + dx, py + dy)
    return best


def narrow_cell(cur, ref, bx, by, pmx, pmy, pcols, prows, r):
    """me_cell with range r: the six centres (coarser MV x2 under the block, left, right, above, below,
    zero), each different one ranked by SAD, both best searched, the runner-up's result kept only
    when strictly cheaper."""
    cents = []
    for ox, oy in ((0, 0), (-1, 0), (1, 0), (0, -1), (0, 1)):
        qx, qy = min(max(bx // 2 + ox, 0), pcols - 1), min(max(by // 2 + oy, 0), prows - 1)
        cents.append((2 * int(pmx[qy * pcols + qx]), 2 * int(pmy[qy * pcols + qx])))
    cents.append((0, 0))
    first = second = None  # (sad, centre)
    for k, c in enumerate(cents):
        if (k > 0 and first and c == first[1]) or (second and c == second[1]):
            continue
        s = sad(cur, ref, bx, by, *c)
        if first is None or s < first[0]:
            if first is not None:
                second = first
            first = (s, c)
        elif second is None or s < second[0]:
            second = (s, c)
    a = search(cur, ref, bx, by, *first[1], r)
    if second is not None:
        b = search(cur, ref, bx, by, *second[1], r)
        if b[0] < a[0]:
            return b[1:]
    return a[1:]


def test_narrow_search_cell_is_me_cell_with_range_two(O):
    rng = np.random.default_rng(43)
    for t in range(60):
        h, w = 48, 64
        cur = rng.integers(0, 256, (h, w), dtype=np.uint8) if t % 2 else np.clip(
            np.add.outer(np.arange(h) * 3, np.arange(w) * 2) + rng.integers(-6, 7, (h, w)), 0, 255).astype(np.uint8)
        ref = np.roll(cur, (int(rng.integers(-4, 5)), int(rng.integers(-4, 5))), axis=(0, 1)).copy()
        pcols, prows = w // 16, h // 16
        pmx = np.ascontiguousarray(rng.integers(-3, 4, pcols * prows).astype(np.int32))
        pmy = np.ascontiguousarray(rng.integers(-3, 4, pcols * prows).astype(np.int32))
        bx, by = int(rng.integers(0, w // 8)), int(rng.integers(0, h // 8))
        mv = np.zeros(2, np.int32)
        O.orc_search_cell_range(ptr(cur), ptr(ref), None, w, h, bx, by, ptr(pmx), ptr(pmy), pcols, prows, 0, 2, ptr(mv))
        assert tuple(mv) == narrow_cell(cur, ref, bx, by, pmx, pmy, pcols, prows, 2), t


def area_rule(level, cols, rows, sb_cols, sb_rows, limits, modes, r):
    """A block is searched full when it lies inside the area some superblock reads - the superblock
    extended by 1 block at level 0, 2 at level 1 - and that superblock keeps r and asks for the full
    search at this level; else narrow."""
    n, reach = 8 >> level, (1, 2)[level]
    full = np.zeros((rows, cols), bool)
    for s in range(sb_cols * sb_rows):
        sy, sx = divmod(s, sb_cols)
        if limits[s] > r and not (modes[s] >> level) & 1:
            full[max(0, n * sy - reach):n * sy + n + reach, max(0, n * sx - reach):n * sx + n + reach] = True
    return (~full).astype(np.uint8)


@pytest.mark.parametrize("w,h", [(352, 288), (202, 122), (1920, 1080)])
def test_narrow_blocks_follow_the_ring_rule(O, w, h):
    rng = np.random.default_rng(47)
    sb_cols, sb_rows = (w + 63) // 64, (h + 63) // 64
    for level in (0, 1):
        pw, ph = (w + 7) & ~7, (h + 7) & ~7
        lw, lh = pw >> level, ph >> level
        cols, rows = (lw + 7) // 8, (lh + 7) // 8
        for _ in range(4):
            limits = np.ascontiguousarray(rng.integers(1, 8, sb_cols * sb_rows).astype(np.uint8))
            modes = np.ascontiguousarray((rng.random(sb_cols * sb_rows) < 0.7).astype(np.uint8) * rng.integers(1, 4, sb_cols * sb_rows).astype(np.uint8))
            for r in range(7):
                out = np.zeros(cols * rows, np.uint8)
                O.orc_narrow_map(level, cols, rows, sb_cols, sb_rows, ptr(limits), ptr(modes), r, ptr(out))
                want = area_rule(level, cols, rows, sb_cols, sb_rows, limits, modes, r)
                assert np.array_equal(out.reshape(rows, cols), want), (level, r)


def test_limits_decide_as_the_nearest_references():
    """Reference limits in the oracle: each superblock's records and cost equal those of a decision
    with its nearest k references; modes of 0 and no restriction decide alike."""
    w, h = 202, 122
    y8 = synth.clip_frames(synth.params(1, 45, w, h), 5)[1]
    lam = 700
    refs = [y8[3], y8[2], y8[1]]
    nsb = ((w + 63) // 64) * ((h + 63) // 64)
    by_k = {k: ol.decide_frame_pruned(y8[4], refs[:k], lam, w, h) for k in (1, 2, 3)}
    limits = np.random.default_rng(49).integers(1, 4, nsb).astype(np.uint8)
    got, cost = ol.decide_frame_pruned(y8[4], refs, lam, w, h, limits=limits)
    for s in range(nsb):
        want, want_cost = by_k[int(limits[s])]
        assert got[s].tobytes() == want[s].tobytes(), s
        assert cost[s] == want_cost[s]
    plain, plain_cost = ol.decide_frame_pruned(y8[4], refs, lam, w, h, modes=np.zeros(nsb, np.uint8))
    assert plain.tobytes() == ol.decide_frame(y8[4], refs, lam, w, h).tobytes()
    assert np.array_equal(plain_cost, by_k[3][1])
