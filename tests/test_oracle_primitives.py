"""Pins the oracle's primitives: against the reference's own primitives (tests/golden/primitives.npz,
written from CelV2's headers by tests/golden/make_golden.py; and live, where oracle/_ref is built),
and against closed forms that a plausible slip (a sign, an index, a dropped term) would break."""
import ctypes
import os

import numpy as np
import pytest

import oracle_lib as ol

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "primitives.npz")


def ptr(a):
    return a.ctypes.data


@pytest.fixture(scope="module")
def g():
    return np.load(GOLDEN, allow_pickle=False)


@pytest.fixture(scope="module")
def O():
    return ol.oracle()


def phases(O, plane):
    h, w = plane.shape
    out = np.zeros((16, h, w), np.uint8)
    for k in range(16):
        for y in range(h):
            for x in range(w):
                out[k, y, x] = O.orc_qpel_sample(ptr(plane), w, h, y, x, k // 4, k % 4)
    return out


# ---- against the reference's vectors --------------------------------------------------------------

def test_wht_golden(g, O):
    got = [O.orc_wht4(ptr(np.ascontiguousarray(r))) for r in g["wht_in"]]
    assert np.array_equal(got, g["wht_out"])


def test_qpel_golden(g, O):
    q = np.ascontiguousarray(g["qpel_plane"])
    ys, xs = np.arange(-5, 25), np.arange(-5, 29)
    for k in range(16):
        got = np.array([[O.orc_qpel_sample(ptr(q), 24, 20, int(y), int(x), k // 4, k % 4) for x in xs] for y in ys])
        assert np.array_equal(got, g["qpel_out"][k]), k


def test_cand_cells_golden(g, O):
    for cols, rows, b, ox, oy, k, want in g["cand_cells"][::7]:
        geo = np.zeros(4, np.int32)
        O.orc_block_geom(int(b), ptr(geo))
        x, y, w, h = [int(v) for v in geo]
        assert O.orc_cand_cell(x + int(ox), y + int(oy), w, h, int(k), int(cols), int(rows)) == want


def test_search_cell_golden(g, O):
    cur, ref = np.ascontiguousarray(g["me_cur"]), np.ascontiguousarray(g["me_ref"])
    ph = phases(O, ref)
    for row in g["me_cells"]:
        bx, by, level0, top = [int(v) for v in row[:4]]
        pmx, pmy = np.ascontiguousarray(row[4:20]), np.ascontiguousarray(row[20:36])
        mv = np.zeros(2, np.int32)
        O.orc_search_cell(ptr(cur), ptr(ref), ptr(ph) if level0 else None, 64, 64, bx, by,
                          None if top else ptr(pmx), None if top else ptr(pmy), 4, 4, level0, ptr(mv))
        assert list(mv) == list(row[36:38]), (bx, by, level0, top)


def test_intra_satd_golden(g, O):
    s = np.ascontiguousarray(g["intra_plane"])
    for x, y, w, h, m, want in g["intra_satd"]:
        assert O.orc_intra_satd(ptr(s), 112, 96, int(x), int(y), int(w), int(h), int(m)) == want, (x, y, w, h, m)


def test_inter_compound_golden(g, O):
    c, r0, r1 = (np.ascontiguousarray(g[k]) for k in ("inter_cur", "inter_r0", "inter_r1"))
    p0, p1 = phases(O, r0), phases(O, r1)
    for x, y, w, h, mx0, my0, mx1, my1, a, b in g["inter_satd"]:
        args = [int(v) for v in (x, y, w, h)]
        assert O.orc_inter_satd(ptr(c), ptr(p0), 80, 64, *args, int(mx0), int(my0)) == a
        assert O.orc_compound_satd(ptr(c), ptr(p0), ptr(p1), 80, 64, *args, int(mx0), int(my0), int(mx1), int(my1)) == b


def test_pyramid_fields_golden(g, O):
    got = ol.fields(O, "orc_", np.ascontiguousarray(g["field_cur"]), np.ascontiguousarray(g["field_ref"]))
    for l in range(5):
        assert np.array_equal(got[l][0], g[f"field_mvx{l}"]), l
        assert np.array_equal(got[l][1], g[f"field_mvy{l}"]), l


# ---- live against the reference build, where it exists --------------------------------------------

# The file test only: the library is loaded inside the tests that use it, never at collection.
needs_ref = pytest.mark.skipif(not os.path.exists(ol.REF_SO), reason="oracle/_ref not built (needs the reference tree)")


@needs_ref
def test_live_small_primitives(O):
    R = ol.ref()
    for m in range(13):
        assert O.orc_mode_bits(m) == R.ref_mode_bits(m)
    for b in range(425):
        a, r = np.zeros(4, np.int32), np.zeros(4, np.int32)
        O.orc_block_geom(b, ptr(a))
        R.ref_block_geom(b, ptr(r))
        assert list(a) == list(r)
        assert O.orc_block_index(*[int(v) for v in a]) == b == R.ref_block_index(*[int(v) for v in r])
    for n in range(1, 8):
        for a in range(7):
            for b in range(7):
                assert O.orc_pair_allowed(a, b, n) == R.ref_pair_allowed(a, b, n)


@needs_ref
def test_live_repeats(O):
    R = ol.ref()
    rng = np.random.default_rng(7)
    for t in range(40):
        cols, rows = int(rng.integers(1, 12)), int(rng.integers(1, 12))
        fx = rng.integers(-2, 3, cols * rows).astype(np.int32)
        fy = rng.integers(-2, 3, cols * rows).astype(np.int32)
        b = int(rng.integers(0, 425))
        geo = np.zeros(4, np.int32)
        O.orc_block_geom(b, ptr(geo))
        x, y, w, h = [int(v) for v in geo]
        for k in range(O.orc_cand_count(x, y, w, h)):
            assert O.orc_cand_repeats(ptr(fx), ptr(fy), cols, rows, x, y, w, h, k) == R.ref_cand_repeats(ptr(fx), ptr(fy), cols, rows, x, y, w, h, k)


@needs_ref
def test_live_intra_pred_every_pixel(O):
    R = ol.ref()
    rng = np.random.default_rng(3)
    s = rng.integers(0, 256, (80, 96), dtype=np.uint8)
    for (x, y, w, h) in [(0, 0, 8, 8), (8, 0, 16, 8), (0, 16, 8, 32), (24, 8, 32, 32), (64, 48, 32, 16), (4, 4, 4, 8)]:
        for m in range(13):
            for i in range(h):
                for j in range(w):
                    assert O.orc_intra_pred(ptr(s), 96, 80, x, y, w, h, m, i, j) == R.ref_intra_pred(ptr(s), 96, 80, x, y, w, h, m, i, j)


@needs_ref
def test_live_fields_random_frames(O):
    R = ol.ref()
    rng = np.random.default_rng(11)
    for (h, w) in [(64, 64), (48, 80), (72, 136)]:
        a = rng.integers(0, 256, (h, w), dtype=np.uint8)
        b = np.roll(a, (int(rng.integers(-6, 7)), int(rng.integers(-6, 7))), axis=(0, 1)).copy()
        b[::7] = rng.integers(0, 256, b[::7].shape, dtype=np.uint8)
        x = ol.fields(O, "orc_", a, b)
        y = ol.fields(R, "ref_", a, b)
        for l in range(5):
            assert np.array_equal(x[l][0], y[l][0]) and np.array_equal(x[l][1], y[l][1]), (h, w, l)


# ---- closed forms ------------------------------------------------------------------------------------

def test_wht_is_hadamard_transform(O):
    H = np.array([[1, 1, 1, 1], [1, -1, 1, -1], [1, 1, -1, -1], [1, -1, -1, 1]])
    rng = np.random.default_rng(0)
    for _ in range(200):
        d = rng.integers(-255, 256, 16).astype(np.int32)
        want = int(np.abs(H @ d.reshape(4, 4) @ H.T).sum())
        assert O.orc_wht4(ptr(d)) == want
    # Parseval bound: the sum of |coefficients| of a 4x4 of differences never exceeds 16 * 4 * 255.
    d = np.full(16, 255, np.int32)
    assert O.orc_wht4(ptr(d)) == 16 * 255


def test_qpel_phase_zero_is_identity_and_matches_separable_filter(O):
    rng = np.random.default_rng(5)
    p = rng.integers(0, 256, (12, 14), dtype=np.uint8)
    taps = np.array([[0, 0, 0, 128, 0, 0, 0, 0], [0, 2, -14, 110, 38, -10, 2, 0],
                     [0, 2, -14, 76, 76, -14, 2, 0], [0, 2, -10, 38, 110, -14, 2, 0]])
    padded = np.pad(p.astype(np.int64), 8, mode="edge")
    for y in range(12):
        for x in range(14):
            assert O.orc_qpel_sample(ptr(p), 14, 12, y, x, 0, 0) == p[y, x]
            for qy in range(4):
                for qx in range(4):
                    win = padded[y + 8 - 3:y + 8 + 5, x + 8 - 3:x + 8 + 5]
                    v = int(taps[qy] @ win @ taps[qx])
                    want = min(255, max(0, (v + 8192) >> 14))
                    assert O.orc_qpel_sample(ptr(p), 14, 12, y, x, qy, qx) == want


def test_pyramid_levels_are_rounded_means(O):
    rng = np.random.default_rng(8)
    y8 = rng.integers(0, 256, (72, 104), dtype=np.uint8)
    lv = ol.pyramid(y8)
    prev = y8.astype(np.int32)
    for l in range(1, 5):
        h, w = prev.shape[0] // 2, prev.shape[1] // 2
        want = (prev[0:2 * h:2, 0:2 * w:2] + prev[0:2 * h:2, 1:2 * w:2] + prev[1:2 * h:2, 0:2 * w:2] + prev[1:2 * h:2, 1:2 * w:2] + 2) >> 2
        assert np.array_equal(lv[l], want), l
        prev = lv[l].astype(np.int32)


def test_round_to_8bit(O):
    v = np.arange(0, 1024, dtype=np.uint16)
    out = np.zeros(1024, np.uint8)
    O.orc_round_plane(ptr(v), ptr(out), 1024, 10)
    assert np.array_equal(out, np.minimum(255, (v.astype(np.int32) + 2) >> 2))


def test_block_geometry_tiles_the_superblock(O):
    # Per level, the squares tile 64x64 once; each square's top+bottom and left+right halves tile it.
    for level in range(4):
        size = 64 >> level
        cover = np.zeros((64, 64), np.int32)
        first = ((1 << (2 * level)) - 1) // 3
        for sq in range(first, first + (1 << (2 * level))):
            geo = [np.zeros(4, np.int32) for _ in range(5)]
            for k in range(5):
                O.orc_block_geom(sq * 5 + k, ptr(geo[k]))
            x, y, w, h = geo[0]
            assert w == h == size
            cover[y:y + h, x:x + w] += 1
            assert geo[1][3] + geo[2][3] == size and geo[3][2] + geo[4][2] == size
    assert O.orc_cand_count(0, 0, 64, 64) == 72 and O.orc_cand_count(4, 0, 4, 8) == 9
