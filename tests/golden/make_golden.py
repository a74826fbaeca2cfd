"""Writes tests/golden/primitives.npz: input/output vectors of CelV2's own decision primitives, and
tests/golden/narrow_search.npz: the reference's me_search at the small ranges of the pruned form's
narrow search.

Run in a tree where the reference is mounted and oracle/_ref is built (`make -C oracle ref`): the
outputs come from the reference's header-only primitives (src/gpu/intra_common.h, inter_common.h,
sb_blocks.h) compiled by oracle/Makefile, never from the oracle. The fixture holds only data.

    python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
import oracle_lib as ol  # noqa: E402


def ptr(a):
    return a.ctypes.data


def plane_noise(rng, h, w, smooth=True):
    if not smooth:
        return rng.integers(0, 256, (h, w), dtype=np.uint8)
    base = rng.integers(30, 220, (h // 8 + 2, w // 8 + 2)).astype(np.float64)
    ys, xs = np.mgrid[0:h, 0:w] / 8.0
    y0, x0 = ys.astype(int), xs.astype(int)
    fy, fx = ys - y0, xs - x0
    v = (base[y0, x0] * (1 - fx) + base[y0, x0 + 1] * fx) * (1 - fy) + (base[y0 + 1, x0] * (1 - fx) + base[y0 + 1, x0 + 1] * fx) * fy
    v += rng.normal(0, 6, (h, w))
    return np.clip(np.rint(v), 0, 255).astype(np.uint8)


def phases_of(R, plane):
    h, w = plane.shape
    out = np.zeros((16, h, w), np.uint8)
    for k in range(16):
        for y in range(h):
            for x in range(w):
                out[k, y, x] = R.ref_qpel_sample(ptr(plane), w, h, y, x, k // 4, k % 4)
    return out


def main():
    R = ol.ref()
    if R is None:
        raise SystemExit("oracle/_ref/libcelv2ref.so is not built (make -C oracle ref)")
    rng = np.random.default_rng(20261203)
    g = {}
    # 4x4 Walsh-Hadamard sums.
    d = rng.integers(-255, 256, (300, 16), dtype=np.int32)
    d[0] = 255
    d[1] = -255
    d[2] = np.tile([255, -255], 8)
    g["wht_in"] = d
    g["wht_out"] = np.array([R.ref_wht4(ptr(np.ascontiguousarray(r))) for r in d], np.int32)
    # Quarter-pel samples on a small plane, positions reaching past every edge.
    q = plane_noise(rng, 20, 24, smooth=False)
    g["qpel_plane"] = q
    ys, xs = np.arange(-5, 25), np.arange(-5, 29)
    out = np.zeros((16, len(ys), len(xs)), np.int32)
    for k in range(16):
        for i, y in enumerate(ys):
            for j, x in enumerate(xs):
                out[k, i, j] = R.ref_qpel_sample(ptr(q), 24, 20, int(y), int(x), k // 4, k % 4)
    g["qpel_out"] = out
    # Candidate cells and repeats.
    rows = []
    for (cols, frows) in ((30, 17), (3, 2), (9, 9)):
        for b in range(425):
            geo = np.zeros(4, np.int32)
            R.ref_block_geom(b, ptr(geo))
            for (ox, oy) in ((0, 0), (64, 64), (128, 0)):
                x, y, w, h = [int(v) for v in geo]
                n = R.ref_cand_count(x + ox, y + oy, w, h)
                for k in range(n):
                    rows.append((cols, frows, b, ox, oy, k, R.ref_cand_cell(x + ox, y + oy, w, h, k, cols, frows)))
    g["cand_cells"] = np.array(rows, np.int32)
    # Motion search of single cells (both centres, refinement at level 0) on 64x64 planes.
    cur = plane_noise(rng, 64, 64)
    ref = np.roll(plane_noise(rng, 64, 64), (3, -5), axis=(0, 1))
    ref[8:40, 8:40] = np.roll(cur, (2, -3), axis=(0, 1))[8:40, 8:40]
    ph = phases_of(R, ref)
    g["me_cur"], g["me_ref"] = cur, ref
    cells = []
    for t in range(48):
        bx, by = int(rng.integers(0, 8)), int(rng.integers(0, 8))
        pcols, prows = 4, 4
        pmx = rng.integers(-6, 7, pcols * prows).astype(np.int32)
        pmy = rng.integers(-6, 7, pcols * prows).astype(np.int32)
        if t % 3 == 0:
            pmx[:] = pmx[0]
            pmy[:] = pmy[0]
        level0 = t % 2
        top = t % 5 == 4
        mv = np.zeros(2, np.int32)
        R.ref_search_cell(ptr(cur), ptr(ref), ptr(ph) if level0 else None, 64, 64, bx, by,
                          None if top else ptr(pmx), None if top else ptr(pmy), pcols, prows, level0, ptr(mv))
        cells.append(np.concatenate([[bx, by, level0, int(top)], pmx, pmy, mv]).astype(np.int32))
    g["me_cells"] = np.array(cells, np.int32)
    # Intra SATDs (every mode) of blocks at frame edges and inside.
    s = plane_noise(rng, 96, 112)
    g["intra_plane"] = s
    rows = []
    for (x, y, w, h) in [(0, 0, 32, 32), (0, 32, 16, 32), (32, 0, 32, 16), (40, 24, 8, 4), (16, 16, 16, 16),
                         (8, 0, 8, 8), (0, 8, 4, 8), (56, 40, 32, 32), (96, 64, 16, 16), (48, 48, 8, 16)]:
        for m in range(13):
            rows.append((x, y, w, h, m, R.ref_intra_satd(ptr(s), 112, 96, x, y, w, h, m)))
    g["intra_satd"] = np.array(rows, np.int64)
    # Inter and compound SATDs at quarter-pel MVs.
    r0, r1 = plane_noise(rng, 64, 80), plane_noise(rng, 64, 80)
    p0, p1 = phases_of(R, r0), phases_of(R, r1)
    c = plane_noise(rng, 64, 80)
    g["inter_cur"], g["inter_r0"], g["inter_r1"] = c, r0, r1
    rows = []
    for t in range(60):
        w, h = [(64, 64), (32, 16), (8, 8), (4, 8), (16, 32), (8, 4)][t % 6]
        x, y = int(rng.integers(0, 80 - w + 1) // 4 * 4), int(rng.integers(0, 64 - h + 1) // 4 * 4)
        mx0, my0, mx1, my1 = [int(v) for v in rng.integers(-90, 91, 4)]
        a = R.ref_inter_satd(ptr(c), ptr(p0), 80, 64, x, y, w, h, mx0, my0)
        b = R.ref_compound_satd(ptr(c), ptr(p0), ptr(p1), 80, 64, x, y, w, h, mx0, my0, mx1, my1)
        rows.append((x, y, w, h, mx0, my0, mx1, my1, a, b))
    g["inter_satd"] = np.array(rows, np.int64)
    # Whole pyramid search of a frame pair (every level's field).
    f0 = plane_noise(rng, 96, 128)
    f1 = np.roll(f0, (2, -3), axis=(0, 1)).copy()
    f1[40:72, 48:96] = np.roll(f0, (-5, 7), axis=(0, 1))[40:72, 48:96]
    g["field_cur"], g["field_ref"] = f1, f0
    flds = ol.fields(R, "ref_", f1, f0)
    for l, (mx, my) in enumerate(flds):
        g[f"field_mvx{l}"], g[f"field_mvy{l}"] = mx, my
    np.savez_compressed(os.path.join(HERE, "primitives.npz"), **g)
    print("wrote", os.path.join(HERE, "primitives.npz"), {k: v.shape for k, v in g.items()})
    narrow(R)


def narrow(R):
    """me_search (src/gpu/inter_common.h:106-119) at ranges 0..4 around centres reaching past every
    edge of a 64x64 pair: rows (bx, by, px, py, range, mvx, mvy, cost)."""
    rng = np.random.default_rng(20261204)
    cur, ref = plane_noise(rng, 64, 64), plane_noise(rng, 64, 64)
    ref[8:40, 16:48] = np.roll(cur, (3, -2), axis=(0, 1))[8:40, 16:48]  # a region whose motion a search finds
    I = __import__("ctypes").c_int
    R.ref_me_search.argtypes = [__import__("ctypes").c_void_p] * 2 + [I] * 7 + [__import__("ctypes").c_void_p]
    rows = []
    for i in range(240):
        bx, by = int(rng.integers(0, 8)), int(rng.integers(0, 8))
        px, py = int(rng.integers(-14, 15)), int(rng.integers(-14, 15))
        rng_ = i % 5
        mv = np.zeros(2, np.int32)
        cost = R.ref_me_search(ptr(cur), ptr(ref), 64, 64, bx, by, px, py, rng_, ptr(mv))
        rows.append([bx, by, px, py, rng_, int(mv[0]), int(mv[1]), cost])
    out = os.path.join(HERE, "narrow_search.npz")
    np.savez_compressed(out, cur=cur, ref=ref, rows=np.array(rows, np.int32))
    print("wrote", out)


if __name__ == "__main__":
    if len(sys.argv) > 1 and sys.argv[1] == "narrow":
        R = ol.ref()
        if R is None:
            raise SystemExit("oracle/_ref/libcelv2ref.so is not built (make -C oracle ref)")
        narrow(R)
    else:
        main()
