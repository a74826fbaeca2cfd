"""Pins the oracle's per-block choice (best_both and best_mode, src/gpu/kernels_ref.cpp:41-104 of
CelV2; that file cannot be built here: it includes AVM's quant_common.h).

Every candidate record of several superblocks is rebuilt here from the oracle's primitives, each of
them pinned on its own against the reference's compiled headers (tests/test_oracle_primitives.py):
intra SATDs of the 13 modes, inter SATDs at each candidate MV, compound SATDs, the candidate cells and
repeats, the level-0 motion fields. The assembly follows kernels_ref.cpp:43-104 as written there:
  - intra: SATD*256 + lambda*(mode_bits + 32 on inter frames), modes in order, first strict minimum,
    blocks of 32x32 and below only;
  - each reference r: SATD*256 + lambda*(3 + r) over its non-repeated candidates, first strict minimum;
  - the best reference (first strict minimum), the runner-up among the others, their compound
    SATD*256 + lambda*(20 + a + b) where compound_pair_ok, taken only when strictly cheaper;
  - the record: costs, MVs x2 (1/8 pel), ref_cost = min(cost >> 8, UINT32_MAX - 1), UINT32_MAX and MV
    0 for references that do not exist.
Known answers that follow from the rules alone (identical references, a source that is the average
of two references, flat pictures) pin the same terms without the rebuild. Each of the mistakes the
oracle can be built with (ORC_PLANT, oracle/oracle.cpp) makes one of these tests fail."""
import ctypes
import os
import subprocess
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

from celv2_amd import synth
import oracle_lib as ol

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
INT64_MAX = 2**63 - 1
U32_MAX = 2**32 - 1
LAM = 799

# ORC_PLANT values of oracle/oracle.cpp and the mistake each plants.
PLANTS = {
    1: "intra blocks of inter frames charged 31 extra bits instead of 32",
    2: "inter bits without the reference index (3 instead of 3 + r)",
    3: "inter bits 4 + r instead of 3 + r",
    4: "a reference's candidates: the last of equal costs instead of the first",
    5: "the runner-up reference: the last of equal costs instead of the first",
    6: "compound bits 20 + a + 2b instead of 20 + a + b",
    7: "ref_cost = cost >> 9 instead of cost >> 8",
    8: "per-reference MVs left in quarter pels",
}


def ptr(a):
    return a.ctypes.data


def geom(O, b):
    g = np.zeros(4, np.int32)
    O.orc_block_geom(b, ptr(g))
    return [int(v) for v in g]


def rebuild(O, src, refs, lam, sb_col, sb_row):
    """The 425 candidate records of superblock (sb_col, sb_row) from the primitives: a list of dicts."""
    h, w = src.shape
    n = len(refs)
    phases = [np.ascontiguousarray(ol.phases(r)) for r in refs]
    fields = [ol.fields(O, "orc_", src, r)[0] for r in refs]
    cols, rows = (fields[0][0].shape[1], fields[0][0].shape[0]) if n else (0, 0)
    fx = [np.ascontiguousarray(f[0].ravel()) for f in fields]
    fy = [np.ascontiguousarray(f[1].ravel()) for f in fields]
    out = []
    for b in range(425):
        bx, by, bw, bh = geom(O, b)
        x, y = 64 * sb_col + bx, 64 * sb_row + by
        rec = {"inter_cost": INT64_MAX, "intra_cost": INT64_MAX, "intra": 0, "mv": (0, 0), "ref": 0, "ref2": 255,
               "mv2": (0, 0), "ref_cost": [U32_MAX] * 7, "ref_mv": [(0, 0)] * 7}
        if b >= 5 or n:  # the 64x64 level has no intra candidates and is scored only on inter frames
            if b >= 5:
                for m in range(13):
                    c = O.orc_intra_satd(ptr(src), w, h, x, y, bw, bh, m) * 256 + lam * (O.orc_mode_bits(m) + (32 if n else 0))
                    if c < rec["intra_cost"]:
                        rec["intra_cost"], rec["intra"] = c, m
            per = []
            for r in range(n):
                best, mv = INT64_MAX, (0, 0)
                for k in range(O.orc_cand_count(x, y, bw, bh)):
                    if O.orc_cand_repeats(ptr(fx[r]), ptr(fy[r]), cols, rows, x, y, bw, bh, k):
                        continue
                    cell = O.orc_cand_cell(x, y, bw, bh, k, cols, rows)
                    mx, my = int(fx[r][cell]), int(fy[r][cell])
                    c = O.orc_inter_satd(ptr(src), ptr(phases[r]), w, h, x, y, bw, bh, mx, my) * 256 + lam * (3 + r)
                    if c < best:
                        best, mv = c, (mx, my)
                per.append((best, mv))
            if n:
                a = min(range(n), key=lambda r: (per[r][0], r))
                others = [r for r in range(n) if r != a]
                runner = min(others, key=lambda r: (per[r][0], r)) if others else -1
                rec["inter_cost"], rec["mv"], rec["ref"] = per[a][0], per[a][1], a
                if runner >= 0 and O.orc_pair_allowed(a, runner, n):
                    (ma, na), (mb, nb) = per[a][1], per[runner][1]
                    c = O.orc_compound_satd(ptr(src), ptr(phases[a]), ptr(phases[runner]), w, h, x, y, bw, bh, ma, na, mb, nb) * 256 \
                        + lam * (20 + a + runner)
                    if c < rec["inter_cost"]:
                        rec["inter_cost"], rec["ref2"], rec["mv2"] = c, runner, per[runner][1]
                for r in range(n):
                    rec["ref_cost"][r] = min(per[r][0] >> 8, U32_MAX - 1)
                    rec["ref_mv"][r] = per[r][1]
        out.append(rec)
    return out


def differences(records, sb):
    """The first fields where an exported superblock differs from rebuilt records, or []."""
    bad = []
    for b, rec in enumerate(records):
        c = sb["cands"][b]
        got = {"inter_cost": int(c["inter_cost"]), "intra_cost": int(c["intra_cost"]), "intra": int(c["intra"]),
               "mv": (int(c["inter"]["mv_col"]), int(c["inter"]["mv_row"])), "ref": int(c["inter"]["ref"]),
               "ref2": int(c["inter"]["ref2"]), "mv2": (int(c["inter"]["mv2_col"]), int(c["inter"]["mv2_row"])),
               "ref_cost": [int(v) for v in c["ref_cost"]],
               "ref_mv": [(int(c["ref_mv_col"][r]), int(c["ref_mv_row"][r])) for r in range(7)]}
        want = dict(rec, mv=(2 * rec["mv"][0], 2 * rec["mv"][1]), mv2=(2 * rec["mv2"][0], 2 * rec["mv2"][1]),
                    ref_mv=[(2 * a, 2 * b_) for a, b_ in rec["ref_mv"]])
        for k in want:
            if got[k] != want[k]:
                bad.append((b, k, got[k], want[k]))
        if len(bad) > 4:
            break
    return bad


def cases():
    """(name, src, refs, lam, superblocks to rebuild) on small frames: key frames and 1, 2, 3, 7 refs."""
    nat = synth.clip_frames(synth.params(1, 12, 128, 64), 8)[1]
    cel = synth.clip_frames(synth.params(4, 5, 128, 64), 8)[1]
    out = [("key frame", nat[7], [], LAM, [(0, 0), (1, 0)])]
    for nrefs in (1, 2, 3, 7):
        out.append((f"natural, {nrefs} refs", nat[7], [nat[6 - i] for i in range(nrefs)], LAM, [(0, 0), (1, 0)]))
    out.append(("cel, 3 refs", cel[7], [cel[6], cel[5], cel[4]], 3 * LAM, [(1, 0)]))
    return out


def check_cases(O):
    fails = []
    for name, src, refs, lam, sbs in cases():
        m = None
        h, w = src.shape
        out = np.zeros(((w + 63) // 64) * ((h + 63) // 64), dtype=ol.SB_DTYPE)
        rc = O.orc_decide_frame(w, h, ptr(src), ol.ptr_array([np.ascontiguousarray(r) for r in refs]), len(refs), lam, 64, ptr(out))
        assert rc == 0
        m = out
        for (c, r) in sbs:
            d = differences(rebuild(ol.oracle(), src, refs, lam, c, r), m[r * ((w + 63) // 64) + c])
            if d:
                fails.append((name, (c, r), d))
    return fails


def test_candidate_records_follow_the_rules():
    fails = check_cases(ol.oracle())
    assert not fails, fails


# ---- known answers ------------------------------------------------------------------------------

def textured(h, w, seed):
    rng = np.random.default_rng(seed)
    t = rng.integers(30, 226, (h, w)).astype(np.int32)
    return ((t + np.roll(t, 1, 0) + np.roll(t, 1, 1)) // 3).astype(np.uint8)


def known_answers(O):
    """The failures of the known-answer checks against oracle library O (empty: all hold)."""
    fails = []
    decide = lambda src, refs, lam: _decide(O, src, refs, lam)
    # Identical references: equal SATDs, so the nearer one wins by exactly lambda (3 vs 4 bits), and a
    # compound of the two predicts the same samples for 21 bits: never taken.
    src, ref = textured(64, 128, 1), textured(64, 128, 2)
    for sb in decide(src, [ref, ref.copy()], LAM):
        c = sb["cands"]
        if not (c["inter"]["ref"] == 0).all() or not (c["inter"]["ref2"] == 255).all():
            fails.append("identical references: reference 0 alone")
        d = c["ref_cost"][:, 1].astype(np.int64) * 256 - c["ref_cost"][:, 0].astype(np.int64) * 256
        if not ((d >= LAM - 256) & (d <= LAM + 256)).all():
            fails.append("identical references: ref_cost apart by lambda / 256")
        if not (c["ref_mv_col"][:, 0] == c["ref_mv_col"][:, 1]).all():
            fails.append("identical references: same MVs")
    # A source that is the rounded average of two references (the same texture 12 above and 12 below):
    # where both references' best MV is 0 the compound predicts the source exactly and costs
    # lambda * (20 + a + b), below either single prediction.
    base = np.clip(textured(64, 128, 3), 20, 235).astype(np.int32)
    r0, r1 = (base + 12).astype(np.uint8), (base - 12).astype(np.uint8)
    src = ((r0.astype(np.int32) + r1 + 1) >> 1).astype(np.uint8)
    for sb in decide(src, [r0, r1, textured(64, 128, 5)], LAM):
        c = sb["cands"]
        i = c["inter"]
        still = (i["mv_col"] == 0) & (i["mv_row"] == 0) & (i["mv2_col"] == 0) & (i["mv2_row"] == 0) & (i["ref2"] != 255)
        pair = np.minimum(i["ref"], i["ref2"]) == 0
        if still.mean() < 0.8 or not pair[still].all():
            fails.append("average of two references: compound of 0 and 1 at MV 0 in most blocks")
        if not (c["inter_cost"][still] == LAM * (20 + i["ref"][still].astype(np.int64) + i["ref2"][still])).all():
            fails.append("average of two references: compound cost lambda * (20 + a + b)")
    # With lambda 0 and the second reference repeated as the third, references 1 and 2 cost the same:
    # where reference 0 costs as much as they do (its offset is the same size), the best is 0, the
    # runner-up the first of the tied two, and their compound at MV 0 is exact.
    for sb in decide(src, [r0, r1, r1.copy()], 0):
        c = sb["cands"]
        i = c["inter"]
        tied = (c["ref_cost"][:, 0] == c["ref_cost"][:, 1]) & (c["ref_mv_col"][:, :2] == 0).all(1) & (c["ref_mv_row"][:, :2] == 0).all(1)
        if tied.mean() < 0.5 or not ((i["ref"][tied] == 0) & (i["ref2"][tied] == 1) & (c["inter_cost"][tied] == 0)).all():
            fails.append("lambda 0, tied runners-up: compound of 0 and the first of them, cost 0")
    # Flat pictures: every SATD is 0, so costs are lambda times bits: intra DC 1 (+32 on inter frames),
    # each reference 3 + r, the 64x64 inter block of reference 0.
    flat = np.full((64, 128), 128, np.uint8)
    key = decide(flat, [], LAM)
    for sb in key:
        c = sb["cands"]
        if not (c["intra_cost"][5:] == LAM).all() or not (c["intra"][5:] == 0).all():
            fails.append("flat key frame: DC at lambda")
    for sb in decide(flat, [flat, flat.copy(), flat.copy()], LAM):
        c = sb["cands"]
        if not (c["intra_cost"][5:] == 33 * LAM).all():
            fails.append("flat inter frame: intra at lambda * 33")
        want = np.array([(3 * LAM) >> 8, (4 * LAM) >> 8, (5 * LAM) >> 8, U32_MAX, U32_MAX, U32_MAX, U32_MAX], np.int64)
        if not (c["ref_cost"].astype(np.int64) == want).all():
            fails.append("flat inter frame: ref_cost (lambda * (3 + r)) >> 8, none past the references")
        if not (c["inter_cost"] == 3 * LAM).all():
            fails.append("flat inter frame: inter at lambda * 3")
    return fails


def _decide(O, src, refs, lam):
    h, w = src.shape
    out = np.zeros(((w + 63) // 64) * ((h + 63) // 64), dtype=ol.SB_DTYPE)
    rc = O.orc_decide_frame(w, h, ptr(src), ol.ptr_array([np.ascontiguousarray(r) for r in refs]), len(refs), lam, 64, ptr(out))
    assert rc == 0
    return out


def test_known_answers():
    assert known_answers(ol.oracle()) == []


# ---- the tests above catch each planted mistake ----------------------------------------------------

def build_plant(k, out_dir):
    so = os.path.join(out_dir, f"liboracle_plant{k}.so")
    cmd = ["g++", "-O2", "-std=c++17", "-fPIC", "-shared", f"-DORC_PLANT={k}", "-o", so,
           os.path.join(ROOT, "oracle", "oracle.cpp"), os.path.join(ROOT, "oracle", "capi.cpp"), "-lpthread"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-2000:]
    lib = ctypes.CDLL(so)
    ol._prims(lib, "orc_")
    lib.orc_decide_frame.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int,
                                     ctypes.c_int64, ctypes.c_int, ctypes.c_void_p]
    return lib


def test_each_planted_mistake_is_caught(tmp_path):
    with ThreadPoolExecutor(max_workers=min(8, os.cpu_count() or 2)) as ex:
        libs = dict(zip(PLANTS, ex.map(lambda k: build_plant(k, str(tmp_path)), PLANTS)))
    missed = []
    for k, lib in libs.items():
        if not check_cases(lib) and not known_answers(lib):
            missed.append(f"{k}: {PLANTS[k]}")
    assert not missed, missed
