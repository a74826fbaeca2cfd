"""The pruned form's motion-search dimension on the GPU (celv2hip_set_search_modes): with
per-superblock search modes (narrow full searches at level 0 and / or 1), alone and together with
reference limits, every record and every superblock's decision cost equal the oracle's under the same
restrictions (oracle.cpp narrow_search: a block is searched narrow only when every superblock that
keeps the reference and reads it asks for it)."""
import numpy as np
import pytest

from celv2_amd import synth
from celv2_amd.records import compare_maps
import oracle_lib as ol

pytestmark = pytest.mark.gpu


def frames(workload, seed, n, w, h):
    return synth.clip_frames(synth.params(workload, seed, w, h), n)[1]


def decide_and_check(gpu, e, y8, f, refs, lam, w, h, modes, limits=None):
    e.set_search_modes(modes)
    e.set_ref_limits(limits)
    e.launch(gpu.RawEngine.luma(y8[f]), [gpu.RawEngine.luma(y8[r]) for r in refs], lam, keys=refs)
    got, got_cost = e.collect(), e.fetch_sb_costs()
    want, want_cost = ol.decide_frame_pruned(y8[f], [y8[r] for r in refs], lam, w, h, limits=limits, modes=modes)
    diff = compare_maps(got, want, e.sb_cols)
    assert diff is None, (None if modes is None else modes.tolist(), diff)
    assert np.array_equal(got_cost, want_cost)
    return got_cost


@pytest.mark.parametrize("w,h,nrefs", [(352, 288, 7), (202, 122, 3)])
def test_search_modes_decide_as_the_restricted_oracle(gpu, w, h, nrefs):
    y8 = frames(1, 31, 9, w, h)
    lam = gpu.lambda_q8(130)
    f, refs = 8, list(range(7, 7 - nrefs, -1))
    e = gpu.RawEngine(w, h)
    nsb = e.sb_cols * e.sb_rows
    rng = np.random.default_rng(17)
    try:
        full = decide_and_check(gpu, e, y8, f, refs, lam, w, h, None)
        for modes in (rng.integers(0, 4, nsb).astype(np.uint8), np.full(nsb, 3, np.uint8), np.full(nsb, 1, np.uint8),
                      np.full(nsb, 2, np.uint8), np.zeros(nsb, np.uint8)):
            cost = decide_and_check(gpu, e, y8, f, refs, lam, w, h, modes)
            if not modes.any():
                assert np.array_equal(cost, full)
        # Together with reference limits: a far reference's blocks follow only the superblocks that keep it.
        for _ in range(2):
            modes = rng.integers(0, 4, nsb).astype(np.uint8)
            limits = rng.integers(1, nrefs + 1, nsb).astype(np.uint8)
            decide_and_check(gpu, e, y8, f, refs, lam, w, h, modes, limits)
        e.set_ref_limits(None)
        with pytest.raises(gpu.EngineError):
            e.set_search_modes(np.full(nsb, 4, np.uint8))
        with pytest.raises(gpu.EngineError):
            e.set_search_modes(np.zeros(nsb + 1, np.uint8))
    finally:
        e.close()


def test_narrow_search_cuts_the_motion_search(gpu):
    """Every superblock narrow at levels 0 and 1 takes less motion-search time than the full search
    (1080p, 7 references, stage events on), and decides at a cost within 1 % of it on this frame.
    The narrow form keeps the centre ranking, the refinement and their dependent loads, so its gain
    is a few per cent of the stage."""
    w, h = 1920, 1080
    y8 = frames(1, 3, 9, w, h)
    lam = gpu.lambda_q8(130)
    refs = list(range(7, 0, -1))
    e = gpu.RawEngine(w, h)
    nsb = e.sb_cols * e.sb_rows
    try:
        e.set_resident(True)
        e.set_timing(True)
        times, cost = {}, {}
        for name, modes in (("full", None), ("narrow", np.full(nsb, 3, np.uint8))):
            e.set_search_modes(modes)
            e.launch(gpu.RawEngine.luma(y8[8]), [gpu.RawEngine.luma(y8[r]) for r in refs], lam, keys=refs)
            e.stats_reset()
            for _ in range(5):
                e.launch(gpu.RawEngine.luma(y8[8]), [gpu.RawEngine.luma(y8[r]) for r in refs], lam, keys=refs)
            e.sync()
            ms, cnt = e.stats()
            times[name] = ms["me"] / cnt["launches"]
            cost[name] = int(e.fetch_sb_costs().sum())
        assert times["narrow"] < times["full"], times
        assert abs(cost["narrow"] / cost["full"] - 1) < 0.01, cost
    finally:
        e.close()
