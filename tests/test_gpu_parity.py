"""HIP engine == oracle, field by field, on the same inputs (tests/test_gpu_intra.cpp:144-165 and
tests/test_gpu_me.cpp:71-194 of CelV2 compare the GPU engine with the CPU rules the same way)."""
import numpy as np
import pytest

from celv2_amd import synth
from celv2_amd.records import compare_maps
import oracle_lib as ol

pytestmark = pytest.mark.gpu


def clip(workload, seed, n, width=0, height=0, **kw):
    p = synth.params(workload, seed, width, height, **kw)
    y16, y8 = synth.clip_frames(p, n)
    return p, y16, y8


def check(h, o, cols):
    diff = compare_maps(h, o, cols)
    assert diff is None, diff


@pytest.mark.parametrize("w,h", [(352, 288), (202, 122), (64, 33), (32, 32)])
def test_intra_frame_matches_oracle(gpu, w, h):
    _, _, y8 = clip(0, 11, 1, w, h)
    lam = gpu.lambda_q8(120)
    e = gpu.HipEngine(w, h)
    got = e.decide(y8[0], [], lam)
    want = ol.decide_frame(y8[0], [], lam, w, h)
    check(got, want, e.sb_cols)


@pytest.mark.parametrize("w,h,q", [(352, 288, 120), (202, 122, 60), (640, 360, 200)])
def test_one_ref_matches_oracle(gpu, w, h, q):
    _, _, y8 = clip(0, 5, 2, w, h)
    lam = gpu.lambda_q8(q)
    e = gpu.HipEngine(w, h)
    got = e.decide(y8[1], [y8[0]], lam)
    want = ol.decide_frame(y8[1], [y8[0]], lam, w, h)
    check(got, want, e.sb_cols)


def test_seven_refs_with_key_reuse(gpu):
    w, h = 320, 192
    _, _, y8 = clip(1, 3, 10, w, h)
    lam = gpu.lambda_q8(90)
    e = gpu.HipEngine(w, h)
    for first in (1, 2):  # the second pass has one new reference, six reused by key
        refs = [y8[i] for i in range(first, first + 7)]
        keys = list(range(first, first + 7))
        e.launch(y8[0], refs, lam=lam, keys=keys)
        got, ready = e.collect()
        assert ready == e.sb_rows
        want = ol.decide_frame(y8[0], refs, lam, w, h)
        check(got, want, e.sb_cols)


@pytest.mark.parametrize("w,h,workload", [(3840, 128, 4), (128, 2160, 1), (3840, 128, 1)])
def test_4k_strips_with_seven_refs(gpu, w, h, workload):
    """The 4K frame's width (60 superblock columns) and height (34 rows) as strips, cel and natural
    content, 7 references, against the oracle's threaded CPU engine."""
    _, y16, y8 = clip(workload, 21, 8, w, h)
    lam = gpu.lambda_q8(130)
    refs = [y8[i] for i in range(6, -1, -1)]
    e = gpu.HipEngine(w, h)
    e.launch(y8[7], refs, lam=lam, keys=list(range(6, -1, -1)))
    got, ready = e.collect()
    assert ready == e.sb_rows
    cpu = ol.CpuEngine(w, h, threads=8)
    cpu.launch(y8[7], refs, lam)
    want, _ = cpu.collect()
    check(got, want, e.sb_cols)


def test_ten_bit_rounded_on_device(gpu):
    w, h = 352, 288
    _, y16, y8 = clip(0, 9, 2, w, h)
    lam = gpu.lambda_q8(130)
    e = gpu.HipEngine(w, h, full_depth=True)
    e.launch(y8[1], [y8[0]], lam=lam, src16=y16[1], refs16=[y16[0]], bit_depth=10)
    got, _ = e.collect()
    want = ol.decide_frame(y8[1], [y8[0]], lam, w, h)
    check(got, want, e.sb_cols)


def test_prepare_and_row_streams(gpu):
    w, h = 640, 360
    _, _, y8 = clip(1, 21, 5, w, h)
    lam = gpu.lambda_q8(110)
    e = gpu.HipEngine(w, h, row_streams=3)
    for i in range(2, 5):
        refs = [y8[i - 1], y8[i - 2]]
        e.launch(y8[i], refs, lam=lam, keys=[i - 1, i - 2], src_key=i)
        if i + 1 < 5:
            e.prepare(y8[i + 1], i + 1)
        got, ready = e.collect()
        assert ready == e.sb_rows
        want = ol.decide_frame(y8[i], refs, lam, w, h)
        check(got, want, e.sb_cols)


def test_resident_device_frames(gpu):
    w, h = 352, 288
    _, y16, y8 = clip(0, 13, 3, w, h)
    lam = gpu.lambda_q8(120)
    e = gpu.RawEngine(w, h)
    e.set_resident(True)
    bufs = [gpu.DeviceBuffer(f) for f in y16]
    try:
        src = gpu.RawEngine.luma(dev16=bufs[2].ptr, bit_depth=10)
        refs = [gpu.RawEngine.luma(dev16=bufs[1].ptr, bit_depth=10), gpu.RawEngine.luma(dev16=bufs[0].ptr, bit_depth=10)]
        e.launch(src, refs, lam, keys=[1, 0])
        got = e.fetch()
    finally:
        for b in bufs:
            b.free()
    want = ol.decide_frame(y8[2], [y8[1], y8[0]], lam, w, h)
    check(got, want, e.sb_cols)


def test_batched_launch_matches_single_launches(gpu):
    """Three engines of different sizes in shared launches (celv2hip_launch_batch) decide exactly as
    each alone; a second batched frame reuses the cached references by key; launched alone after the
    batches, the engines still decide as the ones that never batched."""
    sizes = [(352, 288), (202, 122), (160, 96)]
    clips = [synth.clip_frames(synth.params(0, 11 + i, w, h), 4)[1] for i, (w, h) in enumerate(sizes)]
    lam = gpu.lambda_q8(120)
    batch = [gpu.RawEngine(w, h) for (w, h) in sizes]
    alone = [gpu.RawEngine(w, h) for (w, h) in sizes]
    for f in (2, 3):
        srcs = [gpu.RawEngine.luma(c[f]) for c in clips]
        refs = [[gpu.RawEngine.luma(c[f - 1]), gpu.RawEngine.luma(c[f - 2])] for c in clips]
        keys = [[f - 1, f - 2] for _ in clips]
        gpu.RawEngine.launch_batch(batch, srcs, refs, [lam] * 3, keys=keys, src_keys=[f] * 3)
        got = [e.collect() for e in batch]
        for i, e in enumerate(alone):
            e.launch(srcs[i], refs[i], lam, keys=keys[i], src_key=f)
            want = e.collect()
            diff = compare_maps(got[i], want, e.sb_cols)
            assert diff is None, (f, sizes[i], diff)
    # The batch's engines launched alone afterwards (their own streams, after the batch on another's):
    # the next frame decides as the engines that never batched.
    srcs = [gpu.RawEngine.luma(c[1]) for c in clips]
    refs = [[gpu.RawEngine.luma(c[0])] for c in clips]
    for k in (2, 1, 0):
        batch[k].launch(srcs[k], refs[k], lam, keys=[0])
        batch[k].prepare(srcs[k], 9)
    for k in range(3):
        alone[k].launch(srcs[k], refs[k], lam, keys=[0])
        assert compare_maps(batch[k].collect(), alone[k].collect(), alone[k].sb_cols) is None, sizes[k]
    for e in batch + alone:
        e.close()


def test_reference_limits_decide_as_fewer_references(gpu):
    """celv2hip_set_ref_limits: superblock s decides exactly as a launch with its nearest k[s]
    references (records and decision cost), for a random limit map and for a constant one."""
    w, h = 352, 288
    y8 = synth.clip_frames(synth.params(1, 21, w, h), 9)[1]
    lam = gpu.lambda_q8(130)
    f, refs = 8, list(range(7, 0, -1))
    e = gpu.RawEngine(w, h)
    maps, costs = {}, {}
    for k in range(1, 8):
        e.launch(gpu.RawEngine.luma(y8[f]), [gpu.RawEngine.luma(y8[r]) for r in refs[:k]], lam, keys=refs[:k], src_key=f)
        maps[k], costs[k] = e.collect(), e.fetch_sb_costs()
    nsb = e.sb_cols * e.sb_rows
    rng = np.random.default_rng(5)
    for kmap in (rng.integers(1, 8, nsb).astype(np.uint8), np.full(nsb, 3, np.uint8), np.full(nsb, 1, np.uint8)):
        e.set_ref_limits(kmap)
        e.launch(gpu.RawEngine.luma(y8[f]), [gpu.RawEngine.luma(y8[r]) for r in refs], lam, keys=refs, src_key=f)
        got, got_cost = e.collect(), e.fetch_sb_costs()
        want = np.stack([maps[int(kmap[s])][s] for s in range(nsb)])
        want_cost = np.array([costs[int(kmap[s])][s] for s in range(nsb)])
        assert compare_maps(got, want, e.sb_cols) is None, (kmap.tolist(), compare_maps(got, want, e.sb_cols))
        assert np.array_equal(got_cost, want_cost), kmap.tolist()
    e.set_ref_limits(None)
    e.launch(gpu.RawEngine.luma(y8[f]), [gpu.RawEngine.luma(y8[r]) for r in refs], lam, keys=refs, src_key=f)
    assert compare_maps(e.collect(), maps[7], e.sb_cols) is None
    e.close()


def test_device_frames_through_the_cpp_engine(gpu):
    """celv2::HipEngine::launch_device / prepare_device: 10-bit frames resident in HBM, three frames in a
    row (each next frame prepared, references reused by key), every DecisionMap equal to the oracle's;
    a host frame given as a device one is refused."""
    w, h = 352, 288
    _, y16, y8 = clip(1, 23, 6, w, h)
    lam = gpu.lambda_q8(125)
    bufs = [gpu.DeviceBuffer(f) for f in y16]
    lum = [gpu.RawEngine.luma(dev16=b.ptr, bit_depth=10) for b in bufs]
    e = gpu.HipEngine(w, h)
    try:
        e.prepare_device(lum[3], 3)
        for i in (3, 4, 5):
            refs = [i - 1, i - 2, i - 3]
            e.launch_device(lum[i], [lum[r] for r in refs], lam=lam, keys=refs, src_key=i)
            if i + 1 < 6:
                e.prepare_device(lum[i + 1], i + 1)
            got, ready = e.collect()
            assert ready == e.sb_rows
            check(got, ol.decide_frame(y8[i], [y8[r] for r in refs], lam, w, h), e.sb_cols)
        with pytest.raises(ValueError):
            e.launch_device(gpu.RawEngine.luma(y8[5]), [lum[4]], lam=lam)
    finally:
        e.close()
        for b in bufs:
            b.free()


def test_row_groups_and_their_tuner_leave_the_decisions_alone(gpu):
    """The one-row-stream launch split into 1, 2, 3, 11 (one per row) and the default number of row
    groups gives the oracle's DecisionMap every time; the tuner picks one of the counts it timed,
    leaves the engine at it, and a second engine of the same size finds it without timing."""
    w, h = 448, 704  # 7 x 11 superblocks
    _, _, y8 = clip(1, 31, 4, w, h)
    lam = gpu.lambda_q8(110)
    bufs = [gpu.DeviceBuffer(f) for f in y8]
    lum = [gpu.RawEngine.luma(dev8=b.ptr) for b in bufs]
    want = ol.decide_frame(y8[3], [y8[2], y8[1], y8[0]], lam, w, h)
    engines = [gpu.HipEngine(w, h), gpu.HipEngine(w, h)]
    try:
        e = engines[0]
        with pytest.raises(ValueError):
            e.set_row_groups(-1)
        with pytest.raises(ValueError):
            e.set_row_groups(65)
        for g in (1, 2, 3, 11, 0):
            e.set_row_groups(g)
            e.launch_device(lum[3], [lum[2], lum[1], lum[0]], lam=lam, keys=[2, 1, 0], src_key=3)
            got, ready = e.collect()
            assert ready == e.sb_rows
            check(got, want, e.sb_cols)
        g, timed = e.autotune_device(lum[3], [lum[2], lum[1], lum[0]], lam=lam, keys=[2, 1, 0], src_key=3, frames=2, force=True)
        assert g in timed and len(timed) >= 4 and all(ms > 0 for ms in timed.values())
        assert e.row_groups() == g
        e.launch_device(lum[3], [lum[2], lum[1], lum[0]], lam=lam, keys=[2, 1, 0], src_key=3)
        check(e.collect()[0], want, e.sb_cols)
        g2, timed2 = engines[1].autotune_device(lum[3], [lum[2], lum[1], lum[0]], lam=lam, keys=[2, 1, 0], src_key=3)
        assert (g2, timed2) == (g, {}) and engines[1].row_groups() == g
    finally:
        for x in engines:
            x.close()
        for b in bufs:
            b.free()


def test_graph_replays_decide_as_eager_launches(gpu):
    """Resident launches captured as HIP graphs and replayed (celv2hip_set_graph): eight frames in a
    row, each with one new reference and the next source prepared, alone and three engines batched,
    every replay's decisions equal to an eager engine's on the same frame, with the eager engines'
    own launches and synchronisations between the replays; the graph is updated in place once the
    launches' shape repeats."""
    sizes = [(352, 288), (202, 122), (160, 96)]
    clips = [synth.clip_frames(synth.params(1, 40 + i, w, h), 12)[1] for i, (w, h) in enumerate(sizes)]
    lam = gpu.lambda_q8(125)

    def engines(graph):
        out = [gpu.RawEngine(w, h) for (w, h) in sizes]
        for e in out:
            e.set_resident(True)
            e.set_graph(graph)
        return out

    alone_g, batch_g, eager = engines(True), engines(True), engines(False)
    try:
        for k in range(3):
            for e in (alone_g[k], batch_g[k], eager[k]):
                e.prepare(gpu.RawEngine.luma(clips[k][3]), 3)
        for f in range(3, 11):
            refs = [f - 1, f - 2, f - 3]
            srcs = [gpu.RawEngine.luma(c[f]) for c in clips]
            rl = [[gpu.RawEngine.luma(c[r]) for r in refs] for c in clips]
            want = []
            for k, e in enumerate(eager):
                e.launch(srcs[k], rl[k], lam, keys=refs, src_key=f)
                e.prepare(gpu.RawEngine.luma(clips[k][f + 1]), f + 1)
                want.append(e.fetch())
            for k, e in enumerate(alone_g):
                e.launch(srcs[k], rl[k], lam, keys=refs, src_key=f)
                e.prepare(gpu.RawEngine.luma(clips[k][f + 1]), f + 1)
            gpu.RawEngine.launch_batch(batch_g, srcs, rl, [lam] * 3, keys=[refs] * 3, src_keys=[f] * 3)
            for k, e in enumerate(batch_g):
                e.prepare(gpu.RawEngine.luma(clips[k][f + 1]), f + 1)
            for k in range(3):
                assert compare_maps(alone_g[k].fetch(), want[k], eager[k].sb_cols) is None, (f, sizes[k], "alone")
                assert compare_maps(batch_g[k].fetch(), want[k], eager[k].sb_cols) is None, (f, sizes[k], "batched")
        for e in alone_g:
            launches, builds = e.graph_stats()
            assert launches == 8 and builds <= 3, (launches, builds)
        assert batch_g[0].graph_stats()[0] == 8 and batch_g[0].graph_stats()[1] <= 3
        assert eager[0].graph_stats() == (0, 0)
    finally:
        for e in alone_g + batch_g + eager:
            e.close()


def test_intra_mode_groups_decide_as_the_restricted_oracle(gpu):
    """celv2hip_set_intra_masks: with per-superblock intra mode groups (random, none beyond DC/V/H,
    all), every decision equals the oracle's with the same restriction, on an inter frame and a key
    frame; a source keeps the masks it was scored with (set before prepare(), changed before the
    launch that takes it)."""
    w, h = 352, 288
    _, _, y8 = clip(1, 29, 4, w, h)
    lam = gpu.lambda_q8(115)
    e = gpu.RawEngine(w, h)
    nsb = e.sb_cols * e.sb_rows
    rng = np.random.default_rng(8)
    try:
        for masks in (rng.integers(0, 16, nsb).astype(np.uint8), np.zeros(nsb, np.uint8), np.full(nsb, 15, np.uint8)):
            e.set_intra_masks(masks)
            for refs in ([y8[2], y8[1]], []):
                e.launch(gpu.RawEngine.luma(y8[3]), [gpu.RawEngine.luma(r) for r in refs], lam)
                want = ol.decide_frame(y8[3], refs, lam, w, h, intra_masks=masks)
                check(e.collect(), want, e.sb_cols)
        masks = rng.integers(0, 16, nsb).astype(np.uint8)
        e.set_intra_masks(masks)
        e.prepare(gpu.RawEngine.luma(y8[3]), 3)
        e.set_intra_masks(None)
        e.launch(gpu.RawEngine.luma(y8[3]), [gpu.RawEngine.luma(y8[2])], lam, keys=[2], src_key=3)
        check(e.collect(), ol.decide_frame(y8[3], [y8[2]], lam, w, h, intra_masks=masks), e.sb_cols)
        e.launch(gpu.RawEngine.luma(y8[3]), [gpu.RawEngine.luma(y8[2])], lam, keys=[2])
        check(e.collect(), ol.decide_frame(y8[3], [y8[2]], lam, w, h), e.sb_cols)
        with pytest.raises(gpu.EngineError):
            e.set_intra_masks(np.full(nsb, 16, np.uint8))
    finally:
        e.close()
