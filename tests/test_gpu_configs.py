"""Parity at the workloads' own sizes, and at the largest residuals the SATDs can see.

- A whole 1920x1080 10-bit frame with 7 references from the 1080p generator (frame 26, inside the
  fade to black at frames 24-27), uploaded at full depth and rounded on the device.
- The mixed six-stream workload (640x360, 854x480, 1280x720 x2, 1920x1080, 2560x1440), one frame of
  every stream in one batched launch, 7 references each.
- Two 3840x2160 cel streams (the animation workload) in one batched launch: plain against the
  oracle; learned + batched against the same engines with the network's per-superblock limits set
  explicitly, one stream at a time.
- Extreme residuals: a 0/255 checkerboard against its inverse and against flat pictures, uniform
  noise, black against white (|c - p| = 255 in every sample: a DC coefficient of 16 x 255 in every
  4x4 tile of every candidate, 256 tiles per 64x64 block).

Every map is compared with the oracle's CPU engine record by record (trees, leaves, all 425
candidates per superblock: tests/test_gpu_intra.cpp:144-165 of CelV2 compares its engines the same
way)."""
import os

import numpy as np
import pytest

from celv2_amd import pruner as trained
from celv2_amd import synth
from celv2_amd.records import compare_maps
import oracle_lib as ol

pytestmark = pytest.mark.gpu

THREADS = max(1, min(16, os.cpu_count() or 1))
C2_SIZES = [(640, 360), (854, 480), (1280, 720), (1280, 720), (1920, 1080), (2560, 1440)]


def oracle_map(y8, refs, lam, w, h):
    cpu = ol.CpuEngine(w, h, threads=THREADS)
    cpu.launch(y8, refs, lam)
    return cpu.collect()[0]


def frames(workload, seed, w, h, ts):
    p = synth.params(workload, seed, w, h)
    out = {t: synth.frame(p, t) for t in ts}
    return {t: a for t, (a, _) in out.items()}, {t: b for t, (_, b) in out.items()}


def note(msg):
    print(msg, flush=True)


@pytest.mark.timeout(300)
def test_1080p_fade_frame_with_seven_refs(gpu):
    w, h = 1920, 1080
    src, refs = 26, list(range(25, 18, -1))
    y16, y8 = frames(1, 1, w, h, [src] + refs)
    lam = gpu.lambda_q8(130)
    note("1080p frame 26 (fade), 7 refs, 10-bit upload")
    e = gpu.HipEngine(w, h, full_depth=True)
    e.launch(y8[src], [y8[r] for r in refs], lam=lam, keys=refs, src_key=src, src16=y16[src],
             refs16=[y16[r] for r in refs], bit_depth=10)
    got, ready = e.collect()
    e.close()
    assert ready == (h + 63) // 64
    want = oracle_map(y8[src], [y8[r] for r in refs], lam, w, h)
    diff = compare_maps(got, want, (w + 63) // 64)
    assert diff is None, diff


@pytest.mark.timeout(400)
def test_six_streams_in_one_batched_launch(gpu):
    lam = gpu.lambda_q8(130)
    srcs, refs, keys, clips, engines = [], [], [], [], []
    for k, (w, h) in enumerate(C2_SIZES):
        _, y8 = frames(1, 100 + k, w, h, range(8))
        clips.append(y8)
        engines.append(gpu.RawEngine(w, h))
        srcs.append(gpu.RawEngine.luma(y8[7]))
        refs.append([gpu.RawEngine.luma(y8[r]) for r in range(6, -1, -1)])
        keys.append(list(range(6, -1, -1)))
    note("six streams, one batched launch, 7 refs each")
    gpu.RawEngine.launch_batch(engines, srcs, refs, [lam] * len(engines), keys=keys, src_keys=[7] * len(engines))
    got = [e.collect() for e in engines]
    for e in engines:
        e.close()
    for k, (w, h) in enumerate(C2_SIZES):
        note(f"oracle {w}x{h}")
        want = oracle_map(clips[k][7], [clips[k][r] for r in range(6, -1, -1)], lam, w, h)
        diff = compare_maps(got[k], want, (w + 63) // 64)
        assert diff is None, ((w, h), diff)


@pytest.mark.timeout(600)
def test_two_4k_cel_streams_batched_plain_and_learned(gpu):
    W, H = 3840, 2160
    lam = gpu.lambda_q8(130)
    ref_ids = list(range(6, -1, -1))
    clips = [frames(4, seed, W, H, range(8))[1] for seed in (1, 2)]
    engines = [gpu.RawEngine(W, H) for _ in clips]
    srcs = [gpu.RawEngine.luma(c[7]) for c in clips]
    refs = [[gpu.RawEngine.luma(c[r]) for r in ref_ids] for c in clips]
    note("two 4K cel streams, batched, plain")
    gpu.RawEngine.launch_batch(engines, srcs, refs, [lam, lam], keys=[ref_ids, ref_ids], src_keys=[7, 7])
    plain = [e.collect() for e in engines]
    # Learned + batched: the trained network attached to both engines, one batched launch.
    weights, cfg = trained.load()
    op = trained.operating_point(cfg, 130)
    assert op is not None
    net = gpu.Pruner(max_positions=(W // 64 + 1) * (H // 64 + 1), seed=1)
    net.set_params(weights)
    for e in engines:
        e.set_pruner(net, *op)
    note("two 4K cel streams, batched, learned")
    gpu.RawEngine.launch_batch(engines, srcs, refs, [lam, lam], keys=[ref_ids, ref_ids], src_keys=[7, 7])
    learned = [e.collect() for e in engines]
    limits = [e.fetch_ref_limits() for e in engines]
    costs = [e.fetch_sb_costs() for e in engines]
    masks = [e.fetch_intra_masks() for e in engines]
    modes = [e.fetch_search_modes() for e in engines]
    for i, e in enumerate(engines):
        # The same decisions, one stream alone, with those limits, intra groups and search modes set
        # explicitly.
        e.set_pruner(None)
        e.set_ref_limits(limits[i])
        e.set_intra_masks(masks[i])
        e.set_search_modes(modes[i])
        e.launch(srcs[i], refs[i], lam, keys=ref_ids, src_key=7)
        assert compare_maps(e.collect(), learned[i], e.sb_cols) is None, i
        assert np.array_equal(e.fetch_sb_costs(), costs[i]), i
        assert limits[i].min() >= 1 and limits[i].max() <= 7
        e.close()
    net.close()
    for i, c in enumerate(clips):
        note(f"oracle 4K cel stream {i}")
        want = oracle_map(c[7], [c[r] for r in ref_ids], lam, W, H)
        diff = compare_maps(plain[i], want, (W + 63) // 64)
        assert diff is None, (i, diff)


def extreme_cases(w, h):
    yy, xx = np.mgrid[0:h, 0:w]
    board = np.where((yy + xx) % 2 == 0, 255, 0).astype(np.uint8)
    rng = np.random.default_rng(3)
    noise = rng.integers(0, 256, (h, w), dtype=np.uint8)
    black, white = np.zeros((h, w), np.uint8), np.full((h, w), 255, np.uint8)
    coarse = np.where(((yy // 4) + (xx // 4)) % 2 == 0, 255, 0).astype(np.uint8)
    return [
        ("checkerboard vs inverse and black", board, [255 - board, black]),
        ("checkerboard vs white", board, [white]),
        ("4x4 checkerboard vs its inverse", coarse, [255 - coarse, coarse]),
        ("noise vs other noise", noise, [rng.integers(0, 256, (h, w), dtype=np.uint8), 255 - noise]),
        ("black vs white", black, [white]),
        ("white vs black", white, [black, black]),
        ("checkerboard key frame", board, []),
        ("noise key frame", noise, []),
    ]


@pytest.mark.parametrize("w,h", [(256, 192), (200, 136)])
def test_extreme_residuals(gpu, w, h):
    lam = gpu.lambda_q8(130)
    e = gpu.HipEngine(w, h)
    for name, src, refs in extreme_cases(w, h):
        note(f"extreme residuals {w}x{h}: {name}")
        got = e.decide(src, refs, lam)
        want = ol.decide_frame(src, refs, lam, w, h)
        diff = compare_maps(got, want, e.sb_cols)
        assert diff is None, (name, diff)
    e.close()
