"""Benchmark of the HIP decision engine on one MI355X.

Workload (the reference's engine_bench method, tools/engine_bench.cpp:44-57): a 1920x1080 10-bit
4:2:0 synthetic clip of 32 frames (pan, 12 moving objects, grain, fade to black at frames 24-27);
frame i >= 7 is decided against the 7 source frames before it, nearest first, keys = frame indices,
qindex 130, with the next frame prepare()d behind it. A step is one frame's decision, timed as
engine_bench times it: launch -> prepare -> collect through the C++ engine (celv2::HipEngine, CelV2's
DecisionEngine), the frame's DecisionMap on the host when collect returns. The 10-bit frames are
resident in HBM before timing (the engine rounds them to the 8 bits the decisions read on the device).

Printed: one JSON line with that rate (`value`), the rate with the records left in HBM
(`kernel_only`), the rate with host frames uploaded by launch / prepare (`host_frames_fps`), the
dominant kernel's roofline fraction, stage times measured with HIP events on the engine's stream, and
the CPU engine (the oracle) timed on this host's cores, whose decisions are checked against the GPU's.

    python bench.py [--gpus 1] [--steps 50] [--warmup 5] [--no-cpu]
"""
import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from celv2_amd import engine, gop, pruner, synth  # noqa: E402
from celv2_amd.records import SB_DTYPE, compare_maps  # noqa: E402

METRIC = "1080p10 frames/s on 1 GPU: plain HIP vs learned-pruned at ≤0.5% BD-rate loss"
VALU_PEAK_TOPS = 256 * 64 * 2.4e9 / 1e12  # 256 CUs x 64 lanes x 2.4 GHz, one int32 op per lane per clock
SAD_OPS, SATD_OPS = 3, 7  # int ops per pixel (sub, abs, add; diff, 2 butterfly passes, abs, add)
# The motion search's ceiling: v_sad_u8 takes 4 byte pairs per lane per instruction, so the chip's
# published clock gives 256 CUs x 64 lanes x 4 x 2.4 GHz pixel-SADs per second; a microbenchmark of
# independent v_sad_u8 chains on the MI355X reached 125.4 T/s (profiles/sad_u8_rate.md).
SAD_PEAK_TPX, SAD_MEASURED_TPX = 256 * 64 * 4 * 2.4e9 / 1e12, 125.4
# HBM bytes the motion search's level-0 launch (the dominant kernel launch) moves per 1080p frame with
# 7 references (rocprofv3 FETCH_SIZE and WRITE_SIZE per dispatch, scaled by a 64 MiB device copy:
# profiles/me_traffic_1080p10_7refs.md).
ME_TRAFFIC_BYTES = 342.2e6
# HBM bytes of one forward-transform call over a 1080p frame's candidate blocks (FETCH_SIZE + WRITE_SIZE,
# scaled the same way: profiles/me_traffic_1080p10_7refs.md).
TX_TRAFFIC_BYTES = 142.0e6


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_baseline(y8, frame, refs, lam, gpu_map, threads):
    """The oracle's threaded CPU engine on one frame (test infrastructure, timed here only); with
    gpu_map None, only the time."""
    so = os.path.join(ROOT, "oracle", "liboracle.so")
    O = ctypes.CDLL(so)
    O.orc_engine_create.restype = ctypes.c_void_p
    O.orc_engine_create.argtypes = [ctypes.c_int] * 5
    O.orc_engine_launch.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_int64]
    O.orc_engine_collect.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
    O.orc_engine_destroy.argtypes = [ctypes.c_void_p]
    h, w = y8[0].shape
    e = O.orc_engine_create(w, h, 64, threads, -1)
    arr = (ctypes.c_void_p * len(refs))(*[y8[i].ctypes.data for i in refs])
    out = np.zeros(((w + 63) // 64) * ((h + 63) // 64), dtype=SB_DTYPE)
    t0 = time.perf_counter()
    O.orc_engine_launch(e, w, h, y8[frame].ctypes.data, arr, len(refs), lam)
    rc = O.orc_engine_collect(e, out.ctypes.data, None)
    dt = time.perf_counter() - t0
    O.orc_engine_destroy(e)
    if rc != 0:
        raise RuntimeError("CPU engine failed")
    return dt, (compare_maps(gpu_map, out, (w + 63) // 64) if gpu_map is not None else None)


# The mixed-resolution workload: six streams, one frame of each per step.
C2_SIZES = [(640, 360), (854, 480), (1280, 720), (1280, 720), (1920, 1080), (2560, 1440)]


def bench_batched(NR, lam, steps=20, warmup=3, nf=16):
    """Six streams of 16 10-bit frames (seeded generator, workload 1 at each size), 7 references,
    engine_bench method per stream, frames resident in HBM. Timed twice over the same steps: each
    engine launched on its own (the plain form), then all six in shared launches
    (celv2hip_launch_batch); both prepare() each stream's next frame. Returns the two rates and
    whether the last step's decisions agree between the forms."""
    t = time.perf_counter()
    streams = []
    for k, (w, h) in enumerate(C2_SIZES):
        y16, _ = synth.clip_frames(synth.params(1, 100 + k, w, h), nf)
        streams.append([engine.DeviceBuffer(f) for f in y16])
    log(f"mixed workload: {len(C2_SIZES)} streams x {nf} frames in {time.perf_counter() - t:.1f}s")
    engines = [engine.RawEngine(w, h) for (w, h) in C2_SIZES]
    for e in engines:
        e.set_resident(True)
    lums = [[engine.RawEngine.luma(dev16=b.ptr, bit_depth=10) for b in dev] for dev in streams]
    decidable = list(range(NR, nf))
    sbs = sum(e.sb_cols * e.sb_rows for e in engines)

    def frame(s):
        return decidable[s % len(decidable)], decidable[(s + 1) % len(decidable)]

    def step_plain(s):
        i, nxt = frame(s)
        refs = [i - k for k in range(1, NR + 1)]
        for e, lum in zip(engines, lums):
            e.launch(lum[i], [lum[r] for r in refs], lam, keys=refs, src_key=i)
            e.prepare(lum[nxt], nxt)

    def step_batch(s, groups):
        i, nxt = frame(s)
        refs = [i - k for k in range(1, NR + 1)]
        for g in groups:
            engine.RawEngine.launch_batch([engines[k] for k in g], [lums[k][i] for k in g],
                                          [[lums[k][r] for r in refs] for k in g], [lam] * len(g), keys=[refs] * len(g),
                                          src_keys=[i] * len(g))
        for e, lum in zip(engines, lums):
            e.prepare(lum[nxt], nxt)

    forms = [("plain", step_plain), ("batched", lambda s: step_batch(s, [range(6)])),
             ("batched_2", lambda s: step_batch(s, [(0, 2, 5), (1, 3, 4)])),
             ("batched_3", lambda s: step_batch(s, [(0, 5), (1, 4), (2, 3)]))]
    rates, maps = {}, {}
    for name, step in forms:
        for e, lum in zip(engines, lums):
            e.prepare(lum[frame(0)[0]], frame(0)[0])
        for s in range(warmup):
            step(s)
        for e in engines:
            e.sync()
        t0 = time.perf_counter()
        for s in range(warmup, warmup + steps):
            step(s)
        for e in engines:
            e.sync()
        wall = time.perf_counter() - t0
        rates[name] = (steps * len(engines) / wall, steps * sbs / wall, wall / steps * 1e3)
        maps[name] = [e.fetch() for e in engines]
        log(f"mixed workload, {name}: {rates[name][0]:.1f} frames/s, {rates[name][1]:.0f} superblocks/s, {rates[name][2]:.3f} ms per step")
    # The one-group batched form again with stage events on (first engine: the batch's stages; every
    # engine counts its own centres): the motion search's pixel-SAD rate over the whole batch.
    for e in engines:
        e.set_timing(True)
        e.stats_reset()
    for s in range(warmup + steps, warmup + 2 * steps):
        step_batch(s, [range(6)])
    st = [e.stats() for e in engines]
    for e in engines:
        e.set_timing(False)
    nb = max(1, st[0][1]["launches"])
    roof = me_roofline(sum(c["me_centres_level0"] for _, c in st) / nb, st[0][0]["me_level0"] / nb)
    same = all(compare_maps(a, b, e.sb_cols) is None for f in maps if f != "plain"
               for a, b, e in zip(maps["plain"], maps[f], engines))
    best = max((f for f in rates if f != "plain"), key=lambda f: rates[f][0])
    for e in engines:
        e.close()
    for dev in streams:
        for b in dev:
            b.free()
    return {"workload": "6 streams (640x360, 854x480, 1280x720 x2, 1920x1080, 2560x1440), 16 frames each, 10-bit, "
                        "7 references, engine_bench method per stream, one frame of every stream per step",
            "superblocks_per_step": sbs, "steps": steps,
            "best_grouping": best, "frames_per_s": round(rates[best][0], 2), "superblocks_per_s": round(rates[best][1]),
            "ms_per_step": round(rates[best][2], 4),
            "by_grouping": {f: {"frames_per_s": round(r[0], 2), "ms_per_step": round(r[2], 4)} for f, r in rates.items()},
            "plain_frames_per_s": round(rates["plain"][0], 2), "plain_superblocks_per_s": round(rates["plain"][1]),
            "speedup": round(rates[best][0] / rates["plain"][0], 3), "decisions_equal_plain": same,
            "roofline_batched": roof}


def bench_transforms(y8, frame, steps, warmup, cpu_seconds=0.0):
    """The transform step of the RD pass's trial scoring (celv2hip_fwd_tx) over every candidate block
    of one 1080p frame: the 425 blocks of each of the 510 superblocks, DCT-II at the block's own size
    (8x4 to 64x64), residuals frame `frame` less frame `frame` - 1 (int16, zero below the picture),
    resident in HBM. Per call: the job list uploaded, one kernel, the coefficients (AVM's layout, low
    32x32) and 64-point dropped energies written to HBM. Returns calls/s, the kernel's own time and its
    HBM roofline; cpu_seconds > 0: the oracle's transform on a bounded sample of the same jobs."""
    from celv2_amd import transform as tx

    H8, W8 = y8[frame].shape
    sb_cols, sb_rows = (W8 + 63) // 64, (H8 + 63) // 64
    plane = np.zeros((sb_rows * 64, sb_cols * 64), np.int16)
    plane[:H8, :W8] = y8[frame].astype(np.int16) - y8[frame - 1].astype(np.int16)
    res = engine.DeviceBuffer(plane)
    probe, off = tx.frame_jobs(0, plane.shape[1], sb_cols, sb_rows, 0)
    coeff = engine.DeviceBuffer(np.zeros(int(off[-1]), np.int32))
    dropped = engine.DeviceBuffer(np.zeros(len(probe), np.int64))
    jobs, off = tx.frame_jobs(res.ptr, plane.shape[1], sb_cols, sb_rows, coeff.ptr, dropped.ptr, scale=1.0)
    n = len(jobs)
    sides = np.stack([1 << jobs["w_log2"].astype(np.int64), 1 << jobs["h_log2"].astype(np.int64)])
    res_bytes = int(np.sum(sides[0] * sides[1]) * 2)
    out_bytes = int(off[-1]) * 4 + 8 * n
    list_bytes = n * 48  # the device job records the kernel reads
    for _ in range(warmup):
        tx.fwd_tx(jobs)
    kms = []
    t0 = time.perf_counter()
    for _ in range(steps):
        kms.append(tx.fwd_tx(jobs))
    wall = (time.perf_counter() - t0) / steps
    kernel_s = float(np.median(kms)) / 1e3
    alg = res_bytes + out_bytes + list_bytes
    out = {"workload": f"every candidate block of one 1080p frame ({sb_cols * sb_rows} superblocks x 425 blocks, 8x4 to 64x64, "
                       "DCT-II at the block's size, scale 1.0), residuals of two clip frames resident in HBM",
           "jobs_per_call": n, "calls_per_s": round(1.0 / wall, 2), "jobs_per_s": round(n / wall),
           "ms_per_call": round(wall * 1e3, 4), "kernel_ms": round(kernel_s * 1e3, 4),
           "kernel_jobs_per_s": round(n / kernel_s),
           "roofline": {"bound": "hbm", "kernel": "k_fwd_tx", "achieved": round(alg / kernel_s / 1e9, 1), "peak": 8000.0,
                        "unit": "GB/s", "frac": round(alg / kernel_s / 8e12, 4), "traffic": TX_TRAFFIC_BYTES,
                        "algorithmic_bytes_per_call": alg,
                        "bytes": {"residuals": res_bytes, "coefficients_and_dropped": out_bytes, "job_records": list_bytes}}}
    if cpu_seconds > 0:
        # The oracle's transform (test infrastructure, timed here only) on one thread over the same
        # jobs from the first, a run sized to about cpu_seconds; its coefficients also check the GPU's.
        O = ctypes.CDLL(os.path.join(ROOT, "oracle", "liboracle.so"))
        O.orc_fwd_tx_frame.restype = ctypes.c_longlong
        O.orc_fwd_tx_frame.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_longlong,
                                       ctypes.c_longlong, ctypes.c_double, ctypes.c_void_p]
        want = np.zeros(int(off[-1]), np.int32)
        t0 = time.perf_counter()
        O.orc_fwd_tx_frame(plane.ctypes.data, plane.shape[1], sb_cols, sb_rows, 0, 4250, 1.0, want.ctypes.data)
        per = (time.perf_counter() - t0) / 4250
        m = int(min(n, max(4250, cpu_seconds / per)))
        t0 = time.perf_counter()
        O.orc_fwd_tx_frame(plane.ctypes.data, plane.shape[1], sb_cols, sb_rows, 0, m, 1.0, want.ctypes.data)
        dt = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": round(m / dt), "unit": "jobs/s", "cores": 1, "kind": "port",
                               "sample": f"the first {m} jobs ({m // 425} superblocks) of the same list, the oracle's "
                                         f"transform on one thread in {dt:.1f} s"}
        flat = coeff.read(np.int32, int(off[m]))
        out["sample_equals_oracle"] = bool(np.array_equal(flat, want[:int(off[m])]))
    for b in (res, coeff, dropped):
        b.free()
    log(f"transforms: {n} jobs a call, {out['ms_per_call']} ms a call ({out['kernel_ms']} ms kernel, "
        f"{out['roofline']['achieved']} GB/s algorithmic)")
    return out


def bench_learned(lum, decidable, NR, qindex, steps, warmup, plain_fps):
    """The learned-pruned form on the same workload: the trained network (celv2_amd/pruner) runs inside
    every launch on the engine's coarse search and limits each superblock's references; frames/s over
    the same steps, then, untimed, the decision-cost regret against the plain engine on every
    decidable frame (sum of J_pruned - J_plain over sum of J_plain), the mean reference limit and the
    fallback rate."""
    lam = engine.lambda_q8(qindex)
    w, cfg = pruner.load()
    op = pruner.operating_point(cfg, qindex)
    threshold, band, gth, sth = op if op else (None, None, 0.0, 0.0)
    P = engine.Pruner(max_positions=(1920 // 64 + 1) * (1080 // 64 + 1), seed=1)
    P.set_params(w)
    E = engine.RawEngine(1920, 1080)
    E.set_resident(True)
    if op:
        E.set_pruner(P, threshold, band, gthreshold=gth, sthreshold=sth)

    def step(s):
        i = decidable[s % len(decidable)]
        nxt = decidable[(s + 1) % len(decidable)]
        refs = [i - k for k in range(1, NR + 1)]
        E.launch(lum[i], [lum[r] for r in refs], lam, keys=refs, src_key=i)
        E.prepare(lum[nxt], nxt)

    E.prepare(lum[decidable[0]], decidable[0])
    for s in range(warmup):
        step(s)
    E.stats_reset()
    t0 = time.perf_counter()
    for s in range(warmup, warmup + steps):
        step(s)
    E.sync()
    wall = time.perf_counter() - t0
    fps = steps / wall
    per = {k: round(v, 4) for k, v in stage_times(E, step, warmup + steps, steps).items()}
    roof = me_roofline(stage_times.counts["me_centres_level0"], per["me_level0"])
    # Untimed: inter candidates counted, regret, limits and fallbacks on every decidable frame.
    E.set_counting(True)
    E.stats_reset()
    Plain = engine.RawEngine(1920, 1080)
    Plain.set_resident(True)
    jp, jl, ks = [], [], []
    for i in decidable:
        refs = [i - k for k in range(1, NR + 1)]
        for eng, out in ((Plain, jp), (E, jl)):
            eng.launch(lum[i], [lum[r] for r in refs], lam, keys=refs)
            eng.sync()
            out.append(eng.fetch_sb_costs())
        ks.append(E.fetch_ref_limits())
    _, c2 = E.stats()
    Plain.close()
    E.close()
    P.close()
    jp, jl, ks = (np.stack(a).astype(np.float64) for a in (jp, jl, ks))
    regret = float((jl - jp).sum() / jp.sum())
    nd = len(decidable)
    inter_l = c2["inter_candidates"] / nd
    log(f"learned: {fps:.1f} frames/s ({fps / plain_fps:.2f}x plain), regret {regret * 100:.3f}% of J, mean limit {ks.mean():.2f}, "
        f"fallbacks {c2['pruner_fallbacks'] / ks.size:.4f}; per frame ms {json.dumps(per)}")
    return {"frames_per_s": round(fps, 2), "speedup_vs_plain": round(fps / plain_fps, 3),
            "cost_regret": round(regret, 6), "cost_regret_unit": "sum(J_learned - J_plain) / sum(J_plain) over every decidable frame (engine J = SATD*256 + lambda*bits)",
            "bd_rate": "pending (needs the full encoder and AVM)",
            "mean_reference_limit": round(float(ks.mean()), 3), "fallback_rate": round(c2["pruner_fallbacks"] / ks.size, 5),
            "threshold": threshold, "band": band, "stage_ms_per_frame": per, "roofline": roof,
            "inter_candidates_per_frame": round(inter_l), "compound_blocks_per_frame": round(c2["compound_blocks"] / nd),
            "network": f"conv3x3 {engine.FEATURES}-32, conv3x3 32-32, conv1x1 32-{engine.N_OUT} (6 reference, 3 intra-group, 1 search outputs), "
                       f"BF16 on MFMA, trained on {cfg['train_superblocks']} superblocks of synthetic clips (celv2_amd/pruner/train.py)",
            "search_threshold": sth}


def stage_times(E, step, first, steps):
    """Per-frame stage times (ms) over `steps` more calls of step(s) from s = first, with the stage
    events on for that window only."""
    E.set_timing(True)
    E.stats_reset()
    for s in range(first, first + steps):
        step(s)
    ms, cnt = E.stats()
    E.set_timing(False)
    n = max(1, cnt["launches"])
    per = {k: v / n for k, v in ms.items() if k != "intra_prepare"}
    per["intra"] = ms["intra_prepare"] / n
    stage_times.counts = {k: v / n for k, v in cnt.items() if k != "launches"}
    return per


def me_roofline(centres_per_frame, me_ms):
    """The motion search's level-0 launch: its pixel-SAD rate (289 x 64 per centre searched at level 0,
    as the kernel counts them) over its HIP-event time, as a fraction of the v_sad_u8 peak."""
    achieved = centres_per_frame * 289 * 64 / (me_ms / 1e3) / 1e12
    return {"bound": "valu", "kernel": "k_me_level (level 0)", "achieved": round(achieved, 3), "peak": round(SAD_PEAK_TPX, 2),
            "unit": "T pixel-SAD/s (v_sad_u8)", "frac": round(achieved / SAD_PEAK_TPX, 4), "ms_per_frame": round(me_ms, 4)}


def _timed(step, engines, prepare, steps, warmup):
    """Seconds for `steps` calls of step(s) after `warmup` untimed ones (every engine synchronized on
    both sides); prepare() first."""
    prepare()
    for s in range(warmup):
        step(s)
    for e in engines:
        e.sync()
    t0 = time.perf_counter()
    for s in range(warmup, warmup + steps):
        step(s)
    for e in engines:
        e.sync()
    return time.perf_counter() - t0


def _regret(plain, learned, launch, frames):
    """Cost regret of the learned engines against the plain ones over frames (launch(engine, i) runs
    one frame), the mean reference limit and the fallback rate."""
    jp = jl = 0.0
    ks = []
    for e in learned:
        e.stats_reset()
    for i in frames:
        for ep, el in zip(plain, learned):
            launch(ep, i)
            launch(el, i)
            ep.sync()
            el.sync()
            jp += ep.fetch_sb_costs().astype(np.float64).sum()
            jl += el.fetch_sb_costs().astype(np.float64).sum()
            ks.append(el.fetch_ref_limits())
    fb = sum(e.stats()[1]["pruner_fallbacks"] for e in learned)
    ks = np.concatenate(ks)
    return jl / jp - 1, float(ks.mean()), fb / ks.size


def bench_qindex_sweep(lum, decidable, NR, steps, warmup):
    """The plain and learned forms of the 1080p workload at qindex 100, 130, 160 and 190 (the learned
    form at each qindex's operating point): frames/s of each over the same steps, and the learned
    form's cost regret, mean reference limit and fallback rate over every decidable frame."""
    w, cfg = pruner.load()
    P = engine.Pruner(max_positions=(1920 // 64 + 1) * (1080 // 64 + 1), seed=1)
    P.set_params(w)
    Ep, El = engine.RawEngine(1920, 1080), engine.RawEngine(1920, 1080)
    for e in (Ep, El):
        e.set_resident(True)
    out = {}
    for q in (100, 130, 160, 190):
        lam = engine.lambda_q8(q)
        op = pruner.operating_point(cfg, q)
        El.set_pruner(P, *op) if op else El.set_pruner(None)

        def launch(e, i):
            refs = [i - k for k in range(1, NR + 1)]
            e.launch(lum[i], [lum[r] for r in refs], lam, keys=refs)

        def stepper(e):
            def step(s):
                i, nxt = decidable[s % len(decidable)], decidable[(s + 1) % len(decidable)]
                refs = [i - k for k in range(1, NR + 1)]
                e.launch(lum[i], [lum[r] for r in refs], lam, keys=refs, src_key=i)
                e.prepare(lum[nxt], nxt)
            return step

        first = lambda e: (lambda: e.prepare(lum[decidable[0]], decidable[0]))
        tp = _timed(stepper(Ep), [Ep], first(Ep), steps, warmup)
        tl = _timed(stepper(El), [El], first(El), steps, warmup)
        r, mk, fb = _regret([Ep], [El], launch, decidable)
        out[str(q)] = {"operating_point": op, "plain_frames_per_s": round(steps / tp, 2), "learned_frames_per_s": round(steps / tl, 2),
                       "speedup": round(tp / tl, 3), "cost_regret": round(r, 6), "mean_reference_limit": round(mk, 3),
                       "fallback_rate": round(fb, 5)}
        log(f"qindex {q}: plain {steps / tp:.1f} frames/s, learned {steps / tl:.1f} ({tp / tl:.2f}x), regret {r * 100:.3f}%, mean limit {mk:.2f}")
    for e in (Ep, El):
        e.close()
    P.close()
    return out


def bench_cel(NR, qindex, steps=20, warmup=3, nf=48):
    """Two 3840x2160 10-bit streams of flat-shaded cel content (workload 4: Voronoi cells of a
    16-colour palette, outlines, cuts at frames 16 and 32, motion held every other frame, grain
    sigma 1), 48 frames each, 7 references. Timed: the plain engines launched one by one, then the
    learned engines with both streams in one shared launch (celv2hip_launch_batch); the learned form's
    cost regret against plain over every decidable frame of both streams."""
    t = time.perf_counter()
    W4, H4 = 3840, 2160
    streams = [[engine.DeviceBuffer(f) for f in synth.clip_frames(synth.params(4, seed, W4, H4), nf)[0]] for seed in (1, 2)]
    lums = [[engine.RawEngine.luma(dev16=b.ptr, bit_depth=10) for b in dev] for dev in streams]
    log(f"cel workload: 2 streams x {nf} frames 3840x2160 in {time.perf_counter() - t:.1f}s")
    w, cfg = pruner.load()
    P = engine.Pruner(max_positions=(W4 // 64 + 1) * (H4 // 64 + 1), seed=1)
    P.set_params(w)
    lam = engine.lambda_q8(qindex)
    plain = [engine.RawEngine(W4, H4) for _ in streams]
    learned = [engine.RawEngine(W4, H4) for _ in streams]
    for e in plain + learned:
        e.set_resident(True)
    op = pruner.operating_point(cfg, qindex)
    for e in learned:
        if op:
            e.set_pruner(P, *op)
    decidable = list(range(NR, nf))

    def frame(s):
        return decidable[s % len(decidable)], decidable[(s + 1) % len(decidable)]

    def prep(engines):
        return lambda: [e.prepare(lum[decidable[0]], decidable[0]) for e, lum in zip(engines, lums)]

    def step_plain(s):
        i, nxt = frame(s)
        refs = [i - k for k in range(1, NR + 1)]
        for e, lum in zip(plain, lums):
            e.launch(lum[i], [lum[r] for r in refs], lam, keys=refs, src_key=i)
            e.prepare(lum[nxt], nxt)

    def step_learned(s):
        i, nxt = frame(s)
        refs = [i - k for k in range(1, NR + 1)]
        engine.RawEngine.launch_batch(learned, [lum[i] for lum in lums], [[lum[r] for r in refs] for lum in lums],
                                      [lam] * 2, keys=[refs] * 2, src_keys=[i] * 2)
        for e, lum in zip(learned, lums):
            e.prepare(lum[nxt], nxt)

    tp = _timed(step_plain, plain, prep(plain), steps, warmup)
    tl = _timed(step_learned, learned, prep(learned), steps, warmup)

    def launch(e, i):
        lum = lums[(plain + learned).index(e) % 2]
        refs = [i - k for k in range(1, NR + 1)]
        e.launch(lum[i], [lum[r] for r in refs], lam, keys=refs)

    r, mk, fb = _regret(plain, learned, launch, decidable)
    for e in plain + learned:
        e.close()
    P.close()
    for dev in streams:
        for b in dev:
            b.free()
    sbs = 2 * (W4 // 64) * ((H4 + 63) // 64)
    log(f"cel workload: plain {2 * steps / tp:.1f} frames/s, learned + batched {2 * steps / tl:.1f} ({tp / tl:.2f}x), "
        f"regret {r * 100:.3f}%, mean limit {mk:.2f}")
    return {"workload": "2 streams of 3840x2160 10-bit cel content (workload 4), 48 frames, 7 references, qindex "
                        f"{qindex}; learned form: both streams in one launch", "steps": steps,
            "plain_frames_per_s": round(2 * steps / tp, 2), "learned_batched_frames_per_s": round(2 * steps / tl, 2),
            "superblocks_per_step": sbs, "speedup": round(tp / tl, 3), "cost_regret": round(r, 6),
            "mean_reference_limit": round(mk, 3), "fallback_rate": round(fb, 5),
            "bd_rate": "pending (needs the full encoder and AVM)"}


def bench_gop(lum, qindex, passes=10):
    """The 32-frame clip coded as the frame loop orders it (celv2_amd/gop.py: keyframe, then 16-frame
    mini-GOPs with a hidden anchor and pyramid, 7 references, per-frame qindex deltas bq 16, aq -16,
    level step 8): every coded frame decided against its ranked references at its own qindex, the
    next coded frame prepared behind each launch. One untimed pass, then `passes` timed ones (keys
    fresh per pass, so every pass builds its references as a new sequence would). The learned form
    takes the operating point of each frame's qindex. Then, untimed, the learned form's cost regret
    against plain over every coded frame."""
    nf = len(lum)
    cfg = gop.defaults(7, qindex, 1920, 1080, long_gop=True)
    entries = gop.coded(gop.plan(cfg, nf))
    w, pcfg = pruner.load()
    P = engine.Pruner(max_positions=(1920 // 64 + 1) * (1080 // 64 + 1), seed=1)
    P.set_params(w)
    ops = {e.qindex: pruner.operating_point(pcfg, e.qindex) for e in entries}

    def run(E, base, learned):
        for j, f in enumerate(entries):
            if learned:
                op = ops[f.qindex] if f.refs else None
                E.set_pruner(P, *op) if op else E.set_pruner(None)
            E.launch(lum[f.display], [lum[r] for r in f.refs], engine.lambda_q8(f.qindex),
                     keys=[base + r for r in f.refs], src_key=base + f.display)
            if j + 1 < len(entries):
                E.prepare(lum[entries[j + 1].display], base + entries[j + 1].display)

    rates = {}
    for name, learned in (("plain", False), ("learned", True)):
        E = engine.RawEngine(1920, 1080)
        E.set_resident(True)
        run(E, 0, learned)
        E.sync()
        t0 = time.perf_counter()
        for k in range(passes):
            run(E, 1000 * (k + 1), learned)
        E.sync()
        rates[name] = passes * len(entries) / (time.perf_counter() - t0)
        E.close()
    # Untimed: regret, limits and fallbacks over every coded inter frame.
    Ep, El = engine.RawEngine(1920, 1080), engine.RawEngine(1920, 1080)
    for e in (Ep, El):
        e.set_resident(True)
    jp = jl = 0.0
    ks, fb = [], 0
    El.stats_reset()
    for f in entries:
        if not f.refs:
            continue
        op = ops[f.qindex]
        El.set_pruner(P, *op) if op else El.set_pruner(None)
        for e in (Ep, El):
            e.launch(lum[f.display], [lum[r] for r in f.refs], engine.lambda_q8(f.qindex), keys=list(f.refs))
            e.sync()
        jp += Ep.fetch_sb_costs().astype(np.float64).sum()
        jl += El.fetch_sb_costs().astype(np.float64).sum()
        ks.append(El.fetch_ref_limits())
    fb = El.stats()[1]["pruner_fallbacks"]
    ks = np.concatenate(ks)
    for e in (Ep, El):
        e.close()
    P.close()
    refs_mean = float(np.mean([len(f.refs) for f in entries]))
    log(f"gop: {len(entries)} coded frames (mean {refs_mean:.2f} references): plain {rates['plain']:.1f} frames/s, "
        f"learned {rates['learned']:.1f} ({rates['learned'] / rates['plain']:.2f}x), regret {(jl / jp - 1) * 100:.3f}%, "
        f"mean limit {ks.mean():.2f}")
    return {"workload": f"the same {nf}-frame 1080p 10-bit clip in the frame loop's order: keyframe, 16-frame mini-GOPs "
                        "(hidden anchor, 5 pyramid frames, 10 B frames), up to 7 ranked references, qindex "
                        f"{qindex} with per-frame deltas (anchor -16, pyramid levels 0/+8/+16, B frames +16)",
            "coded_frames": len(entries), "mean_references": round(refs_mean, 3), "passes": passes,
            "qindex_by_type": {t: sorted({f.qindex for f in entries if f.type == t}) for t in "KAMBP" if any(f.type == t for f in entries)},
            "plain_frames_per_s": round(rates["plain"], 2), "learned_frames_per_s": round(rates["learned"], 2),
            "speedup": round(rates["learned"] / rates["plain"], 3), "cost_regret": round(jl / jp - 1, 6),
            "mean_reference_limit": round(float(ks.mean()), 3), "fallback_rate": round(fb / ks.size, 5),
            "operating_points": {str(q): op for q, op in sorted(ops.items())},
            "bd_rate": "pending (needs the full encoder and AVM)"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--frames", type=int, default=32)
    ap.add_argument("--refs", type=int, default=7)
    ap.add_argument("--qindex", type=int, default=130)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-batched", action="store_true", help="skip the mixed-resolution batched form")
    ap.add_argument("--no-learned", action="store_true", help="skip the learned-pruned form")
    ap.add_argument("--no-sweep", action="store_true", help="skip the qindex sweep of the learned form")
    ap.add_argument("--no-cel", action="store_true", help="skip the 4K cel workload (learned + batched)")
    ap.add_argument("--no-gop", action="store_true", help="skip the clip in the frame loop's coding order")
    ap.add_argument("--cpu-frames", type=int, default=3, help="frames of the CPU engine timed on all threads")
    ap.add_argument("--row-groups", type=int, default=0, help="the launch's row groups (0: tuned by timing)")
    ap.add_argument("--no-transforms", action="store_true", help="skip the RD pass's forward transforms")
    ap.add_argument("--no-graph", action="store_true", help="skip the resident steps replayed as HIP graphs")
    ap.add_argument("--calibrate-copy-mib", type=int, default=0,
                    help="first run a device-to-device copy of this many MiB (a known-size kernel for scaling traffic counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: the decision engine runs on a single MI355X (no multi-GPU path is built)",
              file=sys.stderr)
        sys.exit(2)
    if engine.device_count() < 1:
        raise SystemExit("no HIP device")

    W, H, NR = 1920, 1080, args.refs
    p = synth.params(1, args.seed, W, H)
    t = time.perf_counter()
    y16, y8 = synth.clip_frames(p, args.frames)
    log(f"clip: {args.frames} frames {W}x{H} 10-bit in {time.perf_counter() - t:.1f}s")
    lam = engine.lambda_q8(args.qindex)
    dev = [engine.DeviceBuffer(f) for f in y16]  # resident 10-bit luma
    if args.calibrate_copy_mib > 0:
        blk = np.zeros(args.calibrate_copy_mib << 20, np.uint8)
        a, b = engine.DeviceBuffer(blk), engine.DeviceBuffer(blk)
        log(f"calibration copy: {b.copy_from(a)} bytes device to device")
        a.free()
        b.free()
    E = engine.RawEngine(W, H)
    E.set_resident(True)
    lum = [engine.RawEngine.luma(dev16=b.ptr, bit_depth=10) for b in dev]
    decidable = list(range(NR, args.frames))

    def step(s):
        i = decidable[s % len(decidable)]
        nxt = decidable[(s + 1) % len(decidable)]
        refs = [i - k for k in range(1, NR + 1)]
        E.launch(lum[i], [lum[r] for r in refs], lam, keys=refs, src_key=i)
        E.prepare(lum[nxt], nxt)
        return i

    E.prepare(lum[decidable[0]], decidable[0])
    for s in range(args.warmup):
        step(s)
    E.stats_reset()  # synchronizes
    t0 = time.perf_counter()
    last = None
    for s in range(args.warmup, args.warmup + args.steps):
        last = step(s)
    E.sync()
    wall = time.perf_counter() - t0
    _, cnt = E.stats()
    fps = args.steps / wall
    n = cnt["launches"]
    gpu_map = E.fetch()
    # Stage times: the same number of frames again with the stage events on (each event between two
    # kernels costs the stream a few microseconds, so they stay out of the timed window).
    per = stage_times(E, step, args.warmup + args.steps, args.steps)
    log(f"timed {args.steps} frames in {wall:.3f}s: {fps:.1f} frames/s; per frame ms (separate window with stage events) "
        f"{json.dumps({k: round(v, 3) for k, v in per.items()})}")
    log(f"counts over the window: {cnt}")
    # Non-repeated inter candidates per frame, counted on 4 more frames outside the timed window.
    E.set_counting(True)
    E.stats_reset()
    for s_ in range(args.warmup + args.steps, args.warmup + args.steps + 4):
        step(s_)
    E.sync()
    _, cnt4 = E.stats()
    E.set_counting(False)
    inter_cands = cnt4["inter_candidates"] / max(1, cnt4["launches"])

    # The same resident steps eager and as replayed HIP graphs (celv2hip_set_graph): frames/s and the
    # host thread's CPU time to queue a frame (launch + prepare), then the last frame's decisions of
    # both against each other.
    def host_loop(graph, base):
        E.set_graph(graph)
        for s_ in range(base, base + args.warmup):
            step(s_)
        E.sync()
        cpu, t0 = 0.0, time.perf_counter()
        for s_ in range(base + args.warmup, base + args.warmup + args.steps):
            c0 = time.thread_time()
            f_ = step(s_)
            cpu += time.thread_time() - c0
        E.sync()
        return args.steps / (time.perf_counter() - t0), cpu / args.steps * 1e3, f_

    graph_form = None
    if not args.no_graph:
        base = args.warmup + args.steps + 4
        fps_e, cpu_e, f_e = host_loop(False, base)
        map_e = E.fetch()
        fps_g, cpu_g, f_g = host_loop(True, base)
        map_g = E.fetch()
        graph_form = {"frames_per_s": round(fps_g, 2), "host_cpu_ms_per_frame": round(cpu_g, 4),
                      "eager_frames_per_s": round(fps_e, 2), "eager_host_cpu_ms_per_frame": round(cpu_e, 4),
                      "graph_launches_and_builds": list(E.graph_stats()),
                      "decisions_equal_eager": f_e == f_g and compare_maps(map_e, map_g, E.sb_cols) is None}
        E.set_graph(False)
        log(f"graph form: {fps_g:.1f} frames/s, host {cpu_g:.3f} ms a frame; eager {fps_e:.1f} frames/s, host {cpu_e:.3f} ms")

    # The intra dimension of the pruned form at its extreme: every superblock scoring DC, V and H only
    # (celv2hip_set_intra_masks), against every mode: frames/s of the resident steps, the intra stage,
    # and the decision cost it adds (sum of the superblocks' costs over 4 frames).
    intra_groups = None
    if not args.no_learned:
        def cost_sum(masks):
            E.set_intra_masks(masks)
            tot = 0
            for f_ in decidable[-4:]:
                refs = [f_ - k for k in range(1, NR + 1)]
                E.launch(lum[f_], [lum[r] for r in refs], lam, keys=refs)
                E.sync()
                tot += int(E.fetch_sb_costs().sum())
            return tot
        nsb = E.sb_cols * E.sb_rows
        out_ig = {}
        for name, m in (("all_groups", None), ("dc_v_h_only", np.zeros(nsb, np.uint8))):
            E.set_intra_masks(m)
            E.prepare(lum[decidable[0]], decidable[0])
            b0 = args.warmup + 2 * args.steps + 8
            for s_ in range(b0, b0 + args.warmup):
                step(s_)
            E.sync()
            t0 = time.perf_counter()
            for s_ in range(b0 + args.warmup, b0 + args.warmup + args.steps):
                step(s_)
            E.sync()
            fps_m = args.steps / (time.perf_counter() - t0)
            st = stage_times(E, step, b0 + args.warmup + args.steps, args.steps)
            out_ig[name] = {"frames_per_s": round(fps_m, 2), "intra_ms_per_frame": round(st["intra"], 4)}
        c_all, c_g0 = cost_sum(None), cost_sum(np.zeros(nsb, np.uint8))
        E.set_intra_masks(None)
        intra_groups = {**out_ig, "cost_regret": round((c_g0 - c_all) / c_all, 5),
                        "what": "every superblock restricted to DC, V and H against every intra mode, resident steps"}
        log(f"intra groups: DC/V/H only {out_ig['dc_v_h_only']['frames_per_s']} frames/s (intra "
            f"{out_ig['dc_v_h_only']['intra_ms_per_frame']} ms) against {out_ig['all_groups']['frames_per_s']} "
            f"({out_ig['all_groups']['intra_ms_per_frame']} ms), cost +{100 * intra_groups['cost_regret']:.3f} %")

    # Algorithmic integer work per frame (weights of the reference's cost rules, counted as executed).
    sbs = E.sb_cols * E.sb_rows
    cells0 = ((W + 7) // 8) * ((H + 7) // 8)
    cells = 0
    w_, h_ = (W + 7) & ~7, (H + 7) & ~7
    levels_cells = []
    for l in range(5):
        levels_cells.append(((w_ + 7) // 8) * ((h_ + 7) // 8))
        w_, h_ = w_ // 2, h_ // 2
    centre_cells = sum(levels_cells[:4]) * NR  # levels below the top rank 6 centres
    me_sad_px = cnt["me_centres"] / n * 289 * 64 + centre_cells * 6 * 64
    # Level 0 alone (the dominant launch): its centres' windows and its blocks' 6 centre SADs.
    me0_sad_px = cnt["me_centres_level0"] / n * 289 * 64 + levels_cells[0] * NR * 6 * 64
    me_satd_px = cells0 * NR * 18 * 64
    me_ops = SAD_OPS * me_sad_px + SATD_OPS * me_satd_px
    inter_ops = SATD_OPS * cnt["tile_evals"] / n * 4096
    intra_ops = SATD_OPS * sbs * 479232
    # The dominant kernel launch is the motion search's level 0 (about three quarters of the motion
    # search's kernel time, the largest single launch of a frame); its bound is the v_sad_u8 rate.
    dom = max(("me", "inter", "intra", "finalize"), key=lambda k: per[k])
    me_s = per["me"] / 1e3
    me0_s = per["me_level0"] / 1e3
    achieved = me0_sad_px / me0_s / 1e12
    roofline = {"bound": "valu", "kernel": "k_me_level, level 0 (one launch per frame, all 7 references)", "dominant_stage": dom,
                "achieved": round(achieved, 3), "peak": round(SAD_PEAK_TPX, 2), "unit": "T pixel-SAD/s (v_sad_u8)",
                "frac": round(achieved / SAD_PEAK_TPX, 4),
                "traffic": ME_TRAFFIC_BYTES, "traffic_gbps": round(ME_TRAFFIC_BYTES / me0_s / 1e9, 1),
                "measured_ceiling": SAD_MEASURED_TPX, "frac_of_measured_ceiling": round(achieved / SAD_MEASURED_TPX, 4),
                "work_per_frame": me0_sad_px, "ms_per_frame": round(per["me_level0"], 4),
                "five_levels": {"work_per_frame": me_sad_px, "stage_ms_per_frame": round(per["me"], 4),
                                "achieved": round(me_sad_px / me_s / 1e12, 3)},
                "int_ops_model_tops": round(me_ops / me_s / 1e12, 2)}
    total_ops = me_ops + inter_ops + intra_ops
    # Scored candidates (intra mode evaluations + non-repeated inter (block, MV, reference) candidates
    # + compound evaluations) per frame.
    cands = sbs * 420 * 13 + inter_cands + cnt["compound_blocks"] / n

    # The headline: engine_bench's step through the C++ engine (celv2::HipEngine): launch -> prepare ->
    # collect, the frame's DecisionMap on the host when collect returns; frames resident in HBM.
    HE = engine.HipEngine(W, H)

    def bstep(s):
        i = decidable[s % len(decidable)]
        nxt = decidable[(s + 1) % len(decidable)]
        refs = [i - k for k in range(1, NR + 1)]
        HE.launch_device(lum[i], [lum[r] for r in refs], lam=lam, keys=refs, src_key=i)
        HE.prepare_device(lum[nxt], nxt)
        HE.collect(export=False)  # the DecisionMap as an encoder receives it
        return i

    # The launch's row groups: tuned on the first decidable frame (steady-state launch -> prepare ->
    # collect at each candidate count), or as given.
    if args.row_groups > 0:
        HE.set_row_groups(args.row_groups)
        tuning = {"chosen": args.row_groups, "tuned": False}
    else:
        f0 = decidable[0]
        r0 = [f0 - k for k in range(1, NR + 1)]
        g, timed = HE.autotune_device(lum[f0], [lum[r] for r in r0], lam=lam, keys=r0, src_key=f0, frames=5)
        tuning = {"chosen": g, "tuned": True, "median_ms_per_frame": {str(k): round(v, 3) for k, v in sorted(timed.items())}}
        log(f"row groups: {g} (median ms per frame by count: {tuning['median_ms_per_frame']})")
    HE.prepare_device(lum[decidable[0]], decidable[0])
    for s in range(args.warmup):
        bstep(s)
    HE.sync()
    t0 = time.perf_counter()
    for s in range(args.warmup, args.warmup + args.steps):
        blast = bstep(s)
    HE.sync()
    wall_b = time.perf_counter() - t0
    fps_b = args.steps / wall_b
    # The last frame's DecisionMap exported: it must be the device-resident run's decisions of that frame.
    refs_b = [blast - k for k in range(1, NR + 1)]
    HE.launch_device(lum[blast], [lum[r] for r in refs_b], lam=lam, keys=refs_b, src_key=blast)
    bmap, _ = HE.collect()
    E.launch(lum[blast], [lum[r] for r in refs_b], lam, keys=refs_b, src_key=blast)
    boundary_equal = compare_maps(bmap, E.fetch(), E.sb_cols) is None
    HE.close()
    log(f"boundary (launch -> prepare -> collect into the host DecisionMap, frames in HBM): {fps_b:.1f} frames/s; "
        f"kernel only {fps:.1f}; DecisionMap equal to the resident run's: {boundary_equal}")

    # The same steps with host frames (engine_bench's own inputs: 10-bit planes uploaded by launch and
    # prepare; references reused by key).
    HH = engine.HipEngine(W, H, full_depth=True)
    HH.set_row_groups(tuning["chosen"])

    def hstep(s):
        i = decidable[s % len(decidable)]
        nxt = decidable[(s + 1) % len(decidable)]
        refs = [i - k for k in range(1, NR + 1)]
        HH.launch(y8[i], [y8[r] for r in refs], lam=lam, keys=refs, src_key=i, src16=y16[i],
                  refs16=[y16[r] for r in refs], bit_depth=10)
        HH.prepare(y8[nxt], nxt, y16=y16[nxt], bit_depth=10)
        HH.collect(export=False)

    HH.prepare(y8[decidable[0]], decidable[0], y16=y16[decidable[0]], bit_depth=10)
    for s in range(args.warmup):
        hstep(s)
    HH.sync()
    t0 = time.perf_counter()
    for s in range(args.warmup, args.warmup + args.steps):
        hstep(s)
    HH.sync()
    fps_h = args.steps / (time.perf_counter() - t0)
    HH.close()
    log(f"host frames (10-bit uploads by launch / prepare, DecisionMap on the host): {fps_h:.1f} frames/s")

    batched = None if args.no_batched else bench_batched(NR, lam)
    learned = None
    if not args.no_learned and NR == 7:
        learned = bench_learned(lum, decidable, NR, args.qindex, args.steps, args.warmup, fps)
        learned["scored_candidates_per_frame"] = round(sbs * 420 * 13 + learned["inter_candidates_per_frame"] + learned["compound_blocks_per_frame"])
        learned["scored_candidates_per_s"] = round(learned["scored_candidates_per_frame"] * learned["frames_per_s"])
        if not args.no_sweep:
            learned["qindex_sweep"] = bench_qindex_sweep(lum, decidable, NR, args.steps, args.warmup)
    cel = None if args.no_cel or args.no_learned or NR != 7 else bench_cel(NR, args.qindex)
    gop_form = None if args.no_gop or NR != 7 else bench_gop(lum, args.qindex)
    transforms = None if args.no_transforms else bench_transforms(y8, last, 20, 3, 0.0 if args.no_cpu else 10.0)

    cpu = None
    parity = None
    if not args.no_cpu:
        # This host's CPUs as nproc reports them (the process's affinity); the engine's workers are
        # capped at 16, the CPU share a one-GPU box gives a process.
        nproc = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
        threads = max(1, min(16, nproc))
        k0 = decidable.index(last)
        frames_cpu = [decidable[(k0 + k) % len(decidable)] for k in range(max(1, args.cpu_frames))]  # `last` first
        dts = []
        for f in frames_cpu:
            refs = [f - k for k in range(1, NR + 1)]
            want_gpu = gpu_map if f == last else None
            dt, diff = cpu_baseline(y8, f, refs, lam, want_gpu, threads)
            dts.append(dt)
            if f == last:
                parity = diff is None
                if diff:
                    log("PARITY FAILURE vs the CPU engine:", diff)
        dt_all = sum(dts) / len(dts)
        # One thread (as engine_bench runs the CPU engine) on one whole frame.
        dt1, _ = cpu_baseline(y8, last, [last - k for k in range(1, NR + 1)], lam, None, 1)
        cpu = {"value": round(1.0 / dt_all, 4), "unit": "frames/s", "cores": threads, "nproc": nproc, "kind": "port",
               "sample": f"{len(dts)} frames ({', '.join(map(str, frames_cpu))}) of the same 1080p 10-bit 7-reference workload, "
                         f"the oracle's CpuEngine on {threads} threads: {sum(dts):.1f} s",
               "one_thread": {"value": round(1.0 / dt1, 5), "unit": "frames/s", "cores": 1,
                              "sample": f"frame {last}, whole, 1 thread: {dt1:.1f} s"}}
        log(f"cpu engine: {dt_all:.2f} s per frame on {threads} threads (nproc {nproc}), 1 thread {dt1:.1f} s; parity {'ok' if parity else 'FAILED'}")
    for b in dev:
        b.free()

    out = {
        "metric": METRIC,
        "value": round(fps_b, 2),
        "unit": "frames/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall_b / args.steps * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u8",
        "data": "synthetic (seeded 10-bit generator, frames resident in HBM; DecisionMap on the host)",
        "config": {"workload": "1920x1080 10-bit 4:2:0, 32-frame clip, 7 references (previous source frames), qindex 130, "
                               "plain HIP engine, engine_bench method through celv2::HipEngine (frame i decided against i-1..i-7, "
                               "next frame prepared, launch -> prepare -> collect into the host DecisionMap)",
                   "width": W, "height": H, "refs": NR, "qindex": args.qindex, "lambda_q8": lam, "form": "plain"},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "parity_vs_cpu_engine": parity,
        "kernel_only": {"frames_per_s": round(fps, 2), "ms_per_step": round(wall / args.steps * 1e3, 4),
                        "what": "the same steps with the records left in HBM (no download, no DecisionMap)"},
        "host_frames_fps": round(fps_h, 2),
        "boundary_map_equals_resident": boundary_equal,
        "row_groups": tuning,
        "stage_ms_per_frame": {k: round(v, 4) for k, v in per.items()},
        "counts_per_frame": {k: round(v / max(1, n), 1) for k, v in cnt.items() if k != "launches"},
        "scored_candidates_per_frame": round(cands),
        "scored_candidates_per_s": round(cands * fps_b),
        "me_positions_per_s": round(cnt["me_centres"] / n * 289 * fps_b),
        "int_ops_per_frame": total_ops,
        "all_kernels_tops": round(total_ops * fps / 1e12, 3),
        "forms": {"plain": {"frames_per_s": round(fps_b, 2), "kernel_only_frames_per_s": round(fps, 2),
                            "scored_candidates_per_s": round(cands * fps_b), "roofline": me_roofline(cnt["me_centres_level0"] / n, per["me_level0"])},
                  "batched": batched, "learned": learned, "learned_batched_cel_4k": cel, "gop_order": gop_form,
                  "transforms": transforms, "graph": graph_form, "intra_groups": intra_groups},
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
