"""Trains the learned pruner on decision logs of the plain engine.

Data: seeded synthetic 1920x1080 10-bit clips of 32 frames. Most are natural content (workload 1)
with the grain sigma, object speed and pan drawn per clip; some are flat-shaded cel content
(workload 4 at 1080p). Each clip is decided at one qindex from {100, 130, 160, 190}. For every
decidable frame (frame i >= 7 against the 7 source frames before it, nearest first) the engine
decides the frame with its nearest k = 1..7 references and records each superblock's decision cost
J_k (celv2hip_fetch_sb_costs) and, at k = 7, the pruner's features (celv2hip_fetch_features), the
cost with each optional intra mode group left out, and the cost with every motion search narrow at
levels 1 and 0. Each superblock's decision depends only on its own limit, so the total cost of any
per-superblock allocation k(s) follows from these logs exactly (the group and search costs are each
taken alone at 7 references: their effects are summed).

The budget is per sequence (a BD-rate is measured on one sequence): a clip's cost regret is
sum(J_k(s) - J_7) / sum(J_7) over its superblocks. Labels: per clip, the allocation that keeps the
fewest references within the target regret: with a multiplier mu, k*(s) = argmin_k k + mu (J_k(s) -
J_7(s)), mu set by bisection so the clip's label regret equals the target; an intra group, or the
full search, is labelled needed where leaving it out raises the superblock's cost. The network (BF16
on the matrix cores, celv2hip_pruner_*) learns P(k* > j + 1), P(group needed) and P(full search
needed) with AdamW on the GPU.

Operating points: on held-out clips, per qindex, first the highest group threshold and the highest
search threshold whose regret on every held-out clip stays within their shares of the target, then
the (threshold, fallback band) with the fewest references per superblock whose regret stays within
the rest, measured again through the engine itself (the network run inside the launch). Writes
weights.f32 (the FP32 parameters) and pruner.json.

    python -m celv2_amd.pruner.train [--clips 120] [--val-clips 16] [--steps 6000] [--out DIR]
"""
import argparse
import json
import os
import sys
import time
from concurrent.futures import ThreadPoolExecutor

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from celv2_amd import engine, synth  # noqa: E402

W, H, NR, FRAMES = 1920, 1080, 7, 32
QINDICES = (100, 130, 160, 190)


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def clip_spec(seed):
    """The clip of a seed: (synth params, qindex, description). Every fifth clip is cel content."""
    rng = np.random.default_rng(seed)
    q = QINDICES[seed % len(QINDICES)]
    if seed % 5 == 4:
        return synth.params(4, seed, W, H), q, "cel"
    sigma = int(rng.integers(0, 13))
    p = synth.params(1, seed, W, H, sigma_x16=16 * sigma, vmax8=int(rng.integers(24, 193)),
                     pan_x8=int(rng.integers(-32, 33)), pan_y8=int(rng.integers(-16, 17)))
    return p, q, f"natural sigma {sigma}"


def clip_logs(E, seed, y16):
    """(features [n, sb_rows, sb_cols, F], costs [n, sbs, 7], group costs [n, sbs, 3], narrow costs
    [n, sbs]) of a clip's decidable frames: J_k with the nearest k references, J_7 with each of the
    intra mode groups 1..3 left out in turn, and J_7 with every superblock's motion search narrow at
    levels 1 and 0. Reference keys are unique across clips (seed * 1000 + frame), so the engine's
    cached planes never mix clips."""
    lam = engine.lambda_q8(clip_spec(seed)[1])
    dev = [engine.DeviceBuffer(f) for f in y16]
    lum = [engine.RawEngine.luma(dev16=b.ptr, bit_depth=10) for b in dev]
    feats, costs, gcosts, ncosts = [], [], [], []
    nsb = E.sb_cols * E.sb_rows
    for f in range(NR, len(y16)):
        refs = [f - k for k in range(1, NR + 1)]
        keys = [seed * 1000 + r for r in refs]
        J = []
        for k in range(1, NR + 1):
            E.set_features(k == NR)
            E.launch(lum[f], [lum[r] for r in refs[:k]], lam, keys=keys[:k])
            E.sync()
            J.append(E.fetch_sb_costs())
        feats.append(E.fetch_features())
        costs.append(np.stack(J, axis=1))
        E.set_features(False)
        G = []
        for g in range(1, 4):
            E.set_intra_masks(np.full(nsb, 15 & ~(1 << g), np.uint8))
            E.launch(lum[f], [lum[r] for r in refs], lam, keys=keys)
            E.sync()
            G.append(E.fetch_sb_costs())
        E.set_intra_masks(None)
        gcosts.append(np.stack(G, axis=1))
        E.set_search_modes(np.full(nsb, 3, np.uint8))
        E.launch(lum[f], [lum[r] for r in refs], lam, keys=keys)
        E.sync()
        ncosts.append(E.fetch_sb_costs())
        E.set_search_modes(None)
    E.set_features(False)
    for b in dev:
        b.free()
    return np.stack(feats), np.stack(costs), np.stack(gcosts), np.stack(ncosts)


def gather(E, seeds, keep=None):
    """Decision logs of the clips of `seeds`, clips generated on host threads ahead of the GPU; keep
    (a dict) receives each clip's frames."""
    out = {}
    with ThreadPoolExecutor(max_workers=3) as ex:
        futs = {s: ex.submit(lambda s=s: synth.clip_frames(clip_spec(s)[0], FRAMES, threads=4)[0]) for s in seeds}
        for i, s in enumerate(seeds):
            y16 = futs.pop(s).result()
            out[s] = clip_logs(E, s, y16)
            if keep is not None:
                keep[s] = y16
            if i % 10 == 0:
                log(f"  clip {i + 1}/{len(seeds)} (seed {s}, {clip_spec(s)[2]}, qindex {clip_spec(s)[1]})")
    return out


def regret(costs, k):
    """Total cost regret (fraction) and mean references when superblock s keeps k[s]."""
    c = costs.reshape(-1, NR).astype(np.float64)
    kk = k.reshape(-1).astype(np.int64)
    chosen = c[np.arange(c.shape[0]), kk - 1]
    return float((chosen - c[:, NR - 1]).sum() / c[:, NR - 1].sum()), float(kk.mean())


def allocation(costs, mu):
    """argmin over k of k + mu (J_k - J_7) per superblock (the first minimum: fewer references)."""
    c = costs.astype(np.float64)
    obj = np.arange(1, NR + 1) + mu * (c - c[..., NR - 1:NR])
    return (np.argmin(obj, axis=-1) + 1).astype(np.uint8)


def optimal_labels(costs, target):
    """The allocation with the fewest references whose total regret is at most target, and its mu."""
    lo, hi = 0.0, 1.0
    while regret(costs, allocation(costs, hi))[0] > target:
        hi *= 4
    for _ in range(60):
        mid = 0.5 * (lo + hi)
        if regret(costs, allocation(costs, mid))[0] > target:
            lo = mid
        else:
            hi = mid
    return allocation(costs, hi), hi


def group_labels(costs, gcosts):
    """Label bits 3 + g - 1: intra mode group g (1..3) is needed where leaving it out raises the
    superblock's decision cost at 7 references."""
    j7 = costs[..., NR - 1]
    need = gcosts > j7[..., None]
    return (need.astype(np.uint8) << np.array([3, 4, 5], np.uint8)).sum(-1).astype(np.uint8)


def search_labels(costs, ncosts):
    """Label bit 6: the superblock needs the full motion search where the narrow search at levels 1
    and 0 raises its decision cost at 7 references."""
    return ((ncosts > costs[..., NR - 1]).astype(np.uint8) << 6).astype(np.uint8)


def search_regret(logits, costs, ncosts, sthreshold):
    """The cost regret of the narrow search where the search output falls below sthreshold (each
    superblock's cost with every search narrow, at 7 references), and the share of superblocks narrow."""
    s = 1.0 / (1.0 + np.exp(-logits[:, 9].astype(np.float64)))
    narrow = s < sthreshold
    c = costs.reshape(-1, NR)[:, NR - 1].astype(np.float64)
    n = ncosts.reshape(-1).astype(np.float64)
    return float(((n - c) * narrow).sum() / c.sum()), float(narrow.mean())


def search_point(logits, costs, ncosts, budget):
    """The highest search threshold (the most superblocks narrow) whose regret on every clip is within
    budget; logits, costs and ncosts are lists over clips."""
    best = None
    for sth in np.concatenate([[0.005, 0.01, 0.02, 0.03, 0.04], np.arange(0.05, 0.96, 0.05)]):
        rs = [search_regret(z, c, n, sth) for z, c, n in zip(logits, costs, ncosts)]
        if max(r for r, _ in rs) <= budget:
            best = {"sthreshold": round(float(sth), 3), "regret": max(r for r, _ in rs),
                    "narrow_share": float(np.mean([k for _, k in rs]))}
    return best


def group_regret(logits, costs, gcosts, gthreshold):
    """The cost regret of leaving out the groups whose output falls below gthreshold (each group's
    cost taken alone at 7 references: the groups' effects summed), and the groups kept per superblock."""
    s = 1.0 / (1.0 + np.exp(-logits[:, 6:9].astype(np.float64)))
    drop = s < gthreshold
    c = costs.reshape(-1, NR)[:, NR - 1].astype(np.float64)
    g = gcosts.reshape(-1, 3).astype(np.float64)
    return float(((g - c[:, None]) * drop).sum() / c.sum()), float((~drop).sum(1).mean())


def group_point(logits, costs, gcosts, budget):
    """The highest group threshold (the most groups left out) whose regret on every clip is within
    budget; logits, costs and gcosts are lists over clips."""
    best = None
    for gth in np.concatenate([[0.005, 0.01, 0.02, 0.03, 0.04], np.arange(0.05, 0.96, 0.05)]):
        rs = [group_regret(z, c, g, gth) for z, c, g in zip(logits, costs, gcosts)]
        if max(r for r, _ in rs) <= budget:
            best = {"gthreshold": round(float(gth), 3), "regret": max(r for r, _ in rs),
                    "groups_kept": float(np.mean([k for _, k in rs]))}
    return best


def limits(logits, threshold, band):
    """The engine's limit rule (celv2hip_set_pruner) on logits [positions, 6], all 7 references present."""
    s = 1.0 / (1.0 + np.exp(-logits.astype(np.float64)))
    k = np.ones(s.shape[0], np.int64)
    open_ = np.ones(s.shape[0], bool)
    for j in range(s.shape[1]):
        k[open_ & (s[:, j] >= threshold)] = j + 2
        open_ &= s[:, j] >= threshold
    unsure = (np.abs(s - threshold) < band).any(1)
    k[unsure] = NR
    return k, float(unsure.mean())


def operating_point(logits, costs, target):
    """The (threshold, band) with the fewest references whose regret on every clip is within target;
    logits and costs are lists over clips."""
    best = None
    for threshold in np.concatenate([[0.005, 0.01, 0.02, 0.03, 0.04], np.arange(0.05, 0.96, 0.025)]):
        for band in (0.0, 0.025, 0.05, 0.1):
            ks = [limits(z, threshold, band) for z in logits]
            rs = [regret(c, k)[0] for c, (k, _) in zip(costs, ks)]
            n = sum(k.size for k, _ in ks)
            mk = sum(float(k.sum()) for k, _ in ks) / n
            fb = sum(f * k.size for k, f in ks) / n
            if max(rs) <= target and (best is None or mk < best["mean_refs"]):
                best = {"threshold": round(float(threshold), 4), "band": band, "regret": max(rs), "mean_refs": mk, "fallback_rate": fb}
    return best or {"threshold": 1.0, "band": 0.0, "regret": 0.0, "mean_refs": float(NR), "fallback_rate": 0.0}


def engine_check(E, P, op, seeds, clips):
    """An operating point through the engine itself on clips (seed -> y16 frames): (the largest clip
    regret against the plain engine's J_7, mean limit, fallback rate)."""
    gth, sth = op.get("gthreshold", 0.0), op.get("sthreshold", 0.0)
    E.set_pruner(P, op["threshold"], op["band"], gthreshold=gth, sthreshold=sth)
    E.stats_reset()
    worst = 0.0
    ks = []
    for s in seeds:
        num = den = 0.0
        _, q, _ = clip_spec(s)
        lam = engine.lambda_q8(q)
        dev = [engine.DeviceBuffer(f) for f in clips[s]]
        lum = [engine.RawEngine.luma(dev16=b.ptr, bit_depth=10) for b in dev]
        for f in range(NR, len(dev)):
            refs = [f - k for k in range(1, NR + 1)]
            keys = [s * 1000 + r for r in refs]
            E.set_pruner(None)
            E.launch(lum[f], [lum[r] for r in refs], lam, keys=keys)
            E.sync()
            j7 = E.fetch_sb_costs().astype(np.float64)
            E.set_pruner(P, op["threshold"], op["band"], gthreshold=gth, sthreshold=sth)
            E.launch(lum[f], [lum[r] for r in refs], lam, keys=keys)
            E.sync()
            num += (E.fetch_sb_costs() - j7).sum()
            den += j7.sum()
            ks.append(E.fetch_ref_limits())
        for b in dev:
            b.free()
        worst = max(worst, float(num / den))
    fallbacks = E.stats()[1]["pruner_fallbacks"]
    E.set_pruner(None)
    ks = np.concatenate(ks)
    return worst, float(ks.mean()), fallbacks / ks.size


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--clips", type=int, default=120)
    ap.add_argument("--val-clips", type=int, default=16)
    ap.add_argument("--label-target", type=float, default=0.004, help="the total regret of the training labels")
    ap.add_argument("--target", type=float, default=0.004, help="largest total cost regret an operating point may have")
    ap.add_argument("--group-budget", type=float, default=0.001, help="the part of the target the intra groups may take")
    ap.add_argument("--search-budget", type=float, default=0.0005, help="the part of the target the narrow searches may take")
    ap.add_argument("--steps", type=int, default=6000)
    ap.add_argument("--batch", type=int, default=8)
    ap.add_argument("--lr", type=float, default=3e-3)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--out", default=os.path.dirname(os.path.abspath(__file__)))
    ap.add_argument("--logs", default="", help="also write the held-out clips' logs (features, costs) to this .npz")
    args = ap.parse_args()
    train_seeds = list(range(1000, 1000 + args.clips))
    val_seeds = list(range(5000, 5000 + args.val_clips))
    E = engine.RawEngine(W, H)
    E.set_resident(True)
    t0 = time.perf_counter()
    tl = gather(E, train_seeds)
    log(f"train logs: {len(train_seeds)} clips in {time.perf_counter() - t0:.0f}s")
    val_clips = {}
    vl = gather(E, val_seeds, val_clips)
    log(f"validation logs: {len(val_seeds)} clips in {time.perf_counter() - t0:.0f}s")

    # Labels per clip.
    labels = {s: optimal_labels(tl[s][1], args.label_target)[0] | group_labels(tl[s][1], tl[s][2]) | search_labels(tl[s][1], tl[s][3])
              for s in train_seeds}
    for q in QINDICES:
        ss = [s for s in train_seeds if clip_spec(s)[1] == q]
        lab = np.concatenate([labels[s] for s in ss]) if ss else np.zeros(0, np.uint8)
        k = lab & 7
        gneed = [float(((lab >> (3 + g)) & 1).mean()) if lab.size else 0.0 for g in range(3)]
        sneed = float(((lab >> 6) & 1).mean()) if lab.size else 0.0
        log(f"qindex {q}: {len(ss)} clips, label mean refs {k.mean():.3f}, k* histogram {np.bincount(k.ravel(), minlength=8)[1:].tolist()}, "
            f"groups needed {[round(v, 4) for v in gneed]}, full search needed {sneed:.4f}")
    xt = np.concatenate([tl[s][0] for s in train_seeds])
    kt = np.concatenate([labels[s] for s in train_seeds])
    rows, cols = xt.shape[1], xt.shape[2]
    P = engine.Pruner(max_positions=args.batch * rows * cols, seed=args.seed)
    rng = np.random.default_rng(args.seed)
    hist = []
    for step in range(args.steps):
        idx = rng.choice(xt.shape[0], args.batch, replace=False)
        lr = args.lr * (0.05 + 0.95 * 0.5 * (1 + np.cos(np.pi * step / args.steps)))
        hist.append(P.train_step(xt[idx], kt[idx].reshape(args.batch, -1), lr, 1e-4))
        if step % 500 == 0 or step == args.steps - 1:
            log(f"step {step}: loss {np.mean(hist[-100:]):.4f}")

    # Operating points per qindex on the held-out clips, then through the engine.
    ops, report = {}, {}
    for q in QINDICES:
        ss = [s for s in val_seeds if clip_spec(s)[1] == q]
        if not ss:
            continue
        logits = [np.concatenate([P.forward(vl[s][0][i:i + args.batch]) for i in range(0, vl[s][0].shape[0], args.batch)]).reshape(-1, engine.N_OUT)
                  for s in ss]
        costs = [vl[s][1] for s in ss]
        gcosts = [vl[s][2] for s in ss]
        ncosts = [vl[s][3] for s in ss]
        gp = group_point(logits, costs, gcosts, args.group_budget)
        sp = search_point(logits, costs, ncosts, args.search_budget)
        op = operating_point([z[:, :6] for z in logits], costs,
                             args.target - (gp["regret"] if gp else 0.0) - max(0.0, sp["regret"] if sp else 0.0))
        op["gthreshold"] = gp["gthreshold"] if gp else 0.0
        op["sthreshold"] = sp["sthreshold"] if sp else 0.0
        r, mk, fb = engine_check(E, P, op, ss, val_clips)
        best = [optimal_labels(c, args.target)[0] for c in costs]
        # The rule without the network: the nearest k references everywhere, k the fewest within
        # target on every clip.
        nearest = {k: max(regret(c, np.full(c.shape[:2], k))[0] for c in costs) for k in range(1, NR + 1)}
        k_plain = min(k for k in range(1, NR + 1) if nearest[k] <= args.target)
        # An operating point that keeps every reference everywhere is no pruning: the plain engine.
        ops[str(q)] = ({"threshold": op["threshold"], "band": op["band"], "gthreshold": op["gthreshold"], "sthreshold": op["sthreshold"]}
                       if mk < NR - 0.05 else None)
        report[str(q)] = {"val_clips": len(ss), "val_frames": int(sum(c.shape[0] for c in costs)),
                          "worst_clip_regret": r, "mean_refs": mk, "fallback_rate": fb,
                          "simulated_worst_clip_regret": op["regret"], "simulated_mean_refs": op["mean_refs"],
                          "best_allocation_mean_refs": float(np.mean(np.concatenate([b.ravel() for b in best]))),
                          "nearest_k_worst_clip_regret": {str(k): v for k, v in nearest.items()},
                          "nearest_k_within_target": k_plain,
                          "groups": gp or {"gthreshold": 0.0, "regret": 0.0, "groups_kept": 3.0},
                          "search": sp or {"sthreshold": 0.0, "regret": 0.0, "narrow_share": 0.0}}
        log(f"qindex {q}: {json.dumps(ops[str(q)])} worst clip regret {r:.5f}, mean limit {mk:.3f} "
            f"(best allocation {report[str(q)]['best_allocation_mean_refs']:.3f}, nearest-k within target: {k_plain})")
        if args.logs and q == 130:
            np.savez_compressed(args.logs, x=np.concatenate([vl[s][0] for s in ss]).astype(np.float16),
                                c=np.concatenate(costs).astype(np.int64))
    E.close()
    cfg = {"operating_points": ops, "validation": report, "target": args.target, "label_target": args.label_target,
           "budget": "per clip: sum(J_learned - J_7) / sum(J_7) over the clip's superblocks", "steps": args.steps, "batch": args.batch, "lr": args.lr,
           "train_frames": int(xt.shape[0]), "train_superblocks": int(xt.shape[0] * rows * cols),
           "first_loss": float(np.mean(hist[:100])), "final_loss": float(np.mean(hist[-100:])),
           "train_seeds": [train_seeds[0], train_seeds[-1]], "val_seeds": [val_seeds[0], val_seeds[-1]],
           "clips": "1920x1080 10-bit, 32 frames; every fifth seed cel (workload 4), the rest natural (workload 1) "
                    "with grain sigma 0-12, speed and pan drawn per seed; qindex = QINDICES[seed % 4]"}
    os.makedirs(args.out, exist_ok=True)
    P.params().astype("<f4").tofile(os.path.join(args.out, "weights.f32"))
    with open(os.path.join(args.out, "pruner.json"), "w") as f:
        json.dump(cfg, f, indent=1)
    log(json.dumps(cfg))
    P.close()


if __name__ == "__main__":
    main()
