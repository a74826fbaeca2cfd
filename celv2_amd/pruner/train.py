"""Trains the learned pruner on decision logs of the plain engine.

For every decidable frame of seeded synthetic clips (workload 1, 1920x1080 10-bit, 7 references =
the previous source frames), the engine decides the frame with its nearest k = 1..7 references and
records each superblock's decision cost J_k (celv2hip_fetch_sb_costs) and, at k = 7, the pruner's
features (celv2hip_fetch_features). Each superblock's decision depends only on its own limit, so the
total cost of any per-superblock allocation k(s) follows from these logs exactly.

Labels are the allocation that keeps the fewest references for a given total regret,
sum(J_k(s) - J_7) / sum(J_7): with a multiplier mu, k*(s) = argmin_k k + mu (J_k(s) - J_7(s)), mu
set (by bisection) so the training set's label regret equals the target. The network (BF16 on the
matrix cores, celv2hip_pruner_*) learns P(k* > j + 1) with AdamW on the GPU; on held-out clips the
operating point (threshold, fallback band) is the one with the fewest references per superblock whose
regret stays within the target, then measured again through the engine itself. Writes weights.f32
(the FP32 parameters) and pruner.json.

    python -m celv2_amd.pruner.train [--clips 40] [--val-clips 8] [--steps 2000] [--out DIR]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from celv2_amd import engine, synth  # noqa: E402

W, H, NR, QINDEX = 1920, 1080, 7, 130


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def clip_logs(E, seed, frames, lam):
    """(features [n, sb_rows, sb_cols, 10], costs [n, sbs, 7]) of a clip's decidable frames. Reference
    keys are unique across clips (seed * 1000 + frame), so the engine's cached planes never mix clips."""
    y16, _ = synth.clip_frames(synth.params(1, seed, W, H), frames)
    dev = [engine.DeviceBuffer(f) for f in y16]
    lum = [engine.RawEngine.luma(dev16=b.ptr, bit_depth=10) for b in dev]
    feats, costs = [], []
    for f in range(NR, frames):
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
    for b in dev:
        b.free()
    return np.stack(feats), np.stack(costs)


def engine_check(E, P, cfg, seeds, frames, lam, j7):
    """The operating point measured through the engine itself (the network run inside the launch):
    (total cost regret against j7 [n, sbs], mean reference limit, fallback rate)."""
    E.set_features(False)
    E.set_pruner(P, cfg["threshold"], cfg["band"])
    E.stats_reset()
    got, ks = [], []
    for seed in seeds:
        y16, _ = synth.clip_frames(synth.params(1, seed, W, H), frames)
        dev = [engine.DeviceBuffer(f) for f in y16]
        lum = [engine.RawEngine.luma(dev16=b.ptr, bit_depth=10) for b in dev]
        for f in range(NR, frames):
            refs = [f - k for k in range(1, NR + 1)]
            E.launch(lum[f], [lum[r] for r in refs], lam, keys=[seed * 1000 + r for r in refs])
            E.sync()
            got.append(E.fetch_sb_costs())
            ks.append(E.fetch_ref_limits())
        for b in dev:
            b.free()
    fallbacks = E.stats()[1]["pruner_fallbacks"]
    E.set_pruner(None)
    got, ks = np.stack(got).astype(np.float64), np.stack(ks)
    ref = j7.astype(np.float64)
    return float((got - ref).sum() / ref.sum()), float(ks.mean()), fallbacks / ks.size


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


def regret(costs, k):
    """Total cost regret (fraction) and mean references when superblock s keeps k[s]."""
    c = costs.reshape(-1, NR).astype(np.float64)
    kk = k.reshape(-1).astype(np.int64)
    chosen = c[np.arange(c.shape[0]), kk - 1]
    return float((chosen - c[:, NR - 1]).sum() / c[:, NR - 1].sum()), float(kk.mean())


def limits(logits, threshold, band):
    s = 1.0 / (1.0 + np.exp(-logits.astype(np.float64)))
    k = np.ones(s.shape[0], np.int64)
    open_ = np.ones(s.shape[0], bool)
    for j in range(s.shape[1]):
        k[open_ & (s[:, j] >= threshold)] = j + 2
        open_ &= s[:, j] >= threshold
    unsure = (np.abs(s - threshold) < band).any(1)
    k[unsure] = NR
    return k, float(unsure.mean())


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--clips", type=int, default=20)
    ap.add_argument("--val-clips", type=int, default=4)
    ap.add_argument("--frames", type=int, default=16)
    ap.add_argument("--label-target", type=float, default=0.004, help="the total regret of the training labels")
    ap.add_argument("--target", type=float, default=0.004, help="largest total cost regret the operating point may have")
    ap.add_argument("--steps", type=int, default=1500)
    ap.add_argument("--batch", type=int, default=8)
    ap.add_argument("--lr", type=float, default=3e-3)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--out", default=os.path.dirname(os.path.abspath(__file__)))
    ap.add_argument("--logs", default="", help="also write the decision logs (features, costs) to this .npz")
    args = ap.parse_args()
    lam = engine.lambda_q8(QINDEX)
    E = engine.RawEngine(W, H)
    E.set_resident(True)
    t0 = time.perf_counter()
    data = {}
    for name, seeds in (("train", range(1000, 1000 + args.clips)), ("val", range(2000, 2000 + args.val_clips))):
        fs, cs = zip(*[clip_logs(E, s, args.frames, lam) for s in seeds])
        data[name] = (np.concatenate(fs), np.concatenate(cs))
        log(f"{name}: {data[name][0].shape[0]} frames of decision logs ({time.perf_counter() - t0:.1f}s)")
    xt, ct = data["train"]
    xv, cv = data["val"]
    if args.logs:
        np.savez_compressed(args.logs, xt=xt, ct=ct, xv=xv, cv=cv)
    kt, mu = optimal_labels(ct, args.label_target)
    kv = allocation(cv, mu)
    log(f"labels: mu {mu:.3e}, train regret {regret(ct, kt)}, k* histogram {np.bincount(kt.ravel(), minlength=8)[1:].tolist()}")
    rows, cols = xt.shape[1], xt.shape[2]
    P = engine.Pruner(max_positions=args.batch * rows * cols, seed=args.seed)
    rng = np.random.default_rng(args.seed)
    hist = []
    for step in range(args.steps):
        idx = rng.choice(xt.shape[0], args.batch, replace=False)
        lr = args.lr * (0.1 + 0.9 * 0.5 * (1 + np.cos(np.pi * step / args.steps)))
        loss = P.train_step(xt[idx], kt[idx].reshape(args.batch, -1), lr, 1e-4)
        hist.append(loss)
        if step % 250 == 0 or step == args.steps - 1:
            log(f"step {step}: loss {np.mean(hist[-50:]):.4f}")
    logits = np.concatenate([P.forward(xv[i:i + args.batch]) for i in range(0, xv.shape[0], args.batch)]).reshape(-1, 6)
    best = None
    for threshold in np.arange(0.10, 0.91, 0.05):
        for band in (0.0, 0.05, 0.1, 0.15):
            k, fb = limits(logits, threshold, band)
            r, mk = regret(cv, k)
            if r <= args.target and (best is None or mk < best["val_mean_refs"]):
                best = {"threshold": round(float(threshold), 3), "band": band, "val_regret": r, "val_mean_refs": mk,
                        "val_fallback_rate": fb}
    nearest = {k: regret(cv, np.full(cv.shape[:2], k))[0] for k in range(1, NR + 1)}
    label_regret, label_mean = regret(cv, kv)
    cfg = dict(best or {"threshold": 1.0, "band": 0.0}, label_target=args.label_target, label_mu=mu, target=args.target, steps=args.steps,
               train_frames=int(xt.shape[0]), val_frames=int(xv.shape[0]), final_loss=float(np.mean(hist[-50:])),
               first_loss=float(np.mean(hist[:50])), val_label_regret=label_regret, val_label_mean_refs=label_mean,
               val_nearest_k_regret={str(k): v for k, v in nearest.items()}, qindex=QINDEX, workload="1 (1920x1080 10-bit)",
               train_seeds=[1000, 1000 + args.clips - 1], val_seeds=[2000, 2000 + args.val_clips - 1])
    if best is not None:
        r, mk, fb = engine_check(E, P, best, range(2000, 2000 + args.val_clips), args.frames, lam, cv[..., NR - 1])
        cfg.update(engine_val_regret=r, engine_val_mean_refs=mk, engine_val_fallback_rate=fb)
    E.close()
    os.makedirs(args.out, exist_ok=True)
    P.params().astype("<f4").tofile(os.path.join(args.out, "weights.f32"))
    with open(os.path.join(args.out, "pruner.json"), "w") as f:
        json.dump(cfg, f, indent=1)
    log(json.dumps(cfg))
    P.close()


if __name__ == "__main__":
    main()
