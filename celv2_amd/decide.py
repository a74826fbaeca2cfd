"""Decisions for a y4m file on the GPU: the file's frames (celv2_amd/y4m.py) in the frame loop's coding
order (celv2_amd/gop.py), each coded frame decided by the HIP engine against its ranked references at
its own qindex, written to a decision log (celv2_amd/dlog.py).

References are the source frames, as the reference's engine benchmark uses them
(tools/engine_bench.cpp:48-51); an encoder would pass its reconstructions instead.
"""
from . import dlog, engine, gop, y4m


def decide_y4m(path, log_path, refs=7, qindex=130, key=None, long_gop=True, pruner=None, operating_point=None,
               device=0, max_frames=None):
    """Decides every frame of `path`; key: per-frame keyframe flags (None: frame 0 only). With `pruner`
    (an engine.Pruner) and `operating_point(qindex) -> (threshold, band, group threshold, search threshold)
    or None`, each frame runs the pruned form at its qindex's point. Returns the schedule's coded entries."""
    r = y4m.Y4mReader(path)
    frames = []
    for f in r:
        frames.append(f.y.copy())
        if max_frames and len(frames) >= max_frames:
            break
    r.close()
    n = len(frames)
    cfg = gop.defaults(refs, qindex, r.width, r.height, long_gop=long_gop)
    entries = gop.coded(gop.plan(cfg, n, key))
    E = engine.RawEngine(r.width, r.height, device)
    lumas = [E.luma(y) for y in frames]
    try:
        with dlog.Writer(log_path, r.width, r.height) as log:
            for j, e in enumerate(entries):
                op = operating_point(e.qindex) if (pruner is not None and operating_point and e.refs) else None
                E.set_pruner(pruner, *op) if op else E.set_pruner(None)
                lam = engine.lambda_q8(e.qindex)
                E.launch(lumas[e.display], [lumas[k] for k in e.refs], lam, keys=list(e.refs), src_key=e.display)
                if j + 1 < len(entries):
                    E.prepare(lumas[entries[j + 1].display], entries[j + 1].display)
                rec = E.collect()
                log.frame(e.display, e.type, e.qindex, lam, e.refs, rec, E.fetch_sb_costs(), E.fetch_ref_limits())
    finally:
        E.close()
    return entries
