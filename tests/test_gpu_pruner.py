"""The learned pruner on the GPU: the network's forward pass, gradients and AdamW step against the
CPU reference (tests/cnn_ref.py, itself checked by finite differences), training, and its use inside
the engine (limits from the network on the engine's own features; decisions equal to the explicit
per-superblock limits)."""
import numpy as np
import pytest

from celv2_amd import synth
from celv2_amd.records import compare_maps
import cnn_ref as cr
import oracle_lib as ol

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / max(1e-30, np.linalg.norm(b)))


@pytest.fixture(scope="module")
def batch():
    rng = np.random.default_rng(11)
    x = rng.uniform(0, 1, (3, 5, 7, cr.F)).astype(np.float32)
    kstar = (rng.integers(1, 8, 3 * 5 * 7) | rng.integers(0, 16, 3 * 5 * 7) << 3).astype(np.uint8)  # refs + group + search bits
    return x, kstar


def test_forward_matches_cpu_reference(gpu, batch):
    x, _ = batch
    p = gpu.Pruner(max_positions=1024, seed=5)
    got = p.forward(x).reshape(-1, cr.OUT)
    want, _ = cr.forward(p.params(), x)
    assert rel(got, want) < 2e-3, rel(got, want)
    p.close()


def test_forward_on_a_1080p_grid(gpu):
    """The one-launch inference over several 4 x 16 tiles of a 17 x 30 grid (two frames), ragged at
    the bottom and right."""
    rng = np.random.default_rng(12)
    x = rng.uniform(0, 1, (2, 17, 30, cr.F)).astype(np.float32)
    p = gpu.Pruner(max_positions=2 * 17 * 30, seed=13)
    got = p.forward(x).reshape(-1, cr.OUT)
    want, _ = cr.forward(p.params(), x)
    assert rel(got, want) < 2e-3, rel(got, want)
    p.close()


def test_gradients_match_cpu_reference(gpu, batch):
    x, kstar = batch
    p = gpu.Pruner(max_positions=1024, seed=6)
    g, loss = p.gradients(x, kstar)
    want_loss, want = cr.loss_and_grads(p.params(), x, kstar)
    assert abs(loss - want_loss) < 1e-4 * max(1.0, abs(want_loss)), (loss, want_loss)
    for name, lo, hi in (("W1", cr.W1, cr.B1), ("b1", cr.B1, cr.W2), ("W2", cr.W2, cr.B2), ("b2", cr.B2, cr.W3),
                         ("W3", cr.W3, cr.B3), ("b3", cr.B3, cr.NPARAMS)):
        assert rel(g[lo:hi], want[lo:hi]) < 2e-2, (name, rel(g[lo:hi], want[lo:hi]))
    p.close()


def test_adamw_step_matches_cpu(gpu, batch):
    x, kstar = batch
    p = gpu.Pruner(max_positions=1024, seed=7)
    w0 = p.params()
    g, _ = p.gradients(x, kstar)
    p.train_step(x, kstar, 1e-3, 1e-2)
    want, _, _ = cr.adamw(w0.astype(np.float64), g.astype(np.float64), 0.0, 0.0, 1, 1e-3, 1e-2)
    assert np.abs(p.params() - want).max() < 1e-6 + 1e-5 * np.abs(want).max()
    p.close()


def test_training_reduces_the_loss(gpu, batch):
    x, kstar = batch
    p = gpu.Pruner(max_positions=1024, seed=8)
    losses = [p.train_step(x, kstar, 3e-3, 0.0) for _ in range(150)]
    assert losses[-1] < 0.6 * losses[0], (losses[0], losses[-1])
    p.close()


def test_engine_uses_the_network_limits(gpu):
    w, h = 352, 288
    y8 = synth.clip_frames(synth.params(1, 31, w, h), 9)[1]
    lam = gpu.lambda_q8(130)
    f, refs = 8, list(range(7, 0, -1))
    src, rl = gpu.RawEngine.luma(y8[f]), [gpu.RawEngine.luma(y8[r]) for r in refs]
    p = gpu.Pruner(max_positions=1024, seed=9)
    e = gpu.RawEngine(w, h)
    e.set_pruner(p, 0.5, 0.02)
    e.launch(src, rl, lam, keys=refs, src_key=f)
    pruned, pruned_cost = e.collect(), e.fetch_sb_costs()
    kmap = e.fetch_ref_limits()
    feats = e.fetch_features()
    want_k, unsure = cr.limits(p.forward(feats[None]).reshape(-1, cr.OUT), 7, 0.5, 0.02)
    assert np.array_equal(kmap, want_k.astype(np.uint8)), (kmap.tolist(), want_k.tolist())
    e.set_pruner(None)
    e.set_ref_limits(kmap)
    e.launch(src, rl, lam, keys=refs, src_key=f)
    assert compare_maps(e.collect(), pruned, e.sb_cols) is None
    assert np.array_equal(e.fetch_sb_costs(), pruned_cost)
    e.close()
    p.close()


@pytest.mark.parametrize("prepared", [False, True])
def test_engine_uses_the_network_groups(gpu, prepared):
    """With the group and search outputs on (celv2hip_set_pruner_groups, celv2hip_set_pruner_search),
    a launch decides exactly as one with the network's reference limits, intra mode groups and
    motion-search modes given explicitly (celv2hip_set_ref_limits, celv2hip_set_intra_masks,
    celv2hip_set_search_modes), the groups and modes being the reference rules of cnn_ref.groups and
    cnn_ref.search_modes on the same logits; with the source prepared (its intra scores left to the
    launch) and without."""
    w, h = 352, 288
    y8 = synth.clip_frames(synth.params(1, 33, w, h), 9)[1]
    lam = gpu.lambda_q8(140)
    f, refs = 8, list(range(7, 0, -1))
    src, rl = gpu.RawEngine.luma(y8[f]), [gpu.RawEngine.luma(y8[r]) for r in refs]
    p = gpu.Pruner(max_positions=1024, seed=4)
    e = gpu.RawEngine(w, h)
    try:
        for gth, sth in ((0.5, 0.5), (0.45, 0.0)):
            e.set_pruner(p, 0.5, 0.02, gthreshold=gth, sthreshold=sth)
            if prepared:
                e.prepare(src, f)
            e.launch(src, rl, lam, keys=refs, src_key=f if prepared else -1)
            pruned, pruned_cost = e.collect(), e.fetch_sb_costs()
            kmap, modes = e.fetch_ref_limits(), e.fetch_search_modes()
            logits = p.forward(e.fetch_features()[None]).reshape(-1, cr.OUT)
            want_k, unsure = cr.limits(logits, 7, 0.5, 0.02)
            masks = cr.groups(logits, gth, unsure).astype(np.uint8)
            want_modes = cr.search_modes(logits, sth, unsure) if sth > 0 else np.zeros(kmap.size)
            assert np.array_equal(kmap, want_k.astype(np.uint8))
            assert np.array_equal(modes, want_modes.astype(np.uint8))
            if sth > 0:
                assert 0 < np.count_nonzero(modes) < modes.size  # both kinds of superblock present
            e.set_pruner(None)
            e.set_ref_limits(kmap)
            e.set_intra_masks(masks)
            e.set_search_modes(modes)
            e.launch(src, rl, lam, keys=refs)
            assert compare_maps(e.collect(), pruned, e.sb_cols) is None, (gth, sth, prepared)
            assert np.array_equal(e.fetch_sb_costs(), pruned_cost)
            e.set_ref_limits(None)
            e.set_intra_masks(None)
            e.set_search_modes(None)
    finally:
        e.close()
        p.close()


def cpu_features(src2, costs, sb_cols, sb_rows, lam):
    """The features from the level-2 source, each reference's level-2 block costs [r][rows2, cols2]
    and the frame's lambda."""
    n = len(costs)
    rows2, cols2 = costs[0].shape
    h2, w2 = src2.shape
    out = np.zeros((sb_rows, sb_cols, cr.F))
    for sy in range(sb_rows):
        for sx in range(sb_cols):
            f = out[sy, sx]
            blocks = [(min(2 * sy + (q >> 1), rows2 - 1), min(2 * sx + (q & 1), cols2 - 1)) for q in range(4)]
            env = np.array([[c[b] for b in blocks] for c in costs], np.float64)
            e = np.minimum.accumulate(env, axis=0).sum(1)
            l = np.log2(1 + e)
            f[0] = l[-1] / 16
            for k in range(1, n):
                f[k] = min(4.0, 8 * (l[k - 1] - l[k]))
            ys = np.minimum(16 * sy + np.arange(16), h2 - 1)
            xs = np.minimum(16 * sx + np.arange(16), w2 - 1)
            v = src2[np.ix_(ys, xs)].astype(np.float64)
            right = src2[np.ix_(ys, np.minimum(xs + 1, w2 - 1))].astype(np.float64)
            down = src2[np.ix_(np.minimum(ys + 1, h2 - 1), xs)].astype(np.float64)
            mean = v.mean()
            f[7] = mean / 255
            f[8] = np.log2(1 + max(0.0, (v * v).mean() - mean * mean)) / 16
            f[9] = np.log2(1 + 16 * (np.abs(right - v) + np.abs(down - v)).sum() / 512) / 16
            f[10] = np.log2(1 + lam) / 32
    out[..., 11] = out[..., 0].mean()
    out[..., 12] = out[..., 8].mean()
    out[..., 13] = out[..., 1:7].sum(-1).mean()
    return out


@pytest.mark.parametrize("w,h,nrefs", [(352, 288, 7), (202, 122, 3)])
def test_features_match_the_coarse_search(gpu, w, h, nrefs):
    """Features against the oracle's level-2 search costs (on the engine's own level-2 planes and
    level-3 fields, both checked against the oracle elsewhere) and the level-2 source."""
    O = ol.oracle()
    y8 = synth.clip_frames(synth.params(1, 41, w, h), nrefs + 1)[1]
    refs = list(range(nrefs - 1, -1, -1))
    e = gpu.RawEngine(w, h)
    e.set_features(True)
    lam = gpu.lambda_q8(130)
    e.launch(gpu.RawEngine.luma(y8[nrefs]), [gpu.RawEngine.luma(y8[r]) for r in refs], lam, keys=refs)
    e.collect()
    got = e.fetch_features()
    src2 = np.ascontiguousarray(e.fetch_plane(-1, 2))
    h2, w2 = src2.shape
    costs = []
    for r in range(nrefs):
        ref2 = np.ascontiguousarray(e.fetch_plane(r, 2))
        pmx, pmy = (np.ascontiguousarray(a, np.int32) for a in e.fetch_field(r, 3))
        mx2, _ = e.fetch_field(r, 2)
        rows2, cols2 = mx2.shape
        c = np.zeros((rows2, cols2), np.int64)
        mv = np.zeros(2, np.int32)
        for by in range(rows2):
            for bx in range(cols2):
                c[by, bx] = O.orc_search_cell(src2.ctypes.data, ref2.ctypes.data, None, w2, h2, bx, by, pmx.ctypes.data, pmy.ctypes.data,
                                              pmx.shape[1], pmx.shape[0], 0, mv.ctypes.data)
        costs.append(c)
    want = cpu_features(src2, costs, e.sb_cols, e.sb_rows, lam)
    e.close()
    assert np.all(got[..., nrefs:7] == 0)
    err = np.abs(got - want).max(axis=(0, 1))
    assert err.max() < 1e-3, err.tolist()


def test_trained_network_beats_nearest_references(gpu):
    """What the committed network gains over the rule without it. On frames of a clip that neither
    training nor validation used (natural content, seed 77, qindex 130), the engine with the trained
    network and its operating point keeps m references per superblock on average at cost regret R
    against all seven. The nearest-k rule reaches at best the regret of mixing nearest floor(m) and
    ceil(m) over the superblocks in the proportions that average m references, the linear interpolation
    of its regrets at those k; the network's R is lower, and within the 0.5% budget."""
    from celv2_amd import pruner
    w, h, q = 1920, 1080, 130
    y16 = synth.clip_frames(synth.params(1, 77, w, h), 13)[0]
    lam = gpu.lambda_q8(q)
    dev = [gpu.DeviceBuffer(f) for f in y16]
    lum = [gpu.RawEngine.luma(dev16=b.ptr, bit_depth=10) for b in dev]
    weights, cfg = pruner.load()
    p = gpu.Pruner(max_positions=(w // 64 + 1) * (h // 64 + 1), seed=1)
    p.set_params(weights)
    e = gpu.RawEngine(w, h)
    e.set_resident(True)
    jk = np.zeros(8)
    jl, ks = 0.0, []
    for f in range(7, 13):
        refs = [f - k for k in range(1, 8)]
        for k in range(1, 8):
            e.launch(lum[f], [lum[r] for r in refs[:k]], lam, keys=refs[:k])
            e.sync()
            jk[k] += e.fetch_sb_costs().astype(np.float64).sum()
        e.set_pruner(p, *pruner.operating_point(cfg, q))
        e.launch(lum[f], [lum[r] for r in refs], lam, keys=refs)
        e.sync()
        jl += e.fetch_sb_costs().astype(np.float64).sum()
        ks.append(e.fetch_ref_limits())
        e.set_pruner(None)
    e.close()
    p.close()
    for b in dev:
        b.free()
    m = float(np.concatenate(ks).mean())
    regret_l = jl / jk[7] - 1
    lo = int(np.floor(m))
    frac = m - lo
    regret_n = (1 - frac) * (jk[lo] / jk[7] - 1) + frac * (jk[min(lo + 1, 7)] / jk[7] - 1)
    assert regret_l <= 0.005, (regret_l, m)
    assert regret_l < regret_n, (regret_l, regret_n, m)


def test_intra_mode_groups_cut_the_intra_work(gpu):
    """The intra dimension's gain: with only DC/V/H scored in every superblock the 1080p source's intra
    scores take at least 15 % less time than with every group (about 30 % measured; the directional
    groups' scores are the costly ones), and each group left out saves time."""
    w, h, nr = 1920, 1080, 2
    y16, _ = synth.clip_frames(synth.params(1, 3, w, h), 8)
    dev = [gpu.DeviceBuffer(f) for f in y16]
    lum = [gpu.RawEngine.luma(dev16=b.ptr, bit_depth=10) for b in dev]
    e = gpu.RawEngine(w, h)
    e.set_resident(True)
    lam = gpu.lambda_q8(130)
    nsb = e.sb_cols * e.sb_rows
    try:
        def intra_ms(mask):
            e.set_intra_masks(np.full(nsb, mask, np.uint8))
            e.set_timing(True)
            e.stats_reset()
            for s in range(12):
                i = 2 + s % 6
                e.prepare(lum[i], i)
                e.launch(lum[i], [lum[i - 1], lum[i - 2]], lam, keys=[i - 1, i - 2], src_key=i)
            ms, cnt = e.stats()
            e.set_timing(False)
            return ms["intra_prepare"] / max(1, cnt["launches"])

        intra_ms(15)  # warm-up
        full, none_ = intra_ms(15), intra_ms(0)
        no_dir = intra_ms(3)
        assert none_ < 0.85 * full, (full, none_)
        assert no_dir < full, (full, no_dir)
    finally:
        e.close()
        for b in dev:
            b.free()
