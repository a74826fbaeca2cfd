"""The learned pruner on the GPU: the network's forward pass, gradients and AdamW step against the
CPU reference (tests/cnn_ref.py, itself checked by finite differences), training, and its use inside
the engine (limits from the network on the engine's own features; decisions equal to the explicit
per-superblock limits)."""
import numpy as np
import pytest

from celv2_amd import synth
from celv2_amd.records import compare_maps
import cnn_ref as cr

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / max(1e-30, np.linalg.norm(b)))


@pytest.fixture(scope="module")
def batch():
    rng = np.random.default_rng(11)
    x = rng.uniform(0, 1, (3, 5, 7, cr.F)).astype(np.float32)
    kstar = rng.integers(1, 8, 3 * 5 * 7).astype(np.uint8)
    return x, kstar


def test_forward_matches_cpu_reference(gpu, batch):
    x, _ = batch
    p = gpu.Pruner(max_positions=1024, seed=5)
    got = p.forward(x).reshape(-1, 6)
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
    want_k, unsure = cr.limits(p.forward(feats[None]).reshape(-1, 6), 7, 0.5, 0.02)
    assert np.array_equal(kmap, want_k.astype(np.uint8)), (kmap.tolist(), want_k.tolist())
    e.set_pruner(None)
    e.set_ref_limits(kmap)
    e.launch(src, rl, lam, keys=refs, src_key=f)
    assert compare_maps(e.collect(), pruned, e.sb_cols) is None
    assert np.array_equal(e.fetch_sb_costs(), pruned_cost)
    e.close()
    p.close()
