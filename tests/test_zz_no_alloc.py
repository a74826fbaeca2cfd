"""celv2hip_pruner_forward_static (the forward pass that allocates nothing, its scratch a static device
array) against celv2hip_pruner_forward on the same network and features: two calls in a row, the
second on fewer positions, bit for bit; and a call beyond its maximum is refused."""
import numpy as np
import pytest

from celv2_amd import pruner as trained

pytestmark = pytest.mark.gpu


def test_forward_static_matches_forward(gpu):
    weights, _ = trained.load()
    p = gpu.Pruner(max_positions=4096, seed=3)
    p.set_params(weights)
    rng = np.random.default_rng(17)
    try:
        for shape in [(2, 34, 60, 14), (1, 5, 7, 14)]:
            print("forward_static", shape, flush=True)
            feats = rng.standard_normal(shape).astype(np.float32)
            want = p.forward(feats)
            got = p.forward_static(feats)
            assert np.array_equal(got, want), shape
        with pytest.raises(gpu.EngineError):
            p.forward_static(np.zeros((1, 129, 129, 14), np.float32))  # 16,641 positions
        feats = rng.standard_normal((1, 3, 4, 14)).astype(np.float32)
        assert np.array_equal(p.forward_static(feats), p.forward(feats))
    finally:
        p.close()
