"""The CPU reference of the pruner's network is itself checked by central differences (float64, no
rounding), so the GPU's gradients are compared with verified formulas."""
import numpy as np

import cnn_ref as cr


def test_reference_gradients_match_finite_differences():
    rng = np.random.default_rng(3)
    p = rng.normal(0, 0.3, cr.NPARAMS)
    x = rng.normal(0, 1, (2, 3, 4, cr.F))
    kstar = rng.integers(1, 8, 2 * 3 * 4) | rng.integers(0, 16, 2 * 3 * 4) << 3  # refs, groups, search
    loss, g = cr.loss_and_grads(p, x, kstar, rounding=False)
    assert np.isfinite(loss)
    idx = np.concatenate([rng.integers(lo, hi, 6) for lo, hi in
                          ((cr.W1, cr.B1), (cr.B1, cr.W2), (cr.W2, cr.B2), (cr.B2, cr.W3), (cr.W3, cr.B3), (cr.B3, cr.NPARAMS))])
    eps = 1e-6
    for i in idx:
        d = np.zeros(cr.NPARAMS)
        d[i] = eps
        num = (cr.loss_and_grads(p + d, x, kstar, False)[0] - cr.loss_and_grads(p - d, x, kstar, False)[0]) / (2 * eps)
        assert abs(num - g[i]) <= 1e-6 + 1e-4 * abs(num), (i, num, g[i])


def test_bf16_rounding_ties_to_even():
    v = np.array([1.0, 1.00390625, 1.01171875, -2.5, 3.0e-3], np.float32)  # 1 + 2^-8 rounds down, 1 + 3*2^-8 up
    r = cr.bf16(v)
    assert r[0] == 1.0 and r[1] == 1.0 and r[2] == np.float32(1.015625) and r[3] == -2.5
    assert abs(r[4] - 3.0e-3) < 3.0e-3 * 2 ** -8


def test_limits_rule():
    logits = np.log(np.array([[0.9, 0.8, 0.2, 0.9, 0.1, 0.1],    # two leading "needs more": 3
                              [0.1, 0.9, 0.9, 0.9, 0.9, 0.9],    # first below: 1
                              [0.9] * 6,                         # all: 7
                              [0.52, 0.1, 0.1, 0.1, 0.1, 0.1]])  # within the band: fallback
                     / (1 - np.array([[0.9, 0.8, 0.2, 0.9, 0.1, 0.1], [0.1, 0.9, 0.9, 0.9, 0.9, 0.9], [0.9] * 6,
                                      [0.52, 0.1, 0.1, 0.1, 0.1, 0.1]])))
    k, unsure = cr.limits(logits, 7, 0.5, 0.05)
    assert k.tolist() == [3, 1, 7, 7] and unsure.tolist() == [False, False, False, True]
    assert cr.limits(logits, 4, 0.5, 0.05)[0].tolist() == [3, 1, 4, 4]
