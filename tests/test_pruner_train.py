"""Host logic of the pruner's training (celv2_amd/pruner/train.py): labels, regret and the limit rule."""
import numpy as np

import cnn_ref
from celv2_amd.pruner import train


def test_labels_are_the_cheapest_allocation_within_the_target():
    """Against brute force over every allocation of 4 superblocks: the labels keep the regret within
    the target, and with costs convex in k (each further reference gains less) they reach the
    fewest references any allocation has within it."""
    import itertools
    rng = np.random.default_rng(5)
    for trial in range(20):
        j7 = rng.integers(1000, 3000, 4).astype(np.float64)
        gains = np.sort(rng.uniform(0, 0.03, (4, 6)), axis=1)[:, ::-1]  # decreasing gains: convex
        extra = np.concatenate([np.cumsum(gains[:, ::-1], axis=1)[:, ::-1], np.zeros((4, 1))], axis=1)
        costs = (j7[:, None] * (1 + extra))[None]
        target = rng.uniform(0.002, 0.02)
        k, mu = train.optimal_labels(costs, target)
        r, mk = train.regret(costs, k)
        assert r <= target + 1e-12
        best = min(sum(a) for a in itertools.product(range(1, 8), repeat=4)
                   if train.regret(costs, np.array([a]))[0] <= target + 1e-12)
        assert k.sum() >= best
        # the Lagrangian optimum reaches the brute-force optimum up to one superblock's step
        assert k.sum() <= best + 1, (trial, k.tolist(), best)


def test_allocation_keeps_more_references_as_regret_weighs_more():
    rng = np.random.default_rng(6)
    j7 = rng.integers(1000, 5000, (3, 50, 1))
    costs = np.concatenate([j7 + np.sort(rng.integers(0, 400, (3, 50, 6)), axis=-1)[..., ::-1], j7], axis=-1)
    prev = None
    for mu in (0.0, 1e-3, 1e-2, 1e-1, 10.0):
        k = train.allocation(costs, mu).astype(int)
        assert prev is None or np.all(k >= prev)
        prev = k
    assert np.all(train.allocation(costs, 0.0) == 1)


def test_regret_of_all_seven_is_zero_and_nearest_one_is_total_excess():
    rng = np.random.default_rng(3)
    j7 = rng.integers(1000, 5000, (2, 9, 1))
    j = np.concatenate([j7 + rng.integers(0, 300, (2, 9, 6)), j7], axis=-1)
    r, mk = train.regret(j, np.full((2, 9), 7))
    assert r == 0.0 and mk == 7.0
    r1, _ = train.regret(j, np.ones((2, 9), np.int64))
    assert np.isclose(r1, (j[..., 0] - j[..., 6]).sum() / j[..., 6].sum())


def test_limit_rule_matches_the_engine_rule():
    rng = np.random.default_rng(4)
    z = rng.normal(0, 3, (500, 6)).astype(np.float32)
    for threshold, band in ((0.5, 0.0), (0.3, 0.1), (0.7, 0.05)):
        k, fb = train.limits(z, threshold, band)
        want, unsure = cnn_ref.limits(z, 7, threshold, band)
        assert np.array_equal(k, want) and np.isclose(fb, unsure.mean())


def test_search_labels_and_point():
    """Label bit 6 marks exactly the superblocks whose narrow search costs more; the search point is
    the highest threshold within the budget on every clip, and with logits that rank the harmless
    superblocks lowest it makes every harmless superblock narrow and no harmful one."""
    rng = np.random.default_rng(9)
    clips = []
    for _ in range(3):
        j7 = rng.integers(1000, 5000, (4, 20)).astype(np.int64)
        costs = np.repeat(j7[..., None], 7, axis=-1)
        harm = rng.random((4, 20)) < 0.3
        ncosts = j7 + np.where(harm, rng.integers(50, 200, (4, 20)), -rng.integers(0, 3, (4, 20)))
        lab = train.search_labels(costs, ncosts)
        assert np.array_equal(lab >> 6, harm.astype(np.uint8)) and not np.any(lab & 63)
        logits = np.zeros((80, 10))
        logits[:, 9] = np.where(harm.ravel(), 3.0, -3.0)  # sigma 0.95 where harmful, 0.05 elsewhere
        clips.append((logits, costs, ncosts))
    sp = train.search_point(*map(list, zip(*clips)), budget=0.0)
    assert sp is not None and 0.05 < sp["sthreshold"] <= 0.95 and sp["regret"] <= 0.0
    for logits, costs, ncosts in clips:
        r, share = train.search_regret(logits, costs, ncosts, sp["sthreshold"])
        harm = (ncosts > costs[..., 6]).ravel()
        assert np.isclose(share, 1 - harm.mean())
    # A budget that no threshold meets: none chosen, the searches stay full.
    bad = [(np.full((80, 10), -9.0), c, c[..., 6] + 100) for _, c, _ in clips]
    assert train.search_point(*map(list, zip(*bad)), budget=0.0) is None
