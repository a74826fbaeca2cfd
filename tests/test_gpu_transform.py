"""The HIP forward transforms (celv2hip_fwd_tx) against the oracle's (oracle/transform.cpp), bit for
bit: every size from 4x4 to 64x64 on the reference's own residual kinds (tests/test_transform.cpp:
81-100: constant 255, random +-255, random in [-255, 255]) and 10-bit residuals, strides wider than
the block, every kind pair at every size it has, mixed and shuffled in one call, with and without
the dropped energy; and the jobs the library refuses."""
import numpy as np
import pytest

import oracle_lib as ol

pytestmark = pytest.mark.gpu

SIDES = (4, 8, 16, 32, 64)


@pytest.fixture(scope="module")
def tx(gpu):
    from celv2_amd import transform

    return transform


def lcg(seed=12345):
    while True:
        seed = (seed * 1664525 + 1013904223) & 0xFFFFFFFF
        yield seed >> 8


def reference_trials(w, h, gen):
    out = [np.full((h, w), 255, np.int16)]
    out.append(np.array([255 if next(gen) & 1 else -255 for _ in range(w * h)], np.int16).reshape(h, w))
    for _ in range(2):
        out.append(np.array([next(gen) % 511 - 255 for _ in range(w * h)], np.int16).reshape(h, w))
    return out


def check(batch, blocks, kinds, scales):
    got, dropped = batch.results()
    for i, b in enumerate(blocks):
        want, wd = ol.fwd_tx(b, kinds[i][0], kinds[i][1], scales[i])
        assert np.array_equal(got[i], want), (b.shape, kinds[i], scales[i], np.argwhere(got[i] != want)[:4])
        assert dropped[i] == wd, (b.shape, kinds[i], dropped[i], wd)


def test_dct_every_size_matches_the_oracle(tx):
    gen = lcg()
    rng = np.random.default_rng(3)
    blocks, scales = [], []
    for idx, (w, h) in enumerate((w, h) for w in SIDES for h in SIDES):
        for b in reference_trials(w, h, gen):
            blocks.append(b)
            scales.append(0.5 + idx * 0.07)
        blocks.append(rng.integers(-1023, 1024, (h, w)).astype(np.int16))  # 10-bit residuals
        scales.append(1.0 + idx * 0.11)
        blocks.append(np.full((h, w), -1023, np.int16))
        scales.append(2.5)
    kinds = [(tx.DCT, tx.DCT)] * len(blocks)
    batch = tx.TxBatch(blocks, kinds, scales, pad=16)
    try:
        batch.run()
        check(batch, blocks, kinds, scales)
    finally:
        batch.free()


def test_every_kind_pair_mixed_and_shuffled(tx):
    rng = np.random.default_rng(11)
    items = []
    for w in SIDES:
        for h in SIDES:
            for vk in (tx.DCT, tx.DST, tx.FLIPDST, tx.IDT):
                for hk in (tx.DCT, tx.DST, tx.FLIPDST, tx.IDT):
                    if h > tx.MAX_POINTS[vk] or w > tx.MAX_POINTS[hk]:
                        continue
                    items.append((rng.integers(-700, 701, (h, w)).astype(np.int16), (vk, hk), float(rng.uniform(0.3, 3.0))))
    order = rng.permutation(len(items))
    blocks = [items[i][0] for i in order]
    kinds = [items[i][1] for i in order]
    scales = [items[i][2] for i in order]
    batch = tx.TxBatch(blocks, kinds, scales, pad=4)
    try:
        batch.run()
        check(batch, blocks, kinds, scales)
        batch.run()  # the same jobs again: the same values
        check(batch, blocks, kinds, scales)
    finally:
        batch.free()


def test_without_dropped_pointers(tx):
    rng = np.random.default_rng(5)
    blocks = [rng.integers(-255, 256, (h, w)).astype(np.int16) for w, h in ((64, 64), (64, 16), (8, 8), (4, 64))]
    batch = tx.TxBatch(blocks, want_dropped=False)
    try:
        batch.run()
        got, dropped = batch.results()
        assert np.all(dropped == -1)  # untouched
        for g, b in zip(got, blocks):
            assert np.array_equal(g, ol.fwd_tx(b)[0])
    finally:
        batch.free()


def test_refused_jobs_run_nothing(tx):
    blocks = [np.ones((16, 16), np.int16), np.ones((32, 32), np.int16), np.ones((64, 64), np.int16)]
    batch = tx.TxBatch(blocks)
    try:
        bad = []
        j = batch.jobs.copy(); j[1]["vkind"] = tx.DST; bad.append(j)        # DST-VII at 32 points
        j = batch.jobs.copy(); j[2]["hkind"] = tx.IDT; bad.append(j)        # identity at 64 points
        j = batch.jobs.copy(); j[0]["stride"] = 15; bad.append(j)           # stride below the width
        j = batch.jobs.copy(); j[0]["scale"] = 0.0; bad.append(j)           # no scale
        j = batch.jobs.copy(); j[2]["res"] = 0; bad.append(j)               # no residual
        j = batch.jobs.copy(); j[1]["w_log2"] = 7; bad.append(j)            # 128 points
        for jobs in bad:
            with pytest.raises(ValueError):
                tx.fwd_tx(jobs)
        got, dropped = batch.results()
        assert all(np.all(g == 0) for g in got) and np.all(dropped == -1)  # nothing ran
        batch.run()
        check(batch, blocks, [(tx.DCT, tx.DCT)] * 3, [1.0] * 3)
    finally:
        batch.free()
