"""The decision log (celv2_amd/dlog.py): records written and read back equal, and far smaller than the
engine's full records. CPU only (records from the oracle's CPU engine)."""
import numpy as np

from celv2_amd import dlog, synth
from celv2_amd.records import SB_DTYPE
import oracle_lib as ol


def test_log_roundtrip(tmp_path):
    w, h = 192, 136
    _, y8 = synth.clip_frames(synth.params(0, 9, w, h), 3)
    maps = [ol.decide_frame(y8[0], [], 700, w, h), ol.decide_frame(y8[2], [y8[1], y8[0]], 900, w, h)]
    n = maps[0].shape[0]
    rng = np.random.default_rng(1)
    costs = [rng.integers(0, 1 << 40, n), rng.integers(0, 1 << 40, n)]
    limits = [np.zeros(n, np.uint8), np.array([1, 2] * (n // 2) + [2] * (n % 2), np.uint8)]
    p = tmp_path / "d.c2dl"
    with dlog.Writer(p, w, h) as log:
        log.frame(0, "K", 130, 700, [], maps[0], costs[0], limits[0])
        log.frame(2, "B", 146, 900, [1, 0], maps[1], costs[1], limits[1])
    head, frames = dlog.read(p)
    assert head == dict(width=w, height=h, sb_size=64, sb_cols=3, sb_rows=3)
    assert [(f["display"], f["type"], f["qindex"], f["lambda_q8"], f["refs"]) for f in frames] == [
        (0, "K", 130, 700, []), (2, "B", 146, 900, [1, 0])]
    for f, m, c, l in zip(frames, maps, costs, limits):
        for i, (tree, leaves, cost, limit) in enumerate(f["sbs"]):
            assert tree == m["tree"][i][:m["ntree"][i]].tobytes()
            assert leaves.tobytes() == m["blocks"][i][:m["nblocks"][i]].tobytes()
            assert cost == c[i] and limit == l[i]
    assert p.stat().st_size * 30 < 2 * n * SB_DTYPE.itemsize  # the full records are over 30x larger


def test_log_rejects_a_wrong_map(tmp_path):
    import pytest

    with dlog.Writer(tmp_path / "x.c2dl", 64, 64) as log:
        with pytest.raises(ValueError):
            log.frame(0, "K", 130, 700, [], np.zeros(2, SB_DTYPE))
