"""A y4m file decided end to end on the GPU (celv2_amd/decide.py): read by the y4m reader, scheduled
in the frame loop's coding order, decided by the HIP engine, logged; every logged frame's partition
trees and leaves equal the oracle's CPU engine on the same 8-bit planes, references and lambda."""
import numpy as np
import pytest

from celv2_amd import decide, dlog, engine, synth, y4m
import oracle_lib as ol

pytestmark = pytest.mark.gpu


def test_y4m_decided_in_coding_order(gpu, tmp_path):
    w, h, n = 200, 130, 12  # odd padding on both axes (pw 200, ph 136)
    p = synth.params(1, 23, w, h)
    clip = tmp_path / "c.y4m"
    with open(clip, "wb") as f:
        f.write(f"YUV4MPEG2 W{w} H{h} F30:1 Ip C420p10\n".encode())
        for t in range(n):
            fr = synth.frame(p, t, 0, True)[0]  # padded 10-bit luma
            # 4:2:0 payload: the visible luma, then flat chroma planes.
            cw, ch = (w + 1) // 2, (h + 1) // 2
            f.write(b"FRAME\n" + np.ascontiguousarray(fr[:h, :w]).astype("<u2").tobytes()
                    + np.full(2 * cw * ch, 512, "<u2").tobytes())
    log = tmp_path / "c.c2dl"
    entries = decide.decide_y4m(clip, log, refs=7, qindex=130)
    assert [e.type for e in entries[:2]] == ["K", "A"] and len(entries) == n
    frames8 = [f.y.copy() for f in y4m.Y4mReader(clip)]
    head, logged = dlog.read(log)
    assert [f["display"] for f in logged] == [e.display for e in entries]
    cpu = ol.CpuEngine(w, h, threads=8)
    for f in logged:
        print(f"frame {f['display']} {f['type']} q{f['qindex']} refs {f['refs']}", flush=True)
        cpu.launch(frames8[f["display"]], [frames8[r] for r in f["refs"]], engine.lambda_q8(f["qindex"]))
        want, _ = cpu.collect()
        for i, (tree, leaves, cost, limit) in enumerate(f["sbs"]):
            assert tree == want["tree"][i][:want["ntree"][i]].tobytes(), (f["display"], i)
            assert leaves.tobytes() == want["blocks"][i][:want["nblocks"][i]].tobytes(), (f["display"], i)
            assert limit == len(f["refs"])
