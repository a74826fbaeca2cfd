"""Each stage of the HIP engine against its CPU rule (CelV2 checks each Vulkan kernel the same way,
tests/test_vk_kernels.cpp: downsample_matches_ref, subpel_matches_ref, me_matches_ref)."""
import ctypes

import numpy as np
import pytest

from celv2_amd import synth
import oracle_lib as ol

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def launched(gpu):
    out = {}
    for (w, h) in [(352, 288), (202, 122), (64, 33), (200, 72)]:
        p = synth.params(0, 5, w, h)
        y8 = synth.clip_frames(p, 3)[1]
        e = gpu.RawEngine(w, h)
        lam = gpu.lambda_q8(120)
        e.launch(gpu.RawEngine.luma(y8[2]), [gpu.RawEngine.luma(y8[1]), gpu.RawEngine.luma(y8[0])], lam, keys=[1, 0])
        e.collect()
        out[(w, h)] = (e, y8)
    return out


@pytest.mark.parametrize("size", [(352, 288), (202, 122), (64, 33), (200, 72)])
def test_pyramids_match(launched, size):
    """Every level of the source's and a reference's pyramid, border included (each border sample the
    value at its clamped position), at sizes whose coarse levels have odd widths and heights."""
    e, y8 = launched[size]
    want_src = ol.pyramid(y8[2])
    want_ref = ol.pyramid(y8[1])
    for l in range(5):
        assert np.array_equal(e.fetch_plane(-1, l), want_src[l]), ("src", l)
        assert np.array_equal(e.fetch_plane(0, l), want_ref[l]), ("ref", l)
        assert np.array_equal(e.fetch_plane(-1, l, bordered=True), np.pad(want_src[l], 64, mode="edge")), ("src bordered", l)
        assert np.array_equal(e.fetch_plane(0, l, bordered=True), np.pad(want_ref[l], 64, mode="edge")), ("ref bordered", l)


@pytest.mark.parametrize("size", [(352, 288), (202, 122)])
def test_phase_borders_match(launched, size):
    e, y8 = launched[size]
    want = ol.phases(y8[1])
    for k in (1, 6, 15):
        assert np.array_equal(e.fetch_plane(0, 0, k, bordered=True), np.pad(want[k], 64, mode="edge")), k


@pytest.mark.parametrize("size", [(352, 288), (202, 122)])
def test_phases_match(launched, size):
    e, y8 = launched[size]
    want = ol.phases(y8[1])
    for k in range(16):
        got = e.fetch_plane(0, 0, k)
        bad = np.argwhere(got != want[k])
        assert bad.size == 0, (k, bad[:5].tolist())


@pytest.mark.parametrize("size", [(352, 288), (202, 122)])
def test_fields_match(launched, size):
    e, y8 = launched[size]
    O = ol.oracle()
    for r in (0, 1):
        want = ol.fields(O, "orc_", y8[2], y8[1 - r])
        for l in range(4, -1, -1):
            gx, gy = e.fetch_field(r, l)
            bad = np.argwhere((gx != want[l][0]) | (gy != want[l][1]))
            assert bad.size == 0, (r, l, bad[:5].tolist(), [(int(gx[i, j]), int(gy[i, j]), int(want[l][0][i, j]), int(want[l][1][i, j])) for i, j in bad[:5]])


@pytest.mark.parametrize("size", [(352, 288), (202, 122)])
def test_intra_scores_match(launched, size):
    """All 13 modes' SATD of every candidate block of levels 1-3 (block_mode_satd,
    src/gpu/intra_common.h:149-163), frame edges and the partial right/bottom superblocks included."""
    e, y8 = launched[size]
    O = ol.oracle()
    src = np.ascontiguousarray(y8[2])
    H, W = src.shape
    got = e.fetch_intra()
    xywh = (ctypes.c_int * 4)()
    bad = []
    for sb in range(got.shape[0]):
        ox, oy = 64 * (sb % e.sb_cols), 64 * (sb // e.sb_cols)
        for b in range(5, 425):
            O.orc_block_geom(b, xywh)
            x, y, bw, bh = xywh
            for m in range(13):
                want = O.orc_intra_satd(src.ctypes.data, W, H, ox + x, oy + y, bw, bh, m)
                if got[sb, b, m] != want:
                    bad.append((sb, b, m, int(got[sb, b, m]), want))
    assert not bad, (len(bad), bad[:8])
