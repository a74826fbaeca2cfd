"""The y4m reader and writer (include/celv2/y4m.h) against CelV2's own tests of its reader
(tests/test_io.cpp:32-83) and against CelV2's reader itself (src/io/y4m.cpp, built from its own two
files into oracle/_ref by `make -C oracle ref`) on the same files. CPU only."""
import ctypes
import os

import numpy as np
import pytest

from celv2_amd import y4m

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF_IO = os.path.join(ROOT, "oracle", "_ref", "libcelv2ref_io.so")


def write_raw(path, header, payloads):
    with open(path, "wb") as f:
        f.write(header.encode() + b"\n")
        for p in payloads:
            f.write(b"FRAME\n" + p)


def test_roundtrip_16x8(tmp_path):
    """tests/test_io.cpp:32-50: a frame written by write_y4m_frame_bytes reads back equal."""
    y = (np.arange(16 * 8) * 3 % 256).astype(np.uint8).reshape(8, 16)
    u = (100 + np.arange(32)).astype(np.uint8).reshape(4, 8)
    v = (200 - np.arange(32)).astype(np.uint8).reshape(4, 8)
    src = y4m.Frame(16, 8, 8, np.concatenate([y.ravel(), u.ravel(), v.ravel()]), None)
    payload = y4m.frame_bytes(src)
    assert payload == y.tobytes() + u.tobytes() + v.tobytes()
    p = tmp_path / "rt.y4m"
    write_raw(p, "YUV4MPEG2 W16 H8 F30:1 Ip A1:1 C420jpeg", [payload])
    r = y4m.Y4mReader(p)
    assert (r.width, r.height, r.fps_num, r.fps_den, r.bit_depth) == (16, 8, 30, 1, 8)
    g = r.read()
    assert (g.width, g.height, g.pw, g.ph) == (16, 8, 16, 8)
    assert np.array_equal(g.y, y) and np.array_equal(g.u, u) and np.array_equal(g.v, v)
    assert r.read() is None


def test_odd_size_pads_by_edge_replication(tmp_path):
    """tests/test_io.cpp:52-72: 201x121 (chroma 101x61) padded to 208x128."""
    w, h, cw, ch = 201, 121, 101, 61
    pay = bytearray(w * h + 2 * cw * ch)
    for i in range(w * h):
        pay[i] = i % 251
    for i in range(cw * ch):
        pay[w * h + i] = i % 199
        pay[w * h + cw * ch + i] = i % 97
    p = tmp_path / "odd.y4m"
    write_raw(p, "YUV4MPEG2 W201 H121 F25:1 Ip A1:1 C420jpeg", [bytes(pay)])
    g = y4m.Y4mReader(p).read()
    assert (g.pw, g.ph, g.u.shape[1]) == (208, 128, 104)
    assert g.u[60, 100] == pay[w * h + 60 * cw + 100]
    assert g.u[60, 101] == g.u[60, 100]  # right padding replicates the edge
    assert g.u[63, 5] == g.u[60, 5]      # bottom padding replicates the edge
    assert g.y[120, 207] == g.y[120, 200]


@pytest.mark.parametrize("header", ["YUV4MPEG2 W16 H8 F30:1 C444", "YUV4MPEG2 W16 H8 F30:1 C420p12",
                                    "YUV4MPEG2 H8 F30:1", "MPEG2 W16 H8"])
def test_rejects_unsupported(tmp_path, header):
    """tests/test_io.cpp:74-83, and a header without a size or signature."""
    p = tmp_path / "bad.y4m"
    write_raw(p, header, [bytes(16 * 8 * 3)])
    with pytest.raises(y4m.Y4mError):
        y4m.Y4mReader(p)


def ten_bit_clip(w, h, n, seed, max_value=1023):
    rng = np.random.default_rng(seed)
    cw, ch = (w + 1) // 2, (h + 1) // 2
    return [rng.integers(0, max_value + 1, size=w * h + 2 * cw * ch).astype("<u2").tobytes() for _ in range(n)]


def eight_bit_clip(w, h, n, seed):
    rng = np.random.default_rng(seed)
    cw, ch = (w + 1) // 2, (h + 1) // 2
    return [rng.integers(0, 256, size=w * h + 2 * cw * ch).astype(np.uint8).tobytes() for _ in range(n)]


def test_ten_bit_rounds_to_eight_and_rejects_samples_above_the_depth(tmp_path):
    p = tmp_path / "t10.y4m"
    frames = ten_bit_clip(37, 19, 2, 3)
    write_raw(p, "YUV4MPEG2 W37 H19 F30000:1001 C420p10", frames)
    r = y4m.Y4mReader(p)
    assert (r.bit_depth, r.fps_num, r.fps_den) == (10, 30000, 1001)
    g = r.read()
    raw = np.frombuffer(frames[0], "<u2")
    assert np.array_equal(g.hy[:19, :37].ravel(), raw[:37 * 19])
    # round_to_8bit (src/io/y4m.cpp:28-34): min(255, (v + 2) >> 2), on the padded planes too.
    assert np.array_equal(g.y, np.minimum(255, (g.hy.astype(np.int32) + 2) >> 2))
    bad = tmp_path / "t10bad.y4m"
    over = np.frombuffer(ten_bit_clip(8, 8, 1, 4)[0], "<u2").copy()
    over[5] = 1024
    write_raw(bad, "YUV4MPEG2 W8 H8 F30:1 C420p10", [over.tobytes()])
    with pytest.raises(y4m.Y4mError, match="above the bit depth"):
        y4m.Y4mReader(bad).read()


def test_truncated_frame_and_bad_marker(tmp_path):
    p = tmp_path / "trunc.y4m"
    write_raw(p, "YUV4MPEG2 W16 H8 F30:1", [bytes(100)])
    with pytest.raises(y4m.Y4mError, match="truncated"):
        y4m.Y4mReader(p).read()
    q = tmp_path / "marker.y4m"
    with open(q, "wb") as f:
        f.write(b"YUV4MPEG2 W16 H8 F30:1\nFRAMX\n" + bytes(16 * 8 * 3 // 2))
    with pytest.raises(y4m.Y4mError, match="marker"):
        y4m.Y4mReader(q).read()


def ref_read(path, out_depth, max_frames, w, h):
    lib = ctypes.CDLL(REF_IO)
    lib.ref_y4m_read.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
    lib.ref_y4m_error.restype = ctypes.c_char_p
    pw, ph = (w + 7) & ~7, (h + 7) & ~7
    per = pw * ph + 2 * (pw // 2) * (ph // 2)
    p8 = np.zeros(per * max_frames, np.uint8)
    p16 = np.zeros(per * max_frames, np.uint16)
    info = np.zeros(5, np.int32)
    n = lib.ref_y4m_read(str(path).encode(), out_depth, max_frames, p8.ctypes.data, p16.ctypes.data, info.ctypes.data)
    if n < 0:
        raise RuntimeError(lib.ref_y4m_error().decode())
    return n, info, p8.reshape(max_frames, per)[:n], p16.reshape(max_frames, per)[:n]


@pytest.mark.skipif(not os.path.exists(REF_IO), reason="oracle/_ref not built (the reference tree is not mounted)")
@pytest.mark.parametrize("w,h,depth,out_depth", [(64, 48, 8, 0), (201, 121, 8, 0), (37, 19, 10, 0), (202, 122, 10, 8),
                                                 (33, 17, 8, 10), (352, 288, 10, 0), (7, 5, 10, 10)])
def test_matches_the_reference_reader(tmp_path, w, h, depth, out_depth):
    frames = (ten_bit_clip if depth == 10 else eight_bit_clip)(w, h, 3, w * 7 + h)
    p = tmp_path / "clip.y4m"
    write_raw(p, f"YUV4MPEG2 W{w} H{h} F24:1 Ip A0:0 C{'420p10' if depth == 10 else '420mpeg2'} XYSCSS=420", frames)
    n, info, r8, r16 = ref_read(p, out_depth, 4, w, h)
    r = y4m.Y4mReader(p)
    if out_depth:
        r.set_bit_depth(out_depth)
    assert [r.width, r.height, r.bit_depth, r.fps_num, r.fps_den] == list(info)
    got = list(r)
    assert n == len(got) == 3
    for k, g in enumerate(got):
        assert np.array_equal(np.concatenate([g.y.ravel(), g.u.ravel(), g.v.ravel()]), r8[k])
        if g.bit_depth > 8:
            assert np.array_equal(np.concatenate([g.hy.ravel(), g.hu.ravel(), g.hv.ravel()]), r16[k])
        if out_depth in (0, depth):
            assert y4m.frame_bytes(g) == frames[k]  # the writer gives back the file's payload
