"""ctypes bindings of the oracle (oracle/liboracle.so) and, where built, of the reference's own
primitives (oracle/_ref/libcelv2ref.so). Test infrastructure only."""
import ctypes
import os

import numpy as np

from celv2_amd.records import SB_DTYPE, BLOCK_DTYPE

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.path.join(ROOT, "oracle", "liboracle.so")
REF_SO = os.path.join(ROOT, "oracle", "_ref", "libcelv2ref.so")

P = ctypes.c_void_p
I = ctypes.c_int
LL = ctypes.c_longlong
I64 = ctypes.c_int64

_oracle = None
_ref = None


def _prims(lib, prefix):
    f = lambda n: getattr(lib, prefix + n)
    f("wht4").argtypes = [P]
    f("mode_bits").argtypes = [I]
    f("block_geom").argtypes = [I, P]
    f("block_index").argtypes = [I, I, I, I]
    f("cand_count").argtypes = [I, I, I, I]
    f("cand_cell").argtypes = [I] * 7
    f("cand_repeats").argtypes = [P, P] + [I] * 7
    f("pair_allowed").argtypes = [I, I, I]
    f("qpel_sample").argtypes = [P] + [I] * 6
    f("search_cell").argtypes = [P, P, P, I, I, I, I, P, P, I, I, I, P]
    f("intra_satd").argtypes = [P] + [I] * 7
    f("intra_satd").restype = LL
    f("intra_pred").argtypes = [P] + [I] * 9
    f("inter_satd").argtypes = [P, P] + [I] * 8
    f("inter_satd").restype = LL
    f("compound_satd").argtypes = [P, P, P] + [I] * 10
    f("compound_satd").restype = LL
    f("fields").argtypes = [I, I, P, P, P, P, P]
    return lib


def tx_api(lib):
    lib.orc_fwd_tx.argtypes = [P, I, I, I, I, I, ctypes.c_double, P, P]
    return lib


def fwd_tx(res, vkind=0, hkind=0, scale=1.0, lib=None):
    """The oracle's forward transform of one residual block (int16, h x w): (coefficients, min(h, 32)
    x min(w, 32) int32, row = vertical frequency; dropped energy)."""
    lib = lib or oracle()
    res = np.ascontiguousarray(res, dtype=np.int16)
    h, w = res.shape
    out = np.zeros((min(h, 32), min(w, 32)), np.int32)
    dropped = np.zeros(1, np.int64)
    rc = lib.orc_fwd_tx(res.ctypes.data, w, w, h, vkind, hkind, scale, out.ctypes.data, dropped.ctypes.data)
    if rc != 0:
        raise ValueError(f"no {w}x{h} transform of kinds {vkind}/{hkind}")
    return out, int(dropped[0])


def oracle():
    global _oracle
    if _oracle is None:
        if not os.path.exists(ORACLE_SO):
            raise RuntimeError("oracle/liboracle.so is not built (make -C oracle)")
        lib = _prims(ctypes.CDLL(ORACLE_SO), "orc_")
        lib.orc_last_error.restype = ctypes.c_char_p
        lib.orc_decide_frame.argtypes = [I, I, P, P, I, I64, I, P]
        lib.orc_decide_frame_masked.argtypes = [I, I, P, P, I, I64, I, P, P]
        lib.orc_decide_frame_pruned.argtypes = [I, I, P, P, I, I64, I, P, P, P, P, P]
        lib.orc_engine_create.argtypes = [I, I, I, I, I]
        lib.orc_engine_create.restype = P
        lib.orc_engine_destroy.argtypes = [P]
        lib.orc_engine_launch.argtypes = [P, I, I, P, P, I, I64]
        lib.orc_engine_collect.argtypes = [P, P, P]
        lib.orc_decision_threads.argtypes = [I, I]
        lib.orc_round_plane.argtypes = [P, P, LL, I]
        lib.orc_pyramid.argtypes = [I, I, P, I, P, LL, P]
        lib.orc_pyramid.restype = LL
        lib.orc_decide_square.argtypes = [I, I, P, P, P, P, I, I, I, I64, I, I, I, I, I, P, P, P, P]
        lib.orc_decide_square.restype = LL
        tx_api(lib)
        assert lib.orc_sizeof_sb() == SB_DTYPE.itemsize, "celv2_sb layout mismatch"
        _oracle = lib
    return _oracle


def ref():
    """The reference's primitives, or None where oracle/_ref is not built."""
    global _ref
    if _ref is None and os.path.exists(REF_SO):
        lib = _prims(ctypes.CDLL(REF_SO), "ref_")
        lib.ref_half_sample.argtypes = [P, I, I, I, I]
        _ref = lib
    return _ref


def ptr_array(arrays):
    arr = (ctypes.c_void_p * max(1, len(arrays)))()
    for i, a in enumerate(arrays):
        arr[i] = a.ctypes.data
    return arr


def decide_frame(y8, refs, lam, width=None, height=None, sb_size=64, intra_masks=None):
    """The oracle's whole-frame decision (single thread): numpy array of SB_DTYPE records;
    intra_masks: per-superblock intra mode groups (the pruned form's restriction)."""
    ph, pw = y8.shape
    width = width or pw
    height = height or ph
    cols, rows = (width + sb_size - 1) // sb_size, (height + sb_size - 1) // sb_size
    out = np.zeros(cols * rows, dtype=SB_DTYPE)
    refs = [np.ascontiguousarray(r) for r in refs]
    if intra_masks is not None:
        m = np.ascontiguousarray(intra_masks, dtype=np.uint8)
        assert m.size == cols * rows
        rc = oracle().orc_decide_frame_masked(width, height, y8.ctypes.data, ptr_array(refs), len(refs), lam, sb_size,
                                              m.ctypes.data, out.ctypes.data)
    else:
        rc = oracle().orc_decide_frame(width, height, y8.ctypes.data, ptr_array(refs), len(refs), lam, sb_size, out.ctypes.data)
    if rc != 0:
        raise RuntimeError(oracle().orc_last_error().decode())
    return out


def decide_frame_pruned(y8, refs, lam, width=None, height=None, limits=None, modes=None, masks=None, sb_size=64):
    """The oracle's whole-frame decision under the pruned form's restrictions (each per superblock,
    row-major, or None): reference limits, motion-search modes, intra mode groups. Returns (records,
    per-superblock decision costs)."""
    ph, pw = y8.shape
    width = width or pw
    height = height or ph
    cols, rows = (width + sb_size - 1) // sb_size, (height + sb_size - 1) // sb_size
    out = np.zeros(cols * rows, dtype=SB_DTYPE)
    costs = np.zeros(cols * rows, np.int64)
    refs = [np.ascontiguousarray(r) for r in refs]
    keep = []

    def arr(a):
        if a is None:
            return None
        a = np.ascontiguousarray(a, dtype=np.uint8)
        assert a.size == cols * rows
        keep.append(a)
        return a.ctypes.data

    rc = oracle().orc_decide_frame_pruned(width, height, y8.ctypes.data, ptr_array(refs), len(refs), lam, sb_size,
                                          arr(limits), arr(modes), arr(masks), out.ctypes.data, costs.ctypes.data)
    if rc != 0:
        raise RuntimeError(oracle().orc_last_error().decode())
    return out, costs


class CpuEngine:
    """The oracle's threaded engine (CpuEngine, src/gpu/cpu_engine.cpp)."""

    def __init__(self, width, height, sb_size=64, threads=4, fail_row=-1):
        self.width, self.height, self.sb_size = width, height, sb_size
        self.h = oracle().orc_engine_create(width, height, sb_size, threads, fail_row)
        if not self.h:
            raise RuntimeError(oracle().orc_last_error().decode())
        self._keep = None

    def __del__(self):
        if getattr(self, "h", None):
            oracle().orc_engine_destroy(self.h)
            self.h = None

    def launch(self, y8, refs, lam):
        refs = [np.ascontiguousarray(r) for r in refs]
        self._keep = (y8, refs)
        rc = oracle().orc_engine_launch(self.h, self.width, self.height, y8.ctypes.data, ptr_array(refs), len(refs), lam)
        if rc != 0:
            raise RuntimeError(oracle().orc_last_error().decode())

    def collect(self):
        cols = (self.width + self.sb_size - 1) // self.sb_size
        rows = (self.height + self.sb_size - 1) // self.sb_size
        out = np.zeros(cols * rows, dtype=SB_DTYPE)
        ready = ctypes.c_int(0)
        rc = oracle().orc_engine_collect(self.h, out.ctypes.data, ctypes.byref(ready))
        if rc != 0:
            raise RuntimeError(oracle().orc_last_error().decode(), ready.value)
        return out, ready.value


def fields(lib, prefix, y_cur, y_ref):
    """All five levels' fields of cur against ref: list of (mvx, mvy) arrays (rows x cols)."""
    ph, pw = y_cur.shape
    n = 0
    w, h = pw, ph
    for _ in range(5):
        n += ((w + 7) // 8) * ((h + 7) // 8)
        w //= 2
        h //= 2
    mvx = np.zeros(n, np.int32)
    mvy = np.zeros(n, np.int32)
    dims = np.zeros(15, np.int32)
    rc = getattr(lib, prefix + "fields")(pw, ph, y_cur.ctypes.data, y_ref.ctypes.data, mvx.ctypes.data, mvy.ctypes.data, dims.ctypes.data)
    assert rc == n, rc
    out = []
    for l in range(5):
        c, r, o = dims[3 * l : 3 * l + 3]
        out.append((mvx[o : o + c * r].reshape(r, c).copy(), mvy[o : o + c * r].reshape(r, c).copy()))
    return out


def phases(y8):
    """A reference's level 0 with its 15 quarter-pel phase planes (16 planes, oracle)."""
    ph, pw = y8.shape
    total = sum(((pw >> l) * (ph >> l)) for l in range(5)) + 15 * pw * ph + 64
    buf = np.zeros(total, np.uint8)
    dims = np.zeros(10, np.int32)
    n = oracle().orc_pyramid(pw, ph, y8.ctypes.data, 1, buf.ctypes.data, total, dims.ctypes.data)
    assert n > 0
    return buf[: 16 * pw * ph].reshape(16, ph, pw).copy()


def pyramid(y8):
    """The five levels of a source pyramid (oracle)."""
    ph, pw = y8.shape
    total = sum(((pw >> l) * (ph >> l)) for l in range(5)) + 64
    buf = np.zeros(total, np.uint8)
    dims = np.zeros(10, np.int32)
    n = oracle().orc_pyramid(pw, ph, y8.ctypes.data, 0, buf.ctypes.data, total, dims.ctypes.data)
    assert n > 0
    out, off = [], 0
    for l in range(5):
        w, h = dims[2 * l], dims[2 * l + 1]
        out.append(buf[off : off + w * h].reshape(h, w).copy())
        off += w * h
    return out
