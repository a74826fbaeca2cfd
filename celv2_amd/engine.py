"""Python face of libcelv2hip.so (ctypes).

`HipEngine` drives the C++ engine (celv2::HipEngine, CelV2's DecisionEngine interface:
launch / prepare / collect, src/gpu/decision_engine.h:9-51); `RawEngine` drives the C ABI
underneath it (include/celv2hip.h), including the device-resident mode the benchmark times.
Both return decision maps as numpy arrays of `records.SB_DTYPE`. Nothing here falls back to a CPU
path: without the built library or a GPU the constructors raise.
"""
import ctypes
import os

import numpy as np

from .records import SB_DTYPE

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libcelv2hip.so")
_LIB = None

P = ctypes.c_void_p
I = ctypes.c_int
LL = ctypes.c_longlong
FEATURES = 14  # the learned pruner's inputs per superblock (include/celv2hip.h)
N_OUT = 10  # its outputs: 6 reference outputs, 3 intra mode-group outputs, the search output (CELV2HIP_PRUNER_OUTPUTS)


class Luma(ctypes.Structure):
    _fields_ = [("y8", P), ("y16", P), ("bit_depth", ctypes.c_int32), ("on_device", ctypes.c_int32)]


def lib():
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built: run __graft_entry__.build()")
        L = ctypes.CDLL(LIB_PATH)
        L.celv2hip_last_error.restype = ctypes.c_char_p
        L.celv2hip_device_count.argtypes = [ctypes.POINTER(I)]
        L.celv2hip_create.argtypes = [I, I, I, I, ctypes.POINTER(P)]
        L.celv2hip_destroy.argtypes = [P]
        L.celv2hip_device_name.argtypes = [P, ctypes.c_char_p, I]
        L.celv2hip_set_row_streams.argtypes = [P, I]
        L.celv2hip_set_row_groups.argtypes = [P, I]
        L.celv2hip_row_groups.argtypes = [P]
        L.celv2hip_set_resident.argtypes = [P, I]
        L.celv2hip_set_graph.argtypes = [P, I]
        L.celv2hip_graph_stats.argtypes = [P, ctypes.POINTER(LL), ctypes.POINTER(LL)]
        L.celv2hip_prepare.argtypes = [P, ctypes.POINTER(Luma), LL]
        L.celv2hip_launch.argtypes = [P, ctypes.POINTER(Luma), LL, ctypes.POINTER(Luma), ctypes.POINTER(LL), I, ctypes.c_int64]
        L.celv2hip_launch_batch.argtypes = [ctypes.POINTER(P), I, ctypes.POINTER(Luma), ctypes.POINTER(LL),
                                            ctypes.POINTER(ctypes.POINTER(Luma)), ctypes.POINTER(ctypes.POINTER(LL)),
                                            ctypes.POINTER(I), ctypes.POINTER(ctypes.c_int64)]
        L.celv2hip_row_order.argtypes = [P, ctypes.POINTER(I), I]
        L.celv2hip_wait_row.argtypes = [P, I, P]
        L.celv2hip_fetch_row.argtypes = [P, I, P]
        L.celv2hip_sync.argtypes = [P]
        L.celv2hip_stats_reset.argtypes = [P]
        L.celv2hip_set_timing.argtypes = [P, I]
        L.celv2hip_stats.argtypes = [P, ctypes.POINTER(ctypes.c_double), I, ctypes.POINTER(LL), I]
        L.celv2hip_fetch_plane.argtypes = [P, I, I, I, P, ctypes.POINTER(I), ctypes.POINTER(I)]
        L.celv2hip_fetch_plane_bordered.argtypes = [P, I, I, I, P, ctypes.POINTER(I), ctypes.POINTER(I)]
        L.celv2hip_fetch_field.argtypes = [P, I, I, P, P, ctypes.POINTER(I), ctypes.POINTER(I)]
        L.celv2hip_set_ref_limits.argtypes = [P, P, I]
        L.celv2hip_set_intra_masks.argtypes = [P, P, I]
        L.celv2hip_set_search_modes.argtypes = [P, P, I]
        L.celv2hip_fetch_search_modes.argtypes = [P, P, ctypes.c_longlong]
        L.celv2hip_fetch_search_modes.restype = ctypes.c_longlong
        L.celv2hip_set_pruner_search.argtypes = [P, ctypes.c_float]
        L.celv2hip_fetch_intra_masks.argtypes = [P, P, ctypes.c_longlong]
        L.celv2hip_fetch_intra_masks.restype = ctypes.c_longlong
        L.celv2hip_set_pruner_groups.argtypes = [P, ctypes.c_float]
        L.celv2hip_pruner_create.argtypes = [I, I, ctypes.c_uint, ctypes.POINTER(P)]
        L.celv2hip_pruner_destroy.argtypes = [P]
        L.celv2hip_pruner_get_params.argtypes = [P, P, I]
        L.celv2hip_pruner_set_params.argtypes = [P, P, I]
        L.celv2hip_pruner_forward.argtypes = [P, P, I, I, I, P]
        L.celv2hip_pruner_forward_static.argtypes = [P, P, I, I, I, P]
        L.celv2hip_pruner_gradients.argtypes = [P, P, P, I, I, I, P, ctypes.POINTER(ctypes.c_float)]
        L.celv2hip_pruner_train_step.argtypes = [P, P, P, I, I, I, ctypes.c_float, ctypes.c_float, ctypes.POINTER(ctypes.c_float)]
        L.celv2hip_set_pruner.argtypes = [P, P, ctypes.c_float, ctypes.c_float]
        L.celv2hip_set_features.argtypes = [P, I]
        L.celv2hip_fetch_features.argtypes = [P, P, ctypes.c_longlong]
        L.celv2hip_fetch_features.restype = ctypes.c_longlong
        L.celv2hip_fetch_ref_limits.argtypes = [P, P, ctypes.c_longlong]
        L.celv2hip_fetch_ref_limits.restype = ctypes.c_longlong
        L.celv2hip_fetch_sb_costs.argtypes = [P, P, ctypes.c_longlong]
        L.celv2hip_fetch_sb_costs.restype = ctypes.c_longlong
        L.celv2hip_fetch_intra.argtypes = [P, P, ctypes.c_longlong]
        L.celv2hip_fetch_intra.restype = ctypes.c_longlong
        L.celv2hip_device_alloc.argtypes = [I, LL, ctypes.POINTER(P)]
        L.celv2hip_device_free.argtypes = [I, P]
        L.celv2hip_copy_to_device.argtypes = [I, P, P, LL]
        L.celv2hip_copy_on_device.argtypes = [I, P, P, LL]
        L.celv2hip_copy_from_device.argtypes = [I, P, P, LL]
        L.celv2hip_fwd_tx.argtypes = [I, P, I, ctypes.POINTER(ctypes.c_float)]
        L.celv2_engine_last_error.restype = ctypes.c_char_p
        L.celv2_engine_create.argtypes = [I, I, I, I, I]
        L.celv2_engine_create.restype = P
        L.celv2_engine_destroy.argtypes = [P]
        L.celv2_engine_launch.argtypes = [P, P, P, I, P, P, ctypes.POINTER(LL), I, I, LL, LL]
        L.celv2_engine_prepare.argtypes = [P, P, P, I, LL]
        L.celv2_engine_launch_device.argtypes = [P, ctypes.POINTER(Luma), ctypes.POINTER(Luma), ctypes.POINTER(LL), I, I, LL, LL]
        L.celv2_engine_prepare_device.argtypes = [P, ctypes.POINTER(Luma), LL]
        L.celv2_engine_collect.argtypes = [P, P, ctypes.POINTER(I)]
        L.celv2_engine_description.argtypes = [P, ctypes.c_char_p, I]
        L.celv2_engine_sync.argtypes = [P]
        L.celv2_engine_set_row_groups.argtypes = [P, I]
        L.celv2_engine_row_groups.argtypes = [P]
        L.celv2_engine_autotune_device.argtypes = [P, ctypes.POINTER(Luma), ctypes.POINTER(Luma), ctypes.POINTER(LL), I, I, LL,
                                                   LL, I, I, ctypes.POINTER(I), ctypes.POINTER(I),
                                                   ctypes.POINTER(ctypes.c_double), I, ctypes.POINTER(I)]
        L.celv2_lambda_q8.argtypes = [I]
        L.celv2_lambda_q8.restype = ctypes.c_int64
        if L.celv2_sizeof_sb() != SB_DTYPE.itemsize:
            raise RuntimeError("celv2_sb layout mismatch between the library and records.SB_DTYPE")
        L.celv2hip_source_hash.restype = ctypes.c_char_p
        if os.path.isdir(os.path.join(_HERE, "csrc")):
            from .build import source_hash

            built, here = L.celv2hip_source_hash().decode(), source_hash()
            if built != here:
                raise RuntimeError(f"{LIB_PATH} was built from other sources ({built}, sources {here}): "
                                   "run __graft_entry__.build()")
        _LIB = L
    return _LIB


def last_error():
    return lib().celv2hip_last_error().decode()


def device_count():
    n = I(0)
    lib().celv2hip_device_count(ctypes.byref(n))
    return n.value


def lambda_q8(qindex):
    return int(lib().celv2_lambda_q8(qindex))


def _ptrs(arrays):
    arr = (P * max(1, len(arrays)))()
    for i, a in enumerate(arrays):
        arr[i] = a.ctypes.data if a is not None else None
    return arr


class EngineError(RuntimeError):
    pass


class HipEngine:
    """celv2::HipEngine: CelV2's DecisionEngine on the GPU."""

    def __init__(self, width, height, device=0, row_streams=1, full_depth=False):
        self.width, self.height = width, height
        self.sb_cols, self.sb_rows = (width + 63) // 64, (height + 63) // 64
        self.h = lib().celv2_engine_create(width, height, device, row_streams, 1 if full_depth else 0)
        if not self.h:
            raise EngineError(lib().celv2_engine_last_error().decode())
        self._keep = []

    def close(self):
        if getattr(self, "h", None):
            lib().celv2_engine_destroy(self.h)
            self.h = None

    __del__ = close

    def description(self):
        buf = ctypes.create_string_buffer(256)
        lib().celv2_engine_description(self.h, buf, 256)
        return buf.value.decode()

    def launch(self, src8, refs8, qindex=120, lam=-1, keys=None, src_key=-1, src16=None, refs16=None, bit_depth=8):
        refs8 = [np.ascontiguousarray(r) for r in refs8] if refs8 is not None else None
        refs16 = [np.ascontiguousarray(r) for r in refs16] if refs16 is not None else None
        n = len(refs8) if refs8 is not None else len(refs16 or [])
        k = (LL * max(1, n))(*keys) if keys is not None else None
        rc = lib().celv2_engine_launch(
            self.h,
            src8.ctypes.data if src8 is not None else None,
            src16.ctypes.data if src16 is not None else None,
            bit_depth,
            _ptrs(refs8) if refs8 is not None else None,
            _ptrs(refs16) if refs16 is not None else None,
            k,
            n,
            qindex,
            lam,
            src_key,
        )
        if rc == 1:
            raise ValueError(lib().celv2_engine_last_error().decode())
        if rc != 0:
            raise EngineError(lib().celv2_engine_last_error().decode())

    def prepare(self, y8, key, y16=None, bit_depth=8):
        rc = lib().celv2_engine_prepare(self.h, y8.ctypes.data if y8 is not None else None,
                                        y16.ctypes.data if y16 is not None else None, bit_depth, key)
        if rc != 0:
            raise EngineError(lib().celv2_engine_last_error().decode())

    def launch_device(self, src, refs, qindex=120, lam=-1, keys=None, src_key=-1):
        """launch() for frames in device memory: src / refs are RawEngine.luma(dev8=... or dev16=...)."""
        arr = (Luma * max(1, len(refs)))(*refs)
        k = (LL * max(1, len(refs)))(*keys) if keys is not None else None
        rc = lib().celv2_engine_launch_device(self.h, ctypes.byref(src), arr, k, len(refs), qindex, lam, src_key)
        if rc == 1:
            raise ValueError(lib().celv2_engine_last_error().decode())
        if rc != 0:
            raise EngineError(lib().celv2_engine_last_error().decode())

    def prepare_device(self, luma, key):
        if lib().celv2_engine_prepare_device(self.h, ctypes.byref(luma), key) != 0:
            raise EngineError(lib().celv2_engine_last_error().decode())

    def collect(self, export=True):
        """(records, ready rows); export False: the DecisionMap stays in the engine (timing), records None."""
        out = np.zeros(self.sb_cols * self.sb_rows, dtype=SB_DTYPE) if export else None
        ready = I(0)
        rc = lib().celv2_engine_collect(self.h, out.ctypes.data if export else None, ctypes.byref(ready))
        if rc != 0:
            raise EngineError(lib().celv2_engine_last_error().decode())
        return out, ready.value

    def sync(self):
        if lib().celv2_engine_sync(self.h) != 0:
            raise EngineError(lib().celv2_engine_last_error().decode())

    def set_row_groups(self, groups):
        """Row groups of the one-row-stream launch (0: the default); the decisions do not depend on them."""
        if lib().celv2_engine_set_row_groups(self.h, groups) != 0:
            raise ValueError(lib().celv2_engine_last_error().decode())

    def row_groups(self):
        return lib().celv2_engine_row_groups(self.h)

    def autotune_device(self, src, refs, qindex=120, lam=-1, keys=None, src_key=-1, frames=5, force=False):
        """Times steady-state frames (launch -> prepare -> collect) of these device frames at each
        candidate row-group count and keeps the fastest: (chosen count, {count: median ms}); the dict is
        empty when the process had tuned this size, reference count and device already."""
        arr = (Luma * max(1, len(refs)))(*refs)
        k = (LL * max(1, len(refs)))(*keys) if keys is not None else None
        g, n = I(0), I(0)
        cg, cm = (I * 16)(), (ctypes.c_double * 16)()
        rc = lib().celv2_engine_autotune_device(self.h, ctypes.byref(src), arr, k, len(refs), qindex, lam, src_key, frames,
                                                int(force), ctypes.byref(g), cg, cm, 16, ctypes.byref(n))
        if rc == 1:
            raise ValueError(lib().celv2_engine_last_error().decode())
        if rc != 0:
            raise EngineError(lib().celv2_engine_last_error().decode())
        return g.value, {cg[i]: cm[i] for i in range(n.value)}

    def decide(self, src8, refs8, lam, **kw):
        self.launch(src8, refs8, lam=lam, **kw)
        return self.collect()[0]


class RawEngine:
    """The C ABI engine (include/celv2hip.h) for tests of its own behaviour and for timing."""

    def __init__(self, width, height, device=0):
        self.width, self.height, self.device = width, height, device
        self.sb_cols, self.sb_rows = (width + 63) // 64, (height + 63) // 64
        h = P()
        rc = lib().celv2hip_create(width, height, 64, device, ctypes.byref(h))
        if rc != 0:
            raise EngineError(last_error())
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            lib().celv2hip_destroy(self.h)
            self.h = None

    __del__ = close

    def _check(self, rc, what):
        if rc != 0:
            raise EngineError(f"{what}: {last_error()} (status {rc})")

    def set_resident(self, on):
        self._check(lib().celv2hip_set_resident(self.h, 1 if on else 0), "set_resident")

    def set_graph(self, on):
        self._check(lib().celv2hip_set_graph(self.h, 1 if on else 0), "set_graph")

    def graph_stats(self):
        """(launches replayed as graphs, of them the ones that instantiated a graph)."""
        a, b = LL(0), LL(0)
        self._check(lib().celv2hip_graph_stats(self.h, ctypes.byref(a), ctypes.byref(b)), "graph_stats")
        return a.value, b.value

    def set_row_streams(self, n):
        self._check(lib().celv2hip_set_row_streams(self.h, n), "set_row_streams")

    def set_row_groups(self, groups):
        self._check(lib().celv2hip_set_row_groups(self.h, groups), "set_row_groups")

    @staticmethod
    def luma(y8=None, y16=None, bit_depth=8, dev8=None, dev16=None):
        if dev8 is not None or dev16 is not None:
            return Luma(dev8, dev16, bit_depth if dev16 is not None else 8, 1)
        return Luma(y8.ctypes.data if y8 is not None else None, y16.ctypes.data if y16 is not None else None,
                    bit_depth if y16 is not None else 8, 0)

    def prepare(self, luma, key):
        self._check(lib().celv2hip_prepare(self.h, ctypes.byref(luma), key), "prepare")

    def launch(self, src, refs, lam, keys=None, src_key=-1):
        arr = (Luma * max(1, len(refs)))(*refs)
        k = (LL * max(1, len(refs)))(*keys) if keys is not None else None
        self._check(lib().celv2hip_launch(self.h, ctypes.byref(src), src_key, arr, k, len(refs), lam), "launch")

    @staticmethod
    def launch_batch(engines, srcs, refs, lams, keys=None, src_keys=None):
        """One frame of each engine in shared launches (celv2hip_launch_batch)."""
        n = len(engines)
        eng = (P * n)(*[e.h for e in engines])
        sarr = (Luma * n)(*srcs)
        ref_arrs = [(Luma * max(1, len(r)))(*r) for r in refs]
        rp = (ctypes.POINTER(Luma) * n)(*[ctypes.cast(a, ctypes.POINTER(Luma)) for a in ref_arrs])
        key_arrs = [(LL * max(1, len(k)))(*k) for k in keys] if keys is not None else None
        kp = (ctypes.POINTER(LL) * n)(*[ctypes.cast(a, ctypes.POINTER(LL)) for a in key_arrs]) if keys is not None else None
        sk = (LL * n)(*src_keys) if src_keys is not None else None
        nr = (I * n)(*[len(r) for r in refs])
        lm = (ctypes.c_int64 * n)(*lams)
        rc = lib().celv2hip_launch_batch(eng, n, sarr, sk, rp, kp, nr, lm)
        if rc != 0:
            raise EngineError(f"launch_batch: {last_error()}")

    def collect(self):
        out = np.zeros(self.sb_cols * self.sb_rows, dtype=SB_DTYPE)
        rows = (I * self.sb_rows)()
        n = lib().celv2hip_row_order(self.h, rows, self.sb_rows)
        for r in list(rows)[:n]:
            view = out[r * self.sb_cols:(r + 1) * self.sb_cols]
            self._check(lib().celv2hip_wait_row(self.h, r, view.ctypes.data), "wait_row")
        return out

    def fetch(self):
        out = np.zeros(self.sb_cols * self.sb_rows, dtype=SB_DTYPE)
        for r in range(self.sb_rows):
            view = out[r * self.sb_cols:(r + 1) * self.sb_cols]
            self._check(lib().celv2hip_fetch_row(self.h, r, view.ctypes.data), "fetch_row")
        return out

    def sync(self):
        self._check(lib().celv2hip_sync(self.h), "sync")

    def fetch_plane(self, ref, level, phase=0, bordered=False):
        """Interior of a plane of the last launch (ref -1: the source pyramid); bordered: with its
        64-sample border."""
        pw, ph = (self.width + 7) & ~7, (self.height + 7) & ~7
        out = np.zeros((pw + 128) * (ph + 128), np.uint8)
        w, h = I(0), I(0)
        f = lib().celv2hip_fetch_plane_bordered if bordered else lib().celv2hip_fetch_plane
        self._check(f(self.h, ref, level, phase, out.ctypes.data, ctypes.byref(w), ctypes.byref(h)), "fetch_plane")
        return out[: w.value * h.value].reshape(h.value, w.value).copy()

    def fetch_field(self, ref, level):
        pw, ph = (self.width + 7) & ~7, (self.height + 7) & ~7
        n = ((pw + 7) // 8) * ((ph + 7) // 8)
        mx, my = np.zeros(n, np.int32), np.zeros(n, np.int32)
        c, r = I(0), I(0)
        self._check(lib().celv2hip_fetch_field(self.h, ref, level, mx.ctypes.data, my.ctypes.data, ctypes.byref(c), ctypes.byref(r)), "fetch_field")
        k = c.value * r.value
        return mx[:k].reshape(r.value, c.value).copy(), my[:k].reshape(r.value, c.value).copy()

    def fetch_intra(self):
        """Intra SATDs of the last launch's source: (superblocks, 425, 13); blocks 0..4 undefined."""
        n = self.sb_cols * self.sb_rows * 425 * 13
        out = np.zeros(n, np.uint32)
        got = lib().celv2hip_fetch_intra(self.h, out.ctypes.data, n)
        if got < 0:
            self._check(int(got), "fetch_intra")
        return out.reshape(self.sb_cols * self.sb_rows, 425, 13)

    def set_ref_limits(self, k=None):
        """Per-superblock reference limits (uint8, 1..7, row-major) for later launches; None clears."""
        if k is None:
            self._check(lib().celv2hip_set_ref_limits(self.h, None, 0), "set_ref_limits")
            return
        k = np.ascontiguousarray(k, dtype=np.uint8)
        self._check(lib().celv2hip_set_ref_limits(self.h, k.ctypes.data, int(k.size)), "set_ref_limits")

    def set_search_modes(self, m=None):
        """Per-superblock motion-search modes (uint8, bit L narrows level L's full searches, L = 0, 1;
        row-major) for later launches; None: every search full."""
        if m is None:
            self._check(lib().celv2hip_set_search_modes(self.h, None, 0), "set_search_modes")
            return
        m = np.ascontiguousarray(m, dtype=np.uint8)
        self._check(lib().celv2hip_set_search_modes(self.h, m.ctypes.data, int(m.size)), "set_search_modes")

    def set_intra_masks(self, m=None):
        """Per-superblock intra mode groups (uint8 bit sets, 0..15, row-major) for the sources scored
        from now on; None: every group."""
        if m is None:
            self._check(lib().celv2hip_set_intra_masks(self.h, None, 0), "set_intra_masks")
            return
        m = np.ascontiguousarray(m, dtype=np.uint8)
        self._check(lib().celv2hip_set_intra_masks(self.h, m.ctypes.data, int(m.size)), "set_intra_masks")

    def fetch_intra_masks(self):
        """The intra mode groups the last launch's source was scored with (uint8 per superblock)."""
        out = np.zeros(self.sb_cols * self.sb_rows, np.uint8)
        got = lib().celv2hip_fetch_intra_masks(self.h, out.ctypes.data, out.size)
        if got < 0:
            self._check(int(-got), "fetch_intra_masks")
        return out

    def fetch_sb_costs(self):
        out = np.zeros(self.sb_cols * self.sb_rows, np.int64)
        got = lib().celv2hip_fetch_sb_costs(self.h, out.ctypes.data, out.size)
        if got < 0:
            self._check(int(got), "fetch_sb_costs")
        return out

    def set_pruner(self, pruner, threshold=0.5, band=0.0, gthreshold=0.0, sthreshold=0.0):
        """Attaches a pruner (None detaches); gthreshold > 0: it also chooses the intra mode groups;
        sthreshold > 0: and the motion-search modes."""
        self._check(lib().celv2hip_set_pruner(self.h, pruner.h if pruner is not None else None, threshold, band), "set_pruner")
        self._check(lib().celv2hip_set_pruner_groups(self.h, gthreshold if pruner is not None else 0.0), "set_pruner_groups")
        self._check(lib().celv2hip_set_pruner_search(self.h, sthreshold if pruner is not None else 0.0), "set_pruner_search")

    def fetch_search_modes(self):
        """The motion-search modes the last launch's levels 1 and 0 obeyed (uint8 per superblock)."""
        out = np.zeros(self.sb_cols * self.sb_rows, np.uint8)
        got = lib().celv2hip_fetch_search_modes(self.h, out.ctypes.data, out.size)
        if got < 0:
            self._check(int(got), "fetch_search_modes")
        return out

    def set_features(self, on):
        self._check(lib().celv2hip_set_features(self.h, 1 if on else 0), "set_features")

    def fetch_features(self):
        out = np.zeros((self.sb_rows, self.sb_cols, FEATURES), np.float32)
        got = lib().celv2hip_fetch_features(self.h, out.ctypes.data, out.size)
        if got < 0:
            self._check(int(got), "fetch_features")
        return out

    def fetch_ref_limits(self):
        out = np.zeros(self.sb_cols * self.sb_rows, np.uint8)
        got = lib().celv2hip_fetch_ref_limits(self.h, out.ctypes.data, out.size)
        if got < 0:
            self._check(int(got), "fetch_ref_limits")
        return out

    def set_counting(self, on):
        self._check(lib().celv2hip_set_counting(self.h, 1 if on else 0), "set_counting")

    def set_timing(self, on):
        """Stage events for stats() ms (off by default; they cost the stream a few microseconds each)."""
        self._check(lib().celv2hip_set_timing(self.h, 1 if on else 0), "set_timing")

    def stats_reset(self):
        self._check(lib().celv2hip_stats_reset(self.h), "stats_reset")

    def stats(self):
        """(stage ms sums over the window, counts) - see celv2hip_stats."""
        ms = (ctypes.c_double * 8)()
        cnt = (LL * 7)()
        n = lib().celv2hip_stats(self.h, ms, 8, cnt, 7)
        if n < 0:
            raise EngineError(last_error())
        keys = ["build", "me", "inter", "finalize", "copy", "total", "intra_prepare", "me_level0"]
        ckeys = ["launches", "tile_evals", "me_centres", "compound_blocks", "inter_candidates", "pruner_fallbacks",
                 "me_centres_level0"]
        return dict(zip(keys, list(ms))), dict(zip(ckeys, list(cnt)))


class Pruner:
    """The learned pruner's network (celv2hip_pruner_*): features [b, h, w, 10] float32."""

    def __init__(self, max_positions=16384, seed=1, device=0):
        p = P()
        if lib().celv2hip_pruner_create(device, max_positions, seed, ctypes.byref(p)) != 0:
            raise EngineError(last_error())
        self.h = p.value
        self.nparams = lib().celv2hip_pruner_param_count()

    def close(self):
        if getattr(self, "h", None):
            lib().celv2hip_pruner_destroy(self.h)
            self.h = None

    __del__ = close

    def params(self):
        out = np.zeros(self.nparams, np.float32)
        if lib().celv2hip_pruner_get_params(self.h, out.ctypes.data, out.size) != 0:
            raise EngineError(last_error())
        return out

    def set_params(self, w):
        w = np.ascontiguousarray(w, dtype=np.float32)
        if lib().celv2hip_pruner_set_params(self.h, w.ctypes.data, int(w.size)) != 0:
            raise EngineError(last_error())

    def forward(self, feats):
        feats = np.ascontiguousarray(feats, dtype=np.float32)
        b, h, w, _ = feats.shape
        out = np.zeros((b, h, w, N_OUT), np.float32)
        if lib().celv2hip_pruner_forward(self.h, feats.ctypes.data, b, h, w, out.ctypes.data) != 0:
            raise EngineError(last_error())
        return out

    def forward_static(self, feats):
        """forward() through the entry point that allocates nothing (at most 16384 positions)."""
        feats = np.ascontiguousarray(feats, dtype=np.float32)
        b, h, w, _ = feats.shape
        out = np.zeros((b, h, w, N_OUT), np.float32)
        if lib().celv2hip_pruner_forward_static(self.h, feats.ctypes.data, b, h, w, out.ctypes.data) != 0:
            raise EngineError(last_error())
        return out

    def gradients(self, feats, kstar):
        feats = np.ascontiguousarray(feats, dtype=np.float32)
        kstar = np.ascontiguousarray(kstar, dtype=np.uint8)
        b, h, w, _ = feats.shape
        g = np.zeros(self.nparams, np.float32)
        loss = ctypes.c_float()
        if lib().celv2hip_pruner_gradients(self.h, feats.ctypes.data, kstar.ctypes.data, b, h, w, g.ctypes.data, ctypes.byref(loss)) != 0:
            raise EngineError(last_error())
        return g, loss.value

    def train_step(self, feats, kstar, lr, weight_decay):
        feats = np.ascontiguousarray(feats, dtype=np.float32)
        kstar = np.ascontiguousarray(kstar, dtype=np.uint8)
        b, h, w, _ = feats.shape
        loss = ctypes.c_float()
        if lib().celv2hip_pruner_train_step(self.h, feats.ctypes.data, kstar.ctypes.data, b, h, w, lr, weight_decay, ctypes.byref(loss)) != 0:
            raise EngineError(last_error())
        return loss.value


class DeviceBuffer:
    """A device allocation holding a copy of a numpy array (frames resident in HBM)."""

    def __init__(self, array, device=0):
        self.device = device
        self.nbytes = array.nbytes
        p = P()
        if lib().celv2hip_device_alloc(device, array.nbytes, ctypes.byref(p)) != 0:
            raise EngineError(last_error())
        self.ptr = p.value
        if lib().celv2hip_copy_to_device(device, self.ptr, np.ascontiguousarray(array).ctypes.data, array.nbytes) != 0:
            raise EngineError(last_error())

    def read(self, dtype, count=-1, offset=0):
        """The buffer's bytes from `offset` as a numpy array of dtype (count items; -1: to the end)."""
        dt = np.dtype(dtype)
        n = (self.nbytes - offset) // dt.itemsize if count < 0 else count
        out = np.empty(n, dt)
        if n and lib().celv2hip_copy_from_device(self.device, out.ctypes.data, self.ptr + offset, out.nbytes) != 0:
            raise EngineError(last_error())
        return out

    def copy_from(self, other):
        """Device-to-device copy of other's bytes into this buffer (a copy kernel of known size)."""
        n = min(self.nbytes, other.nbytes)
        if lib().celv2hip_copy_on_device(self.device, self.ptr, other.ptr, n) != 0:
            raise EngineError(last_error())
        return n

    def free(self):
        if self.ptr:
            lib().celv2hip_device_free(self.device, self.ptr)
            self.ptr = None
