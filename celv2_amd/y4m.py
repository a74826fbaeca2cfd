"""y4m input (ctypes over the celv2_y4m_* C ABI of libcelv2hip.so, include/celv2/y4m.h): 4:2:0 at 8 or
10 bits, frames padded to a multiple of 8 by edge replication, with the 8-bit planes the decisions
read (CelV2's Y4mReader, src/io/y4m.cpp:56-158)."""
import ctypes

import numpy as np

from . import engine

_BOUND = False


def _lib():
    global _BOUND
    L = engine.lib()
    if not _BOUND:
        L.celv2_y4m_open.argtypes = [ctypes.c_char_p, ctypes.POINTER(ctypes.c_void_p), ctypes.POINTER(ctypes.c_int32)]
        L.celv2_y4m_set_bit_depth.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.celv2_y4m_read.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.POINTER(ctypes.c_int32)]
        L.celv2_y4m_close.argtypes = [ctypes.c_void_p]
        L.celv2_y4m_close.restype = None
        L.celv2_y4m_last_error.restype = ctypes.c_char_p
        L.celv2_y4m_frame_bytes.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                            ctypes.c_void_p, ctypes.c_size_t, ctypes.POINTER(ctypes.c_size_t)]
        _BOUND = True
    return L


class Y4mError(RuntimeError):
    pass


def _check(rc, what):
    if rc == 1:
        raise ValueError(f"{what}: {_lib().celv2_y4m_last_error().decode()}")
    if rc != 0:
        raise Y4mError(f"{what}: {_lib().celv2_y4m_last_error().decode()}")


class Frame:
    """One frame: padded planes y, u, v (uint8) and, above 8 bits, hy, hu, hv (uint16)."""

    def __init__(self, width, height, bit_depth, p8, p16):
        self.width, self.height, self.bit_depth = width, height, bit_depth
        pw, ph = (width + 7) & ~7, (height + 7) & ~7
        self.pw, self.ph = pw, ph
        n, c = pw * ph, (pw // 2) * (ph // 2)
        self.y, self.u, self.v = p8[:n].reshape(ph, pw), p8[n:n + c].reshape(ph // 2, pw // 2), p8[n + c:].reshape(ph // 2, pw // 2)
        self.hy = self.hu = self.hv = None
        if p16 is not None:
            self.hy, self.hu, self.hv = p16[:n].reshape(ph, pw), p16[n:n + c].reshape(ph // 2, pw // 2), p16[n + c:].reshape(ph // 2, pw // 2)


class Y4mReader:
    def __init__(self, path):
        info = (ctypes.c_int32 * 5)()
        h = ctypes.c_void_p()
        _check(_lib().celv2_y4m_open(str(path).encode(), ctypes.byref(h), info), "y4m open")
        self.h = h
        self.width, self.height, self._depth, self.fps_num, self.fps_den = list(info)

    @property
    def bit_depth(self):
        return self._depth

    def set_bit_depth(self, d):
        _check(_lib().celv2_y4m_set_bit_depth(self.h, d), "y4m set_bit_depth")
        self._depth = d

    def read(self):
        """The next Frame, or None at the end of the stream."""
        pw, ph = (self.width + 7) & ~7, (self.height + 7) & ~7
        size = pw * ph + 2 * (pw // 2) * (ph // 2)
        p8 = np.empty(size, np.uint8)
        p16 = np.empty(size, np.uint16) if self._depth > 8 else None
        got = ctypes.c_int32()
        _check(_lib().celv2_y4m_read(self.h, p8.ctypes.data, p16.ctypes.data if p16 is not None else None,
                                     ctypes.byref(got)), "y4m read")
        return Frame(self.width, self.height, self._depth, p8, p16) if got.value else None

    def __iter__(self):
        while True:
            f = self.read()
            if f is None:
                return
            yield f

    def close(self):
        if getattr(self, "h", None):
            _lib().celv2_y4m_close(self.h)
            self.h = None

    __del__ = close


def frame_bytes(frame):
    """The frame's visible area as one y4m frame payload (write_y4m_frame_bytes)."""
    p8 = np.ascontiguousarray(np.concatenate([frame.y.ravel(), frame.u.ravel(), frame.v.ravel()]))
    p16 = None
    if frame.bit_depth > 8:
        p16 = np.ascontiguousarray(np.concatenate([frame.hy.ravel(), frame.hu.ravel(), frame.hv.ravel()]))
    cap = p8.size * 2
    out = np.empty(cap, np.uint8)
    n = ctypes.c_size_t()
    _check(_lib().celv2_y4m_frame_bytes(p8.ctypes.data, p16.ctypes.data if p16 is not None else None, frame.width,
                                        frame.height, frame.bit_depth, out.ctypes.data, cap, ctypes.byref(n)), "y4m frame_bytes")
    return out[:n.value].tobytes()
