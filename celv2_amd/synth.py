"""Synthetic clips for the decision-engine workloads (ctypes over libcelv2synth.so).

See include/celv2synth.h for the generator; `clip_frames` returns the padded 10-bit luma and the
8-bit luma the decisions read, as numpy arrays.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None


class SynthParams(ctypes.Structure):
    _fields_ = [
        ("kind", ctypes.c_int32),
        ("seed", ctypes.c_uint64),
        ("width", ctypes.c_int32),
        ("height", ctypes.c_int32),
        ("bit_depth", ctypes.c_int32),
        ("base_period", ctypes.c_int32),
        ("nobjects", ctypes.c_int32),
        ("obj_min", ctypes.c_int32),
        ("obj_max", ctypes.c_int32),
        ("vmax8", ctypes.c_int32),
        ("pan_x8", ctypes.c_int32),
        ("pan_y8", ctypes.c_int32),
        ("rotate", ctypes.c_int32),
        ("ellipses", ctypes.c_int32),
        ("sigma_x16", ctypes.c_int32),
        ("fade_first", ctypes.c_int32),
        ("cut_every", ctypes.c_int32),
        ("hold_odd", ctypes.c_int32),
    ]


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "lib", "libcelv2synth.so")
        if not os.path.exists(path):
            raise RuntimeError(f"{path} is not built: run __graft_entry__.build()")
        _LIB = ctypes.CDLL(path)
        _LIB.celv2_synth_config.argtypes = [ctypes.c_int, ctypes.c_uint64, ctypes.c_int, ctypes.c_int, ctypes.POINTER(SynthParams)]
        _LIB.celv2_synth_frame.argtypes = [ctypes.POINTER(SynthParams), ctypes.c_int] + [ctypes.c_void_p] * 4 + [ctypes.c_int]
    return _LIB


def params(workload, seed, width=0, height=0, **overrides):
    p = SynthParams()
    if lib().celv2_synth_config(workload, seed, width, height, ctypes.byref(p)) != 0:
        raise ValueError(f"unknown workload {workload}")
    for k, v in overrides.items():
        setattr(p, k, v)
    return p


def padded(width, height):
    return (width + 7) & ~7, (height + 7) & ~7


def frame(p, t, threads=0, with16=True):
    """(y16, y8) of frame t: y16 the padded luma at p.bit_depth (None if with16 is False), y8 rounded."""
    pw, ph = padded(p.width, p.height)
    y8 = np.empty((ph, pw), dtype=np.uint8)
    y16 = np.empty((ph, pw), dtype=np.uint16) if with16 else None
    rc = lib().celv2_synth_frame(
        ctypes.byref(p), t, y16.ctypes.data if with16 else None, None, None, y8.ctypes.data, threads
    )
    if rc != 0:
        raise RuntimeError("synth frame failed")
    return y16, y8


def clip_frames(p, n, threads=0, with16=True):
    out16, out8 = [], []
    for t in range(n):
        a, b = frame(p, t, threads, with16)
        out16.append(a)
        out8.append(b)
    return out16, out8
