"""The frame loop's coding schedule (ctypes over celv2_gop_plan in libcelv2hip.so, include/celv2/gop.h)
and a driver that runs the HIP engine over it.

`plan()` gives the stream's entries in order: keyframes, lone frames and mini-GOPs (hidden anchor,
hidden pyramid frames, shown B frames, show-existing entries), each coded frame's qindex and ranked
references (ChunkCoder::code, src/chunk_coder.cpp:155-616; the slots of CodecState::begin_frame,
src/state/codec_state.cpp:476-498). `drive()` launches the engine on every coded frame the way
Pipeline::code does (src/pipeline.cpp:125-177): the frame's references with their display indices as
keys, lambda_q8 of the frame's qindex, and the next coded frame prepared behind the launch.
"""
import ctypes
from dataclasses import dataclass, field
from typing import List

import numpy as np

from . import engine

I32 = ctypes.c_int32
MAX_REFS = 7


class _Config(ctypes.Structure):
    _fields_ = [(n, I32) for n in ("refs", "bframes", "pyramid", "parallel_b", "bq", "aq", "mq", "kq",
                                   "level_step", "level_step16", "qindex")]


class _Frame(ctypes.Structure):
    _fields_ = [(n, I32) for n in ("display", "type", "coded", "shown", "reference", "keep_shown", "level",
                                   "qindex_delta", "qindex", "refresh", "nrefs")] + [("refs", I32 * MAX_REFS)]


_BOUND = False


def _lib():
    global _BOUND
    L = engine.lib()
    if not _BOUND:
        L.celv2_gop_defaults.argtypes = [ctypes.POINTER(_Config), I32, I32, I32, I32, I32]
        L.celv2_gop_defaults.restype = None
        L.celv2_gop_plan.argtypes = [ctypes.POINTER(_Config), I32, ctypes.c_void_p, ctypes.POINTER(_Frame), I32,
                                     ctypes.POINTER(I32)]
        L.celv2_gop_last_error.restype = ctypes.c_char_p
        _BOUND = True
    return L


@dataclass
class Entry:
    display: int
    type: str  # K keyframe, P lone frame, A anchor, M pyramid frame, B shown B frame
    coded: bool  # False: the show-existing entry of a hidden frame
    shown: bool
    reference: bool
    keep_shown: bool
    level: int
    qindex_delta: int
    qindex: int
    refresh: int
    refs: List[int] = field(default_factory=list)

    def as_dict(self):
        return dict(display=self.display, type=self.type, coded=int(self.coded), shown=int(self.shown),
                    reference=int(self.reference), keep_shown=int(self.keep_shown), level=self.level,
                    qindex_delta=self.qindex_delta, qindex=self.qindex, refresh=self.refresh, refs=list(self.refs))


def defaults(refs, qindex, width, height, long_gop=False):
    """celv2enc's derived settings for a sequence (src/main.cpp:739-740) as a dict."""
    c = _Config()
    _lib().celv2_gop_defaults(ctypes.byref(c), refs, qindex, width, height, 1 if long_gop else 0)
    return {n: getattr(c, n) for n, _ in _Config._fields_}


def plan(config, frames, key=None):
    """The stream's entries for `frames` frames; key: per-frame keyframe flags (None: frame 0 only)."""
    c = _Config(**{n: int(config[n]) for n, _ in _Config._fields_})
    keys = None
    if key is not None:
        keys = np.ascontiguousarray(np.asarray(key, dtype=np.uint8))
        if keys.size != frames:
            raise ValueError("one key flag per frame")
    cap = 2 * frames + 1
    out = (_Frame * cap)()
    n = I32()
    rc = _lib().celv2_gop_plan(ctypes.byref(c), frames, keys.ctypes.data if keys is not None else None, out, cap,
                               ctypes.byref(n))
    if rc != 0:
        raise ValueError(_lib().celv2_gop_last_error().decode())
    return [Entry(f.display, chr(f.type), bool(f.coded), bool(f.shown), bool(f.reference), bool(f.keep_shown),
                  f.level, f.qindex_delta, f.qindex, f.refresh, list(f.refs[:f.nrefs])) for f in out[:n.value]]


def coded(entries):
    """The coded frames of a schedule, in coding order."""
    return [e for e in entries if e.coded]


def drive(eng, frames8, entries, lam_of=engine.lambda_q8, collect=True):
    """Runs a RawEngine over the coded frames of `entries`; frames8[d] is frame d's padded 8-bit luma
    (the references are the source frames, as the reference's engine benchmark uses them). Returns
    {display: decision map} when collect, else runs to the last launch and syncs."""
    todo = coded(entries)
    lumas = {}

    def luma(d):
        if d not in lumas:
            lumas[d] = eng.luma(frames8[d])
        return lumas[d]

    maps = {}
    for i, e in enumerate(todo):
        eng.launch(luma(e.display), [luma(r) for r in e.refs], lam_of(e.qindex), keys=list(e.refs),
                   src_key=e.display)
        if i + 1 < len(todo):
            eng.prepare(luma(todo[i + 1].display), todo[i + 1].display)
        if collect:
            maps[e.display] = eng.collect()
    if not collect:
        eng.sync()
    return maps
