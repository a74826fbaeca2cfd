"""A compact decision log: each coded frame's schedule entry and, per superblock, the chosen partition
tree, its leaves, the decision's cost and the reference limit it was decided under.

The engine's exported record (`celv2_sb`, 41,256 bytes per superblock) carries all 425 candidate
blocks for the RD pass; a log for training and analysis keeps what was chosen: at most 1,889 bytes
per superblock (12 + 85 tree entries + 128 leaves of 14 bytes), far less where the tree is shallow.
Binary, little-endian, in the spirit of CelV2's CELV2_TX_DUMP records (src/core/rd_pass.cpp:172-214):

    file header  "C2DL", version u32 = 1, width, height, sb_size, sb_cols, sb_rows (i32 each), 4 bytes 0
    per frame    display i32, type u8 (K P A M B), qindex u8, nrefs u8, 0 u8, lambda_q8 i64,
                 refs i16 x 7 (display indices, -1 past nrefs), 2 bytes 0, payload bytes u32,
                 then per superblock in raster order:
                   ntree u8, nblocks u8, limit u8, 0 u8, cost i64,
                   tree u8 x ntree (Part: 0 None, 1 Horz, 2 Vert, 3 Split; preorder),
                   leaves x nblocks (14 bytes each: the celv2_block layout of include/celv2hip.h)
"""
import struct

import numpy as np

from .records import BLOCK_DTYPE, SB_DTYPE

MAGIC = b"C2DL"
VERSION = 1
_FILE = struct.Struct("<4sI6i")
_FRAME = struct.Struct("<iBBBBq7h2xI")
_SB = struct.Struct("<BBBBq")


class Writer:
    def __init__(self, path, width, height, sb_size=64):
        self.sb_cols, self.sb_rows = (width + sb_size - 1) // sb_size, (height + sb_size - 1) // sb_size
        self.f = open(path, "wb")
        self.f.write(_FILE.pack(MAGIC, VERSION, width, height, sb_size, self.sb_cols, self.sb_rows, 0))

    def frame(self, display, type_, qindex, lambda_q8, refs, records, costs=None, limits=None):
        """records: the frame's exported map (SB_DTYPE); costs / limits per superblock (None: 0 /
        every reference)."""
        n = self.sb_cols * self.sb_rows
        if records.dtype != SB_DTYPE or records.shape != (n,):
            raise ValueError("records must be the frame's SB_DTYPE map")
        costs = np.zeros(n, np.int64) if costs is None else np.asarray(costs, np.int64)
        limits = np.full(n, len(refs), np.uint8) if limits is None else np.asarray(limits, np.uint8)
        parts = []
        for i in range(n):
            nt, nb = int(records["ntree"][i]), int(records["nblocks"][i])
            parts.append(_SB.pack(nt, nb, int(limits[i]), 0, int(costs[i])))
            parts.append(records["tree"][i][:nt].tobytes())
            parts.append(records["blocks"][i][:nb].tobytes())
        payload = b"".join(parts)
        r = list(refs) + [-1] * (7 - len(refs))
        self.f.write(_FRAME.pack(display, ord(type_), qindex, len(refs), 0, lambda_q8, *r, len(payload)))
        self.f.write(payload)

    def close(self):
        if self.f:
            self.f.close()
            self.f = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


def read(path):
    """(header dict, list of frame dicts); each frame's "sbs" is a list of (tree bytes, leaves array of
    BLOCK_DTYPE, cost, limit)."""
    data = open(path, "rb").read()
    magic, version, width, height, sb_size, sb_cols, sb_rows, _ = _FILE.unpack_from(data, 0)
    if magic != MAGIC or version != VERSION:
        raise ValueError("not a version-1 decision log")
    head = dict(width=width, height=height, sb_size=sb_size, sb_cols=sb_cols, sb_rows=sb_rows)
    frames, off = [], _FILE.size
    while off < len(data):
        display, t, q, nrefs, _, lam, *rest = _FRAME.unpack_from(data, off)
        refs, nbytes = rest[:7], rest[7]
        off += _FRAME.size
        end = off + nbytes
        sbs = []
        for _ in range(sb_cols * sb_rows):
            nt, nb, limit, _, cost = _SB.unpack_from(data, off)
            off += _SB.size
            tree = data[off:off + nt]
            off += nt
            leaves = np.frombuffer(data, BLOCK_DTYPE, count=nb, offset=off).copy()
            off += nb * BLOCK_DTYPE.itemsize
            sbs.append((tree, leaves, cost, limit))
        if off != end:
            raise ValueError("decision log frame size mismatch")
        frames.append(dict(display=display, type=chr(t), qindex=q, lambda_q8=lam, refs=list(refs[:nrefs]), sbs=sbs))
    return head, frames
