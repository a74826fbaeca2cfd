"""The RD pass's forward transforms on the GPU (celv2hip_fwd_tx, include/celv2hip.h).

`fwd_tx(jobs)` runs a list of `celv2_tx_job` records (TX_JOB) whose residuals and outputs are in
device memory; `TxBatch` packs host residual blocks into one device buffer, runs them and reads the
coefficients back, for tests and measurement. A job is one call of the transform step of CelV2's
tx_cost (src/core/rd_pass.cpp:231-300): fwd_dct's coefficients (src/core/transform.cpp:170-214), in
AVM's layout with only min(w, 32) x min(h, 32) kept, and dropped_energy for a 64-point side.
"""
import ctypes

import numpy as np

from .engine import DeviceBuffer, EngineError, last_error, lib

DCT, DST, FLIPDST, IDT = 0, 1, 2, 3
MAX_POINTS = {DCT: 64, DST: 16, FLIPDST: 16, IDT: 32}

TX_JOB = np.dtype(
    {
        "names": ["res", "coeff", "dropped", "scale", "stride", "w_log2", "h_log2", "vkind", "hkind"],
        "formats": ["<u8", "<u8", "<u8", "<f8", "<i4", "u1", "u1", "u1", "u1"],
        "offsets": [0, 8, 16, 24, 32, 36, 37, 38, 39],
        "itemsize": 40,
    }
)


def log2_side(n):
    if n not in (4, 8, 16, 32, 64):
        raise ValueError(f"no transform of {n} points")
    return n.bit_length() - 1


def kept_shape(w, h):
    return min(h, 32), min(w, 32)


def fwd_tx(jobs, device=0):
    """Transforms every TX_JOB record on `device` (returns when done) and returns the kernel's own
    milliseconds; ValueError for a job the library refuses (nothing runs then)."""
    jobs = np.ascontiguousarray(jobs, dtype=TX_JOB)
    ms = ctypes.c_float(0)
    rc = lib().celv2hip_fwd_tx(device, jobs.ctypes.data if len(jobs) else None, len(jobs), ctypes.byref(ms))
    if rc == 1:
        raise ValueError(last_error())
    if rc != 0:
        raise EngineError(last_error())
    return ms.value


class TxBatch:
    """Residual blocks (int16 arrays, h x w) packed into one device buffer, `pad` extra samples per
    row (a stride wider than the block), with room for their coefficients and dropped energies."""

    def __init__(self, blocks, kinds=None, scales=None, pad=0, device=0, want_dropped=True):
        self.device = device
        self.shapes = [b.shape for b in blocks]
        n = len(blocks)
        kinds = kinds or [(DCT, DCT)] * n
        scales = scales if scales is not None else [1.0] * n
        offs, parts, pos = [], [], 0
        for b in blocks:
            h, w = b.shape
            plane = np.zeros((h, w + pad), np.int16)
            plane[:, :w] = b
            offs.append(pos)
            parts.append(plane.ravel())
            pos += plane.size
        self.res = DeviceBuffer(np.concatenate(parts) if parts else np.zeros(1, np.int16), device)
        sizes = [int(np.prod(kept_shape(w, h))) for h, w in self.shapes]
        self.coff = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
        self.coeff = DeviceBuffer(np.zeros(max(1, int(self.coff[-1])), np.int32), device)
        self.dropped = DeviceBuffer(np.full(max(1, n), -1, np.int64), device)
        self.jobs = np.zeros(n, TX_JOB)
        for i, ((h, w), o) in enumerate(zip(self.shapes, offs)):
            j = self.jobs[i]
            j["res"] = self.res.ptr + 2 * o
            j["coeff"] = self.coeff.ptr + 4 * int(self.coff[i])
            j["dropped"] = self.dropped.ptr + 8 * i if want_dropped else 0
            j["scale"] = scales[i]
            j["stride"] = w + pad
            j["w_log2"], j["h_log2"] = log2_side(w), log2_side(h)
            j["vkind"], j["hkind"] = kinds[i]

    def run(self):
        return fwd_tx(self.jobs, self.device)

    def results(self):
        """([coefficients of job i, min(h, 32) x min(w, 32) int32], dropped energies int64)."""
        flat = self.coeff.read(np.int32)
        out = [flat[self.coff[i]:self.coff[i + 1]].reshape(kept_shape(w, h)) for i, (h, w) in enumerate(self.shapes)]
        return out, self.dropped.read(np.int64, len(self.shapes))

    def free(self):
        for b in (self.res, self.coeff, self.dropped):
            b.free()


def superblock_blocks(sb=64):
    """The 425 candidate blocks of a superblock (x, y, w, h): every square of 64, 32, 16 and 8 samples,
    then its two horizontal and two vertical halves (the engine's candidate blocks, CelV2's
    src/gpu/sb_blocks.h)."""
    out = []
    s = sb
    while s >= 8:
        for y in range(0, sb, s):
            for x in range(0, sb, s):
                out += [(x, y, s, s), (x, y, s, s // 2), (x, y + s // 2, s, s // 2), (x, y, s // 2, s), (x + s // 2, y, s // 2, s)]
        s //= 2
    return np.array(out, np.int64)


def frame_jobs(res_ptr, stride, sb_cols, sb_rows, coeff_ptr, dropped_ptr=0, scale=1.0, kinds=(DCT, DCT)):
    """TX_JOB records transforming every candidate block of every superblock of a residual plane
    (int16 at res_ptr, `stride` samples a row, sb_rows * 64 rows), block by block in superblock raster
    order; coefficients packed from coeff_ptr, dropped energies (8 bytes each) from dropped_ptr (0:
    none). Returns (jobs, coefficient offsets in int32 units, one more than the jobs)."""
    geo = superblock_blocks()
    nb = len(geo)
    sx, sy = np.meshgrid(np.arange(sb_cols), np.arange(sb_rows))
    ox = np.repeat(sx.ravel() * 64, nb) + np.tile(geo[:, 0], sb_cols * sb_rows)
    oy = np.repeat(sy.ravel() * 64, nb) + np.tile(geo[:, 1], sb_cols * sb_rows)
    w = np.tile(geo[:, 2], sb_cols * sb_rows)
    h = np.tile(geo[:, 3], sb_cols * sb_rows)
    kept = np.minimum(w, 32) * np.minimum(h, 32)
    off = np.concatenate([[0], np.cumsum(kept)])
    jobs = np.zeros(len(w), TX_JOB)
    jobs["res"] = res_ptr + 2 * (oy * stride + ox)
    jobs["coeff"] = coeff_ptr + 4 * off[:-1]
    jobs["dropped"] = dropped_ptr + 8 * np.arange(len(w)) if dropped_ptr else 0
    jobs["scale"] = scale
    jobs["stride"] = stride
    jobs["w_log2"] = np.log2(w).astype(np.uint8)
    jobs["h_log2"] = np.log2(h).astype(np.uint8)
    jobs["vkind"], jobs["hkind"] = kinds
    return jobs, off
