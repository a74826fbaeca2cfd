"""The oracle's forward transforms (oracle/transform.cpp) pinned by the mathematics, before they
check the HIP kernel:

- DCT-II (fwd_dct, src/core/transform.cpp:170-214): against the double-precision DCT-II by its
  cosine sums at every size from 4 to 64 on the reference's own residual kinds and seed (constant
  255, random +-255, random in [-255, 255]; scale 0.5 + 0.07 per size), each coefficient within +-1,
  the bar of the reference's tests/test_transform.cpp:81-100; a constant block's AC coefficients 0
  and its DC sqrt(w h) c scale;
- dropped_energy (src/core/rd_pass.cpp:268-276): against the double sum of the squares a 64-point
  transform leaves out;
- DST-VII, its flip and the identity: the 4-point matrix against AV1's published sinpi constants
  (1321, 2482, 3344, 3803 in 1/4096, a sqrt(2) gain), every size orthonormal through the oracle
  itself (Parseval on random blocks), the flip as the transform of the reversed rows, the identity
  as the scaled residual. Parity with AVM's integer ADST and identity is unpinned (AVM's tables are
  absent).

Each planted mistake (ORC_TX_PLANT in oracle/transform.cpp) fails one of these checks."""
import ctypes
import os
import subprocess

import numpy as np
import pytest

import oracle_lib as ol

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SIZES = [(w, h) for w in (4, 8, 16, 32, 64) for h in (4, 8, 16, 32, 64)]  # every shape the kernel takes
DCT, DST, FLIP, IDT = 0, 1, 2, 3


def lcg(seed=12345):
    """The reference test's generator (tests/test_transform.cpp:75-79): seed * 1664525 + 1013904223,
    the value seed >> 8."""
    while True:
        seed = (seed * 1664525 + 1013904223) & 0xFFFFFFFF
        yield seed >> 8


def residuals(w, h, gen):
    """The reference's four trials: constant 255, random signs of 255, two uniform in [-255, 255]."""
    out = [np.full((h, w), 255, np.int16)]
    out.append(np.array([255 if next(gen) & 1 else -255 for _ in range(w * h)], np.int16).reshape(h, w))
    for _ in range(2):
        out.append(np.array([next(gen) % 511 - 255 for _ in range(w * h)], np.int16).reshape(h, w))
    return out


def dct_basis(n):
    """Orthonormal DCT-II rows, k = frequency: sqrt((k ? 2 : 1) / n) cos(pi (2i + 1) k / (2n))."""
    k, i = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    return np.where(k == 0, np.sqrt(1.0 / n), np.sqrt(2.0 / n)) * np.cos(np.pi * (2 * i + 1) * k / (2.0 * n))


def exact_dct(r, scale):
    h, w = r.shape
    return scale * dct_basis(h) @ r.astype(np.float64) @ dct_basis(w).T


def check_dct(lib):
    """Failures of the DCT-II checks (empty when the oracle passes)."""
    bad = []
    gen = lcg()
    for idx, (w, h) in enumerate(SIZES):
        scale = 0.5 + idx * 0.07
        for t, r in enumerate(residuals(w, h, gen)):
            got, dropped = ol.fwd_tx(r, DCT, DCT, scale, lib)
            ex = exact_dct(r, scale)
            want = np.rint(ex[: min(h, 32), : min(w, 32)])
            worst = int(np.max(np.abs(got - want)))
            if worst > 1:
                bad.append(f"{w}x{h} trial {t}: a coefficient {worst} off the double DCT")
            if t == 0:
                dc = 255 * np.sqrt(w * h) * scale
                if abs(got[0, 0] - dc) > 1 or np.count_nonzero(got.ravel()[1:]):
                    bad.append(f"{w}x{h}: constant block DC {got[0, 0]} (want {dc:.2f}) or AC not all 0")
            if w > 32 or h > 32:
                true = float(np.sum(ex**2) - np.sum(ex[: min(h, 32), : min(w, 32)] ** 2))
                kept = ex[: min(h, 32), : min(w, 32)]
                slack = float(np.sum(2 * np.abs(kept) + 1)) + 1
                if abs(dropped - true) > slack:
                    bad.append(f"{w}x{h} trial {t}: dropped {dropped} vs {true:.0f}")
            elif dropped != 0:
                bad.append(f"{w}x{h}: dropped {dropped} below 64 points")
    return bad


def check_matrix_kinds(lib):
    bad = []
    rng = np.random.default_rng(4)
    # AV1's 4-point ADST constants: sinpi(j) = round(4096 * 2 sqrt(2) / 3 * sin(j pi / 9)), the
    # matrix rows k = 0..3 over samples i = 0..3 as the published kernel arranges them.
    s1, s2, s3, s4 = 1321, 2482, 3344, 3803
    av1 = np.array([[s1, s2, s3, s4], [s3, s3, 0, -s3], [s4, -s1, -s3, s2], [s2, -s4, s3, -s1]], np.float64) / (4096 * np.sqrt(2))
    r = rng.integers(-255, 256, (4, 4)).astype(np.int16)
    got, _ = ol.fwd_tx(r, DST, DST, 1.0, lib)
    want = av1 @ r.astype(np.float64) @ av1.T
    if np.max(np.abs(got - want)) > 1.5:
        bad.append(f"4x4 DST-VII against AV1's constants: {np.max(np.abs(got - want)):.2f} off")
    for w in (4, 8, 16, 32, 64):
        for h in (4, 8, 16, 32, 64):
            for vk in (DCT, DST, FLIP, IDT):
                for hk in (DCT, DST, FLIP, IDT):
                    ok = all(n <= {DCT: 64, DST: 16, FLIP: 16, IDT: 32}[k] for k, n in ((vk, h), (hk, w)))
                    r = rng.integers(-255, 256, (h, w)).astype(np.int16)
                    if not ok:
                        try:
                            ol.fwd_tx(r, vk, hk, 1.0, lib)
                            bad.append(f"{w}x{h} kinds {vk}/{hk} accepted")
                        except ValueError:
                            pass
                        continue
                    if (vk, hk) == (DCT, DCT):
                        continue
                    got, _ = ol.fwd_tx(r, vk, hk, 1.0, lib)
                    # Orthonormal: the energy stays (within the rounding of each coefficient).
                    e_in, e_out = float(np.sum(r.astype(np.float64) ** 2)), float(np.sum(got.astype(np.float64) ** 2))
                    if max(w, h) <= 32 and abs(e_in - e_out) > np.sum(np.abs(got)) + got.size:
                        bad.append(f"{w}x{h} kinds {vk}/{hk}: energy {e_out:.0f} vs {e_in:.0f}")
                    if vk == FLIP or hk == FLIP:  # the flip is the transform of the reversed samples
                        rr = r[::-1] if vk == FLIP else r
                        rr = rr[:, ::-1] if hk == FLIP else rr
                        plain, _ = ol.fwd_tx(rr, DST if vk == FLIP else vk, DST if hk == FLIP else hk, 1.0, lib)
                        if not np.array_equal(plain, got):
                            bad.append(f"{w}x{h} kinds {vk}/{hk}: not the reversed samples' transform")
                    if vk == IDT and hk == IDT and not np.array_equal(got, r[:32, :32]):
                        bad.append(f"{w}x{h}: identity is not the residual")
    return bad


def test_dct_matches_the_double_dct_within_one():
    assert check_dct(ol.oracle()) == []


def test_adst_flip_and_identity_forms():
    assert check_matrix_kinds(ol.oracle()) == []


def test_wider_stride_reads_the_same_samples():
    lib = ol.oracle()
    rng = np.random.default_rng(9)
    for w, h in ((16, 8), (64, 32), (4, 64)):
        big = rng.integers(-1023, 1024, (h, w + 24)).astype(np.int16)
        out = np.zeros((min(h, 32), min(w, 32)), np.int32)
        d = np.zeros(1, np.int64)
        assert lib.orc_fwd_tx(big.ctypes.data, w + 24, w, h, 0, 0, 1.3, out.ctypes.data, d.ctypes.data) == 0
        want, dw = ol.fwd_tx(big[:, :w], 0, 0, 1.3, lib)
        assert np.array_equal(out, want) and d[0] == dw


PLANTS = {1: "an 8-point Lee factor from the wrong index", 2: "odd outputs summed with the wrong sign",
          3: "the DC column normalised as an AC one"}


def build_plant(k, out_dir):
    so = os.path.join(out_dir, f"libtx_plant{k}.so")
    cmd = ["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-ffp-contract=off", f"-DORC_TX_PLANT={k}", "-o", so,
           os.path.join(ROOT, "oracle", "transform.cpp")]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-2000:]
    return ol.tx_api(ctypes.CDLL(so))


@pytest.mark.parametrize("k", sorted(PLANTS))
def test_each_planted_mistake_is_caught(tmp_path, k):
    lib = build_plant(k, str(tmp_path))
    assert check_dct(lib), f"plant {k} ({PLANTS[k]}) passed the DCT checks"
