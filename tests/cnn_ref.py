"""CPU reference of the learned pruner's network (celv2_amd/csrc/kernels/cnn.hip), test
infrastructure only: the same layers and parameter layout, float64 arithmetic, and (with
rounding=True) BF16 rounding of every GEMM operand exactly where the GPU rounds - so the GPU's
outputs and gradients differ from it only by FP32 accumulation order."""
import numpy as np

F, HID, REF_OUT, GROUP_OUT, SEARCH_OUT = 14, 32, 6, 3, 1
OUT = REF_OUT + GROUP_OUT + SEARCH_OUT
W1 = 0
B1 = W1 + 9 * F * HID
W2 = B1 + HID
B2 = W2 + 9 * HID * HID
W3 = B2 + HID
B3 = W3 + HID * OUT
NPARAMS = B3 + OUT


def bf16(x):
    """Round to the nearest BF16 (ties to even), as float32."""
    u = np.ascontiguousarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return u.astype(np.uint32).view(np.float32)


def im2col(x):
    """[b, h, w, c] -> [b h w, 9 c], tap-major (tap = (dy + 1) * 3 + dx + 1), zero outside the grid."""
    b, h, w, c = x.shape
    pad = np.pad(x, ((0, 0), (1, 1), (1, 1), (0, 0)))
    cols = [pad[:, 1 + dy:1 + dy + h, 1 + dx:1 + dx + w, :] for dy in (-1, 0, 1) for dx in (-1, 0, 1)]
    return np.concatenate(cols, axis=-1).reshape(b * h * w, 9 * c)


def col2im(dcol, b, h, w, c):
    """The transpose of im2col: [b h w, 9 c] -> [b, h, w, c]."""
    d = dcol.reshape(b, h, w, 9, c)
    out = np.zeros((b, h + 2, w + 2, c))
    for tap in range(9):
        dy, dx = tap // 3 - 1, tap % 3 - 1
        out[:, 1 + dy:1 + dy + h, 1 + dx:1 + dx + w, :] += d[:, :, :, tap, :]
    return out[:, 1:-1, 1:-1, :]


def unpack(p, rnd):
    r = bf16 if rnd else (lambda v: np.asarray(v, np.float64))
    return (r(p[W1:B1]).reshape(9 * F, HID).astype(np.float64), p[B1:W2].astype(np.float64),
            r(p[W2:B2]).reshape(9 * HID, HID).astype(np.float64), p[B2:W3].astype(np.float64),
            r(p[W3:B3]).reshape(HID, OUT).astype(np.float64), p[B3:].astype(np.float64))


def forward(p, x, rounding=True):
    """Logits [b h w, 6] and the activations backward() needs."""
    r = bf16 if rounding else (lambda v: v)
    w1, b1, w2, b2, w3, b3 = unpack(p, rounding)
    b, h, w, _ = x.shape
    col1 = np.asarray(r(im2col(x)), np.float64)
    z1 = np.maximum(col1 @ w1 + b1, 0.0)
    col2 = np.asarray(r(im2col(z1.reshape(b, h, w, HID))), np.float64)
    z2 = np.maximum(col2 @ w2 + b2, 0.0)
    z2r = np.asarray(r(z2), np.float64)
    z3 = z2r @ w3 + b3
    return z3, (col1, z1, col2, z2, z2r, (b, h, w))


def labels(kstar):
    """Targets per position: reference outputs j: (label & 7) > j + 1; group outputs g: bit 3 + g;
    the search output: bit 6."""
    lab = np.asarray(kstar).reshape(-1, 1).astype(np.int64)
    refs = (lab & 7) > np.arange(REF_OUT).reshape(1, -1) + 1
    groups = (lab >> (3 + np.arange(GROUP_OUT).reshape(1, -1))) & 1
    search = (lab >> 6) & 1
    return np.concatenate([refs, groups.astype(bool), search.astype(bool)], axis=1).astype(np.float64)


def loss_and_grads(p, x, kstar, rounding=True):
    r = bf16 if rounding else (lambda v: v)
    w1, b1, w2, b2, w3, b3 = unpack(p, rounding)
    z3, (col1, z1, col2, z2, z2r, (b, h, w)) = forward(p, x, rounding)
    y = labels(kstar)
    m = z3.size
    loss = float(np.mean(np.maximum(z3, 0) - z3 * y + np.log1p(np.exp(-np.abs(z3)))))
    dz3 = (1.0 / (1.0 + np.exp(-z3)) - y) / m
    dz3r = np.asarray(r(dz3), np.float64)
    g = np.zeros(NPARAMS)
    g[W3:B3] = (z2r.T @ dz3r).ravel()
    g[B3:] = dz3.sum(0)
    dz2 = (dz3r @ w3.T) * (z2 > 0)
    dz2r = np.asarray(r(dz2), np.float64)
    g[W2:B2] = (col2.T @ dz2r).ravel()
    g[B2:W3] = dz2.sum(0)
    dz1 = col2im(dz2r @ w2.T, b, h, w, HID).reshape(-1, HID) * (z1 > 0)
    dz1r = np.asarray(r(dz1), np.float64)
    g[W1:B1] = (col1.T @ dz1r).ravel()
    g[B1:W2] = dz1.sum(0)
    return loss, g


def adamw(p, g, m1, m2, step, lr, wd, b1=0.9, b2=0.999, eps=1e-8):
    m1 = b1 * m1 + (1 - b1) * g
    m2 = b2 * m2 + (1 - b2) * g * g
    upd = (m1 / (1 - b1 ** step)) / (np.sqrt(m2 / (1 - b2 ** step)) + eps) + wd * p
    return p - lr * upd, m1, m2


def limits(logits, nrefs, threshold, band):
    """The engine's per-superblock reference limit from the network's outputs."""
    s = 1.0 / (1.0 + np.exp(-np.asarray(logits, np.float64)))[:, :REF_OUT]
    k = np.ones(s.shape[0], np.int64)
    open_ = np.ones(s.shape[0], bool)
    for j in range(REF_OUT):
        hit = open_ & (s[:, j] >= threshold)
        k[hit] = j + 2
        open_ &= s[:, j] >= threshold
    unsure = (np.abs(s - threshold) < band).any(1)
    k[unsure] = 7
    return np.minimum(k, nrefs), unsure


def search_modes(logits, sthreshold, unsure):
    """The engine's motion-search modes from the search output: 0 (full) where sigma(output 9) >=
    sthreshold or the superblock fell back, else 3 (narrow at levels 1 and 0)."""
    s = 1.0 / (1.0 + np.exp(-np.asarray(logits, np.float64)))[:, REF_OUT + GROUP_OUT]
    return np.where(unsure | (s >= sthreshold), 0, 3)


def groups(logits, gthreshold, unsure):
    """The engine's intra mode groups from the group outputs: bit 0 always, bit g when
    sigma(output 5 + g) >= gthreshold, every group where the superblock fell back."""
    s = 1.0 / (1.0 + np.exp(-np.asarray(logits, np.float64)))[:, REF_OUT:]
    m = np.ones(s.shape[0], np.int64)
    for g in range(GROUP_OUT):
        m |= (s[:, g] >= gthreshold).astype(np.int64) << (g + 1)
    m[unsure] = 15
    return m
