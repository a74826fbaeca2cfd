"""numpy views of the flat decision-map records of include/celv2hip.h.

`SB_DTYPE` matches `celv2_sb` byte for byte (BlockDecision / BlockCandidates / SbDecision of
CelV2's src/gpu/decision_map.h:22-78, flattened); `compare_maps` reports the first difference
between two exported maps, field by field, the way the reference's tests compare DecisionMaps
(tests/test_gpu_intra.cpp:144-165: trees, leaves and every candidate).
"""
import numpy as np

SB_CAP = 256
SB_BLOCKS = 425
MAX_REFS = 7
NUM_MODES = 13

BLOCK_DTYPE = np.dtype(
    {
        "names": ["w4", "h4", "inter", "intra", "mv_row", "mv_col", "ref", "ref2", "mv2_row", "mv2_col"],
        "formats": ["u1", "u1", "u1", "u1", "<i2", "<i2", "u1", "u1", "<i2", "<i2"],
        "offsets": [0, 1, 2, 3, 4, 6, 8, 9, 10, 12],
        "itemsize": 14,
    }
)

CANDS_DTYPE = np.dtype(
    {
        "names": ["inter_cost", "intra_cost", "inter", "intra", "ref_cost", "ref_mv_col", "ref_mv_row"],
        "formats": ["<i8", "<i8", BLOCK_DTYPE, "u1", ("<u4", (MAX_REFS,)), ("<i2", (MAX_REFS,)), ("<i2", (MAX_REFS,))],
        "offsets": [0, 8, 16, 30, 32, 60, 74],
        "itemsize": 88,
    }
)

SB_DTYPE = np.dtype(
    {
        "names": ["ntree", "nblocks", "ncands", "tree", "blocks", "cands"],
        "formats": ["<i4", "<i4", "<i4", ("u1", (SB_CAP,)), (BLOCK_DTYPE, (SB_CAP,)), (CANDS_DTYPE, (SB_BLOCKS,))],
        "offsets": [0, 4, 8, 16, 272, 3856],
        "itemsize": 41256,
    }
)


def empty_map(n_sbs):
    return np.zeros(n_sbs, dtype=SB_DTYPE)


def _block_tuple(b):
    return tuple(int(b[f]) for f in BLOCK_DTYPE.names)


def sb_canonical(sb):
    """The meaningful content of one superblock record (entries past the counts ignored)."""
    nt, nb, nc = int(sb["ntree"]), int(sb["nblocks"]), int(sb["ncands"])
    tree = bytes(sb["tree"][:nt])
    blocks = sb["blocks"][:nb].tobytes()
    cands = sb["cands"][:nc].tobytes()
    return nt, nb, nc, tree, blocks, cands


def compare_maps(a, b, sb_cols=None):
    """None when equal, else a text naming the first differing superblock and field, with a count of
    differing candidates per field over the whole map."""
    if a.shape != b.shape:
        return f"sizes differ: {a.shape} vs {b.shape}"
    first = None
    per_field = {}
    for i in range(a.shape[0]):
        x, y = a[i], b[i]
        where = f"sb {i}" + (f" (col {i % sb_cols}, row {i // sb_cols})" if sb_cols else "")
        msg = None
        for f in ("ntree", "nblocks", "ncands"):
            if int(x[f]) != int(y[f]) and msg is None:
                msg = f"{where}: {f} {int(x[f])} vs {int(y[f])}"
        nt, nb, nc = int(x["ntree"]), int(x["nblocks"]), int(x["ncands"])
        if msg is None and not np.array_equal(x["tree"][:nt], y["tree"][:nt]):
            msg = f"{where}: tree {list(x['tree'][:nt])} vs {list(y['tree'][:nt])}"
        if msg is None:
            for k in range(nb):
                if x["blocks"][k].tobytes() != y["blocks"][k].tobytes():
                    msg = f"{where}: leaf {k} {_block_tuple(x['blocks'][k])} vs {_block_tuple(y['blocks'][k])}"
                    break
        nc = min(int(x["ncands"]), int(y["ncands"]))
        if nc and x["cands"][:nc].tobytes() != y["cands"][:nc].tobytes():
            for k in range(nc):
                p, q = x["cands"][k], y["cands"][k]
                bad = [f for f in CANDS_DTYPE.names if np.asarray(p[f]).tobytes() != np.asarray(q[f]).tobytes()]
                for f in bad:
                    per_field[f] = per_field.get(f, 0) + 1
                if bad and (msg is None or "candidate" not in msg):
                    msg = f"{where}: candidate block {k}: " + "; ".join(f"{f} {p[f]} vs {q[f]}" for f in bad)
        if msg and first is None:
            first = msg
    if first is None:
        return None
    return first + (f" | differing candidates per field: {per_field}" if per_field else "")


def leaf_rects(sb, ox=0, oy=0, size=64):
    """Leaf rectangles (x, y, w, h) from the preorder tree, as tests/test_gpu_me.cpp:87-97 walks it."""
    tree = list(sb["tree"][: int(sb["ntree"])])
    out, pos = [], [0]

    def walk(x, y, w, h):
        p = tree[pos[0]]
        pos[0] += 1
        if p == 0:
            out.append((x, y, w, h))
        elif p == 1:
            walk(x, y, w, h // 2)
            walk(x, y + h // 2, w, h // 2)
        elif p == 2:
            walk(x, y, w // 2, h)
            walk(x + w // 2, y, w // 2, h)
        else:
            for q in range(4):
                walk(x + (q & 1) * w // 2, y + (q >> 1) * h // 2, w // 2, h // 2)

    walk(ox, oy, size, size)
    return out
