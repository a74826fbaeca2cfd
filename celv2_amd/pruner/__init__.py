"""The learned pruner's trained parameters and operating point (weights.f32, pruner.json), written by
celv2_amd.pruner.train."""
import json
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def load(directory=HERE):
    """(parameters float32, config dict) of a trained pruner."""
    w = np.fromfile(os.path.join(directory, "weights.f32"), dtype="<f4")
    with open(os.path.join(directory, "pruner.json")) as f:
        return w, json.load(f)


def operating_point(cfg, qindex):
    """(threshold, band, group threshold, search threshold) chosen for the trained qindex nearest to
    `qindex`, or None where no threshold keeps every held-out clip within the budget short of keeping
    all references (run plain there). A group threshold of 0 scores every intra mode group, a search
    threshold of 0 keeps every motion search full."""
    ops = cfg["operating_points"]
    q = min(ops, key=lambda k: abs(int(k) - qindex))
    if not ops[q]:
        return None
    return (ops[q]["threshold"], ops[q]["band"], ops[q].get("gthreshold", 0.0), ops[q].get("sthreshold", 0.0))
