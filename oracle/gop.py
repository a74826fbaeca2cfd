"""ORACLE — test infrastructure, not product code: the frame loop's coding schedule in plain Python.

Used only by tests/ to check celv2_gop_plan (celv2_amd/csrc/host/gop.cpp). Each step cites the
reference it restates:
  - the loop over keyframes, lone frames and mini-GOPs: src/chunk_coder.cpp:478-616 (keyframes 479-516, the mini-GOP's length 517-519, lone frames 544-550);
  - the keyframe delta of a short shot: src/chunk_coder.cpp:486-501;
  - mids_of (the pyramid): src/chunk_coder.cpp:209-224;
  - code_minigop's order and deltas: src/chunk_coder.cpp:323-437;
  - the slots a frame refreshes: src/state/codec_state.cpp:476-498 (8 slots, seq_header.cpp:87);
  - the references handed to the engine: src/pipeline.cpp:298-309 over CodecState's ranked hints
    (codec_state.h:84-88: ranked by display distance). AV2's tie rule is not in the reference tree;
    this restatement ranks the earlier of two equally distant frames first.
Frames are dicts with the fields of celv2_gop_frame (include/celv2/gop.h).

Pinning (tests/test_gop.py): hand-derived schedules of the C1 sequence (32 frames, 7 references,
16-frame mini-GOPs), the reference's own test structures (tests/test_random_decisions.cpp:205-236,
nearest past reference first; tests/test_cli.cpp:1044-1046, anchors at 8 and 16 with bframes 7)
and invariants over a sweep (every frame shown once in display order, references coded before,
distinct and nearest first).
"""
import math

SLOTS = 8


def defaults(refs, qindex, width, height, long_gop=False):
    """src/main.cpp:739-740 and LoopConfig's initialisers (src/chunk_coder.h:86-92)."""
    if refs < 2:
        bframes = 0
    elif refs >= 5 and qindex >= 100:
        bframes = 15 if long_gop else 7
    else:
        bframes = 3
    return dict(refs=refs, bframes=bframes, pyramid=1, parallel_b=int(width * height <= 2 * 3840 * 2160),
                bq=16, aq=-16, mq=0, kq=0, level_step=8, level_step16=-1, qindex=qindex)


def _round_half_away(x):
    return int(math.floor(abs(x) + 0.5)) * (1 if x >= 0 else -1)


def plan(cfg, frames, key=None):
    key = [i == 0 for i in range(frames)] if key is None else [bool(k) for k in key]
    if frames and not key[0]:
        raise ValueError("an inter frame needs a preceding keyframe")
    refs_n = cfg["refs"]
    hints = [-1] * SLOTS
    state = {"since_key": 0}
    out = []

    def clampq(d):
        return max(1, min(255, cfg["qindex"] + d))

    def references(display):
        distinct = sorted(set(h for h in hints if h >= 0), key=lambda h: (abs(display - h), h))
        return distinct[:refs_n]

    def code(display, typ, is_key, shown, reference, keep_shown, level, delta):
        f = dict(display=display, type=typ, coded=1, shown=int(shown), reference=int(reference or is_key),
                 keep_shown=int(keep_shown), level=level, qindex_delta=delta, qindex=clampq(delta),
                 refs=[] if is_key else references(display))
        all_slots, turns = (1 << SLOTS) - 1, (1 << refs_n) - 1
        if is_key:
            state["since_key"] = 0
            flags = all_slots
        elif reference:
            state["since_key"] += 1
            slot = state["since_key"] % refs_n
            for i in range(refs_n):
                if hints[i] < hints[slot]:
                    slot = i
            flags = (1 << slot) | ((all_slots & ~turns) if shown else 0)
        else:
            flags = (all_slots & ~turns) if (shown and keep_shown) else 0
        f["refresh"] = flags
        for i in range(SLOTS):
            if flags >> i & 1:
                hints[i] = display
        out.append(f)

    def show_existing(display, typ, level, delta):
        out.append(dict(display=display, type=typ, coded=0, shown=1, reference=0, keep_shown=0, level=level,
                        qindex_delta=delta, qindex=clampq(delta), refresh=0, refs=[]))

    def mids_of(m):
        mids = []

        def split(lo, hi, level):
            if hi - lo - 1 < 3 or refs_n < 3 or len(mids) + 2 >= refs_n:
                return
            mid = (lo + hi) // 2
            mids.append((mid, level))
            split(lo, mid, level + 1)
            split(mid, hi, level + 1)

        if cfg["pyramid"]:
            split(-1, m - 1, 1)
        return mids

    def minigop(start, m):
        step = cfg["level_step16"] if m > 8 and cfg["level_step16"] >= 0 else cfg["level_step"]
        mids = mids_of(m)
        code(start + m - 1, "A", False, False, True, True, 0, cfg["aq"])
        deltas = {}
        for pos, level in mids:
            deltas[pos] = (level, cfg["mq"] + (level - 1) * step)
            code(start + pos, "M", False, False, True, True, level, deltas[pos][1])
        leaves = [i for i in range(m - 1) if i not in deltas]
        for i in range(m - 1):
            if i in deltas:
                show_existing(start + i, "M", *deltas[i])
            else:
                keep = (not cfg["parallel_b"]) or i == leaves[-1]
                code(start + i, "B", False, True, False, keep, 0, cfg["bq"])
        show_existing(start + m - 1, "A", 0, cfg["aq"])

    i = 0
    while i < frames:
        if key[i]:
            ahead = 0
            while ahead < 8 and ahead < cfg["bframes"] + 1 and i + ahead + 1 < frames and not key[i + ahead + 1]:
                ahead += 1
            kq = cfg["kq"]
            if ahead < 8 and kq < 0:
                kq = _round_half_away(kq * (ahead + 1) / 9.0)
            code(i, "K", True, True, True, True, 0, kq)
            i += 1
            continue
        m = 0
        while m < cfg["bframes"] + 1 and i + m < frames and not key[i + m]:
            m += 1
        if m <= 1:
            code(i, "P", False, True, True, True, 0, 0)
        else:
            minigop(i, m)
        i += m
    return out
