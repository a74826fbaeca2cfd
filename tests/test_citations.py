"""Every `path:line` citation of the reference in this repository points inside the cited file, and a
line that names one of the reference's functions next to a single citation of the file defining it
cites the lines that hold the definition. Runs where the reference tree is mounted (it is read here
as text, never at run time of the product)."""
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF = "/root/reference"
pytestmark = pytest.mark.skipif(not os.path.isdir(os.path.join(REF, "src", "gpu")), reason="reference tree not mounted")

CITE = re.compile(r"((?:[\w./-]+/)?[\w.-]+\.(?:h|cpp|c|slang|py|md))((?::\d+(?:-\d+)?)(?:,\s*\d+(?:-\d+)?)*)")
SKIP = {"SURVEY.md", "VERDICT.md", "ADVICE.md", "BASELINE.md"}
# Functions of the reference whose definitions a citation next to their name must cover.
SYMBOLS = [
    "subpel_tap", "subpel_px", "mv_cells", "mv_candidates", "candidate_mv_index", "repeated_candidate", "downsample_px",
    "sad8", "me_search", "satd8_q", "me_refine", "me_cell", "compound_pair_ok", "compound_bits", "block_compound_satd",
    "block_inter_satd", "edge_above", "edge_left", "dr_derivative", "dir_pred", "intra_pred", "dc_value", "hadamard4x4",
    "block_mode_satd", "sb_block_geom", "sb_block_index", "best_both", "best_mode", "decide_square", "build_pyramid_ref",
    "me_level_rows_ref", "round_to_8bit", "make_frame", "pad_frame", "sb_out_to_decision", "decision_threads",
    "write_y4m_frame_bytes", "vk_phase_layout",
]


def reference_index():
    by_name = {}
    for d, _, files in os.walk(REF):
        if "/.git" in d:
            continue
        for f in files:
            by_name.setdefault(f, []).append(os.path.relpath(os.path.join(d, f), REF))
    return by_name


def resolve(name, by_name):
    cands = by_name.get(os.path.basename(name), [])
    exact = [p for p in cands if p.endswith(name.lstrip("./"))]
    return (exact or cands or [None])[0]


def definitions(path, symbol):
    """(first, last) lines (1-based) of each definition of `symbol` in the file: where `symbol(` starts
    a definition (not a call or a declaration) to the line closing its body."""
    out = []
    with open(os.path.join(REF, path), errors="ignore") as f:
        lines = f.read().splitlines()
    for i, line in enumerate(lines, 1):
        if re.search(r"\b" + symbol + r"\s*\(", line) and re.match(r"^\s*(CELV2_HD|static|[\w:<>,*&\s]+)\s+[\w:*&]*\b" + symbol + r"\s*\(", line):
            tail = line.rstrip()
            if tail.endswith(";") and "{" not in tail:
                continue
            depth, end, opened = 0, i, False
            for j in range(i - 1, len(lines)):
                depth += lines[j].count("{") - lines[j].count("}")
                opened = opened or "{" in lines[j]
                if opened and depth <= 0:
                    end = j + 1
                    break
            out.append((i, end))
    return out


def repo_files():
    files = subprocess.run(["git", "ls-files"], cwd=ROOT, capture_output=True, text=True).stdout.split()
    return [f for f in files if f not in SKIP and not f.endswith((".npz", ".f32", ".json"))]


def citations():
    by_name = reference_index()
    for f in repo_files():
        try:
            text = open(os.path.join(ROOT, f), errors="ignore").read()
        except (IsADirectoryError, FileNotFoundError):
            continue
        for ln, line in enumerate(text.splitlines(), 1):
            found = [(m.group(1), m.group(2)) for m in CITE.finditer(line)]
            for name, spans in found:
                path = resolve(name, by_name)
                if path:
                    yield f, ln, line, name, path, spans, len(found)


def ranges(spans):
    for part in spans.lstrip(":").split(","):
        a, _, b = part.strip().partition("-")
        yield int(a), int(b or a)


def test_citations_inside_the_cited_files():
    bad = []
    for f, ln, _, name, path, spans, _ in citations():
        n = sum(1 for _ in open(os.path.join(REF, path), errors="ignore"))
        for a, b in ranges(spans):
            if a < 1 or b > n or b < a:
                bad.append(f"{f}:{ln}: {name}{spans} ({path} has {n} lines)")
    assert not bad, "\n".join(bad)


def test_named_functions_cite_their_definitions():
    bad, checked = [], 0
    for f, ln, line, name, path, spans, count in citations():
        if count != 1:
            continue
        named = [s for s in SYMBOLS if re.search(r"\b" + s + r"\b", line)]
        if len(named) != 1:
            continue
        defs = definitions(path, named[0])
        if not defs:
            continue  # the function is defined in another file than the one cited
        checked += 1
        rs = list(ranges(spans))
        # A cited range holds the definition's first line, or lies inside its body (a part of it).
        if not any(a <= d <= b or (d <= a and b <= e) for d, e in defs for a, b in rs):
            bad.append(f"{f}:{ln}: {named[0]} cited as {name}{spans}, defined at {path}:{defs}")
    assert checked > 10
    assert not bad, "\n".join(bad)
