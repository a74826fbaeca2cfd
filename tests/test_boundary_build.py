"""The drop-in boundary builds as INTEGRATION.md describes, and the standalone build's RowGate keeps
the reference's semantics (CPU only)."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF_SRC = "/root/reference/src"


@pytest.mark.skipif(not os.path.isdir(os.path.join(REF_SRC, "gpu")), reason="reference tree not mounted")
def test_hip_engine_compiles_inside_the_celv2_tree():
    """celv2::HipEngine against CelV2's own decision_engine.h / decision_map.h / y4m.h / kernels_ref.h
    (-DCELV2_IN_TREE, CelV2's src/ and this repository's include/ on the include path)."""
    cmd = ["g++", "-std=c++17", "-fsyntax-only", "-Wall", "-DCELV2_IN_TREE", f"-I{REF_SRC}", f"-I{os.path.join(ROOT, 'include')}",
           os.path.join(ROOT, "celv2_amd", "csrc", "host", "hip_engine.cpp")]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-4000:]


def test_standalone_hip_engine_compiles():
    cmd = ["g++", "-std=c++17", "-fsyntax-only", "-Wall", f"-I{os.path.join(ROOT, 'include')}",
           os.path.join(ROOT, "celv2_amd", "csrc", "host", "hip_engine.cpp")]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-4000:]


def test_row_gate_semantics(tmp_path):
    exe = str(tmp_path / "row_gate_check")
    cmd = ["g++", "-std=c++17", "-O1", "-Wall", f"-I{os.path.join(ROOT, 'include')}",
           os.path.join(ROOT, "tests", "contract", "row_gate_check.cpp"),
           os.path.join(ROOT, "celv2_amd", "csrc", "host", "row_gate.cpp"), "-lpthread", "-o", exe]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-4000:]
    r = subprocess.run([exe], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stdout + r.stderr
