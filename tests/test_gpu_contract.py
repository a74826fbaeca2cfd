"""The DecisionEngine contract on celv2::HipEngine (C++), as CelV2's own GPU tests check its engines:
tests/contract/engine_contract.cpp runs every case once in its own process (the first engines it
creates are created on 8 threads at once) and reports one line per case; each test below asserts
one case. The program is built by __graft_entry__.build() (make -C tests/contract)."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PROGRAM = os.path.join(ROOT, "tests", "bin", "engine_contract")
CASES = [
    "concurrent_creation",
    "launch_collect_matches_decide",
    "prepare_copies_its_frame",
    "rows_publish_before_collect_ends",
    "rejects_bad_sizes",
    "exception_mapping",
    "three_engines_match_cpu",
]


@pytest.fixture(scope="module")
def results(gpu):
    if not os.path.exists(PROGRAM):
        pytest.fail(f"{PROGRAM} is not built: run __graft_entry__.build()")
    print("engine_contract: running all cases", flush=True)
    r = subprocess.run([PROGRAM], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=400)
    out = {}
    for line in r.stdout.splitlines():
        parts = line.split(" ", 3)
        if len(parts) >= 3 and parts[0] == "CASE" and parts[2] != "start":
            out[parts[1]] = (parts[2], parts[3] if len(parts) > 3 else "")
    return r.returncode, r.stdout, out


@pytest.mark.timeout(450)
@pytest.mark.parametrize("case", CASES)
def test_contract(results, case):
    rc, text, out = results
    assert case in out, f"{case} did not finish (exit status {rc}):\n{text[-3000:]}"
    status, detail = out[case]
    assert status == "ok", f"{case}: {detail}"
