"""Builds the native libraries in-tree.

  celv2_amd/lib/libcelv2hip.so    HIP kernels for gfx950 + the C++ engine + C ABI (hipcc)
  celv2_amd/lib/libcelv2synth.so  synthetic clip generator (g++)
  celv2_amd/lib/sad_rate          v_sad_u8 issue-rate microbenchmark (the motion search's ceiling)

`python -m celv2_amd.build` or `build()` compiles every time it is called.
"""
import hashlib
import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "lib")
OBJ = os.path.join(HERE, "csrc", "_build")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
# Packed FP32 instructions stay off: the decision kernels are integer, and hand-off of float code to
# v_pk_*_f32 is not wanted anywhere in this library.
HIP_FLAGS = [
    "-O3",
    "-std=c++17",
    f"--offload-arch={ARCH}",
    "-fPIC",
    "-Xclang",
    "-target-feature",
    "-Xclang",
    "-packed-fp32-ops",
    # MFMA results straight into VGPRs (no v_accvgpr_read after each product).
    "-mllvm",
    "-amdgpu-mfma-vgpr-form",
    "-Wall",
    "-Wno-unused-function",
    f"-I{os.path.join(ROOT, 'include')}",
]
HOST_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-Wall", f"-I{os.path.join(ROOT, 'include')}"]

KERNELS = ["planes", "me", "intra", "inter", "finalize", "pruner", "cnn", "transform"]
HOST = ["engine", "hip_engine", "row_gate", "engine_capi", "gop", "y4m", "tx_capi"]


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines() if "-fp32-ops" not in l and "-access-mode" not in l and "recognized feature" not in l)
    if r.returncode != 0:
        raise RuntimeError("build failed: " + " ".join(cmd) + "\n" + out)
    return out


def source_hash():
    """A digest of every native source of libcelv2hip.so (csrc/ and include/), embedded in the library
    at build time; engine.lib() refuses a library whose digest differs from the sources beside it."""
    paths = []
    for base in (CSRC, os.path.join(ROOT, "include")):
        for d, _, files in os.walk(base):
            if os.path.basename(d) == "_build":
                continue
            paths += [os.path.join(d, f) for f in files if f.endswith((".hip", ".cpp", ".h"))]
    h = hashlib.sha256()
    for p in sorted(paths, key=lambda p: os.path.relpath(p, ROOT)):
        h.update(os.path.relpath(p, ROOT).encode())
        with open(p, "rb") as f:
            h.update(f.read())
    return h.hexdigest()[:16]


def build_hip():
    os.makedirs(OBJ, exist_ok=True)
    os.makedirs(LIB, exist_ok=True)
    jobs = []
    for k in KERNELS:
        jobs.append([HIPCC, *HIP_FLAGS, "-c", "-x", "hip", os.path.join(CSRC, "kernels", k + ".hip"), "-o", os.path.join(OBJ, k + ".o")])
    digest = f'-DCELV2_SOURCE_HASH="{source_hash()}"'
    for h in HOST:
        jobs.append([HIPCC, *HOST_FLAGS, digest, "-c", os.path.join(CSRC, "host", h + ".cpp"), "-o", os.path.join(OBJ, h + ".o")])
    with ThreadPoolExecutor(max_workers=min(8, os.cpu_count() or 4)) as ex:
        for out in ex.map(_run, jobs):
            if out.strip():
                print(out, file=sys.stderr)
    objs = [os.path.join(OBJ, n + ".o") for n in KERNELS + HOST]
    tmp = os.path.join(LIB, "libcelv2hip.so.tmp")
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs, "-lpthread"])
    os.replace(tmp, os.path.join(LIB, "libcelv2hip.so"))


def build_synth():
    os.makedirs(LIB, exist_ok=True)
    tmp = os.path.join(LIB, "libcelv2synth.so.tmp")
    _run(["g++", "-O2", "-std=c++17", "-fPIC", "-Wall", f"-I{os.path.join(ROOT, 'include')}", "-shared", "-o", tmp,
          os.path.join(CSRC, "synth", "synth.cpp"), "-lpthread"])
    os.replace(tmp, os.path.join(LIB, "libcelv2synth.so"))


def build_probes():
    os.makedirs(LIB, exist_ok=True)
    tmp = os.path.join(LIB, "sad_rate.tmp")
    _run([HIPCC, *[f for f in HIP_FLAGS if f != "-fPIC"], "-o", tmp, os.path.join(CSRC, "probes", "sad_rate.hip")])
    os.replace(tmp, os.path.join(LIB, "sad_rate"))


def build_oracle():
    """The test oracle (oracle/liboracle.so) and, where the reference tree is mounted, oracle/_ref."""
    _run(["make", "-C", os.path.join(ROOT, "oracle"), "-s", "-B", "liboracle.so"])
    if os.path.isdir("/root/reference/src/gpu"):
        _run(["make", "-C", os.path.join(ROOT, "oracle"), "-s", "-B", "ref"])


def build_all():
    build_synth()
    build_oracle()
    build_hip()
    build_probes()


if __name__ == "__main__":
    build_all()
    print("built", os.listdir(LIB))
