"""Builds libfmftsaxs_amd.so in-tree: host C11 sources with gcc, HIP kernels
with hipcc for gfx950, linked by hipcc into one shared library.

    python -m libfmftsaxs_amd.build
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libfmftsaxs_amd.so")
INCLUDE = os.path.join(ROOT, "include")

HOST_SRC = ["host/fsx_math.c", "host/fsx_api.c", "host/fsx_score.c", "host/fsx_spf.c",
            "host/fsx_front.c", "host/fsx_naive.c"]
HIP_SRC = ["hip/fsx_kernels.hip"]

CFLAGS = ["-std=c11", "-O2", "-fPIC", "-ffp-contract=off", "-fno-fast-math", "-Wall",
          "-Wno-unused-function", "-D_POSIX_C_SOURCE=200809L", "-I" + INCLUDE, "-I" + CSRC]
HIPFLAGS = ["-O3", "-fPIC", "--offload-arch=gfx950", "-ffp-contract=off", "-std=c++17",
            "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
            "-Wno-unused-function", "-I" + INCLUDE, "-I" + CSRC]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + r.stdout + r.stderr)
        raise RuntimeError("build step failed: " + cmd[0])
    # the host pass of hipcc warns about the GPU-only feature flag; keep other output
    err = "\n".join(l for l in r.stderr.splitlines() if "packed-fp32-ops" not in l
                    and "is not a recognized feature" not in l)
    if err.strip():
        sys.stderr.write(err + "\n")


def build(verbose=False):
    os.makedirs(OBJ, exist_ok=True)
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    objs = []
    for s in HOST_SRC:
        o = os.path.join(OBJ, os.path.basename(s).replace(".c", ".o"))
        _run(["gcc"] + CFLAGS + ["-c", os.path.join(CSRC, s), "-o", o])
        objs.append(o)
    for s in HIP_SRC:
        o = os.path.join(OBJ, os.path.basename(s).replace(".hip", ".o"))
        _run([hipcc] + HIPFLAGS + ["-c", os.path.join(CSRC, s), "-o", o])
        objs.append(o)
    tmp = LIB + ".tmp"
    _run([hipcc, "-shared", "-fPIC", "--offload-arch=gfx950", "-o", tmp] + objs
         + ["-lpthread", "-lm", "-lgcc"])
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    return LIB


if __name__ == "__main__":
    build(verbose=True)
