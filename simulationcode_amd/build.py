"""Builds libsc for gfx950 in-tree (simulationcode_amd/_lib/libsc.so).

One hipcc invocation over the HIP sources.  Flags:
  -ffp-contract=off  every FP64 expression rounds as written (bit-exact pair
                     predicates and reference operation order)
  -Xclang -target-feature -Xclang -packed-fp32-ops
                     no packed FP32 instructions are generated
"""
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OUT_DIR = os.path.join(HERE, "_lib")
OUT = os.path.join(OUT_DIR, "libsc.so")
SOURCES = ["sc_kernels.hip", "sc_replica.hip", "sc_api.hip", "sc_batch.hip"]
ARCH = os.environ.get("SC_OFFLOAD_ARCH", "gfx950")


def hipcc():
    for cand in (shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def build(verbose=False):
    os.makedirs(OUT_DIR, exist_ok=True)
    srcs = [os.path.join(CSRC, s) for s in SOURCES if os.path.exists(os.path.join(CSRC, s))]
    tmp = OUT + ".tmp"
    cmd = [hipcc(), f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-shared",
           "-ffp-contract=off", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
           "-Wno-unused-command-line-argument",
           "-I", os.path.join(ROOT, "include"), "-I", CSRC, *srcs, "-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError("libsc build failed")
    os.replace(tmp, OUT)
    return OUT


def build_oracle():
    odir = os.path.join(ROOT, "oracle")
    res = subprocess.run(["make", "-s", "-B", "-C", odir], capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError("oracle build failed")


if __name__ == "__main__":
    print(build(verbose=True))
