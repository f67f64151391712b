"""Builds libcostuum_amd.so (the HIP/gfx950 library) in-tree with hipcc."""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
LIB = os.path.join(HERE, "libcostuum_amd.so")
SOURCES = ["csrc/ctm_kernels.hip", "csrc/ctm_observables.hip",
           "csrc/ctm_average.hip", "csrc/ctm_api.hip",
           "csrc/ctm_host.hip", "csrc/ctm_tangent.hip", "csrc/ctm_mixed.hip",
           "csrc/ctm_plan.hip"]
HEADERS = ["csrc/ctm_device.hpp", "csrc/ctm_kernels.hpp", "csrc/ctm_gj.hpp",
           "csrc/ctm_plan.hpp",
           "csrc/ctm_mixed.hpp",
           "csrc/ctm_observables.hpp", "csrc/ctm_average.hpp",
           os.path.join("..", "include", "costuum_amd.h")]

HIPCC_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    # packed FP32 instructions stay disabled in every HIP build
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def hipcc():
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", "hipcc"):
        if cand and (os.path.isabs(cand) and os.path.exists(cand) or
                     not os.path.isabs(cand)):
            return cand
    return "hipcc"


def build(verbose=False):
    srcs = [os.path.join(HERE, s) for s in SOURCES]
    tmp = LIB + ".tmp"
    cmd = [hipcc()] + HIPCC_FLAGS + ["-I", os.path.join(ROOT, "include"),
                                     "-o", tmp] + srcs
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError("hipcc failed for libcostuum_amd.so")
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
