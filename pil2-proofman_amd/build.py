"""Build libpil2_commit_hip.so (gfx950) in-tree with hipcc.

Usage: python pil2-proofman_amd/build.py
The library lands next to this file; tests, smoke() and bench.py load it from
there and fail loudly if it is absent.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libpil2_commit_hip.so")
SOURCES = ["ntt_kernels.hip", "poseidon2_kernels.hip", "fri_kernels.hip", "starks_api_hip.hip"]
ARCH = os.environ.get("PIL2HIP_ARCH", "gfx950")

FLAGS = [
    "--offload-arch=" + ARCH,
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    "-Wall",
    "-Wno-unused-function",
    # keep the compiler from forming packed FP32 instructions (none are used here)
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def build(verbose=False):
    cmd = ["hipcc"] + FLAGS + [os.path.join(CSRC, s) for s in SOURCES] + ["-o", LIB + ".tmp"]
    if verbose:
        print(" ".join(cmd))
    out = subprocess.run(cmd, capture_output=True, text=True)
    if out.returncode != 0:
        raise RuntimeError("hipcc failed:\n" + out.stdout + out.stderr)
    os.replace(LIB + ".tmp", LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
