"""Build libpil2_commit_hip.so (gfx950) in-tree with hipcc.

Usage: python pil2-proofman_amd/build.py
The library lands next to this file; tests, smoke() and bench.py load it from
there and fail loudly if it is absent.
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libpil2_commit_hip.so")
SOURCES = ["ntt_kernels.hip", "poseidon2_kernels.hip", "fri_kernels.hip", "starks_api_hip.hip", "serve_engine.hip"]
ARCH = os.environ.get("PIL2HIP_ARCH", "gfx950")

FLAGS = [
    "--offload-arch=" + ARCH,
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-Wall",
    "-Wno-unused-function",
    # keep the compiler from forming packed FP32 instructions (none are used here)
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd))
    out = subprocess.run(cmd, capture_output=True, text=True)
    if out.returncode != 0:
        raise RuntimeError("hipcc failed:\n" + " ".join(cmd) + "\n" + out.stdout + out.stderr)


def build(verbose=False):
    """Compile each source to an object in parallel, then link the shared library."""
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    objs = [os.path.join(objdir, os.path.splitext(s)[0] + ".o") for s in SOURCES]
    cmds = [["hipcc"] + FLAGS + ["-c", os.path.join(CSRC, s), "-o", o] for s, o in zip(SOURCES, objs)]
    with ThreadPoolExecutor(max_workers=len(cmds)) as ex:
        for f in [ex.submit(_run, c, verbose) for c in cmds]:
            f.result()
    _run(["hipcc", "--offload-arch=" + ARCH, "-shared", "-fPIC"] + objs + ["-o", LIB + ".tmp"], verbose)
    os.replace(LIB + ".tmp", LIB)
    build_c_host(verbose)
    return LIB


C_HOST_SRC = os.path.join(os.path.dirname(HERE), "tests", "c_host", "commit_host.c")
C_HOST_BIN = os.path.join(os.path.dirname(HERE), "tests", "c_host", "commit_host")


def build_c_host(verbose=False):
    """A plain C program linked against the library through the public header (the
    way a non-Python host such as proofman's Rust FFI links it); the tests run it."""
    _run(["gcc", "-std=c99", "-O2", "-Wall", "-Wextra", "-o", C_HOST_BIN + ".tmp", C_HOST_SRC, "-L" + HERE,
          "-lpil2_commit_hip", "-Wl,-rpath,$ORIGIN/../../pil2-proofman_amd"], verbose)
    os.replace(C_HOST_BIN + ".tmp", C_HOST_BIN)
    return C_HOST_BIN


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
