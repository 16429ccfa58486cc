"""Builds the package's HIP libraries (gfx950) in-tree with hipcc:
libgpu_masstree.so (the index) and libhbm_probe.so (random-line read ceiling used by
the benchmark)."""
import os
import subprocess
import sys

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
REPO_DIR = os.path.dirname(PKG_DIR)
LIB_PATH = os.path.join(PKG_DIR, "libgpu_masstree.so")
SOURCES = [os.path.join(PKG_DIR, "csrc", "gpu_masstree_capi.hip")]
PROBE_PATH = os.path.join(PKG_DIR, "libhbm_probe.so")
PROBE_SOURCES = [os.path.join(PKG_DIR, "csrc", "hbm_probe.hip")]

HIPCC_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    # no packed fp32 instructions (none are needed; keep them out of the build)
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def hipcc():
    rocm = os.environ.get("ROCM_PATH", "/opt/rocm")
    cand = os.path.join(rocm, "bin", "hipcc")
    return cand if os.path.exists(cand) else "hipcc"


def _compile(sources, out, verbose):
    cmd = [hipcc()] + HIPCC_FLAGS + ["-I", os.path.join(REPO_DIR, "include"), "-o",
                                      out + ".tmp"] + sources
    if verbose:
        print(" ".join(cmd), file=sys.stderr)
    subprocess.check_call(cmd)
    os.replace(out + ".tmp", out)


def build(verbose=False):
    """Compiles every HIP source of the package (both libraries)."""
    _compile(SOURCES, LIB_PATH, verbose)
    _compile(PROBE_SOURCES, PROBE_PATH, verbose)
    return LIB_PATH


if __name__ == "__main__":
    print(build(verbose=True))
