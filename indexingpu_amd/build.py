"""Builds libgpu_masstree.so (HIP, gfx950) in-tree with hipcc."""
import os
import subprocess
import sys

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
REPO_DIR = os.path.dirname(PKG_DIR)
LIB_PATH = os.path.join(PKG_DIR, "libgpu_masstree.so")
SOURCES = [os.path.join(PKG_DIR, "csrc", "gpu_masstree_capi.hip")]

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


def build(verbose=False):
    """Compiles every HIP source of the package into libgpu_masstree.so."""
    cmd = [hipcc()] + HIPCC_FLAGS + ["-I", os.path.join(REPO_DIR, "include"), "-o",
                                      LIB_PATH + ".tmp"] + SOURCES
    if verbose:
        print(" ".join(cmd), file=sys.stderr)
    subprocess.check_call(cmd)
    os.replace(LIB_PATH + ".tmp", LIB_PATH)
    return LIB_PATH


if __name__ == "__main__":
    print(build(verbose=True))
