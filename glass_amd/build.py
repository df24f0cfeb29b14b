"""Build libglass_hip.so in-tree with hipcc for gfx950.

    python -m glass_amd.build        (or glass_amd.build.build())
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libglass_hip.so")
SOURCES = ["runtime.hip", "fd_plain.hip", "nwip.hip", "fd_batch.hip", "fisher_fd.hip", "fd_derivs.hip",
           "wdm_ucb.hip", "wdm_transform.hip", "logl_batch.hip"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = [
    "--offload-arch=gfx950", "-O3", "-fPIC", "-shared", "-std=c++17",
    # packed-FP32 instructions are kept out of every kernel
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
    # keep loop-invariant values inside the loops that use them: hoisting
    # them (FP64 constants, LDS addresses) out of the per-template loops
    # costs ~60 VGPRs and halves occupancy of the template kernels
    "-mllvm", "-disable-machine-licm",
]
EXTRA = os.environ.get("GLASS_HIP_EXTRA_FLAGS", "").split()


def build(verbose=False):
    srcs = [os.path.join(CSRC, s) for s in SOURCES]
    tmp = LIB + ".tmp"
    cmd = [HIPCC] + FLAGS + EXTRA + srcs + ["-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
