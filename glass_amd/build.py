"""Build libglass_hip.so in-tree with hipcc for gfx950.

    python -m glass_amd.build        (or glass_amd.build.build())
"""
import os
import shutil
import subprocess
import sys
import tempfile
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libglass_hip.so")
SOURCES = ["runtime.hip", "fd_plain.hip", "nwip.hip", "fd_batch.hip", "fisher_fd.hip", "fd_derivs.hip", "fd_adjoint.hip",
           "wdm_ucb.hip", "wdm_transform.hip", "logl_batch.hip", "noise.hip",
           "eigen.hip", "ucb_sampler.hip", "fstat.hip", "ucb_io.hip"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = [
    "--offload-arch=gfx950", "-O3", "-fPIC", "-std=c++17",
    # packed-FP32 instructions are kept out of every kernel
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
    # keep loop-invariant values inside the loops that use them: hoisting
    # them (FP64 constants, LDS addresses) out of the per-template loops
    # costs ~60 VGPRs and halves occupancy of the template kernels
    "-mllvm", "-disable-machine-licm",
]
EXTRA = os.environ.get("GLASS_HIP_EXTRA_FLAGS", "").split()
# the LDC HDF5 readers: host code, linked against libhdf5 (HDF5_DIR/include,
# HDF5_DIR/lib; the image's copy lives under /opt/conda)
LDC_LIB = os.path.join(HERE, "libglass_ldc.so")
HDF5_DIR = os.environ.get("HDF5_DIR", "/opt/conda")


def build_ldc(verbose=False):
    """Builds libglass_ldc.so (g++, -lhdf5 with an rpath to HDF5_DIR/lib).
    Returns its path, or None when hdf5.h is not under HDF5_DIR."""
    inc, libdir = os.path.join(HDF5_DIR, "include"), os.path.join(HDF5_DIR, "lib")
    if not os.path.exists(os.path.join(inc, "hdf5.h")):
        print(f"libglass_ldc.so not built: no hdf5.h under {inc}", flush=True)
        return None
    tmp = LDC_LIB + ".tmp"
    cmd = ["g++", "-O2", "-fPIC", "-shared", "-std=c++17", "-Wall", "-I" + inc,
           os.path.join(CSRC, "ldc_hdf5.cpp"), "-o", tmp, "-L" + libdir, "-Wl,-rpath," + libdir, "-lhdf5"]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LDC_LIB)
    return LDC_LIB


def build(verbose=False, jobs=None):
    """Compiles every source to an object in parallel (a temporary folder),
    then links libglass_hip.so next to this file."""
    jobs = jobs or min(8, os.cpu_count() or 1)
    tmpd = tempfile.mkdtemp(prefix="glass_hip_build_")
    try:
        def cc(src):
            obj = os.path.join(tmpd, os.path.splitext(src)[0] + ".o")
            cmd = [HIPCC] + FLAGS + EXTRA + ["-c", os.path.join(CSRC, src), "-o", obj]
            if verbose:
                print(" ".join(cmd), flush=True)
            r = subprocess.run(cmd, capture_output=True, text=True)
            if r.returncode:
                sys.stderr.write(r.stderr)
                raise subprocess.CalledProcessError(r.returncode, cmd)
            return obj
        with ThreadPoolExecutor(jobs) as ex:
            objs = list(ex.map(cc, SOURCES))
        tmp = LIB + ".tmp"
        cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC"] + objs + ["-o", tmp]
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.run(cmd, check=True)
        os.replace(tmp, LIB)
    finally:
        shutil.rmtree(tmpd, ignore_errors=True)
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    build_ldc(verbose=True)
    sys.exit(0)
