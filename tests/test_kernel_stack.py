"""Every kernel of libglass_hip.so has a statically sized stack.

A kernel whose call graph the compiler cannot bound (recursion, for one, as
a recursive constexpr helper becomes when it is called outside a constant
expression) is marked `.amdhsa_uses_dynamic_stack 1`; the runtime then
reserves only the descriptor's fixed private segment per lane and deeper
frames run past it into unallocated scratch.  This compiles each source's
device code for gfx950 with the library's flags (hipcc, no GPU needed) and
checks every kernel descriptor.
"""
import os
import re
import shutil
import subprocess
from concurrent.futures import ThreadPoolExecutor

import pytest

from glass_amd import build

pytestmark = pytest.mark.skipif(shutil.which(build.HIPCC) is None and not os.path.exists(build.HIPCC),
                                reason="hipcc not available")


def _device_asm(src, out):
    flags = [f for f in build.FLAGS if f != "-fPIC"]
    cmd = [build.HIPCC] + flags + ["--cuda-device-only", "-S", "-o", out, os.path.join(build.CSRC, src)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed on {src}: {r.stderr[-2000:]}")
    return open(out).read()


def test_no_kernel_uses_a_dynamic_stack(tmp_path):
    def one(src):
        asm = _device_asm(src, str(tmp_path / (src + ".s")))
        kernels = re.findall(r"\.amdhsa_kernel (\S+)\n(.*?)\.end_amdhsa_kernel", asm, re.S)
        bad = [k for k, body in kernels if re.search(r"\.amdhsa_uses_dynamic_stack\s+1", body)]
        return src, len(kernels), bad

    with ThreadPoolExecutor(min(8, os.cpu_count() or 1)) as ex:
        res = list(ex.map(one, build.SOURCES))
    total = sum(n for _, n, _ in res)
    bad = [(s, k) for s, _, b in res for k in b]
    assert total > 100, f"only {total} kernels found"
    assert not bad, f"kernels with a dynamic stack: {bad}"
