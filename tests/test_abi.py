"""The C-ABI library loads and exports every symbol declared in include/*.h
(no compute calls: runs without a GPU)."""
import ctypes
import glob
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    syms = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        txt = open(h).read()
        txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
        for m in re.finditer(r"\b(glass_hip_\w+)\s*\(", txt):
            syms.add(m.group(1))
    return syms


def test_library_exports_declared_symbols():
    import glass_amd
    path = glass_amd.LIB_PATH
    assert os.path.exists(path), "build the library first (python -m glass_amd.build)"
    L = ctypes.CDLL(path)
    syms = declared_symbols()
    assert len(syms) >= 8
    missing = [s for s in sorted(syms) if not hasattr(L, s)]
    assert not missing, missing


def test_headers_compile_as_c(tmp_path):
    """Every public header is plain C (no torch or HIP types)."""
    import subprocess
    src = tmp_path / "t.c"
    heads = sorted(os.path.basename(h) for h in glob.glob(os.path.join(ROOT, "include", "*.h")))
    src.write_text("".join(f'#include "{h}"\n' for h in heads) + "int main(void){return 0;}\n")
    subprocess.run(["gcc", "-std=c99", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"),
                    "-c", str(src), "-o", str(tmp_path / "t.o")], check=True)
