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


def test_batch_plan_refuses_windows_wider_than_the_template_slots():
    """A window of W >= 8192 bins would let ucb_bandwidth ask for 8192-bin
    templates, beyond the 4096-bin slots: the plan is refused before any
    device work (host argument check, no GPU needed)."""
    import glass_amd
    L = glass_amd.lib()
    ip = ctypes.POINTER(ctypes.c_int)
    dp = ctypes.POINTER(ctypes.c_double)
    o = glass_amd.analytic_orbit()
    for W in (8192, 20000, 16):
        qmin = (ctypes.c_int * 1)(1000)
        gw = (ctypes.c_int * 1)(W)
        data = (ctypes.c_double * 2)()
        cg = (ctypes.c_int * 1)(0)
        cn = (ctypes.c_int * 1)(1)
        sp = (ctypes.c_double * 8)()
        plan = ctypes.c_void_p()
        rc = L.glass_hip_fd_batch_create_ex(ctypes.byref(plan), ctypes.byref(o), b"sangria",
                                            ctypes.c_double(62914560.0), ctypes.c_double(0.0), 1,
                                            ctypes.cast(qmin, ip), ctypes.cast(gw, ip), ctypes.cast(data, dp),
                                            ctypes.cast(data, dp), 1, ctypes.cast(cg, ip), ctypes.cast(cn, ip),
                                            8, ctypes.cast(sp, dp), 16)
        assert rc == -1, (W, rc)
        assert "outside 32..8191" in glass_amd.last_error()


def test_fstat_density_refused_when_the_rejection_sampler_cannot_accept():
    """glass_hip_ucb_sampler_set_fstat checks the proposal density on the
    host before touching the sampler: an all-zero table, a negative or NaN
    cell, or maxp above the table's maximum is refused."""
    import glass_amd
    L = glass_amd.lib()
    dp = ctypes.POINTER(ctypes.c_double)
    L.glass_hip_ucb_sampler_set_fstat.argtypes = [ctypes.c_void_p, ctypes.c_int, dp, ctypes.c_int,
                                                  ctypes.c_double, ctypes.c_int, ctypes.c_int,
                                                  ctypes.c_double, ctypes.c_int, ctypes.c_void_p]
    L.glass_hip_ucb_sampler_set_fstat.restype = ctypes.c_int

    def call(vals, maxp):
        t = (ctypes.c_double * len(vals))(*vals)
        rc = L.glass_hip_ucb_sampler_set_fstat(None, 0, ctypes.cast(t, dp), len(vals), 1e-8, 1, 1,
                                               maxp, 1000, None)
        return rc, glass_amd.last_error()

    for vals, maxp, msg in (([0.0] * 8, 1.0, "maxp"), ([1.0, -1.0, 2.0], 2.0, "cell 1"),
                            ([1.0, float("nan")], 1.0, "cell 1"), ([1.0, 3.0], 3.5, "maxp"),
                            ([1.0, float("inf")], 1.0, "cell 1")):
        rc, err = call(vals, maxp)
        assert rc == -1 and msg in err, (vals, maxp, rc, err)
    # a valid table passes the check and then needs a sampler
    rc, err = call([1.0, 3.0, 0.5], 3.0)
    assert rc == -1 and "bad arguments" in err
