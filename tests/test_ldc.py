"""LDC HDF5 readers (libglass_ldc.so) on files written by
tests/golden/ldc_writer.c with known contents.

The reference holds no HDF5 file, so parity rests on what the layout and
the mathematics fix: X, Y, Z read back bit for bit from the named compound
members (in any member order, extra members ignored) or from columns 1..3
of the RADLER matrix; A, E, T against the closed forms (Z-X)/sqrt2,
(X-2Y+Z)/sqrt6, (X+Y+Z)/sqrt3 to the reference's 15-digit constants; delta
= t[1]-t[0] (row 1 - row 0 of column 0); the RADLER sample count Nrow/2
(the size the reference's arrays have).  Malformed files raise.
"""
import ctypes
import os
import subprocess

import numpy as np
import pytest

HDF5 = os.environ.get("HDF5_DIR", "/opt/conda")
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HAVE = os.path.exists(os.path.join(HDF5, "include", "hdf5.h"))
pytestmark = pytest.mark.skipif(not HAVE, reason="no HDF5 headers (HDF5_DIR)")


@pytest.fixture(scope="module")
def ldc():
    from glass_amd import build, ldc
    if not os.path.exists(ldc.LIB_PATH):
        build.build_ldc()
    ldc.lib()
    return ldc


@pytest.fixture(scope="module")
def writer(tmp_path_factory):
    d = tmp_path_factory.mktemp("ldcw")
    so = str(d / "libldcw.so")
    subprocess.run(["gcc", "-O1", "-fPIC", "-shared", "-I" + os.path.join(HDF5, "include"),
                    os.path.join(ROOT, "tests", "golden", "ldc_writer.c"), "-o", so,
                    "-L" + os.path.join(HDF5, "lib"), "-Wl,-rpath," + os.path.join(HDF5, "lib"), "-lhdf5"],
                   check=True)
    W = ctypes.CDLL(so)
    dp = ctypes.POINTER(ctypes.c_double)
    W.ldc_write_compound.argtypes = [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                     ctypes.POINTER(ctypes.c_char_p), ctypes.POINTER(dp)]
    W.ldc_write_matrix.argtypes = [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_int, ctypes.c_int, dp]
    W.ldc_write_int_compound.argtypes = [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_int,
                                         ctypes.POINTER(ctypes.c_int)]
    return W


def write_compound(W, path, dset, cols, ncol=0):
    names = list(cols)
    arrs = [np.ascontiguousarray(cols[k], dtype=np.float64) for k in names]
    N = len(arrs[0])
    cn = (ctypes.c_char_p * len(names))(*[k.encode() for k in names])
    dp = ctypes.POINTER(ctypes.c_double)
    cp = (dp * len(arrs))(*[a.ctypes.data_as(dp) for a in arrs])
    assert W.ldc_write_compound(str(path).encode(), dset.encode(), N, ncol, len(names), cn, cp) == 0


def series(N, seed=0, dt=5.0, t0=0.0):
    rng = np.random.default_rng(seed)
    t = t0 + dt * np.arange(N)
    # strain-like magnitudes with a wide dynamic range
    X, Y, Z = (rng.standard_normal(N) * 10.0 ** rng.uniform(-22, -18, N) for _ in range(3))
    return t, X, Y, Z


def check_aet(r):
    s2, s6, s3 = np.sqrt(2.0), np.sqrt(6.0), np.sqrt(3.0)
    A = (r.Z - r.X) / s2
    E = (r.X - 2 * r.Y + r.Z) / s6
    T = (r.X + r.Y + r.Z) / s3
    for got, ref in ((r.A, A), (r.E, E), (r.T, T)):
        scale = np.abs(r.X) + 2 * np.abs(r.Y) + np.abs(r.Z)
        assert np.all(np.abs(got - ref) <= 4e-16 * scale)
    # a transposed channel or a flipped sign would miss by O(1)
    assert np.max(np.abs(r.A - (r.X - r.Z) / s2)) > 1e-3 * np.max(np.abs(A))


def test_sangria_layout(ldc, writer, tmp_path):
    t, X, Y, Z = series(4099, seed=1, dt=5.0, t0=1234.5)
    p = tmp_path / "sangria.h5"
    write_compound(writer, p, "/obs/tdi", {"t": t, "X": X, "Y": Y, "Z": Z})
    r = ldc.read_tdi(p)
    assert r.N == 4099
    assert np.array_equal(r.X, X) and np.array_equal(r.Y, Y) and np.array_equal(r.Z, Z)
    assert r.delta == t[1] - t[0] == 5.0
    check_aet(r)


def test_members_by_name_in_any_order(ldc, writer, tmp_path):
    t, X, Y, Z = series(257, seed=2, dt=15.0)
    extra = np.full(257, 7.0)
    p = tmp_path / "sky.h5"
    # the sky components (glass_data.c:705-760) sit under /sky/<name>/tdi
    write_compound(writer, p, "/sky/dgb/tdi", {"Z": Z, "extra": extra, "t": t, "Y": Y, "X": X})
    r = ldc.read_tdi(p, "/sky/dgb/tdi")
    assert np.array_equal(r.X, X) and np.array_equal(r.Y, Y) and np.array_equal(r.Z, Z)
    assert r.delta == 15.0
    check_aet(r)


def test_single_column_2d_shape(ldc, writer, tmp_path):
    t, X, Y, Z = series(64, seed=3)
    p = tmp_path / "n1.h5"
    write_compound(writer, p, "/obs/tdi", {"t": t, "X": X, "Y": Y, "Z": Z}, ncol=1)
    r = ldc.read_tdi(p)
    assert r.N == 64 and np.array_equal(r.Z, Z)


def test_integer_members_are_converted(ldc, writer, tmp_path):
    N = 10
    v = np.arange(4 * N, dtype=np.int32).reshape(N, 4)
    v[:, 0] = 10 * np.arange(N) + 3
    p = tmp_path / "int.h5"
    assert writer.ldc_write_int_compound(str(p).encode(), b"/obs/tdi", N,
                                         v.ctypes.data_as(ctypes.POINTER(ctypes.c_int))) == 0
    r = ldc.read_tdi(p)
    assert np.array_equal(r.X, v[:, 1].astype(float)) and np.array_equal(r.Z, v[:, 3].astype(float))
    assert r.delta == 10.0


def test_radler_layout(ldc, writer, tmp_path):
    Nrow, Ncol = 1002, 5
    rng = np.random.default_rng(4)
    M = rng.standard_normal((Nrow, Ncol)) * 1e-20
    M[:, 0] = 15.0 * np.arange(Nrow) + 0.25
    p = tmp_path / "radler.h5"
    assert writer.ldc_write_matrix(str(p).encode(), b"/H5LISA/PreProcess/TDIdata", Nrow, Ncol,
                                   M.ctypes.data_as(ctypes.POINTER(ctypes.c_double))) == 0
    r = ldc.read_radler_tdi(p)
    assert r.N == Nrow // 2
    assert np.array_equal(r.X, M[:Nrow // 2, 1]) and np.array_equal(r.Y, M[:Nrow // 2, 2])
    assert np.array_equal(r.Z, M[:Nrow // 2, 3])
    assert r.delta == M[1, 0] - M[0, 0] == 15.0
    check_aet(r)


def test_errors(ldc, writer, tmp_path):
    with pytest.raises(ldc.LdcError, match="cannot open"):
        ldc.read_tdi(tmp_path / "missing.h5")
    t, X, Y, Z = series(16, seed=5)
    p = tmp_path / "e.h5"
    write_compound(writer, p, "/obs/tdi", {"t": t, "X": X, "Y": Y})
    with pytest.raises(ldc.LdcError, match="no member Z"):
        ldc.read_tdi(p)
    with pytest.raises(ldc.LdcError, match="no dataset"):
        ldc.read_tdi(p, "/sky/mbhb/tdi")
    write_compound(writer, p, "/obs/tdi", {"t": t[:1], "X": X[:1], "Y": Y[:1], "Z": Z[:1]})
    with pytest.raises(ldc.LdcError, match="fewer than 2 rows"):
        ldc.read_tdi(p)
    write_compound(writer, p, "/obs/tdi", {"t": t, "X": X, "Y": Y, "Z": Z}, ncol=3)
    with pytest.raises(ldc.LdcError, match="more than one column"):
        ldc.read_tdi(p)
    M = np.zeros((8, 3))
    assert writer.ldc_write_matrix(str(p).encode(), b"/H5LISA/PreProcess/TDIdata", 8, 3,
                                   M.ctypes.data_as(ctypes.POINTER(ctypes.c_double))) == 0
    with pytest.raises(ldc.LdcError, match="fewer than 4 columns"):
        ldc.read_radler_tdi(p)
    # a wrong Nsamples is refused before anything is written
    write_compound(writer, p, "/obs/tdi", {"t": t, "X": X, "Y": Y, "Z": Z})
    out = [np.zeros(15) for _ in range(6)]
    dp = ctypes.POINTER(ctypes.c_double)
    d = ctypes.c_double()
    rc = ldc.lib().glass_ldc_read_tdi(os.fsencode(p), b"/obs/tdi", 15, *[a.ctypes.data_as(dp) for a in out],
                                      ctypes.byref(d))
    assert rc == -1 and b"holds 16" in ldc.lib().glass_ldc_last_error()
    assert all(not a.any() for a in out)


def test_exports_match_header(ldc):
    import re
    hdr = open(os.path.join(ROOT, "include", "glass_ldc.h")).read()
    names = set(re.findall(r"\b(glass_ldc_\w+)\s*\(", hdr))
    assert len(names) == 5
    for n in names:
        assert hasattr(ldc.lib(), n), n


def test_reference_side_shims_compile_and_run(ldc, writer, tmp_path):
    """glass_amd/shim/glass_ldc_shims.c compiles as C against struct TDI's
    fields (utils/src/glass_lisa.h:116-144) and a restatement of alloc_tdi
    (glass_lisa.c:1281-1298), links against libglass_ldc.so and, run on a
    written file, fills struct TDI as read_tdi does."""
    sub = tmp_path / "tdi.h"
    sub.write_text("struct TDI { double *X, *Y, *Z, *A, *E, *T; int Nchannel; int N; double delta; };\n"
                   "void alloc_tdi(struct TDI *tdi, int N, int Nchannel);\n")
    main = tmp_path / "main.c"
    main.write_text(r'''
#include <stdio.h>
#include <stdlib.h>
void alloc_tdi(struct TDI *t, int N, int Nchannel)
{
    t->N = N; t->Nchannel = Nchannel;
    t->X = calloc(N, sizeof(double)); t->Y = calloc(N, sizeof(double)); t->Z = calloc(N, sizeof(double));
    t->A = calloc(N, sizeof(double)); t->E = calloc(N, sizeof(double)); t->T = calloc(N, sizeof(double));
}
void LISA_Read_HDF5_LDC_TDI(struct TDI *tdi, char *fileName, const char *dataName);
int main(int argc, char **argv)
{
    struct TDI t;
    LISA_Read_HDF5_LDC_TDI(&t, argv[1], "/obs/tdi");
    printf("%d %.17g %.17g %.17g %.17g\n", t.N, t.delta, t.X[t.N - 1], t.E[1], t.T[2]);
    return 0;
}
''')
    inc = os.path.join(ROOT, "include")
    shim = os.path.join(ROOT, "glass_amd", "shim", "glass_ldc_shims.c")
    libdir = os.path.dirname(ldc.LIB_PATH)
    exe = str(tmp_path / "a.out")
    subprocess.run(["gcc", "-std=gnu99", "-Wall", "-Werror", "-DGLASS_HIP_SHIM_SUBSET", "-include", str(sub),
                    "-I", inc, shim, str(main), "-L", libdir, "-lglass_ldc", "-Wl,-rpath," + libdir, "-o", exe],
                   check=True)
    t, X, Y, Z = series(33, seed=6, dt=2.5)
    p = tmp_path / "s.h5"
    write_compound(writer, p, "/obs/tdi", {"t": t, "X": X, "Y": Y, "Z": Z})
    out = subprocess.run([exe, str(p)], capture_output=True, text=True, check=True).stdout.split()
    r = ldc.read_tdi(p)
    assert int(out[0]) == 33 and float(out[1]) == 2.5
    assert float(out[2]) == X[-1] and float(out[3]) == r.E[1] and float(out[4]) == r.T[2]
    # a missing file ends the program with the reference's exit(1)
    bad = subprocess.run([exe, str(tmp_path / "none.h5")], capture_output=True, text=True)
    assert bad.returncode == 1 and "LISA_Read_HDF5_LDC_TDI" in bad.stderr
