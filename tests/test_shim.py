"""The reference-side shim file (glass_amd/shim/glass_hip_shims.c) compiles
as C against the fields it reads (declared here with the reference's field
types: struct Orbit, TDI, Noise, Source, Data, Model and Wavelets of
utils/src/glass_lisa.h, glass_data.h, glass_model.h, glass_wavelet.h) and
links against libglass_hip.so with every library symbol it calls resolved.
Nothing is run (no GPU); the *_cpu originals and LAPACK's
matrix_eigenstuff are stubs of the right signature."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

SUBSET = r"""
struct Orbit;
void analytic_orbits(struct Orbit *orbit, double t, double *x, double *y, double *z);
struct Orbit { int Norb; double L, fstar, ecc, R; double *t; double **x, **y, **z;
               void (*orbit_function)(struct Orbit *, double, double *, double *, double *); };
struct TDI { double *X, *Y, *Z, *A, *E, *T; int Nchannel; int N; };
struct Noise { int N; int Nchannel; double ***C; double ***invC; double *logdetC; };
struct Wavelets { int NF, NT; double dt, df, A; int kmin, kmax, imin, imax; };
struct Source { double *params; struct TDI *tdi; double **fisher_matrix, **fisher_evectr;
                double *fisher_evalue; int *list; int Nlist; };
struct Data { int N, NFFT, Nchannel; double T, t0; struct Wavelets *wdm; int qmin; struct TDI *tdi;
              char format[16]; };
struct Model { struct Noise *noise; struct TDI *tdi; int *list; int Nlist; };
double snr_cpu(struct Source *source, struct Noise *noise);
void invert_noise_covariance_matrix_cpu(struct Noise *noise);
void ucb_fisher_cpu(struct Orbit *orbit, struct Data *data, struct Source *source, struct Noise *noise);
void matrix_eigenstuff(double **matrix, double **evector, double *evalue, int N);
"""

STUBS = r"""
#include <stdlib.h>
void analytic_orbits(struct Orbit *o, double t, double *x, double *y, double *z) { abort(); }
double snr_cpu(struct Source *s, struct Noise *n) { abort(); }
void invert_noise_covariance_matrix_cpu(struct Noise *n) { abort(); }
void ucb_fisher_cpu(struct Orbit *o, struct Data *d, struct Source *s, struct Noise *n) { abort(); }
void matrix_eigenstuff(double **m, double **v, double *e, int N) { abort(); }
int main(void) { return 0; }
"""


def test_shim_compiles_and_links(tmp_path):
    import glass_amd
    sub = tmp_path / "ref_subset.h"
    sub.write_text(SUBSET)
    stubs = tmp_path / "stubs.c"
    stubs.write_text(STUBS)
    inc = os.path.join(ROOT, "include")
    shim = os.path.join(ROOT, "glass_amd", "shim", "glass_hip_shims.c")
    cc = ["gcc", "-std=gnu99", "-Wall", "-Werror", "-DGLASS_HIP_SHIM_SUBSET", "-include", str(sub), "-I", inc]
    subprocess.run(cc + ["-c", shim, "-o", str(tmp_path / "shim.o")], check=True)
    subprocess.run(["gcc", "-std=gnu99", "-include", str(sub), "-c", str(stubs), "-o", str(tmp_path / "stubs.o")],
                   check=True)
    libdir = os.path.dirname(glass_amd.LIB_PATH)
    r = subprocess.run(["gcc", str(tmp_path / "shim.o"), str(tmp_path / "stubs.o"), "-L", libdir, "-lglass_hip",
                        "-Wl,-rpath," + libdir, "-Wl,--allow-shlib-undefined", "-o", str(tmp_path / "a.out")],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    # every glass_hip_* the shim calls is exported by the library
    nm = subprocess.run(["nm", "-u", str(tmp_path / "shim.o")], capture_output=True, text=True, check=True).stdout
    called = sorted({ln.split()[-1] for ln in nm.splitlines() if "glass_hip_" in ln})
    assert len(called) >= 15, called
    import ctypes
    L = ctypes.CDLL(glass_amd.LIB_PATH)
    assert all(hasattr(L, s) for s in called), [s for s in called if not hasattr(L, s)]
