"""The wavelet response stage divides by a correctly rounded reciprocal with
one FMA correction (q = a y, q' = q + (a - b q) y with y = RN(1/b)), in
place of a division, on its serial spline sweeps and phase unwrap, and so
do the light-time delays k.x / c of every slow sample (FD, tangent and
wavelet templates); the templates are claimed unchanged bit for bit.  Markstein's theorem says q'
is the correctly rounded a/b (round to nearest, no underflow or overflow);
this checks it against IEEE division, compiled with the system C compiler's
fma, on random operands across 120 binades, on divisors near 1-2 (the
spline pivots' scale), on near-exact quotients and on quotients next to
representable values, and on the unwrap's |u - v| / 2 pi.
"""
import ctypes
import subprocess

import numpy as np
import pytest

SRC = r'''
#include <math.h>
#include <stdint.h>
#include <string.h>
static uint64_t s = 88172645463325252ull;
static uint64_t xr(void) { s ^= s << 13; s ^= s >> 7; s ^= s << 17; return s; }
static double rnd(int emin, int emax) {
    uint64_t m = xr() & ((1ull << 52) - 1);
    int e = emin + (int)(xr() % (uint64_t)(emax - emin + 1));
    uint64_t bits = ((uint64_t)(e + 1023) << 52) | m;
    double d; memcpy(&d, &bits, 8);
    return (xr() & 1) ? -d : d;
}
static double div_by_recip(double a, double b, double y) {
    double q = a * y;
    return fma(fma(-q, b, a), y, q);
}
long check(long n) {
    long bad = 0;
    const double PI2 = 6.283185307179586, INV = 1.0 / PI2;
    for (long k = 0; k < n; k++) {
        double a = rnd(-60, 60), b = fabs(rnd(-60, 60));
        switch (k % 5) {
        case 1: b = 1.0 + (double)(xr() % 1000000) * 1e-6; break;
        case 2: a = rnd(-20, 20) * b; break;
        case 3: {
            uint64_t m = xr() & ((1ull << 52) - 1), bits = (1023ull << 52) | m;
            memcpy(&b, &bits, 8);
            a = nextafter(b * (double)(1 + (xr() % 1000)), INFINITY);
            break;
        }
        case 4: {
            double d = fabs(rnd(-30, 25));
            if (div_by_recip(d, PI2, INV) != d / PI2) bad++;
            continue;
        }
        }
        volatile double y = 1.0 / b;
        if (div_by_recip(a, b, y) != a / b) bad++;
    }
    return bad;
}
/* the light-time quotient k.x / c: numerators across the positions' range
   (|k.x| up to ~2e11 m, down to cancellation near 0) */
long check_c(long n) {
    long bad = 0;
    const double C = 299792458.;
    volatile double yv = 1.0 / C;
    const double y = yv;
    for (long k = 0; k < n; k++) {
        double a = rnd(-40, 45);
        if (div_by_recip(a, C, y) != a / C) bad++;
    }
    return bad;
}
'''


@pytest.fixture(scope="module")
def lib(tmp_path_factory):
    d = tmp_path_factory.mktemp("rd")
    (d / "rd.c").write_text(SRC)
    so = d / "rd.so"
    subprocess.run(["gcc", "-O2", "-ffp-contract=off", "-shared", "-fPIC", str(d / "rd.c"), "-o", str(so), "-lm"],
                   check=True)
    L = ctypes.CDLL(str(so))
    L.check.restype = ctypes.c_long
    L.check.argtypes = [ctypes.c_long]
    L.check_c.restype = ctypes.c_long
    L.check_c.argtypes = [ctypes.c_long]
    return L


def test_reciprocal_fma_division_is_correctly_rounded(lib):
    assert lib.check(10_000_000) == 0


def test_light_time_quotient_is_correctly_rounded(lib):
    assert lib.check_c(20_000_000) == 0


def test_the_check_sees_a_plain_reciprocal_product(tmp_path):
    """Without the correction, a * RN(1/b) differs from a/b often: the check
    above is not vacuous."""
    rng = np.random.default_rng(0)
    a = rng.uniform(1, 2, 100000)
    b = rng.uniform(1, 2, 100000)
    assert np.mean(a * (1.0 / b) != a / b) > 0.05
