/*
 * orc_fft.c -- oracle FFT (test infrastructure only).
 *
 * Restates the arithmetic of the reference's vendored kissfft so that the
 * oracle's slow-series and WDM transforms round the way the reference does:
 *   - factorisation: radix 4 first, then 2, then odd primes 3,5,7,...
 *     (external/kissfft/kiss_fft.c:307-329)
 *   - recursive decimation in time with input stride fstride
 *     (kiss_fft.c:245-302)
 *   - butterflies radix 2/3/4/5/generic with the reference's operation order
 *     (kiss_fft.c:15-240)
 *   - twiddles w_k = exp(-+ 2 pi i k / n) from cos/sin of the phase formed as
 *     -2*pi*k/n (kiss_fft.c:356-363)
 *   - real transforms by packing n real samples as n/2 complex values and a
 *     split step with super-twiddles exp(-i pi ((k+1)/ncfft + 1/2))
 *     (external/kissfft/kiss_fftr.c:20-153)
 * Pinned bitwise against kissfft compiled from the reference sources
 * (oracle/_ref) and against a brute-force DFT in the tests.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "glass_oracle.h"

typedef struct { double r, i; } cpx;

typedef struct
{
    int n, inverse;
    int nfac;
    int radix[64];   /* p_s   */
    int sub[64];     /* m_s = n / (p_1 ... p_s) */
    cpx *tw;
} plan_t;

static void plan_factor(plan_t *pl)
{
    int n = pl->n, p = 4, s = 0;
    double lim = floor(sqrt((double)n));
    do
    {
        while (n % p)
        {
            if (p == 4) p = 2;
            else if (p == 2) p = 3;
            else p += 2;
            if (p > lim) p = n;
        }
        n /= p;
        pl->radix[s] = p;
        pl->sub[s] = n;
        s++;
    } while (n > 1);
    pl->nfac = s;
}

static int plan_init(plan_t *pl, int n, int inverse)
{
    pl->n = n;
    pl->inverse = inverse;
    pl->tw = (cpx *)malloc(sizeof(cpx) * (size_t)n);
    if (!pl->tw) return -1;
    const double pi = 3.141592653589793238462643383279502884197169399375105820974944;
    for (int k = 0; k < n; k++)
    {
        double ph = -2 * pi * k / n;
        if (inverse) ph *= -1;
        pl->tw[k].r = cos(ph);
        pl->tw[k].i = sin(ph);
    }
    plan_factor(pl);
    return 0;
}

static void plan_free(plan_t *pl) { free(pl->tw); }

static inline cpx cmul(cpx a, cpx b)
{
    cpx m;
    m.r = a.r * b.r - a.i * b.i;
    m.i = a.r * b.i + a.i * b.r;
    return m;
}

static void radix2(cpx *F, size_t fs, const plan_t *pl, int m)
{
    const cpx *w = pl->tw;
    for (int u = 0; u < m; u++)
    {
        cpx t = cmul(F[u + m], w[u * fs]);
        F[u + m].r = F[u].r - t.r;
        F[u + m].i = F[u].i - t.i;
        F[u].r += t.r;
        F[u].i += t.i;
    }
}

static void radix4(cpx *F, size_t fs, const plan_t *pl, size_t m)
{
    const cpx *w = pl->tw;
    for (size_t u = 0; u < m; u++)
    {
        cpx a1 = cmul(F[u + m], w[u * fs]);
        cpx a2 = cmul(F[u + 2 * m], w[2 * u * fs]);
        cpx a3 = cmul(F[u + 3 * m], w[3 * u * fs]);
        cpx d02, s13, d13;
        d02.r = F[u].r - a2.r;  d02.i = F[u].i - a2.i;
        F[u].r += a2.r;         F[u].i += a2.i;
        s13.r = a1.r + a3.r;    s13.i = a1.i + a3.i;
        d13.r = a1.r - a3.r;    d13.i = a1.i - a3.i;
        F[u + 2 * m].r = F[u].r - s13.r;
        F[u + 2 * m].i = F[u].i - s13.i;
        F[u].r += s13.r;
        F[u].i += s13.i;
        if (pl->inverse)
        {
            F[u + m].r = d02.r - d13.i;      F[u + m].i = d02.i + d13.r;
            F[u + 3 * m].r = d02.r + d13.i;  F[u + 3 * m].i = d02.i - d13.r;
        }
        else
        {
            F[u + m].r = d02.r + d13.i;      F[u + m].i = d02.i - d13.r;
            F[u + 3 * m].r = d02.r - d13.i;  F[u + 3 * m].i = d02.i + d13.r;
        }
    }
}

static void radix3(cpx *F, size_t fs, const plan_t *pl, size_t m)
{
    const cpx *w = pl->tw;
    cpx e3 = w[fs * m];
    for (size_t u = 0; u < m; u++)
    {
        cpx b1 = cmul(F[u + m], w[u * fs]);
        cpx b2 = cmul(F[u + 2 * m], w[2 * u * fs]);
        cpx s, d;
        s.r = b1.r + b2.r;  s.i = b1.i + b2.i;
        d.r = b1.r - b2.r;  d.i = b1.i - b2.i;
        F[u + m].r = F[u].r - s.r * 0.5;
        F[u + m].i = F[u].i - s.i * 0.5;
        d.r *= e3.i;
        d.i *= e3.i;
        F[u].r += s.r;
        F[u].i += s.i;
        F[u + 2 * m].r = F[u + m].r + d.i;
        F[u + 2 * m].i = F[u + m].i - d.r;
        F[u + m].r -= d.i;
        F[u + m].i += d.r;
    }
}

static void radix5(cpx *F, size_t fs, const plan_t *pl, int m)
{
    const cpx *w = pl->tw;
    cpx ya = w[fs * m], yb = w[fs * 2 * m];
    for (int u = 0; u < m; u++)
    {
        cpx *f0 = F + u, *f1 = F + u + m, *f2 = F + u + 2 * m, *f3 = F + u + 3 * m, *f4 = F + u + 4 * m;
        cpx x0 = *f0;
        cpx x1 = cmul(*f1, w[u * fs]);
        cpx x2 = cmul(*f2, w[2 * u * fs]);
        cpx x3 = cmul(*f3, w[3 * u * fs]);
        cpx x4 = cmul(*f4, w[4 * u * fs]);
        cpx s14, d14, s23, d23, p, q, p2, q2;
        s14.r = x1.r + x4.r;  s14.i = x1.i + x4.i;
        d14.r = x1.r - x4.r;  d14.i = x1.i - x4.i;
        s23.r = x2.r + x3.r;  s23.i = x2.i + x3.i;
        d23.r = x2.r - x3.r;  d23.i = x2.i - x3.i;
        f0->r += s14.r + s23.r;
        f0->i += s14.i + s23.i;
        p.r = x0.r + s14.r * ya.r + s23.r * yb.r;
        p.i = x0.i + s14.i * ya.r + s23.i * yb.r;
        q.r = d14.i * ya.i + d23.i * yb.i;
        q.i = -(d14.r * ya.i) - d23.r * yb.i;
        f1->r = p.r - q.r;  f1->i = p.i - q.i;
        f4->r = p.r + q.r;  f4->i = p.i + q.i;
        p2.r = x0.r + s14.r * yb.r + s23.r * ya.r;
        p2.i = x0.i + s14.i * yb.r + s23.i * ya.r;
        q2.r = -(d14.i * yb.i) + d23.i * ya.i;
        q2.i = d14.r * yb.i - d23.r * ya.i;
        f2->r = p2.r + q2.r;  f2->i = p2.i + q2.i;
        f3->r = p2.r - q2.r;  f3->i = p2.i - q2.i;
    }
}

static void radixg(cpx *F, size_t fs, const plan_t *pl, int m, int p)
{
    const cpx *w = pl->tw;
    int n = pl->n;
    cpx *tmp = (cpx *)malloc(sizeof(cpx) * (size_t)p);
    for (int u = 0; u < m; u++)
    {
        for (int a = 0, k = u; a < p; a++, k += m) tmp[a] = F[k];
        for (int a = 0, k = u; a < p; a++, k += m)
        {
            int ti = 0;
            F[k] = tmp[0];
            for (int b = 1; b < p; b++)
            {
                ti += (int)fs * k;
                if (ti >= n) ti -= n;
                cpx t = cmul(tmp[b], w[ti]);
                F[k].r += t.r;
                F[k].i += t.i;
            }
        }
    }
    free(tmp);
}

/* stage s: output F has length p_s*m_s; input read with stride fs */
static void work(cpx *F, const cpx *in, size_t fs, int s, const plan_t *pl)
{
    const int p = pl->radix[s], m = pl->sub[s];
    if (m == 1)
    {
        for (int a = 0; a < p; a++) F[a] = in[(size_t)a * fs];
    }
    else
    {
        for (int a = 0; a < p; a++) work(F + (size_t)a * m, in + (size_t)a * fs, fs * p, s + 1, pl);
    }
    switch (p)
    {
        case 2: radix2(F, fs, pl, m); break;
        case 3: radix3(F, fs, pl, m); break;
        case 4: radix4(F, fs, pl, m); break;
        case 5: radix5(F, fs, pl, m); break;
        default: radixg(F, fs, pl, m, p); break;
    }
}

int orc_fft(const double *in, double *out, int n, int inverse)
{
    plan_t pl;
    if (n < 1 || plan_init(&pl, n, inverse)) return -1;
    cpx *buf = (cpx *)malloc(sizeof(cpx) * (size_t)n);
    work(buf, (const cpx *)in, 1, 0, &pl);
    memcpy(out, buf, sizeof(cpx) * (size_t)n);
    free(buf);
    plan_free(&pl);
    return 0;
}

void orc_forward_complex_fft(double *data, int n) { orc_fft(data, data, n, 0); }

void orc_inverse_complex_fft(double *data, int n)
{
    orc_fft(data, data, n, 1);
    for (int i = 0; i < 2 * n; i++) data[i] = data[i] / n;
}

static void super_twiddles(cpx *st, int ncfft, int inverse)
{
    for (int i = 0; i < ncfft / 2; i++)
    {
        double ph = -3.14159265358979323846264338327 * ((double)(i + 1) / ncfft + .5);
        if (inverse) ph *= -1;
        st[i].r = cos(ph);
        st[i].i = sin(ph);
    }
}

int orc_fftr(const double *timedata, double *freqdata, int n)
{
    if (n & 1) return -1;
    int nc = n / 2;
    cpx *tmp = (cpx *)malloc(sizeof(cpx) * (size_t)nc);
    cpx *st = (cpx *)malloc(sizeof(cpx) * (size_t)(nc / 2 + 1));
    cpx *F = (cpx *)freqdata;
    super_twiddles(st, nc, 0);
    orc_fft(timedata, (double *)tmp, nc, 0);
    double dr = tmp[0].r, di = tmp[0].i;
    F[0].r = dr + di;
    F[nc].r = dr - di;
    F[nc].i = F[0].i = 0;
    for (int k = 1; k <= nc / 2; k++)
    {
        cpx a = tmp[k], b, e, o, t;
        b.r = tmp[nc - k].r;
        b.i = -tmp[nc - k].i;
        e.r = a.r + b.r;  e.i = a.i + b.i;
        o.r = a.r - b.r;  o.i = a.i - b.i;
        t = cmul(o, st[k - 1]);
        F[k].r = (e.r + t.r) * 0.5;
        F[k].i = (e.i + t.i) * 0.5;
        F[nc - k].r = (e.r - t.r) * 0.5;
        F[nc - k].i = (t.i - e.i) * 0.5;
    }
    free(st);
    free(tmp);
    return 0;
}

int orc_fftri(const double *freqdata, double *timedata, int n)
{
    if (n & 1) return -1;
    int nc = n / 2;
    const cpx *F = (const cpx *)freqdata;
    cpx *tmp = (cpx *)malloc(sizeof(cpx) * (size_t)nc);
    cpx *st = (cpx *)malloc(sizeof(cpx) * (size_t)(nc / 2 + 1));
    super_twiddles(st, nc, 1);
    tmp[0].r = F[0].r + F[nc].r;
    tmp[0].i = F[0].r - F[nc].r;
    for (int k = 1; k <= nc / 2; k++)
    {
        cpx a = F[k], b, e, d, o;
        b.r = F[nc - k].r;
        b.i = -F[nc - k].i;
        e.r = a.r + b.r;  e.i = a.i + b.i;
        d.r = a.r - b.r;  d.i = a.i - b.i;
        o = cmul(d, st[k - 1]);
        tmp[k].r = e.r + o.r;        tmp[k].i = e.i + o.i;
        tmp[nc - k].r = e.r - o.r;   tmp[nc - k].i = e.i - o.i;
        tmp[nc - k].i *= -1;
    }
    orc_fft((const double *)tmp, timedata, nc, 1);
    free(st);
    free(tmp);
    return 0;
}
