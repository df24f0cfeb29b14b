// noise.hip -- instrument-noise model, noise likelihoods and the batched
// instrument-noise MCMC on the device.
//
// Reference (paths relative to the reference root):
//   generate_instrument_noise_model          noise/src/glass_noise_model.c:389-545
//   generate_instrument_noise_model_wavelet  noise/src/glass_noise_model.c:561-610
//   generate_full_covariance_matrix          noise/src/glass_noise_model.c:894-907
//   invert_noise_covariance_matrix           utils/src/glass_math.c:240-289
//   noise_log_likelihood                     noise/src/glass_noise_model.c:1104-1134
//   my_noise_log_likelihood_wavelet          noise/src/glass_noise_model.c:1182-1206
//   my_noise_log_likelihood                  noise/src/glass_noise_model.c:1275-1306
//   noise_instrument_model_mcmc              noise/src/glass_noise_sampler.c:395-592
//   noise_ptmcmc                             noise/src/glass_noise_sampler.c:70-117
//   rand_r_U_0_1 / rand_r_N_0_1              utils/src/glass_utils.c:19-35
//
// Batched layout.  One data segment of N bins is shared by every chain (the
// reference's noise sampler runs all its chains on one data set).  What does
// not depend on a chain is computed once per bin at plan creation: the four
// transfer-function factors of the model (acc, oms, -8 sin^2 x, cos x) and
// the six data products |X|^2 .. Re(Y conj Z).  A chain's likelihood then
// reads 80 bytes per bin (+48 with fixed extra covariance) and does the
// 3x3 model, its inverse and the quadratic form in registers; chains are
// processed in tiles of kTile per workgroup so those bytes are read once per
// tile.  Chains live in ladders of ntemp temperature slots; slot s of ladder
// l holds state index[s] (a permutation within the ladder, changed by PT
// swaps), its own random-number state and temperature, as chain->r[ic],
// chain->temperature[ic] and chain->index[ic] of the reference.
//
// Inverse in the fused likelihood.  invert_noise_covariance_matrix forms
// each entry as sign(num) exp(log|num| - log det) with log det = log d1 +
// log d2 + log d3 from the pivots d1, d2, d3.  The fused kernel keeps the
// same pivots and log-determinant and takes the entries as num / (d1 d2 d3):
// three logarithms per chain and bin instead of fifteen.  The two forms differ
// by the rounding of exp(log|num| - log det), a relative ~|log det| x 1e-16
// (below 1e-13 for LISA covariances); the drop-in inverse
// (glass_hip_invert_noise_covariance) keeps the reference's form.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <vector>

#include "../../include/glass_hip_noise.h"
#include "reduce.h"
#include "rng.h"
#include "runtime.h"

namespace glass {

namespace {

constexpr double kAccUnits = 2.81837551648e-19;  // 1/(2 pi c)^2
constexpr double kOmsUnits = 4.39256635604e-16;  // (2 pi / c)^2
constexpr double kAccFmax = 0.01;                // Hz
constexpr int kNThr = 256;
#ifndef GH_NOISE_TILE
#define GH_NOISE_TILE 4
#endif
#ifndef GH_NOISE_WAVES
#define GH_NOISE_WAVES 1
#endif
// chains per workgroup of the fused likelihood: each bin's transfer factors
// and data products are read once per tile; the tile's chains are computed
// side by side, so their temporaries set the register count
constexpr int kTile = GH_NOISE_TILE;

struct Basis
{
    double acc, oms, tdi, cosx;
};

// transfer-function factors of one bin, in the reference's operation order
// (noise/src/glass_noise_model.c:421-445)
__device__ inline Basis noise_basis(double f, double fstar)
{
    Basis b;
    double x = f / fstar, f2 = f * f;
    b.cosx = cos(x);
    double s = sin(x);
    b.tdi = s * s;
    b.acc = (f < kAccFmax) ? kAccUnits / f2 * (1.0 + pow(0.4e-3 / f, 2)) * (1.0 + pow(f / 8.0e-3, 4)) : 0.0;
    b.oms = kOmsUnits * f2 * (1.0 + pow(2.0e-3 / f, 4));
    b.tdi *= -8.0;
    return b;
}

// Three-channel model of link parameters p = sacc[6], soms[6] (links 12, 21,
// 23, 32, 13, 31), normalised by 4: c = XX, YY, ZZ, XY, XZ, YZ
// (noise/src/glass_noise_model.c:468-544).
__device__ inline void model_c3(const double* p, const Basis& b, bool lowf, double* c)
{
    double O12 = p[6] * b.oms, O21 = p[7] * b.oms, O23 = p[8] * b.oms;
    double O32 = p[9] * b.oms, O13 = p[10] * b.oms, O31 = p[11] * b.oms;
    double cxx = 0., cyy = 0., czz = 0., cxy = 0., cxz = 0., cyz = 0.;
    cxx += (O12 + O21 + O13 + O31) * .25;
    cyy += (O23 + O32 + O21 + O12) * .25;
    czz += (O31 + O13 + O32 + O23) * .25;
    cxy += (O12 + O21) * .5;
    cxz += (O13 + O31) * .5;
    cyz += (O32 + O23) * .5;
    if (lowf)
    {
        double S12 = p[0] * b.acc, S21 = p[1] * b.acc, S23 = p[2] * b.acc;
        double S32 = p[3] * b.acc, S13 = p[4] * b.acc, S31 = p[5] * b.acc;
        cxx += ((S12 + S13) + ((S21 + S31)) * b.cosx * b.cosx);
        cyy += ((S23 + S21) + ((S32 + S12)) * b.cosx * b.cosx);
        czz += ((S31 + S32) + ((S13 + S23)) * b.cosx * b.cosx);
        cxy += 2. * (S12 + S21);
        cxz += 2. * (S13 + S31);
        cyz += 2. * (S32 + S23);
    }
    cxx *= -2. * b.tdi;
    cyy *= -2. * b.tdi;
    czz *= -2. * b.tdi;
    cxy *= b.tdi * b.cosx;
    cxz *= b.tdi * b.cosx;
    cyz *= b.tdi * b.cosx;
    c[0] = cxx / 4.0;
    c[1] = cyy / 4.0;
    c[2] = czz / 4.0;
    c[3] = cxy / 4.0;
    c[4] = cxz / 4.0;
    c[5] = cyz / 4.0;
}

// XYZnoise_FF / AEnoise_FF (utils/src/glass_lisa.c:1121-1145)
__device__ inline double xyz_ff(double fstar, double f, double Spm, double Sop)
{
    double x = f / fstar, cx = cos(x), s = sin(x);
    return 16. * (s * s) * (2. * (1.0 + cx * cx) * Spm + Sop);
}
__device__ inline double ae_ff(double fstar, double f, double Spm, double Sop)
{
    double x = f / fstar, cx = cos(x), c2x = cos(2. * x), s = sin(x);
    return 8. * (s * s) * (2. * Spm * (3. + 2. * cx + c2x) + Sop * (2. + cx));
}

struct Links
{
    double p[12];  // sacc[6], soms[6]
};

// generate_instrument_noise_model for Nchannel 1, 2, 3; C in place (entries
// the reference does not write stay as they are; its final /4 also touches
// C[1][0] of two channels)
__global__ void inst_model_kernel(double fstar, const double* __restrict__ f, int N, int nch, Links L,
                                  double* __restrict__ C)
{
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= N) return;
    const size_t P = (size_t)N;
    auto cp = [&](int i, int j) -> double& { return C[((size_t)i * nch + j) * P + n]; };
    const double fn = f[n];
    if (fn == 0.0)
    {
        if (nch > 0) cp(0, 0) = 0.0;
        if (nch > 1) cp(0, 1) = 0.0;
        if (nch > 2)
        {
            cp(0, 2) = 0.0;
            cp(1, 2) = 0.0;
            cp(2, 2) = 0.0;
        }
        return;
    }
    const Basis b = noise_basis(fn, fstar);
    if (nch == 3)
    {
        double c[6];
        model_c3(L.p, b, fn < kAccFmax, c);
        cp(0, 0) = c[0];
        cp(1, 1) = c[1];
        cp(2, 2) = c[2];
        cp(0, 1) = c[3];
        cp(0, 2) = c[4];
        cp(1, 2) = c[5];
        cp(1, 0) = c[3];
        cp(2, 0) = c[4];
        cp(2, 1) = c[5];
        return;
    }
    const double Spm = L.p[0] * b.acc, Sop = L.p[6] * b.oms;
    if (nch == 1)
    {
        cp(0, 0) = xyz_ff(fstar, fn, Spm, Sop) / 4.0;
    }
    else
    {
        double v = ae_ff(fstar, fn, Spm, Sop);
        cp(0, 0) = v / 4.0;
        cp(1, 1) = v / 4.0;
        cp(0, 1) = 0.0;
        cp(1, 0) = cp(1, 0) / 4.0;
    }
}

// generate_instrument_noise_model_wavelet: layer-centre frequencies
// (imin + k) df, three channels, divided by ND, planes of stride ldc
__global__ void inst_model_wavelet_kernel(double fstar, int imin, double df, int Na, double inv_ND, Links L,
                                          double* __restrict__ C, int ldc)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= Na) return;
    const double fk = (double)(imin + k) * df;
    double c[6];
    if (fk == 0.0)
    {
        // the grid model leaves C[1][1] (and the lower triangle) of an f = 0
        // layer at its freshly allocated zero
        for (int i = 0; i < 6; i++) c[i] = 0.0;
    }
    else
    {
        const Basis b = noise_basis(fk, fstar);
        model_c3(L.p, b, fk < kAccFmax, c);
    }
    const size_t P = (size_t)ldc;
    const int ii[6] = {0, 4, 8, 1, 2, 5}, jj[6] = {0, 4, 8, 3, 6, 7};
    for (int e = 0; e < 6; e++)
    {
        double v = c[e] * inv_ND;
        C[ii[e] * P + k] = v;
        C[jj[e] * P + k] = v;
    }
}

// per-bin my_noise_log_likelihood term with a given inverse (no constant)
__global__ __launch_bounds__(kNThr) void nlogl_given_kernel(const double2* __restrict__ X, const double2* __restrict__ Y,
                                                            const double2* __restrict__ Z,
                                                            const double* __restrict__ iC,
                                                            const double* __restrict__ ld, int N, int chunk,
                                                            double* __restrict__ part)
{
    const int b0 = blockIdx.x * chunk, b1 = min(N, b0 + chunk);
    const size_t P = (size_t)N;
    double acc = 0.0;
    for (int n = b0 + threadIdx.x; n < b1; n += kNThr)
    {
        double2 x = X[n], y = Y[n], z = Z[n];
        double xx = x.x * x.x + x.y * x.y, yy = y.x * y.x + y.y * y.y, zz = z.x * z.x + z.y * z.y;
        double xy = x.x * y.x + x.y * y.y, xz = x.x * z.x + x.y * z.y, yz = y.x * z.x + y.y * z.y;
        double t = 0.0;
        t -= xx * iC[0 * P + n];
        t -= yy * iC[4 * P + n];
        t -= zz * iC[8 * P + n];
        t -= xy * iC[1 * P + n] * 2;
        t -= xz * iC[2 * P + n] * 2;
        t -= yz * iC[5 * P + n] * 2;
        t -= ld[n];
        acc += t;
    }
    acc = block_sum<kNThr>(acc);
    if (threadIdx.x == 0) part[blockIdx.x] = acc;
}

// per-pixel my_noise_log_likelihood_wavelet term (no constant); pixels
// k = i + j NT of Nlayer layers, planes of stride ldc
__global__ __launch_bounds__(kNThr) void wnlogl_given_kernel(const double* __restrict__ X, const double* __restrict__ Y,
                                                             const double* __restrict__ Z,
                                                             const double* __restrict__ iC,
                                                             const double* __restrict__ ld, int n, int ldc,
                                                             int chunk, double* __restrict__ part)
{
    const int b0 = blockIdx.x * chunk, b1 = min(n, b0 + chunk);
    const size_t P = (size_t)ldc;
    double acc = 0.0;
    for (int k = b0 + threadIdx.x; k < b1; k += kNThr)
    {
        double wx = X[k], wy = Y[k], wz = Z[k];
        double s = 0.0;
        s -= 0.5 * iC[0 * P + k] * (wx * wx);
        s -= 0.5 * iC[4 * P + k] * (wy * wy);
        s -= 0.5 * iC[8 * P + k] * (wz * wz);
        s -= iC[1 * P + k] * (wx * wy);
        s -= iC[2 * P + k] * (wx * wz);
        s -= iC[5 * P + k] * (wy * wz);
        s -= 0.5 * ld[k];
        acc += s;
    }
    acc = block_sum<kNThr>(acc);
    if (threadIdx.x == 0) part[blockIdx.x] = acc;
}

__global__ void sum_parts_kernel(const double* __restrict__ part, int np, double c, double* out)
{
    if (threadIdx.x == 0 && blockIdx.x == 0)
    {
        double s = 0.0;
        for (int i = 0; i < np; i++) s += part[i];
        *out = s + c;
    }
}

// ------------------------------------------------------- batched plan ---

// per bin: transfer factors and data products
__global__ void noise_prep_kernel(const double* __restrict__ f, const double2* __restrict__ X,
                                  const double2* __restrict__ Y, const double2* __restrict__ Z, int N, double fstar,
                                  double* __restrict__ basis, double* __restrict__ prod)
{
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= N) return;
    const size_t P = (size_t)N;
    const Basis b = noise_basis(f[n], fstar);
    basis[0 * P + n] = b.acc;
    basis[1 * P + n] = b.oms;
    basis[2 * P + n] = b.tdi;
    basis[3 * P + n] = b.cosx;
    double2 x = X[n], y = Y[n], z = Z[n];
    prod[0 * P + n] = x.x * x.x + x.y * x.y;
    prod[1 * P + n] = y.x * y.x + y.y * y.y;
    prod[2 * P + n] = z.x * z.x + z.y * z.y;
    prod[3 * P + n] = x.x * y.x + x.y * y.y;
    prod[4 * P + n] = x.x * z.x + x.y * z.y;
    prod[5 * P + n] = y.x * z.x + y.y * z.y;
}

// Fused model -> (+ fixed covariance) -> inverse -> likelihood for kTile
// chains per workgroup over one chunk of bins; partial sums part[chunk][i].
// Every bin must have f > 0 (checked at plan creation).
__global__ __launch_bounds__(kNThr, GH_NOISE_WAVES) void noise_logl_kernel(int N, int chunk, const double* __restrict__ basis,
                                                           const double* __restrict__ prod,
                                                           const double* __restrict__ cadd,
                                                           const double* __restrict__ params, int n,
                                                           double* __restrict__ part)
{
    // the tile's parameters are read with workgroup-uniform addresses:
    // scalar loads into SGPRs (through LDS they took vector registers,
    // 12 per chain, held across the bin loop)
    const int c0 = blockIdx.y * kTile;
    const int nc = min(kTile, n - c0);
    const double* sp = params + (size_t)c0 * 12;
    const int b0 = blockIdx.x * chunk, b1 = min(N, b0 + chunk);
    const size_t P = (size_t)N;
    double acc[kTile];
#pragma unroll
    for (int c = 0; c < kTile; c++) acc[c] = 0.0;
    for (int j = b0 + threadIdx.x; j < b1; j += kNThr)
    {
        Basis b;
        b.acc = basis[0 * P + j];
        b.oms = basis[1 * P + j];
        b.tdi = basis[2 * P + j];
        b.cosx = basis[3 * P + j];
        const bool lowf = b.acc != 0.0;
        double d[6], a[6];
#pragma unroll
        for (int e = 0; e < 6; e++) d[e] = prod[e * P + j];
        if (cadd)
        {
#pragma unroll
            for (int e = 0; e < 6; e++) a[e] = cadd[e * P + j];
        }
#pragma unroll
        for (int c = 0; c < kTile; c++)
        {
            if (c >= nc) break;
            double m[6];
            model_c3(sp + 12 * c, b, lowf, m);
            if (cadd)
            {
#pragma unroll
                for (int e = 0; e < 6; e++) m[e] += a[e];
            }
            const double cxx = m[0], cyy = m[1], czz = m[2], cxy = m[3], cxz = m[4], cyz = m[5];
            // the pivots of the reference's elimination; log det C and
            // 1/det from their product (one logarithm and three divisions
            // per chain and bin; det ~ 1e-120 for LISA noise, far from the
            // double range's ends)
            const double d1 = cxx, i1 = 1.0 / d1;
            const double d2 = cyy - cxy * cxy * i1;
            const double t = cyz - cxy * cxz * i1;
            const double d3 = (czz - cxz * cxz * i1) - (t * t) / d2;
            const double det = d1 * d2 * d3;
            const double ldet = log(det);
            const double r = 1.0 / det;
            const double ixx = (cyy * czz - cyz * cyz) * r, iyy = (czz * cxx - cxz * cxz) * r;
            const double izz = (cxx * cyy - cxy * cxy) * r, ixy = (cxz * cyz - czz * cxy) * r;
            const double ixz = (cxy * cyz - cxz * cyy) * r, iyz = (cxy * cxz - cxx * cyz) * r;
            double s = 0.0;
            s -= d[0] * ixx;
            s -= d[1] * iyy;
            s -= d[2] * izz;
            s -= d[3] * ixy * 2;
            s -= d[4] * ixz * 2;
            s -= d[5] * iyz * 2;
            s -= ldet;
            acc[c] += s;
        }
    }
#pragma unroll
    for (int c = 0; c < kTile; c++)
    {
        double v = block_sum<kNThr>(acc[c]);
        if (threadIdx.x == 0 && c < nc) part[(size_t)blockIdx.x * n + c0 + c] = v;
    }
}

// One wave per chain: lane j sums the chunks k = j, j + 64, ... in order,
// then a fixed shuffle tree adds the lanes (reproducible run to run; the
// chunk partials of a small plan number in the hundreds, which a serial
// sum per chain would read one dependent load at a time).
__global__ __launch_bounds__(256) void noise_finish_kernel(const double* __restrict__ part, int nchunk, int n,
                                                           double cst, double* __restrict__ out)
{
    const int i = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
    if (i >= n) return;  // uniform over the wave
    double s = 0.0;
    for (int k = lane; k < nchunk; k += 64) s += part[(size_t)k * n + i];
    s = wave_sum(s);
    if (lane == 0) out[i] = s + cst;
}

// a link index; U = 1 exactly would give 6, one past the six links
__device__ inline int link_draw(unsigned* s)
{
    int i = (int)(rand_U(s) * 6.0);
    return i > 5 ? 5 : i;
}

// -------------------------------------------------- MCMC step kernels ---

struct NoiseState
{
    double* x;      // [nslot][12] by state
    double* xlogl;  // [nslot] by state
    double* y;      // [nslot][12] by slot (trial model of the slot's state)
    double* ylogl;  // [nslot] by slot
    double* logH;   // [nslot]
    double* logPy;  // [nslot]
    unsigned* seed;
    const double* temp;
    int* index;
    int* nacc;
    int* swaps;
    int nslot;
};

// copy_instrument_model(model_x, model_y) and the accumulators of one call
__global__ void noise_begin_kernel(NoiseState S)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= S.nslot) return;
    const int st = S.index[s];
    for (int k = 0; k < 12; k++) S.y[(size_t)s * 12 + k] = S.x[(size_t)st * 12 + k];
    S.ylogl[s] = S.xlogl[st];
    S.logH[s] = 0.0;
    S.logPy[s] = 0.0;
}

// proposal of noise/src/glass_noise_sampler.c:463-566
__global__ void noise_propose_kernel(NoiseState S)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= S.nslot) return;
    const double Sacc = 9.00e-30, Soms = 2.25e-22;
    const double Sacc_min = 9.00e-30 / 10, Sacc_max = 9.00e-30 * 10;
    const double Soms_min = 2.25e-22 / 10, Soms_max = 2.25e-22 * 10;
    unsigned seed = S.seed[s];
    const int st = S.index[s];
    const double* xa = S.x + (size_t)st * 12;
    const double* xo = xa + 6;
    double* ya = S.y + (size_t)s * 12;
    double* yo = ya + 6;
    double logPy = S.logPy[s];
    double scale;
    if (rand_U(&seed) > 0.75)
        scale = 1;
    else if (rand_U(&seed) > 0.5)
        scale = 0.1;
    else if (rand_U(&seed) > 0.25)
        scale = 0.01;
    else
        scale = 0.001;
    const int type = (int)(rand_U(&seed) * 3.);
    if (type == 0)
    {
        int i = link_draw(&seed);
        ya[i] = xa[i] + scale * Sacc * rand_N(&seed);
        yo[i] = xo[i] + scale * Soms * rand_N(&seed);
        if (i % 2 == 0)
            yo[i + 1] = yo[i];
        else
            yo[i - 1] = yo[i];
        if (ya[i] < Sacc_min || ya[i] > Sacc_max) logPy = -INFINITY;
        if (yo[i] < Soms_min || yo[i] > Soms_max) logPy = -INFINITY;
    }
    else if (type == 1)
    {
        int i = link_draw(&seed), j;
        do
            j = link_draw(&seed);
        while (i == j);
        double aj = scale * Sacc * rand_N(&seed);
        double oj = scale * Soms * rand_N(&seed);
        if (abs(i - j) == 1)
        {
            ya[i] = xa[i] + aj;
            ya[j] = xa[j] - aj;
            yo[i] = xo[i] + oj;
        }
        else
        {
            ya[i] = xa[i] + aj;
            ya[j] = xa[j] + aj;
            yo[i] = xo[i] + oj;
            yo[j] = xo[j] - oj;
        }
        if (i % 2 == 0)
            yo[i + 1] = yo[i];
        else
            yo[i - 1] = yo[i];
        if (j % 2 == 0)
            yo[j + 1] = yo[j];
        else
            yo[j - 1] = yo[j];
        if (ya[i] < Sacc_min || ya[i] > Sacc_max) logPy = -INFINITY;
        if (ya[j] < Sacc_min || ya[j] > Sacc_max) logPy = -INFINITY;
        if (yo[i] < Soms_min || yo[i] > Soms_max) logPy = -INFINITY;
        if (yo[j] < Soms_min || yo[j] > Soms_max) logPy = -INFINITY;
    }
    else if (type == 2)
    {
        // correlation_matrix of :436-461: +1 where both links or neither
        // are in {12, 23, 31} (indices 0, 2, 5), -1 otherwise
        const int grp[6] = {1, 0, 1, 0, 0, 1};
        double jv[6];
        for (int i = 0; i < 6; i++)
        {
            jv[i] = scale * Sacc * rand_N(&seed);
            ya[i] = xa[i];
        }
        for (int i = 0; i < 6; i++)
        {
            for (int j = 0; j < 6; j++) ya[i] += (grp[i] == grp[j] ? 1.0 : -1.0) * jv[j];
            if (ya[i] < Sacc_min || ya[i] > Sacc_max) logPy = -INFINITY;
        }
    }
    // type 3 (U = 1 exactly) matches no case of the reference's switch
    S.logPy[s] = logPy;
    S.seed[s] = seed;
}

// A move's likelihood sums, Hastings ratio and accept (:569-585) in one
// launch: one wave per slot adds the slot's chunk partials as
// noise_finish_kernel does, then its lane 0 takes the accept step.  logH and
// logPy carry over the ten moves of one call, as in the reference, where
// both are declared outside its loop.
__global__ __launch_bounds__(256) void noise_finish_accept_kernel(const double* __restrict__ part, int nchunk,
                                                                  double cst, NoiseState S)
{
    const int s = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
    if (s >= S.nslot) return;  // uniform over the wave
    double v = 0.0;
    for (int k = lane; k < nchunk; k += 64) v += part[(size_t)k * S.nslot + s];
    v = wave_sum(v);
    if (lane != 0) return;
    S.ylogl[s] = v + cst;
    unsigned seed = S.seed[s];
    const int st = S.index[s];
    const double logPy = S.logPy[s];
    double logH = S.logH[s];
    if (logPy > -INFINITY) logH += (S.ylogl[s] - S.xlogl[st]) / S.temp[s];
    logH += logPy - 0.0;
    const double loga = log(rand_U(&seed));
    if (logH > loga)
    {
        for (int k = 0; k < 12; k++) S.x[(size_t)st * 12 + k] = S.y[(size_t)s * 12 + k];
        S.xlogl[st] = S.ylogl[s];
        S.nacc[s] += 1;
    }
    S.logH[s] = logH;
    S.seed[s] = seed;
}

// noise_ptmcmc over each ladder of ntemp slots, one thread per ladder
__global__ void noise_ptswap_kernel(NoiseState S, int nladder, int ntemp)
{
    const int l = blockIdx.x * blockDim.x + threadIdx.x;
    if (l >= nladder) return;
    const int o = l * ntemp;
    for (int b = ntemp - 1; b > 0; b--)
    {
        const int a = b - 1;
        const int olda = S.index[o + a], oldb = S.index[o + b];
        const double heat1 = S.temp[o + a], heat2 = S.temp[o + b];
        const double logL1 = S.xlogl[olda], logL2 = S.xlogl[oldb];
        unsigned seed = S.seed[o + a];
        if (rand_U(&seed) < 1.0)
        {
            const double dlogL = logL2 - logL1;
            const double H = (heat2 - heat1) / (heat2 * heat1);
            const double alpha = exp(dlogL * H);
            const double beta = rand_U(&seed);
            if (alpha >= beta)
            {
                S.index[o + a] = oldb;
                S.index[o + b] = olda;
                S.swaps[o + a] += 1;
            }
        }
        S.seed[o + a] = seed;
    }
}

// device rand_r stream, for checking it against the C library
__global__ void rand_stream_kernel(unsigned seed, int n, double* out_u, double* out_n)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    unsigned s = seed;
    for (int i = 0; i < n; i++) out_u[i] = rand_U(&s);
    for (int i = 0; i < n; i++) out_n[i] = rand_N(&s);
}

}  // namespace

}  // namespace glass

using namespace glass;

// ---------------------------------------------------------- drop-ins ---

static int staged_call(size_t in_bytes, size_t out_bytes, ThreadCtx** pcx)
{
    int rc;
    if ((rc = ctx_get(pcx))) return rc;
    size_t need = (in_bytes > out_bytes ? in_bytes : out_bytes) + 4096;
    if ((rc = (*pcx)->dbuf.ensure(in_bytes + out_bytes + 4096))) return rc;
    if ((rc = (*pcx)->hbuf.ensure(need))) return rc;
    return (*pcx)->dpart.ensure(sizeof(double) * 1024);
}

extern "C" int glass_hip_instrument_noise_model(double fstar, const double* f, int N, int nch, const double* sacc,
                                                const double* soms, double* C)
{
    if (N <= 0 || nch < 1 || nch > 3 || !f || !sacc || !soms || !C || !(fstar > 0))
    {
        set_error("instrument_noise_model: bad arguments (N %d, nch %d)", N, nch);
        return GLASS_HIP_EINVAL;
    }
    ThreadCtx* cx;
    const size_t pb = sizeof(double) * (size_t)N, cb = pb * nch * nch;
    int rc;
    if ((rc = staged_call(pb + cb, cb, &cx))) return rc;
    double* df = (double*)cx->dbuf.p;
    double* dC = df + N;
    memcpy(cx->hbuf.p, f, pb);
    memcpy((char*)cx->hbuf.p + pb, C, cb);
    GH_CHECK(hipMemcpyAsync(df, cx->hbuf.p, pb + cb, hipMemcpyHostToDevice, cx->stream));
    Links L;
    for (int i = 0; i < 6; i++)
    {
        L.p[i] = sacc[i];
        L.p[6 + i] = soms[i];
    }
    inst_model_kernel<<<(N + 255) / 256, 256, 0, cx->stream>>>(fstar, df, N, nch, L, dC);
    GH_CHECK(hipGetLastError());
    GH_CHECK(hipMemcpyAsync(cx->hbuf.p, dC, cb, hipMemcpyDeviceToHost, cx->stream));
    GH_CHECK(hipStreamSynchronize(cx->stream));
    memcpy(C, cx->hbuf.p, cb);
    return GLASS_HIP_OK;
}

extern "C" int glass_hip_instrument_noise_model_wavelet(double fstar, const double* f, int N, double df, int ND,
                                                        const double* sacc, const double* soms, double* C,
                                                        int* nactive)
{
    if (N <= 0 || !f || !sacc || !soms || !C || !(df > 0) || ND <= 0 || !(fstar > 0))
    {
        set_error("instrument_noise_model_wavelet: bad arguments");
        return GLASS_HIP_EINVAL;
    }
    const int imin = (int)round(f[0] / df);
    const int imax = (int)round(f[N - 1] / df) + 1;
    const int Na = imax - imin;
    if (Na <= 0 || Na > N)
    {
        set_error("instrument_noise_model_wavelet: %d active layers for %d model bins", Na, N);
        return GLASS_HIP_EINVAL;
    }
    ThreadCtx* cx;
    const size_t cb = sizeof(double) * 9 * (size_t)N;
    int rc;
    if ((rc = staged_call(cb, cb, &cx))) return rc;
    double* dC = (double*)cx->dbuf.p;
    memcpy(cx->hbuf.p, C, cb);
    GH_CHECK(hipMemcpyAsync(dC, cx->hbuf.p, cb, hipMemcpyHostToDevice, cx->stream));
    Links L;
    for (int i = 0; i < 6; i++)
    {
        L.p[i] = sacc[i];
        L.p[6 + i] = soms[i];
    }
    inst_model_wavelet_kernel<<<(Na + 255) / 256, 256, 0, cx->stream>>>(fstar, imin, df, Na, 1.0 / (double)ND, L,
                                                                         dC, N);
    GH_CHECK(hipGetLastError());
    GH_CHECK(hipMemcpyAsync(cx->hbuf.p, dC, cb, hipMemcpyDeviceToHost, cx->stream));
    GH_CHECK(hipStreamSynchronize(cx->stream));
    memcpy(C, cx->hbuf.p, cb);
    if (nactive) *nactive = Na;
    return GLASS_HIP_OK;
}

static int red_grid(int n, int* chunk)
{
    int nb = (n + 4 * kNThr - 1) / (4 * kNThr);
    if (nb < 1) nb = 1;
    if (nb > 1024) nb = 1024;
    *chunk = (n + nb - 1) / nb;
    return nb;
}

static double given_logl(const double* X, const double* Y, const double* Z, const double* invC, const double* ld,
                         int N, int mode, int NT, int Nlayer, int Ntot, int ldc)
{
    // mode 0: my_noise_log_likelihood, 1: noise_log_likelihood, 2: wavelet
    ThreadCtx* cx;
    const bool wav = mode == 2;
    const int n = wav ? NT * Nlayer : N;
    const int P = wav ? ldc : N;
    const size_t chb = sizeof(double) * (size_t)(wav ? n : 2 * n);
    const size_t icb = sizeof(double) * 9 * (size_t)P, ldb = sizeof(double) * (size_t)P;
    if (staged_call(3 * chb + icb + ldb, 64, &cx)) return NAN;
    char* h = (char*)cx->hbuf.p;
    memcpy(h, X, chb);
    memcpy(h + chb, Y, chb);
    memcpy(h + 2 * chb, Z, chb);
    memcpy(h + 3 * chb, invC, icb);
    memcpy(h + 3 * chb + icb, ld, ldb);
    char* d = (char*)cx->dbuf.p;
    if (hipMemcpyAsync(d, h, 3 * chb + icb + ldb, hipMemcpyHostToDevice, cx->stream) != hipSuccess)
    {
        set_error("noise likelihood: staging copy failed");
        return NAN;
    }
    int chunk;
    const int nb = red_grid(n, &chunk);
    double* part = (double*)cx->dpart.p;
    double* dres = (double*)(d + 3 * chb + icb + ldb);
    const double* diC = (const double*)(d + 3 * chb);
    const double* dld = (const double*)(d + 3 * chb + icb);
    double cst;
    if (wav)
    {
        wnlogl_given_kernel<<<nb, kNThr, 0, cx->stream>>>((const double*)d, (const double*)(d + chb),
                                                          (const double*)(d + 2 * chb), diC, dld, n, P, chunk, part);
        cst = -0.5 * 3 * Ntot * log(2 * M_PI);
    }
    else
    {
        nlogl_given_kernel<<<nb, kNThr, 0, cx->stream>>>((const double2*)d, (const double2*)(d + chb),
                                                         (const double2*)(d + 2 * chb), diC, dld, n, chunk, part);
        cst = mode == 0 ? -3.0 * n * log(2 * M_PI) : 0.0;
    }
    sum_parts_kernel<<<1, 64, 0, cx->stream>>>(part, nb, cst, dres);
    double out = NAN;
    if (hipGetLastError() != hipSuccess ||
        hipMemcpyAsync(cx->hbuf.p, dres, sizeof(double), hipMemcpyDeviceToHost, cx->stream) != hipSuccess ||
        hipStreamSynchronize(cx->stream) != hipSuccess)
    {
        set_error("noise likelihood: kernel failed");
        return NAN;
    }
    memcpy(&out, cx->hbuf.p, sizeof(double));
    return out;
}

extern "C" double glass_hip_my_noise_log_likelihood(const double* X, const double* Y, const double* Z,
                                                    const double* invC, const double* logdetC, int N)
{
    if (N <= 0 || !X || !Y || !Z || !invC || !logdetC)
    {
        set_error("my_noise_log_likelihood: bad arguments");
        return NAN;
    }
    return given_logl(X, Y, Z, invC, logdetC, N, 0, 0, 0, 0, 0);
}

extern "C" double glass_hip_noise_log_likelihood(const double* X, const double* Y, const double* Z,
                                                 const double* invC, const double* logdetC, int N)
{
    if (N <= 0 || !X || !Y || !Z || !invC || !logdetC)
    {
        set_error("noise_log_likelihood: bad arguments");
        return NAN;
    }
    return given_logl(X, Y, Z, invC, logdetC, N, 1, 0, 0, 0, 0);
}

extern "C" double glass_hip_my_noise_log_likelihood_wavelet(const double* X, const double* Y, const double* Z,
                                                            const double* invC, const double* logdetC, int NT,
                                                            int Nlayer, int Ntot, int ldc)
{
    if (NT <= 0 || Nlayer <= 0 || ldc < NT * Nlayer || !X || !Y || !Z || !invC || !logdetC)
    {
        set_error("my_noise_log_likelihood_wavelet: bad arguments");
        return NAN;
    }
    return given_logl(X, Y, Z, invC, logdetC, 0, 2, NT, Nlayer, Ntot, ldc);
}

// ------------------------------------------------------ batched plan ---

struct glass_hip_noise_plan
{
    int device = 0;
    int N = 0, nladder = 0, ntemp = 0, nslot = 0;
    double fstar = 0.0;
    int nchunk = 0, chunk = 0;
    std::vector<void*> allocs;
    double *basis = nullptr, *prod = nullptr, *cadd = nullptr, *part = nullptr;
    double *x = nullptr, *xlogl = nullptr, *y = nullptr, *ylogl = nullptr, *logH = nullptr, *logPy = nullptr,
           *temp = nullptr;
    unsigned* seed = nullptr;
    int *index = nullptr, *nacc = nullptr, *swaps = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    int g_calls = -1, g_swap = -1;
};

static int nalloc(glass_hip_noise_plan* P, void* pp, size_t bytes)
{
    void** p = (void**)pp;
    if (hipMalloc(p, bytes ? bytes : 8) != hipSuccess)
    {
        set_error("noise plan: hipMalloc of %zu bytes failed", bytes);
        return GLASS_HIP_ERUNTIME;
    }
    P->allocs.push_back(*p);
    return GLASS_HIP_OK;
}

static NoiseState state_of(glass_hip_noise_plan* P)
{
    return NoiseState{P->x,    P->xlogl, P->y,     P->ylogl, P->logH,  P->logPy, P->seed,
                      P->temp, P->index, P->nacc, P->swaps, P->nslot};
}

static void chunking(glass_hip_noise_plan* P)
{
    // about 2048 workgroups over the chain tiles of a full plan
    const int ntile = (P->nslot + kTile - 1) / kTile;
    int want = (2048 + ntile - 1) / ntile;
    int maxc = (P->N + kNThr - 1) / kNThr;
    if (want < 1) want = 1;
    if (want > maxc) want = maxc;
    P->chunk = (P->N + want - 1) / want;
    P->nchunk = (P->N + P->chunk - 1) / P->chunk;
}

static int launch_logl(glass_hip_noise_plan* P, const double* params, int n, double* out, hipStream_t s)
{
    dim3 grid(P->nchunk, (n + kTile - 1) / kTile);
    noise_logl_kernel<<<grid, kNThr, 0, s>>>(P->N, P->chunk, P->basis, P->prod, P->cadd, params, n, P->part);
    noise_finish_kernel<<<(n + 3) / 4, 256, 0, s>>>(P->part, P->nchunk, n, -3.0 * P->N * log(2 * M_PI), out);
    GH_CHECK(hipGetLastError());
    return GLASS_HIP_OK;
}

static int drop_graph(glass_hip_noise_plan* P)
{
    if (P->gexec) GH_CHECK(hipGraphExecDestroy(P->gexec));
    if (P->graph) GH_CHECK(hipGraphDestroy(P->graph));
    P->gexec = nullptr;
    P->graph = nullptr;
    P->g_calls = P->g_swap = -1;
    return GLASS_HIP_OK;
}

extern "C" int glass_hip_noise_plan_create(glass_hip_noise_plan** out, int N, int nladder, int ntemp, double fstar,
                                           const double* f, const double* X, const double* Y, const double* Z,
                                           const double* Cadd, void* stream)
{
    if (!out || N <= 0 || nladder <= 0 || ntemp <= 0 || (long long)nladder * ntemp > (1 << 24) || !f || !X || !Y ||
        !Z || !(fstar > 0))
    {
        set_error("noise_plan_create: bad arguments (N %d, ladders %d, temperatures %d)", N, nladder, ntemp);
        return GLASS_HIP_EINVAL;
    }
    for (int n = 0; n < N; n++)
        if (!(f[n] > 0))
        {
            set_error("noise_plan_create: bin %d has frequency %g; the fused likelihood needs f > 0", n, f[n]);
            return GLASS_HIP_EINVAL;
        }
    *out = nullptr;
    auto* P = new glass_hip_noise_plan();
    GH_CHECK(hipGetDevice(&P->device));
    P->N = N;
    P->nladder = nladder;
    P->ntemp = ntemp;
    P->nslot = nladder * ntemp;
    P->fstar = fstar;
    chunking(P);
    hipStream_t s = (hipStream_t)stream;
    const size_t pb = sizeof(double) * (size_t)N;
    const int S = P->nslot;
    int rc = 0;
    double *df = nullptr, *dd = nullptr;
    if ((rc = nalloc(P, &P->basis, 4 * pb)) || (rc = nalloc(P, &P->prod, 6 * pb)) ||
        (Cadd && (rc = nalloc(P, &P->cadd, 6 * pb))) ||
        (rc = nalloc(P, &P->part, sizeof(double) * (size_t)P->nchunk * S)) ||
        (rc = nalloc(P, &P->x, sizeof(double) * 12 * S)) || (rc = nalloc(P, &P->xlogl, sizeof(double) * S)) ||
        (rc = nalloc(P, &P->y, sizeof(double) * 12 * S)) || (rc = nalloc(P, &P->ylogl, sizeof(double) * S)) ||
        (rc = nalloc(P, &P->logH, sizeof(double) * S)) || (rc = nalloc(P, &P->logPy, sizeof(double) * S)) ||
        (rc = nalloc(P, &P->temp, sizeof(double) * S)) || (rc = nalloc(P, &P->seed, sizeof(unsigned) * S)) ||
        (rc = nalloc(P, &P->index, sizeof(int) * S)) || (rc = nalloc(P, &P->nacc, sizeof(int) * S)) ||
        (rc = nalloc(P, &P->swaps, sizeof(int) * S)) || (rc = nalloc(P, &df, pb)) || (rc = nalloc(P, &dd, 6 * pb)))
    {
        glass_hip_noise_plan_destroy(P);
        return rc;
    }
    // staging: f, X, Y, Z (interleaved complex) and the fixed covariance
    std::vector<double> cadd6;
    if (Cadd)
    {
        cadd6.resize(6 * (size_t)N);
        const int src[6] = {0, 4, 8, 1, 2, 5};  // XX YY ZZ XY XZ YZ of the 3x3 planes
        for (int e = 0; e < 6; e++) memcpy(&cadd6[(size_t)e * N], Cadd + (size_t)src[e] * N, pb);
    }
    hipError_t e1 = hipMemcpyAsync(df, f, pb, hipMemcpyHostToDevice, s);
    hipError_t e2 = hipMemcpyAsync(dd, X, 2 * pb, hipMemcpyHostToDevice, s);
    hipError_t e3 = hipMemcpyAsync(dd + 2 * (size_t)N, Y, 2 * pb, hipMemcpyHostToDevice, s);
    hipError_t e4 = hipMemcpyAsync(dd + 4 * (size_t)N, Z, 2 * pb, hipMemcpyHostToDevice, s);
    hipError_t e5 = Cadd ? hipMemcpyAsync(P->cadd, cadd6.data(), 6 * pb, hipMemcpyHostToDevice, s) : hipSuccess;
    if (e1 || e2 || e3 || e4 || e5)
    {
        set_error("noise_plan_create: upload failed");
        glass_hip_noise_plan_destroy(P);
        return GLASS_HIP_ERUNTIME;
    }
    noise_prep_kernel<<<(N + 255) / 256, 256, 0, s>>>(df, (const double2*)dd, (const double2*)(dd + 2 * (size_t)N),
                                                      (const double2*)(dd + 4 * (size_t)N), N, fstar, P->basis,
                                                      P->prod);
    if (hipGetLastError() != hipSuccess || hipStreamSynchronize(s) != hipSuccess)
    {
        set_error("noise_plan_create: preparation kernel failed");
        glass_hip_noise_plan_destroy(P);
        return GLASS_HIP_ERUNTIME;
    }
    *out = P;
    return GLASS_HIP_OK;
}

extern "C" int glass_hip_noise_plan_destroy(glass_hip_noise_plan* P)
{
    if (!P) return GLASS_HIP_OK;
    (void)hipSetDevice(P->device);
    if (P->gexec) (void)hipGraphExecDestroy(P->gexec);
    if (P->graph) (void)hipGraphDestroy(P->graph);
    (void)hipDeviceSynchronize();
    for (void* p : P->allocs) (void)hipFree(p);
    delete P;
    return GLASS_HIP_OK;
}

extern "C" int glass_hip_noise_logl(glass_hip_noise_plan* P, int n, const double* d_params, double* d_logL,
                                    void* stream)
{
    if (!P || n < 0 || n > P->nslot || (n && (!d_params || !d_logL)))
    {
        set_error("noise_logl: bad arguments (n %d, capacity %d)", n, P ? P->nslot : 0);
        return GLASS_HIP_EINVAL;
    }
    if (n == 0) return GLASS_HIP_OK;
    GH_CHECK(hipSetDevice(P->device));
    return launch_logl(P, d_params, n, d_logL, (hipStream_t)stream);
}

extern "C" int glass_hip_noise_set_state(glass_hip_noise_plan* P, const double* params, const unsigned* seeds,
                                         const double* temperature, void* stream)
{
    if (!P || !params || !seeds || !temperature)
    {
        set_error("noise_set_state: bad arguments");
        return GLASS_HIP_EINVAL;
    }
    GH_CHECK(hipSetDevice(P->device));
    hipStream_t s = (hipStream_t)stream;
    const int S = P->nslot;
    std::vector<int> idx(S), zero(S, 0);
    for (int i = 0; i < S; i++) idx[i] = i;
    GH_CHECK(hipMemcpyAsync(P->x, params, sizeof(double) * 12 * S, hipMemcpyHostToDevice, s));
    GH_CHECK(hipMemcpyAsync(P->seed, seeds, sizeof(unsigned) * S, hipMemcpyHostToDevice, s));
    GH_CHECK(hipMemcpyAsync(P->temp, temperature, sizeof(double) * S, hipMemcpyHostToDevice, s));
    GH_CHECK(hipMemcpyAsync(P->index, idx.data(), sizeof(int) * S, hipMemcpyHostToDevice, s));
    GH_CHECK(hipMemcpyAsync(P->nacc, zero.data(), sizeof(int) * S, hipMemcpyHostToDevice, s));
    GH_CHECK(hipMemcpyAsync(P->swaps, zero.data(), sizeof(int) * S, hipMemcpyHostToDevice, s));
    int rc = launch_logl(P, P->x, S, P->xlogl, s);
    if (rc) return rc;
    GH_CHECK(hipStreamSynchronize(s));
    return GLASS_HIP_OK;
}

extern "C" int glass_hip_noise_get_state(glass_hip_noise_plan* P, double* params, double* logL, unsigned* seeds,
                                         int* index, int* nacc, int* swaps, void* stream)
{
    if (!P)
    {
        set_error("noise_get_state: null plan");
        return GLASS_HIP_EINVAL;
    }
    GH_CHECK(hipSetDevice(P->device));
    hipStream_t s = (hipStream_t)stream;
    const int S = P->nslot;
    if (params) GH_CHECK(hipMemcpyAsync(params, P->x, sizeof(double) * 12 * S, hipMemcpyDeviceToHost, s));
    if (logL) GH_CHECK(hipMemcpyAsync(logL, P->xlogl, sizeof(double) * S, hipMemcpyDeviceToHost, s));
    if (seeds) GH_CHECK(hipMemcpyAsync(seeds, P->seed, sizeof(unsigned) * S, hipMemcpyDeviceToHost, s));
    if (index) GH_CHECK(hipMemcpyAsync(index, P->index, sizeof(int) * S, hipMemcpyDeviceToHost, s));
    if (nacc) GH_CHECK(hipMemcpyAsync(nacc, P->nacc, sizeof(int) * S, hipMemcpyDeviceToHost, s));
    if (swaps) GH_CHECK(hipMemcpyAsync(swaps, P->swaps, sizeof(int) * S, hipMemcpyDeviceToHost, s));
    GH_CHECK(hipStreamSynchronize(s));
    return GLASS_HIP_OK;
}

static int enqueue_sweep(glass_hip_noise_plan* P, int ncalls, int swap, hipStream_t s)
{
    NoiseState S = state_of(P);
    const int g = (P->nslot + 255) / 256;
    for (int c = 0; c < ncalls; c++)
    {
        noise_begin_kernel<<<g, 256, 0, s>>>(S);
        for (int mc = 0; mc < 10; mc++)
        {
            noise_propose_kernel<<<g, 256, 0, s>>>(S);
            dim3 grid(P->nchunk, (P->nslot + kTile - 1) / kTile);
            noise_logl_kernel<<<grid, kNThr, 0, s>>>(P->N, P->chunk, P->basis, P->prod, P->cadd, P->y, P->nslot,
                                                     P->part);
            noise_finish_accept_kernel<<<(P->nslot + 3) / 4, 256, 0, s>>>(P->part, P->nchunk,
                                                                          -3.0 * P->N * log(2 * M_PI), S);
        }
    }
    if (swap) noise_ptswap_kernel<<<(P->nladder + 63) / 64, 64, 0, s>>>(S, P->nladder, P->ntemp);
    GH_CHECK(hipGetLastError());
    return GLASS_HIP_OK;
}

extern "C" int glass_hip_noise_mcmc(glass_hip_noise_plan* P, int ncalls, int swap, int graph, void* stream)
{
    if (!P || ncalls < 0 || ncalls > 1000)
    {
        set_error("noise_mcmc: bad arguments (calls %d)", ncalls);
        return GLASS_HIP_EINVAL;
    }
    GH_CHECK(hipSetDevice(P->device));
    hipStream_t s = (hipStream_t)stream;
    if (!graph) return enqueue_sweep(P, ncalls, swap ? 1 : 0, s);
    if (!P->gexec || P->g_calls != ncalls || P->g_swap != (swap ? 1 : 0))
    {
        if (int rc = drop_graph(P)) return rc;
        hipStream_t cs;
        GH_CHECK(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking));
        GH_CHECK(hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal));
        int rc = enqueue_sweep(P, ncalls, swap ? 1 : 0, cs);
        hipGraph_t gr = nullptr;
        hipError_t ec = hipStreamEndCapture(cs, &gr);
        (void)hipStreamDestroy(cs);
        if (rc) return rc;
        if (ec != hipSuccess)
        {
            set_error("noise_mcmc: graph capture failed: %s", hipGetErrorString(ec));
            return GLASS_HIP_ERUNTIME;
        }
        P->graph = gr;
        GH_CHECK(hipGraphInstantiate(&P->gexec, gr, nullptr, nullptr, 0));
        P->g_calls = ncalls;
        P->g_swap = swap ? 1 : 0;
    }
    GH_CHECK(hipGraphLaunch(P->gexec, s));
    return GLASS_HIP_OK;
}

extern "C" int glass_hip_noise_ptswap(glass_hip_noise_plan* P, void* stream)
{
    return glass_hip_noise_mcmc(P, 0, 1, 0, stream);
}

extern "C" int glass_hip_noise_plan_info(glass_hip_noise_plan* P, int* nchunk, int* chunk, size_t* bytes)
{
    if (!P)
    {
        set_error("noise_plan_info: null plan");
        return GLASS_HIP_EINVAL;
    }
    if (nchunk) *nchunk = P->nchunk;
    if (chunk) *chunk = P->chunk;
    if (bytes)
    {
        const size_t S = P->nslot, pb = sizeof(double) * (size_t)P->N;
        *bytes = 10 * pb + (P->cadd ? 6 * pb : 0) + sizeof(double) * (size_t)P->nchunk * S +
                 S * (sizeof(double) * 30 + sizeof(unsigned) + 3 * sizeof(int)) + 7 * pb;
    }
    return GLASS_HIP_OK;
}

extern "C" int glass_hip_rand_stream(unsigned seed, int n, double* u, double* g)
{
    if (n <= 0 || n > (1 << 20) || !u || !g)
    {
        set_error("rand_stream: bad arguments");
        return GLASS_HIP_EINVAL;
    }
    ThreadCtx* cx;
    int rc;
    const size_t b = sizeof(double) * (size_t)n;
    if ((rc = staged_call(0, 2 * b, &cx))) return rc;
    double* d = (double*)cx->dbuf.p;
    rand_stream_kernel<<<1, 64, 0, cx->stream>>>(seed, n, d, d + n);
    GH_CHECK(hipGetLastError());
    GH_CHECK(hipMemcpyAsync(cx->hbuf.p, d, 2 * b, hipMemcpyDeviceToHost, cx->stream));
    GH_CHECK(hipStreamSynchronize(cx->stream));
    memcpy(u, cx->hbuf.p, b);
    memcpy(g, (char*)cx->hbuf.p + b, b);
    return GLASS_HIP_OK;
}
