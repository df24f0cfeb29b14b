// fstat.hip -- the F-statistic of the UCB sampler's F-statistic proposal,
// batched over grid cells.
//
// Reference (paths relative to the reference root):
//   get_Fstat_logL                 ucb/src/glass_ucb_fstatistic.c:646-703
//   get_filters (the 4 filters)    ucb/src/glass_ucb_fstatistic.c:231-348
//   get_N, get_M, init_M_matrix    ucb/src/glass_ucb_fstatistic.c:172-517
//   calc_Fstat_logL, calc_a_i      ucb/src/glass_ucb_fstatistic.c:599-644
//   get_F_params                   ucb/src/glass_ucb_fstatistic.c:525-597
//   setup_ / build_fstatistic_proposal  ucb/src/glass_ucb_proposal.c:916-1180
//
// One workgroup per grid cell (f0, fdot, theta, phi): the four filter
// templates A^i (ucb_waveform at amplitude 2, iota = pi/2 and the four
// (psi, phase) pairs) go through the batched template pipeline (slow series
// -> LDS FFTs -> TDI) one after another and stay in registers, one bin per
// thread; then the projections N^i = (s|A^i) and the Gram M^ij = (A^i|A^j)
// for X and for A+E are summed over the workgroup, and thread 0 inverts the
// two 4x4 Grams, forms F = N^T M^-1 N / 2 and the maximum-likelihood
// extrinsic parameters.  Cells are bucketed by their filter length
// N = ucb_bandwidth(f0, fdot, cos theta, A = 1e-22) into one kernel per
// power-of-two class, as the proposal evaluations are.
//
// The 4x4 inverse is Gauss-Jordan elimination with partial pivoting in
// place of LAPACKE dgetrf/dgetri (utils/src/glass_math.c:501-523); the two
// agree to rounding.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <vector>

#include "../../include/glass_hip_fstat.h"
#include "fd_waveform.h"
#include "reduce.h"
#include "runtime.h"

namespace glass {

int parse_format(const char* format);

namespace {

constexpr int kFsClass = 8;  // filter lengths 16 .. 2048
constexpr int kPosNF = 4096;

struct FstatDev
{
    const double2* X;
    const double2* A;
    const double2* E;
    const double* C00;
    int NFFT, qmin0;
    double T, t0, L, fstar;
    int fmt;
    const double* pos;
    const double2* tw;
    double fp[4][4];  // per filter: ln A, cos iota, psi, phase
};

struct Cells
{
    const double* cell;  // [n][4]: f0, fdot, theta, phi
    double* lx;          // [n] F-statistic of X
    double* lae;         // [n] F-statistic of A, E
    double* fpar;        // [n][4]: A, iota, psi, phase of the A,E fit
};

// Gauss-Jordan inverse with partial pivoting of a 4x4 matrix (in place)
__host__ __device__ inline void invert4(double m[4][4])
{
#pragma clang fp contract(off)
    double a[4][8];
    for (int i = 0; i < 4; i++)
        for (int j = 0; j < 4; j++)
        {
            a[i][j] = m[i][j];
            a[i][4 + j] = i == j ? 1.0 : 0.0;
        }
    for (int c = 0; c < 4; c++)
    {
        int p = c;
        for (int r = c + 1; r < 4; r++)
            if (fabs(a[r][c]) > fabs(a[p][c])) p = r;
        if (p != c)
            for (int j = 0; j < 8; j++)
            {
                double t = a[c][j];
                a[c][j] = a[p][j];
                a[p][j] = t;
            }
        const double piv = a[c][c];
        for (int j = 0; j < 8; j++) a[c][j] = a[c][j] / piv;
        for (int r = 0; r < 4; r++)
        {
            if (r == c) continue;
            const double f = a[r][c];
            for (int j = 0; j < 8; j++) a[r][j] = a[r][j] - f * a[c][j];
        }
    }
    for (int i = 0; i < 4; i++)
        for (int j = 0; j < 4; j++) m[i][j] = a[i][4 + j];
}

__host__ __device__ inline int sgn(double v) { return v < 0 ? -1 : (v > 0 ? 1 : 0); }

// calc_Fstat_logL for one channel set (the reference's term order)
__host__ __device__ inline double fstat_value(const double M[4][4], const double* N)
{
#pragma clang fp contract(off)
    return 0.5 * (M[0][0] * N[0] * N[0] + 2 * M[0][1] * N[0] * N[1] + 2 * M[0][2] * N[0] * N[2] +
                  2 * M[0][3] * N[0] * N[3] + M[1][1] * N[1] * N[1] + 2 * M[1][2] * N[1] * N[2] +
                  2 * M[1][3] * N[1] * N[3] + M[2][2] * N[2] * N[2] + 2 * M[2][3] * N[2] * N[3] +
                  M[3][3] * N[3] * N[3]);
}

// calc_a_i + get_F_params for the A,E channels
__device__ inline void fstat_params(const double M[4][4], const double* N, double* out)
{
#pragma clang fp contract(off)
    double a[4];
    for (int i = 0; i < 4; i++) a[i] = M[i][0] * N[0] + M[i][1] * N[1] + M[i][2] * N[2] + M[i][3] * N[3];
    const double a1p4 = a[0] + a[3], a2m3 = a[1] - a[2], a1m4 = a[0] - a[3], a2p3 = a[1] + a[2];
    const double s1 = sqrt(a1p4 * a1p4 + a2m3 * a2m3), s2 = sqrt(a1m4 * a1m4 + a2p3 * a2p3);
    const double Ap = s1 + s2, Ac = s1 - s2;
    double psi = atan2(Ap * a[3] - Ac * a[0], -(Ac * a[1] + Ap * a[2]));
    psi *= 0.5;
    const double c = (double)sgn(sin(2.0 * psi));
    const double r = sqrt(Ap * Ap - Ac * Ac);
    out[0] = 0.5 * (Ap + r);
    out[1] = acos(-Ac / (Ap + r));
    out[2] = psi;
    out[3] = atan2(c * (Ap * a[3] - Ac * a[0]), -c * (Ac * a[2] + Ap * a[1]));
}

// filter length of each cell and the per-class lists.  Slots are reserved
// per block (an LDS histogram, then one global atomic per class per
// block): the cells of a segment mostly share one filter length, and one
// global atomic per cell on that class's counter serialises the launch.
__global__ __launch_bounds__(256) void fstat_align_kernel(int n, const double* __restrict__ cell, FstatDev D,
                                                          int* __restrict__ cls, int* __restrict__ list, int cap,
                                                          int* __restrict__ err)
{
    __shared__ int hist[kFsClass + 1], base[kFsClass + 1];
    if (threadIdx.x <= kFsClass) hist[threadIdx.x] = 0;
    __syncthreads();
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    int k = -1, slot = 0;
    if (e < n)
    {
        const double f0 = cell[(size_t)e * 4], fdot = cell[(size_t)e * 4 + 1], theta = cell[(size_t)e * 4 + 2];
        const int N = ucb_bandwidth(D.L, D.fstar, f0, fdot, cos(theta), 1.e-22, D.T, D.NFFT);
        k = 31 - __clz(N) - 4;  // log2(N) - 4
        if (k < 0 || k >= kFsClass || (N & (N - 1))) k = kFsClass;
        slot = atomicAdd(&hist[k], 1);
    }
    __syncthreads();
    if (threadIdx.x < kFsClass && hist[threadIdx.x]) base[threadIdx.x] = atomicAdd(&cls[threadIdx.x], hist[threadIdx.x]);
    if (threadIdx.x == kFsClass && hist[kFsClass]) atomicAdd(err, hist[kFsClass]);
    __syncthreads();
    if (k >= 0 && k < kFsClass) list[(size_t)k * cap + base[k] + slot] = e;
}

template <int N, int NTHR>
__global__ __launch_bounds__(NTHR) void fstat_kernel(const int* __restrict__ list, const int* __restrict__ count,
                                                     FstatDev D, Cells C)
{
    using S = FdShape<N, NTHR>;
    __shared__ double2 lds[S::LDS];
    if ((int)blockIdx.x >= *count) return;
    const int e = list[blockIdx.x];
    const double f0 = C.cell[(size_t)e * 4], fdot = C.cell[(size_t)e * 4 + 1];
    const double theta = C.cell[(size_t)e * 4 + 2], phi = C.cell[(size_t)e * 4 + 3];
    const long q = (long)(f0 * D.T);
    double2 hx[S::PER][4], ha[S::PER][4], he[S::PER][4];
    // the four filters' constants, one lane each (the same code on four lanes
    // of the first wave: one setup's latency instead of four in turn)
    __shared__ UcbConst sc4[4];
    if (threadIdx.x < 4)
    {
        const int fi = threadIdx.x;
        double p[8] = {f0 * D.T,    cos(theta),   phi,          D.fp[fi][0],
                       D.fp[fi][1], D.fp[fi][2], D.fp[fi][3], fdot * (D.T * D.T)};
        ucb_setup(p, 8, D.T, D.t0, D.L, D.fstar, sc4[fi]);
    }
    __syncthreads();
    for (int fi = 0; fi < 4; fi++)
    {
        const UcbConst& k = sc4[fi];
        double2 dl[S::PER][6];
        fd_link_spectra<N, NTHR>(k, lds, D.pos, kPosNF / N, D.tw, kTwN / N, dl);
#pragma unroll
        for (int kk = 0; kk < S::PER; kk++)
        {
            const int m = threadIdx.x + kk * NTHR;
            if (m >= N) continue;
            double2 v[5];
            fd_tdi<N>(k, m, D.fmt, dl[kk], v);
            hx[kk][fi] = v[0];
            ha[kk][fi] = v[3];
            he[kk][fi] = v[4];
        }
    }
    // projections and Gram sums (get_N, get_M), the reference's term order
    double nx[4] = {0, 0, 0, 0}, nae[4] = {0, 0, 0, 0}, mx[10], mae[10];
    for (int i = 0; i < 10; i++) mx[i] = mae[i] = 0.0;
    const long qo = q - D.qmin0;
    {
#pragma clang fp contract(off)
#pragma unroll
        for (int kk = 0; kk < S::PER; kk++)
        {
            const int m = threadIdx.x + kk * NTHR;
            const long kb = qo + m - N / 2;
            if (m >= N || !(kb > 0 && kb < D.NFFT)) continue;
            const double c = D.C00[kb];
            const double2 dX = D.X[kb], dA = D.A[kb], dE = D.E[kb];
            for (int i = 0; i < 4; i++)
            {
                nx[i] += (dX.x * hx[kk][i].x + dX.y * hx[kk][i].y) / c;
                nae[i] += (dA.x * ha[kk][i].x + dA.y * ha[kk][i].y + dE.x * he[kk][i].x + dE.y * he[kk][i].y) / c;
            }
            int t = 0;
            for (int i = 0; i < 4; i++)
                for (int j = i; j < 4; j++, t++)
                {
                    mx[t] += 4.0 * (hx[kk][i].x * hx[kk][j].x + hx[kk][i].y * hx[kk][j].y) / c;
                    mae[t] += 4.0 * (ha[kk][i].x * ha[kk][j].x + ha[kk][i].y * ha[kk][j].y) / c +
                              4.0 * (he[kk][i].x * he[kk][j].x + he[kk][i].y * he[kk][j].y) / c;
                }
        }
    }
    // the 28 workgroup sums in one pass (one barrier instead of 56)
    constexpr int B = NTHR < 64 ? 64 : NTHR;
    __shared__ double tot[28];
    {
        double v[28];
#pragma unroll
        for (int i = 0; i < 4; i++)
        {
            v[i] = nx[i];
            v[4 + i] = nae[i];
        }
#pragma unroll
        for (int i = 0; i < 10; i++)
        {
            v[8 + i] = mx[i];
            v[18 + i] = mae[i];
        }
        block_sum_many<B, 28>(v, tot);
    }
    if (threadIdx.x != 0) return;
    for (int i = 0; i < 4; i++)
    {
        nx[i] = tot[i];
        nae[i] = tot[4 + i];
    }
    for (int i = 0; i < 10; i++)
    {
        mx[i] = tot[8 + i];
        mae[i] = tot[18 + i];
    }
    double MX[4][4], MA[4][4];
    int t = 0;
    for (int i = 0; i < 4; i++)
        for (int j = i; j < 4; j++, t++)
        {
            MX[i][j] = MX[j][i] = mx[t];
            MA[i][j] = MA[j][i] = mae[t];
        }
    for (int i = 0; i < 4; i++)
    {
        nx[i] *= 4.0;
        nae[i] *= 4.0;
    }
    invert4(MX);
    invert4(MA);
    C.lx[e] = fstat_value(MX, nx);
    C.lae[e] = fstat_value(MA, nae);
    fstat_params(MA, nae, C.fpar + (size_t)e * 4);
}

}  // namespace

}  // namespace glass

using namespace glass;

struct glass_hip_fstat
{
    int device = 0;
    FstatDev D{};
    int cap = 0;
    std::vector<void*> allocs;
    double* cells = nullptr;
    double *lx = nullptr, *lae = nullptr, *fpar = nullptr;
    int *cls = nullptr, *list = nullptr;
};

template <typename T>
static int falloc(glass_hip_fstat* F, T** p, size_t n)
{
    void* q = nullptr;
    if (hipMalloc(&q, sizeof(T) * (n ? n : 1)) != hipSuccess)
    {
        set_error("fstat: hipMalloc of %zu bytes failed", sizeof(T) * (n ? n : 1));
        return GLASS_HIP_ERUNTIME;
    }
    F->allocs.push_back(q);
    *p = (T*)q;
    return GLASS_HIP_OK;
}

extern "C" int glass_hip_fstat_destroy(glass_hip_fstat* F)
{
    if (!F) return GLASS_HIP_OK;
    (void)hipSetDevice(F->device);
    (void)hipDeviceSynchronize();
    for (void* p : F->allocs) (void)hipFree(p);
    delete F;
    return GLASS_HIP_OK;
}

extern "C" int glass_hip_fstat_create(glass_hip_fstat** out, const glass_hip_orbit* orbit, const char* format,
                                      double T, double t0, int NFFT, int qmin0, const double* X, const double* A,
                                      const double* E, const double* C00, int max_cells, void* stream)
{
    if (!out || !orbit || !format || NFFT <= 0 || !X || !A || !E || !C00 || max_cells <= 0 || !(T > 0))
    {
        set_error("fstat_create: bad arguments");
        return GLASS_HIP_EINVAL;
    }
    const int fmt = parse_format(format);
    if (fmt < 0)
    {
        set_error("fstat_create: unsupported format '%s'", format);
        return GLASS_HIP_EINVAL;
    }
    *out = nullptr;
    auto* F = new glass_hip_fstat();
    GH_CHECK(hipGetDevice(&F->device));
    hipStream_t s = (hipStream_t)stream;
    FstatDev& D = F->D;
    D.NFFT = NFFT;
    D.qmin0 = qmin0;
    D.T = T;
    D.t0 = t0;
    D.L = orbit->L;
    D.fstar = orbit->fstar;
    D.fmt = fmt;
    // the four filters of get_filters: amplitude 2, iota = pi/2 and
    // (psi, phase) = (0, 0), (pi/4, pi), (0, 3 pi/2), (pi/4, pi/2), mapped
    // to the waveform's ln A and cos iota on the host as the reference does
    const double psi[4] = {0.0, M_PI_4, 0.0, M_PI_4}, ph[4] = {0.0, M_PI, 3.0 * M_PI_2, M_PI_2};
    for (int i = 0; i < 4; i++)
    {
        D.fp[i][0] = log(2.0);
        D.fp[i][1] = cos(M_PI_2);
        D.fp[i][2] = psi[i];
        D.fp[i][3] = ph[i];
    }
    F->cap = max_cells;
    const size_t nb = sizeof(double) * 2 * (size_t)NFFT;
    double *dX = nullptr, *dA = nullptr, *dE = nullptr, *dC = nullptr;
    int rc;
    if ((rc = falloc(F, &dX, 2 * (size_t)NFFT)) || (rc = falloc(F, &dA, 2 * (size_t)NFFT)) ||
        (rc = falloc(F, &dE, 2 * (size_t)NFFT)) || (rc = falloc(F, &dC, NFFT)) ||
        (rc = falloc(F, &F->cells, 4 * (size_t)max_cells)) || (rc = falloc(F, &F->lx, max_cells)) ||
        (rc = falloc(F, &F->lae, max_cells)) || (rc = falloc(F, &F->fpar, 4 * (size_t)max_cells)) ||
        (rc = falloc(F, &F->cls, kFsClass + 1)) || (rc = falloc(F, &F->list, (size_t)kFsClass * max_cells)))
    {
        glass_hip_fstat_destroy(F);
        return rc;
    }
    const double* pos = nullptr;
    const double2* tw = nullptr;
    if ((rc = positions(*orbit, T, t0, kPosNF, &pos)) || (rc = twiddles(&tw)))
    {
        glass_hip_fstat_destroy(F);
        return rc;
    }
    D.pos = pos;
    D.tw = tw;
    D.X = (const double2*)dX;
    D.A = (const double2*)dA;
    D.E = (const double2*)dE;
    D.C00 = dC;
    if (hipMemcpyAsync(dX, X, nb, hipMemcpyHostToDevice, s) || hipMemcpyAsync(dA, A, nb, hipMemcpyHostToDevice, s) ||
        hipMemcpyAsync(dE, E, nb, hipMemcpyHostToDevice, s) ||
        hipMemcpyAsync(dC, C00, sizeof(double) * NFFT, hipMemcpyHostToDevice, s) || hipStreamSynchronize(s))
    {
        set_error("fstat_create: upload failed");
        glass_hip_fstat_destroy(F);
        return GLASS_HIP_ERUNTIME;
    }
    *out = F;
    return GLASS_HIP_OK;
}

__global__ void fs_zero_kernel(int* p, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

__global__ void fs_nan_kernel(int n, double* lx, double* lae, double* fpar)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    lx[i] = lae[i] = NAN;
    for (int k = 0; k < 4; k++) fpar[(size_t)i * 4 + k] = NAN;
}

static int fstat_launch(glass_hip_fstat* F, int n, const double* cells, double* lx, double* lae, double* fpar,
                        hipStream_t s)
{
    fs_zero_kernel<<<1, 64, 0, s>>>(F->cls, kFsClass + 1);
    fs_nan_kernel<<<(n + 255) / 256, 256, 0, s>>>(n, lx, lae, fpar);
    fstat_align_kernel<<<(n + 255) / 256, 256, 0, s>>>(n, cells, F->D, F->cls, F->list, F->cap,
                                                       F->cls + kFsClass);
    // the class counts decide the grids: only non-empty classes launch, one
    // workgroup per cell (an empty class's full grid of immediately
    // exiting workgroups costs 80-160 us)
    int cnt[kFsClass + 1];
    GH_CHECK(hipMemcpyAsync(cnt, F->cls, sizeof(cnt), hipMemcpyDeviceToHost, s));
    GH_CHECK(hipStreamSynchronize(s));
    Cells C{cells, lx, lae, fpar};
#define FS_L(kk, NN)                                                                                            \
    if (cnt[kk] > 0)                                                                                            \
        fstat_kernel<NN, fd_threads(NN)><<<cnt[kk], fd_threads(NN), 0, s>>>(F->list + (size_t)kk * F->cap,     \
                                                                            F->cls + kk, F->D, C);
    FS_L(0, 16) FS_L(1, 32) FS_L(2, 64) FS_L(3, 128) FS_L(4, 256) FS_L(5, 512) FS_L(6, 1024) FS_L(7, 2048)
#undef FS_L
    GH_CHECK(hipGetLastError());
    return GLASS_HIP_OK;
}

extern "C" int glass_hip_fstat_eval(glass_hip_fstat* F, int n, const double* d_cells, double* d_logL_X,
                                    double* d_logL_AE, double* d_Fparams, void* stream)
{
    if (!F || n < 0 || n > F->cap || (n && (!d_cells || !d_logL_X || !d_logL_AE || !d_Fparams)))
    {
        set_error("fstat_eval: bad arguments (n %d, capacity %d)", n, F ? F->cap : 0);
        return GLASS_HIP_EINVAL;
    }
    if (n == 0) return GLASS_HIP_OK;
    GH_CHECK(hipSetDevice(F->device));
    return fstat_launch(F, n, d_cells, d_logL_X, d_logL_AE, d_Fparams, (hipStream_t)stream);
}

// the fdot of a chirp mass, ucb_fdot (ucb/src/glass_ucb_waveform.c:307-314)
static double ucb_fdot_host(double Mc, double f0)
{
    const double TSUN = 4.92569043916e-6;
    double M = Mc * TSUN;
    double Q = 19.2;
    return Q * pow(pow(M_PI, 8) * pow(M, 5) * pow(f0, 11), 1. / 3.);
}

static int eval_host_cells(glass_hip_fstat* F, const std::vector<double>& cells, std::vector<double>& lae,
                           hipStream_t s)
{
    const int n = (int)(cells.size() / 4);
    lae.assign(n, 0.0);
    for (int o = 0; o < n; o += F->cap)
    {
        const int m = std::min(F->cap, n - o);
        GH_CHECK(hipMemcpyAsync(F->cells, cells.data() + (size_t)o * 4, sizeof(double) * 4 * m,
                                hipMemcpyHostToDevice, s));
        if (int rc = fstat_launch(F, m, F->cells, F->lx, F->lae, F->fpar, s)) return rc;
        GH_CHECK(hipMemcpyAsync(lae.data() + o, F->lae, sizeof(double) * m, hipMemcpyDeviceToHost, s));
        int err = 0;
        GH_CHECK(hipMemcpyAsync(&err, F->cls + kFsClass, sizeof(int), hipMemcpyDeviceToHost, s));
        GH_CHECK(hipStreamSynchronize(s));
        if (err)
        {
            set_error("fstat: %d cells with a filter length outside 16..2048", err);
            return GLASS_HIP_EINVAL;
        }
    }
    return GLASS_HIP_OK;
}

extern "C" int glass_hip_fstat_proposal(glass_hip_fstat* F, double fmin, double fmax, int n_theta, int n_phi,
                                        double* tensor, int* n_f_out, double* d_f_out, double* norm_out,
                                        double* maxp_out, void* stream)
{
    if (!F || !(fmax > fmin) || n_theta <= 0 || n_phi <= 0)
    {
        set_error("fstat_proposal: bad arguments");
        return GLASS_HIP_EINVAL;
    }
    GH_CHECK(hipSetDevice(F->device));
    hipStream_t s = (hipStream_t)stream;
    const double T = F->D.T;
    // setup_fstatistic_proposal (ucb/src/glass_ucb_proposal.c:928-936)
    const int n_f = (int)((fmax - fmin) * T);
    if (n_f <= 0)
    {
        set_error("fstat_proposal: empty band");
        return GLASS_HIP_EINVAL;
    }
    const double d_f = (int)((fmax - fmin) * T) / (double)n_f / T;
    const double d_theta = 2. / (double)n_theta;
    const double d_phi = 6.283185307179586 / (double)n_phi;
    if (n_f_out) *n_f_out = n_f;
    if (d_f_out) *d_f_out = d_f;
    if (!tensor) return GLASS_HIP_OK;
    // build_fstatistic_proposal (:1086-1180): every cell at fdot = 0, then
    // for cells with logL_AE > 1 the chirp-mass scan of fdot
    const size_t ncell = (size_t)n_f * n_theta * n_phi;
    std::vector<double> cells(4 * ncell);
    for (int i = 0; i < n_f; i++)
    {
        double f = fmin + i * d_f;
        f += d_f / 2;
        for (int j = 0; j < n_theta; j++)
            for (int k = 0; k < n_phi; k++)
            {
                const size_t c = ((size_t)i * n_theta + j) * n_phi + k;
                cells[4 * c] = f;
                cells[4 * c + 1] = 0.0;
                cells[4 * c + 2] = acos((-1. + (double)j * d_theta + d_theta / 2));
                cells[4 * c + 3] = (double)k * d_phi + d_phi / 2;
            }
    }
    std::vector<double> l0;
    if (int rc = eval_host_cells(F, cells, l0, s)) return rc;
    const int n_fdot = 10;
    const double d_fdot = (1.1 - 0.1) / n_fdot;
    std::vector<size_t> loud;
    for (size_t c = 0; c < ncell; c++)
        if (l0[c] > 1) loud.push_back(c);
    std::vector<double> cells2(4 * loud.size() * n_fdot);
    for (size_t u = 0; u < loud.size(); u++)
        for (int nn = 0; nn < n_fdot; nn++)
        {
            const double* src = &cells[4 * loud[u]];
            double* dst = &cells2[4 * (u * n_fdot + nn)];
            dst[0] = src[0];
            dst[1] = ucb_fdot_host(0.1 + nn * d_fdot, src[0]);
            dst[2] = src[2];
            dst[3] = src[3];
        }
    std::vector<double> l1;
    if (!loud.empty())
        if (int rc = eval_host_cells(F, cells2, l1, s)) return rc;
    for (size_t c = 0; c < ncell; c++) tensor[c] = l0[c] * l0[c];
    for (size_t u = 0; u < loud.size(); u++)
    {
        double lmax = l0[loud[u]];
        for (int nn = 0; nn < n_fdot; nn++)
        {
            const double v = l1[u * n_fdot + nn];
            if (v > lmax) lmax = v;
            if (v < 1) break;
        }
        tensor[loud[u]] = lmax * lmax;
    }
    double norm = 0.0;
    for (size_t c = 0; c < ncell; c++) norm += tensor[c];
    norm = (n_f * n_theta * n_phi) / norm;
    double maxp = -1.e60;
    for (size_t c = 0; c < ncell; c++)
    {
        tensor[c] *= norm;
        if (tensor[c] > maxp) maxp = tensor[c];
    }
    if (norm_out) *norm_out = norm;
    if (maxp_out) *maxp_out = maxp;
    return GLASS_HIP_OK;
}

extern "C" int glass_hip_get_Fstat_logL(const glass_hip_orbit* orbit, const char* format, double T, double t0,
                                        int NFFT, int qmin0, const double* X, const double* A, const double* E,
                                        const double* C00, double f0, double fdot, double theta, double phi,
                                        double* logL_X, double* logL_AE, double* Fparams)
{
    if (!logL_X || !logL_AE || !Fparams)
    {
        set_error("get_Fstat_logL: null output");
        return GLASS_HIP_EINVAL;
    }
    ThreadCtx* cx;
    int rc;
    if ((rc = ctx_get(&cx))) return rc;
    glass_hip_fstat* F = nullptr;
    if ((rc = glass_hip_fstat_create(&F, orbit, format, T, t0, NFFT, qmin0, X, A, E, C00, 1, cx->stream)))
        return rc;
    const double cell[4] = {f0, fdot, theta, phi};
    double res[6];
    if (hipMemcpyAsync(F->cells, cell, sizeof(cell), hipMemcpyHostToDevice, cx->stream) != hipSuccess)
        rc = GLASS_HIP_ERUNTIME;
    if (!rc) rc = fstat_launch(F, 1, F->cells, F->lx, F->lae, F->fpar, cx->stream);
    int err = 0;
    if (!rc && (hipMemcpyAsync(res, F->lx, sizeof(double), hipMemcpyDeviceToHost, cx->stream) ||
                hipMemcpyAsync(res + 1, F->lae, sizeof(double), hipMemcpyDeviceToHost, cx->stream) ||
                hipMemcpyAsync(res + 2, F->fpar, sizeof(double) * 4, hipMemcpyDeviceToHost, cx->stream) ||
                hipMemcpyAsync(&err, F->cls + kFsClass, sizeof(int), hipMemcpyDeviceToHost, cx->stream) ||
                hipStreamSynchronize(cx->stream)))
    {
        set_error("get_Fstat_logL: kernel or copy failed");
        rc = GLASS_HIP_ERUNTIME;
    }
    glass_hip_fstat_destroy(F);
    if (rc) return rc;
    if (err)
    {
        set_error("get_Fstat_logL: filter length outside 16..2048");
        return GLASS_HIP_EINVAL;
    }
    *logL_X = res[0];
    *logL_AE = res[1];
    for (int i = 0; i < 4; i++) Fparams[i] = res[2 + i];
    return GLASS_HIP_OK;
}
