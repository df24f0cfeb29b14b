// wdm_ucb.hip -- galactic-binary templates in the WDM wavelet domain on
// gfx950: ucb_waveform_wavelet (ucb/src/glass_ucb_waveform.c:1061-1240),
// wavelet_nwip-based inner products, snr_wavelet and the wavelet
// likelihood, as drop-ins and as a batched device API.
//
// Layout of the computation (one template per workgroup, two kernels):
//  K1 resp_kernel   : on the Norb orbit knots -- barycentre amplitude and
//                     phase splines, phase at spacecraft 1, the 24 delayed
//                     one-arm terms of X,Y,Z (LISA_spline_response), their
//                     amplitude/phase (extract_amplitude_and_phase,
//                     unwrap_phase) and the six TDI splines.  The serial
//                     parts (tridiagonal sweeps, sign-flip scan, unwrap)
//                     run one spline / channel per lane, side by side.
//  K2 layer_kernel  : the heterodyned series at NT*(L+1) samples (LDS),
//                     a packed real FFT (X+iY in one complex FFT, Z as
//                     half-length), per layer the Meyer-windowed band and
//                     an inverse FFT of NT, the pixel list, then the fused
//                     inner products against data and invC.
// Everything that depends only on the orbit and the time grids (knots,
// spline x-terms, spacecraft geometry at the knots, spacecraft 1 at the
// downsampled times, interval indices) is tabulated once on the host with
// the reference's operation order and lives on the device.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <mutex>
#include <vector>

#include "../../include/glass_hip_wdm.h"
#include "fft_lds.h"
#include "reduce.h"
#include "runtime.h"
#include "wdm_common.h"

namespace glass {
namespace wdm {

// ------------------------------------------------------------ TDI table --

__constant__ TdiTerm c_tdi[8] = {
    {0, 2, 1, +1, 0, 1, -1, 1, 2, 2.0, 2.0},
    {1, 2, 1, -1, 0, 2, +1, 1, 2, 1.0, 2.0},
    {2, 1, 2, +1, 1, 1, -1, 0, 1, 1.0, 2.0},
    {0, 1, 0, -1, 0, 1, +1, 0, 2, 2.0, 0.0},
    {0, 2, 0, +1, 1, 2, -1, 1, 1, 2.0, 0.0},
    {2, 1, 0, -1, 1, 1, +1, 0, 1, 1.0, 0.0},
    {1, 2, 0, +1, 0, 2, -1, 1, 2, 1.0, 0.0},
    {0, 0, 0, -1, 0, 2, +1, 1, 1, 0.0, 0.0},
};

// ----------------------------------------------------- host: splines ----

// Natural cubic spline coefficients c = y0|y1|y2|y3 (N each), the
// recurrence of spline_coefficients (utils/src/glass_math.c:98-178).
void spline_host(int N, const double* x, const double* y, double* c)
{
#pragma clang fp contract(off)
    std::vector<double> dx(N, 0.0), alpha(N, 0.0), l(N, 0.0), z(N, 0.0), mu(N, 0.0);
    double *y0 = c, *y1 = c + N, *y2 = c + 2 * N, *y3 = c + 3 * N;
    dx[0] = x[1] - x[0];
    for (int i = 1; i < N - 1; i++)
    {
        dx[i] = x[i + 1] - x[i];
        alpha[i] = 3.0 * ((y[i + 1] - y[i]) / dx[i] - (y[i] - y[i - 1]) / dx[i - 1]);
    }
    l[0] = l[N - 1] = 1.0;
    for (int i = 1; i < N - 1; i++)
    {
        l[i] = 2.0 * (x[i + 1] - x[i - 1]) - dx[i - 1] * mu[i - 1];
        mu[i] = dx[i] / l[i];
        z[i] = (alpha[i] - dx[i - 1] * z[i - 1]) / l[i];
    }
    y0[N - 1] = y[N - 1];
    y1[N - 1] = y2[N - 1] = y3[N - 1] = 0.0;
    for (int i = N - 2; i >= 0; i--)
    {
        y0[i] = y[i];
        y2[i] = z[i] - mu[i] * y2[i + 1];
        y1[i] = (y0[i + 1] - y0[i]) / dx[i] - dx[i] * (y2[i + 1] + 2.0 * y2[i]) / 3.0;
        y3[i] = (y2[i + 1] - y2[i]) / (3.0 * dx[i]);
    }
}

// x-only terms of the recurrence: dx | l | mu | 1/l (N each; the
// reciprocals correctly rounded, for spline_sweep's division)
static void spline_prep_host(int N, const double* x, double* p)
{
#pragma clang fp contract(off)
    double *dx = p, *l = p + N, *mu = p + 2 * N, *rl = p + 3 * N;
    for (int i = 0; i < 4 * N; i++) p[i] = 0.0;
    dx[0] = x[1] - x[0];
    for (int i = 1; i < N - 1; i++) dx[i] = x[i + 1] - x[i];
    l[0] = l[N - 1] = 1.0;
    for (int i = 1; i < N - 1; i++)
    {
        l[i] = 2.0 * (x[i + 1] - x[i - 1]) - dx[i - 1] * mu[i - 1];
        mu[i] = dx[i] / l[i];
    }
    for (int i = 0; i < N; i++) rl[i] = 1.0 / l[i];
}

// ------------------------------------------------------ host: geometry --

static double incomplete_beta(double a, double b, double x)
{
#pragma clang fp contract(off)
    // incomplete_beta_function (utils/src/glass_math.c:1214-1242): the
    // Gauss series 2F1(a, 1-b; a+1; x) to |term| <= 1e-8
    double aa = a, bb = 1 - b, cc = a + 1;
    double term = aa * bb * x / cc;
    double value = 1.0 + term;
    int n = 1;
    while (fabs(term) > 1e-8)
    {
        aa++, bb++, cc++, n++;
        term *= aa * bb * x / cc / n;
        value += term;
    }
    double B = exp(lgamma(a) + lgamma(b) - lgamma(a + b));
    return pow(x, a) * value / B / a;
}

// build_wdm_filter_freq (utils/src/glass_wavelet.c:305-349), forward
void wdm_window(int Nf, int Nt, double A, double* phif)
{
#pragma clang fp contract(off)
    const double DeltaOmega = M_PI / Nf;
    const double domega = 2 * DeltaOmega / (double)Nt;
    for (int i = 0; i < Nt / 2 + 1; i++)
    {
        double omega = domega * i;
        double B = DeltaOmega - 2 * A;
        double insDOM = 1.0 / sqrt(DeltaOmega);
        double z = 0.0;
        if (fabs(omega) >= A && fabs(omega) < A + B)
        {
            double x = (fabs(omega) - A) / B;
            double y = incomplete_beta(6.0, 6.0, x);
            z = insDOM * cos(y * M_PI_2);
        }
        if (fabs(omega) < A) z = insDOM;
        phif[i] = z;
    }
    double norm = 0.0;
    for (int i = 0; i < Nt / 2 + 1; i++)
    {
        if (i == 0) norm += phif[i] * phif[i];
        else norm += 2 * phif[i] * phif[i];
    }
    norm *= domega * M_1_PI;
    norm = sqrt(norm);
    norm *= Nf / 2.0;
    for (int i = 0; i < Nt / 2 + 1; i++) phif[i] /= norm;
}

// Device tables of one (wdm grid, orbit knots, Tobs, t0) combination.
struct Geo
{
    int Norb, NT, NF, kmin, kmax, imin, imax;
    double Tobs, t0, dt, df, A, ot0, ods;
    const double* ot;      // Norb orbit knots
    const double* oc;      // 9 x 4 x Norb position splines (row 3*sc+coord)
    const double* prep_o;  // 4 x Norb  dx | l | mu | 1/l on the orbit knots
    const double* tssb;    // Norb response times t0 + i Tobs/(Norb-1)
    const double* prep_s;  // 4 x Norb  on the response times
    const double* resp;    // Norb x 24: x[3][3]/c, n[3][3], L[3] at tssb
    // per layer count L (slot L; a forced count Lf in slot kForcedSlot):
    const double* ds[kForcedSlot + 1];   // N_ds x 5: t, r0x, r0y, r0z, dx
    const int* dsi[kForcedSlot + 1];     // response-time interval
    const double* win[kForcedSlot + 1];  // window of L+1 layers
    int Lf;                // forced layer count of every template (0: ucb_wavelet_layers' own)
    const double2* tw;     // 4096-entry twiddles
    const double2* tw3;    // 12288-entry twiddles
    const double2* tw16;   // 16384-entry twiddles (template series of NT >= 1024)
    const double2* tw32;   // 32768-entry twiddles (forced layer counts)
};

struct GeoHost
{
    Geo g;
    std::vector<void*> bufs;
    ~GeoHost()
    {
        for (void* p : bufs) (void)hipFree(p);
    }
};

template <typename T>
static int upload(GeoHost& G, const std::vector<T>& v, const T** out)
{
    void* d = nullptr;
    GH_CHECK(hipMalloc(&d, sizeof(T) * std::max<size_t>(v.size(), 1)));
    G.bufs.push_back(d);
    GH_CHECK(hipMemcpy(d, v.data(), sizeof(T) * v.size(), hipMemcpyHostToDevice));
    *out = (const T*)d;
    return GLASS_HIP_OK;
}

static double spline_host_at(const std::vector<double>& x, const double* c, int N, double t)
{
#pragma clang fp contract(off)
    int n = knot_index(x.data(), N, t);
    double dx = t - x[n];
    return c[n] + c[N + n] * dx + c[2 * N + n] * dx * dx + c[3 * N + n] * dx * dx * dx;
}

// Fills the tables.  pos: 9 x Norb positions (row 3*sc + coord); Lf > 0
// adds the tables of that forced layer count.
static int geo_build(GeoHost& G, const glass_hip_wdm& w, int Norb, const double* ot_in,
                     const double* pos, double Tobs, double t0, int Lf)
{
#pragma clang fp contract(off)
    Geo& g = G.g;
    const int N = Norb;
    g.Norb = N;
    g.NT = w.NT;
    g.NF = w.NF;
    g.kmin = w.kmin;
    g.kmax = w.kmax;
    g.imin = w.imin;
    g.imax = w.imax;
    g.Tobs = Tobs;
    g.t0 = t0;
    g.dt = w.dt;
    g.df = w.df;
    g.A = w.A;
    std::vector<double> ot(ot_in, ot_in + N);
    g.ot0 = ot[0];
    g.ods = ot[1] - ot[0];
    std::vector<double> oc(9 * 4 * N), prep_o(4 * N), tssb(N), prep_s(4 * N), resp(24 * N, 0.0);
    for (int s = 0; s < 9; s++) spline_host(N, ot.data(), pos + s * N, oc.data() + 4 * s * N);
    spline_prep_host(N, ot.data(), prep_o.data());
    const double dts = Tobs / (double)(N - 1);
    for (int i = 0; i < N; i++) tssb[i] = t0 + i * dts;
    spline_prep_host(N, tssb.data(), prep_s.data());
    if (tssb[0] < ot[0] || tssb[N - 1] > ot[N - 1])
    {
        set_error("wavelet waveform: response times [%g, %g] outside the orbit knots [%g, %g]",
                  tssb[0], tssb[N - 1], ot[0], ot[N - 1]);
        return GLASS_HIP_EINVAL;
    }
    // spacecraft geometry at the response times (LISA_spline_response,
    // utils/src/glass_lisa.c:976-1016)
    for (int m = 0; m < N; m++)
    {
        double x[3][3], n[3][3], L[3];
        for (int i = 0; i < 3; i++)
            for (int d = 0; d < 3; d++)
                x[i][d] = spline_host_at(ot, oc.data() + 4 * (3 * i + d) * N, N, tssb[m]) / CLIGHT;
        for (int i = 0; i < 3; i++)
        {
            n[0][i] = x[1][i] - x[2][i];
            n[1][i] = x[2][i] - x[0][i];
            n[2][i] = x[0][i] - x[1][i];
        }
        for (int i = 0; i < 3; i++)
        {
            L[i] = 0.0;
            for (int j = 0; j < 3; j++) L[i] += n[i][j] * n[i][j];
            L[i] = sqrt(L[i]);
        }
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) n[i][j] /= L[i];
        double* r = resp.data() + 24 * m;
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++)
            {
                r[3 * i + j] = x[i][j];
                r[9 + 3 * i + j] = n[i][j];
            }
        for (int i = 0; i < 3; i++) r[18 + i] = L[i];
    }
    int rc;
    if ((rc = upload(G, ot, &g.ot)) || (rc = upload(G, oc, &g.oc)) || (rc = upload(G, prep_o, &g.prep_o)) ||
        (rc = upload(G, tssb, &g.tssb)) || (rc = upload(G, prep_s, &g.prep_s)) ||
        (rc = upload(G, resp, &g.resp)))
        return rc;
    // per layer count L: the downsampled times, spacecraft 1 there, the
    // response-time interval and offset of each sample, and the window
    g.Lf = Lf;
    for (int slot = 1; slot <= (Lf > 0 ? kForcedSlot : kMaxLayers); slot++)
    {
        const int L = slot == kForcedSlot ? Lf : slot;
        const int Nds = w.NT * (L + 1);
        const double dt_ds = w.dt / (double)(L + 1);
        std::vector<double> ds(5 * (size_t)Nds);
        std::vector<int> dsi(Nds);
        for (int i = 0; i < Nds; i++)
        {
            double t = t0 + i * dt_ds;
            if (t < tssb[0] || t > tssb[N - 1] || t < ot[0] || t > ot[N - 1])
            {
                set_error("wavelet waveform: sample time %g of the %d-layer grid (NT*dt = %g) lies "
                          "outside [t0, t0+Tobs] = [%g, %g]",
                          t, L, w.NT * w.dt, tssb[0], tssb[N - 1]);
                return GLASS_HIP_EINVAL;
            }
            double* e = ds.data() + 5 * (size_t)i;
            e[0] = t;
            for (int d = 0; d < 3; d++) e[1 + d] = spline_host_at(ot, oc.data() + 4 * d * N, N, t) / CLIGHT;
            int k = knot_index(tssb.data(), N, t);
            dsi[i] = k;
            e[4] = t - tssb[k];
        }
        std::vector<double> win(w.NT / 2 + 1);
        wdm_window(L + 1, w.NT, w.A, win.data());
        if ((rc = upload(G, ds, &g.ds[slot])) || (rc = upload(G, dsi, &g.dsi[slot])) ||
            (rc = upload(G, win, &g.win[slot])))
            return rc;
    }
    g.ds[0] = nullptr;
    g.dsi[0] = nullptr;
    g.win[0] = nullptr;
    if (Lf <= 0)
    {
        g.ds[kForcedSlot] = nullptr;
        g.dsi[kForcedSlot] = nullptr;
        g.win[kForcedSlot] = nullptr;
    }
    g.tw32 = nullptr;
    if ((rc = twiddles(&g.tw)) || (rc = twiddles3(&g.tw3)) || (rc = twiddles16(&g.tw16)) ||
        (Lf > 0 && (rc = twiddles32(&g.tw32))))
        return rc;
    return GLASS_HIP_OK;
}

// Validates the grid for the template kernels.
static int check_grid(const glass_hip_wdm* w, int Norb)
{
    if (!w) return set_error("NULL wdm"), GLASS_HIP_EINVAL;
    if (w->NT != 256 && w->NT != 512 && w->NT != 1024 && w->NT != 2048 && w->NT != 4096)
    {
        set_error("wavelet templates: NT = %d unsupported (this build computes templates for NT = 256, 512, "
                  "1024, 2048 or 4096)",
                  w->NT);
        return GLASS_HIP_EINVAL;
    }
    if (w->NF < 4 || w->dt <= 0 || w->df <= 0)
        return set_error("wavelet templates: bad grid NF=%d dt=%g df=%g", w->NF, w->dt, w->df), GLASS_HIP_EINVAL;
    if (Norb < 4 || Norb > kMaxKnots)
        return set_error("wavelet templates: %d orbit knots (supported 4..%d)", Norb, kMaxKnots),
               GLASS_HIP_EINVAL;
    return GLASS_HIP_OK;
}

// ------------------------------------------------------- K1: response ---

// Per-template scratch (global): header, barycentre phase spline (4N),
// TDI splines (6 x 4N: amp X, phase X, amp Y, phase Y, amp Z, phase Z).
__host__ __device__ inline size_t scratch_doubles(int N) { return 64 + 28 * (size_t)N; }

constexpr int kRespThreads = 256;

// LDS of K1 in doubles: ssb 8N | M,Mf 6N | As 3N | flip,pjump 6N | phase_sc N.
// The sweep scratch of the splines lives in regions dead by then: the two
// barycentre sweeps use M,Mf (not yet written), the six TDI sweeps use ssb
// (z) and As + flip (y2), dead after the signed amplitude/phase step.
// Then the sign-flip scan's knot flags: per channel three bit masks (local
// minimum, the d2 and d3 tests) of kMaxKnots bits.
constexpr int kFlagWords = kMaxKnots / 64;
__host__ __device__ inline size_t resp_lds_bytes(int N)
{
    return sizeof(double) * 24 * (size_t)N + sizeof(Hdr) + sizeof(unsigned long long) * 9 * kFlagWords;
}

// Per-template constants, one thread per template, into each template's
// scratch header (read by K1 and K2).
__global__ void __launch_bounds__(256) hdr_kernel(Geo g, int n, const double* __restrict__ params,
                                                 double* __restrict__ scratch)
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n) return;
    Hdr h;
    make_hdr(params + 8 * (size_t)e, g.Tobs, g.df, g.A, &h);
    if (g.Lf > 0) h.L = g.Lf;  // forced layer count: same first layer and heterodyne
    *(Hdr*)(scratch + scratch_doubles(g.Norb) * (size_t)e) = h;
}

// The right-hand side alpha_i (1 <= i <= N-2) of the natural spline
// through y (knots described by prep), into z[i]: independent per knot, so
// the whole workgroup computes them before one lane's sweep.
__device__ inline void spline_alpha(int N, const double* __restrict__ prep, const double* y, double* z, int i)
{
#pragma clang fp contract(off)
    const double* dx = prep;
    z[i] = 3.0 * ((y[i + 1] - y[i]) / dx[i] - (y[i] - y[i - 1]) / dx[i - 1]);
}

// One lane's tridiagonal sweep of that spline, z holding the alpha_i of
// spline_alpha: writes z and y2 (N each).  Only the recurrence itself is
// serial; its division by l_i goes through the precomputed 1/l_i.
__device__ inline void spline_sweep(int N, const double* __restrict__ prep, double* z, double* y2)
{
#pragma clang fp contract(off)
    const double *dx = prep, *l = prep + N, *mu = prep + 2 * N, *rl = prep + 3 * N;
    z[0] = 0.0;
    double zp = 0.0;
#pragma unroll 4
    for (int i = 1; i < N - 1; i++)
    {
        const double alpha = z[i];
        zp = div_by_recip(alpha - dx[i - 1] * zp, l[i], rl[i]);
        z[i] = zp;
    }
    z[N - 1] = 0.0;
    double yn = 0.0;
    y2[N - 1] = 0.0;
#pragma unroll 4
    for (int i = N - 2; i >= 0; i--)
    {
        yn = z[i] - mu[i] * yn;
        y2[i] = yn;
    }
}

// y1, y3 of knot i from y0 and y2
__device__ inline void spline_finish(int N, int i, const double* __restrict__ prep, const double* y0,
                                     const double* y2, double* y1o, double* y3o)
{
#pragma clang fp contract(off)
    if (i == N - 1)
    {
        *y1o = 0.0;
        *y3o = 0.0;
        return;
    }
    const double dx = prep[i];
    *y1o = (y0[i + 1] - y0[i]) / dx - dx * (y2[i + 1] + 2.0 * y2[i]) / 3.0;
    *y3o = (y2[i + 1] - y2[i]) / (3.0 * dx);
}

// a[i] of a 3-element register array for a runtime i in 0..2, by selects
// (an indexed private array would live in scratch memory)
__device__ inline double pick3(const double (&a)[3], int i) { return i == 0 ? a[0] : (i == 1 ? a[1] : a[2]); }

// four waves per SIMD (128 VGPRs): four 256-thread workgroups per CU,
// as many as the 24N-double LDS footprint allows
__global__ void __launch_bounds__(kRespThreads) __attribute__((amdgpu_waves_per_eu(4, 8))) resp_kernel(Geo g, int n, const double* __restrict__ params,
                                                            double* __restrict__ scratch)
{
#pragma clang fp contract(off)
    const int e = blockIdx.x;
    if (e >= n) return;
    extern __shared__ double lds[];
    const int N = g.Norb;
    const int tid = threadIdx.x;
    double* ssb = lds;               // amp y0..y3 | phase y0..y3
    double* Mx = ssb + 8 * N;        // M X,Y,Z | Mf X,Y,Z
    double* As = Mx + 6 * N;         // |M| X,Y,Z
    double* fl = As + 3 * N;         // flip X,Y,Z | pjump X,Y,Z
    double* phsc = fl + 6 * N;       // phase at spacecraft 1
    Hdr* H = (Hdr*)(phsc + N);
    unsigned long long* flg = (unsigned long long*)(H + 1);  // 9 x kFlagWords
    double* scr = scratch + scratch_doubles(N) * (size_t)e;
    (void)params;

    // the header hdr_kernel wrote
    for (int i = tid; i < (int)(sizeof(Hdr) / sizeof(double)); i += kRespThreads) ((double*)H)[i] = scr[i];
    __syncthreads();
    const Hdr h = *H;
    // barycentre amplitude and phase at the orbit knots
    for (int i = tid; i < N; i += kRespThreads)
    {
        const double t = g.ot[i];
        ssb[i] = ucb_amplitude(t, h.A0, h.f0, h.fdot);
        ssb[4 * N + i] = ucb_phase(t, h.f0, h.phi0, h.fdot);
    }
    __syncthreads();
    for (int it = tid; it < 2 * (N - 2); it += kRespThreads)
    {
        const int sp = it / (N - 2), i = 1 + it % (N - 2);
        spline_alpha(N, g.prep_o, ssb + 4 * N * sp, Mx + N * sp, i);
    }
    __syncthreads();
    if (tid < 2) spline_sweep(N, g.prep_o, Mx + N * tid, ssb + 4 * N * tid + 2 * N);
    __syncthreads();
    for (int i = tid; i < 2 * N; i += kRespThreads)
    {
        const int s = i / N, k = i % N;
        double* c = ssb + 4 * N * s;
        spline_finish(N, k, g.prep_o, c, c + 2 * N, c + N + k, c + 3 * N + k);
    }
    __syncthreads();
    const double* cph = ssb + 4 * N;
    // export the barycentre phase spline for the layer kernel
    for (int i = tid; i < 4 * N; i += kRespThreads) scr[64 + i] = cph[i];
    // phase at spacecraft 1 (LISA_spacecraft_to_barycenter_time, flag -1)
    for (int m = tid; m < N; m += kRespThreads)
    {
        const double* r = g.resp + 24 * m;
        double kdotr = 0.0;
        for (int i = 0; i < 3; i++) kdotr += h.k[i] * r[i];
        const double tsc = g.tssb[m] - kdotr;
        const int j = even_index(tsc, g.ot0, g.ods, N);
        phsc[m] = spline_at(cph, N, j, tsc - g.ot[j]);
    }
    // the X, Y, Z responses: items (sample m, channel ch)
    for (int it = tid; it < 3 * N; it += kRespThreads)
    {
        const int m = it / 3, ch = it % 3;
        const double* r = g.resp + 24 * m;
        const double* x = r;          // x[i][j] at 3i+j
        const double* nn = r + 9;
        const double* L = r + 18;
        double kr[3], kn[3], App[3], Apm[3], Acp[3], Acm[3];
        for (int i = 0; i < 3; i++)
        {
            kr[i] = 0.0;
            for (int j = 0; j < 3; j++) kr[i] += h.k[j] * x[3 * i + j];
        }
        for (int i = 0; i < 3; i++)
        {
            kn[i] = 0.0;
            for (int j = 0; j < 3; j++) kn[i] += h.k[j] * nn[3 * i + j];
        }
        for (int i = 0; i < 3; i++)
        {
            double dp = 0.0, dc = 0.0;
            for (int j = 0; j < 3; j++)
                for (int l = 0; l < 3; l++)
                {
                    dp += (nn[3 * i + j] * nn[3 * i + l]) * h.ep[3 * j + l];
                    dc += (nn[3 * i + j] * nn[3 * i + l]) * h.ec[3 * j + l];
                }
            App[i] = dp / (1.0 + kn[i]);
            Apm[i] = dp / (1.0 - kn[i]);
            Acp[i] = dc / (1.0 + kn[i]);
            Acm[i] = dc / (1.0 - kn[i]);
        }
        const double t = g.tssb[m];
        double M = 0.0, Mf = 0.0;
#pragma unroll 1
        for (int q = 0; q < 8; q++)
        {
            const TdiTerm T = c_tdi[q];
            // vertex permutation of channel ch: vtx[i] = (ch + i) mod 3
            const int vv = (ch + T.v) % 3, v1 = (ch + T.i1) % 3, v2 = (ch + T.i2) % 3;
            const double te = t - pick3(kr, vv) - T.m1 * L[v1] - T.m2 * L[v2];
            const int j = even_index(te, g.ot0, g.ods, N);
            const double dx = te - g.ot[j];
            const double a = spline_at(ssb, N, j, dx);
            const double ph = spline_at(cph, N, j, dx);
            // correctly rounded: the channel sums cancel by (f/f*)^2, which
            // would magnify one-ulp differences of sin/cos
            double sp, cp;
            sincos_cr(ph, &sp, &cp);
            const double hp = a * (h.Ap * h.c2p * cp + h.Ac * h.s2p * sp);
            const double hc = a * (h.Ac * h.c2p * sp - h.Ap * h.s2p * cp);
            const double hpf = a * (-h.Ap * h.c2p * sp + h.Ac * h.s2p * cp);
            const double hcf = a * (h.Ac * h.c2p * cp + h.Ap * h.s2p * sp);
            const int j1 = (ch + T.j1) % 3, j2 = (ch + T.j2) % 3;
            const double P1 = T.pm1 ? pick3(App, j1) : pick3(Apm, j1), Q1 = T.pm1 ? pick3(Acp, j1) : pick3(Acm, j1);
            const double P2 = T.pm2 ? pick3(App, j2) : pick3(Apm, j2), Q2 = T.pm2 ? pick3(Acp, j2) : pick3(Acm, j2);
            double v;
            v = hp * P1 + hc * Q1;
            M = T.s1 > 0 ? M + v : M - v;
            v = hp * P2 + hc * Q2;
            M = T.s2 > 0 ? M + v : M - v;
            v = hpf * P1 + hcf * Q1;
            Mf = T.s1 > 0 ? Mf + v : Mf - v;
            v = hpf * P2 + hcf * Q2;
            Mf = T.s2 > 0 ? Mf + v : Mf - v;
        }
        Mx[ch * N + m] = M;
        Mx[(3 + ch) * N + m] = Mf;
        As[ch * N + m] = sqrt(M * M + Mf * Mf);
    }
    __syncthreads();
    // sign flips of the amplitude (extract_amplitude_and_phase,
    // utils/src/glass_math.c:1244-1296).  The tests at each knot (a local
    // minimum of |M|, then |d2/d1| < 0.1 and |d3/d1| < 0.1) need no state:
    // every wave evaluates 64 knots of a channel at once into ballot masks;
    // one lane per channel then walks the knots in the reference's order
    // (a d2 hit skips the next knot), holding the running flip and phase
    // jump in registers.
    {
        const int NWd = (N + 63) / 64, lane = tid & 63;
        for (int b = tid >> 6; b < 3 * NWd; b += kRespThreads / 64)
        {
            const int ch = b / NWd, w = b % NWd, i = w * 64 + lane;
            bool mn = false, t2 = false, t3 = false;
            if (i >= 1 && i <= N - 2)
            {
                const double* A = As + ch * N;
                mn = A[i] < A[i - 1] && A[i] < A[i + 1];
                if (mn)
                {
                    const double d1 = A[i + 1] + A[i - 1] - 2.0 * A[i];
                    const double d2 = -A[i + 1] + A[i - 1] - 2.0 * A[i];
                    const double d3 = -A[i + 1] + A[i - 1] + 2.0 * A[i];
                    t2 = fabs(d2 / d1) < 0.1;
                    t3 = fabs(d3 / d1) < 0.1;
                }
            }
            const unsigned long long m0 = __ballot(mn), m2 = __ballot(t2), m3 = __ballot(t3);
            if (lane == 0)
            {
                flg[(ch * 3 + 0) * kFlagWords + w] = m0;
                flg[(ch * 3 + 1) * kFlagWords + w] = m2;
                flg[(ch * 3 + 2) * kFlagWords + w] = m3;
            }
        }
    }
    __syncthreads();
    if (tid < 3)
    {
        const unsigned long long* fm = flg + tid * 3 * kFlagWords;
        double* flip = fl + tid * N;
        double* pj = fl + (3 + tid) * N;
        flip[0] = 1.0;
        pj[0] = 0.0;
        double pf = 1.0, pp = 0.0;  // flip[i-1], pj[i-1]
        int wc = -1;
        unsigned long long wm = 0, w2 = 0, w3 = 0;
        int i = 1;
        do
        {
            double cf = pf, cp = pp;  // flip[i] = flip[i-1]
            const int w = i >> 6;
            if (w != wc)
            {
                wc = w;
                wm = fm[w];
                w2 = fm[kFlagWords + w];
                w3 = fm[2 * kFlagWords + w];
            }
            const unsigned long long bit = 1ull << (i & 63);
            if (wm & bit)
            {
                const bool b2 = (w2 & bit) != 0, b3 = (w3 & bit) != 0;
                if (b2)
                {
                    // flip[i+1] = -flip[i], pj[i+1] = pj[i] + pi; i++
                    flip[i] = cf;
                    pj[i] = cp;
                    pf = cf;
                    pp = cp;
                    cf = -1.0 * cf;
                    cp = cp + M_PI;
                    i++;
                }
                if (b3)
                {
                    // flip[i] = -flip[i-1], pj[i] = pj[i-1] + pi
                    cf = -1.0 * pf;
                    cp = pp + M_PI;
                }
            }
            flip[i] = cf;
            pj[i] = cp;
            pf = cf;
            pp = cp;
            i++;
        } while (i < N - 1);
        flip[N - 1] = flip[N - 2];
        pj[N - 1] = pj[N - 2];
    }
    __syncthreads();
    // signed amplitude and phase relative to the spacecraft-1 carrier
    for (int it = tid; it < 3 * N; it += kRespThreads)
    {
        const int ch = it / N, i = it % N;
        const double M = Mx[ch * N + i], Mf = Mx[(3 + ch) * N + i];
        const double v = remainder(phsc[i], PI2);
        Mx[ch * N + i] = fl[ch * N + i] * As[ch * N + i];
        Mx[(3 + ch) * N + i] = -atan2_cr(Mf, M) + fl[(3 + ch) * N + i] - v;
    }
    __syncthreads();
    // unwrap_phase (utils/src/glass_math.c:1034-1052)
    if (tid < 3)
    {
        constexpr double kInvPI2 = 1.0 / PI2;  // correctly rounded (compile time)
        double* ph = Mx + (3 + tid) * N;
        double v = ph[0];
#pragma unroll 4
        for (int i = 0; i < N; i++)
        {
            double u = ph[i];
            const double q = rint(div_by_recip(fabs(u - v), PI2, kInvPI2));
            if (q > 0.0)
            {
                if (v > u) u += q * PI2;
                else u -= q * PI2;
            }
            v = u;
            ph[i] = u;
        }
    }
    __syncthreads();
    // six TDI splines on the response times: spline s = 2*ch + (0 amp, 1 phase)
    double* y2s = As;  // 6N over As and the flip half of fl
    for (int it = tid; it < 6 * (N - 2); it += kRespThreads)
    {
        const int sp = it / (N - 2), i = 1 + it % (N - 2);
        const int ch = sp >> 1, ph = sp & 1;
        spline_alpha(N, g.prep_s, Mx + (3 * ph + ch) * N, ssb + sp * N, i);
    }
    __syncthreads();
    if (tid < 6) spline_sweep(N, g.prep_s, ssb + tid * N, y2s + tid * N);
    __syncthreads();
    for (int it = tid; it < 6 * N; it += kRespThreads)
    {
        const int s = it / N, i = it % N;
        const int ch = s >> 1, ph = s & 1;
        const double* y0 = Mx + (3 * ph + ch) * N;
        const double* y2 = y2s + s * N;
        double y1, y3;
        spline_finish(N, i, g.prep_s, y0, y2, &y1, &y3);
        double* c = scr + 64 + 4 * N + 4 * (size_t)N * s;
        c[i] = y0[i];
        c[N + i] = y1;
        c[2 * N + i] = y2[i];
        c[3 * N + i] = y3;
    }
}

// ---------------------------------------------------------- K2: layers --

// 512 threads: the kernel's 74 KB of LDS (NT 512) admits two workgroups
// per CU, so 512-thread workgroups give four waves per SIMD where 256-thread
// ones gave two (111 VGPRs allow four)
constexpr int kLayerThreads = 512;

struct LayerArgs
{
    const double* data;   // 3 planes of npix (NULL: no inner products)
    const double* invC;   // 6 planes XX,YY,ZZ,XY,XZ,YZ of npix
    int npix;
    int* nwave;           // [n]
    int* list;            // [n x vcap] (compact pixel list)
    double* values;       // [n x 3 x vcap]
    double* dlogL;        // [n] (with data)
    double* snr;          // [n] (with data)
    double2* big;         // long-series kernel: nbig global scratch slices of bigslice double2
    int nbig;
    int vcap;             // pixels per template: 3 NT (NT max(3, Lf) with a forced layer count)
    size_t bigslice;
};

// X, Y (packed in U) or Z (half-length V) spectrum at bin k, 0 < k < Nds/2
template <int NDS>
__device__ inline double2 spectrum(int ch, int k, const double2* U, const double2* V, const double2* tw,
                                   int twstride)
{
    if (ch < 2)
    {
        const double2 a = U[k], b = U[(NDS - k) % NDS];
        // X = (U[k] + conj U[N-k]) / 2 ; Y = (U[k] - conj U[N-k]) / (2i)
        if (ch == 0) return make_double2(0.5 * (a.x + b.x), 0.5 * (a.y - b.y));
        return make_double2(0.5 * (a.y + b.y), -0.5 * (a.x - b.x));
    }
    constexpr int M = NDS / 2;
    const double2 a = V[k % M], b = V[(M - k % M) % M];
    const double2 E = make_double2(0.5 * (a.x + b.x), 0.5 * (a.y - b.y));
    const double2 O = make_double2(0.5 * (a.y + b.y), -0.5 * (a.x - b.x));
    const double2 w = tw[k * twstride];  // exp(-2 pi i k / NDS)
    return make_double2(E.x + (w.x * O.x - w.y * O.y), E.y + (w.x * O.y + w.y * O.x));
}

// Exclusive prefix sum of v over the workgroup in thread order, by wave
// scans (__shfl_up) and one pass over the NTHR/64 wave totals; ws holds
// NTHR/64 + 1 ints of LDS and ends with the total in ws[NTHR/64].  Ends
// with a barrier.
template <int NTHR>
__device__ inline int block_exclusive_scan(int v, int* ws)
{
    static_assert(NTHR % 64 == 0, "wave64 workgroups");
    constexpr int NW = NTHR / 64;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    int x = v;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1)
    {
        const int y = __shfl_up(x, d, 64);
        if (lane >= d) x += y;
    }
    if (lane == 63) ws[w] = x;
    __syncthreads();
    if (threadIdx.x == 0)
    {
        int sum = 0;
        for (int i = 0; i < NW; i++)
        {
            const int t = ws[i];
            ws[i] = sum;
            sum += t;
        }
        ws[NW] = sum;
    }
    __syncthreads();
    return ws[w] + x - v;
}

template <int NT, int L>
__device__ void layer_body(const Geo& g, const Hdr& h, int e, const LayerArgs& a, double2* U, double2* V,
                           double2* B, int* cnt, const double* __restrict__ scr)
{
    constexpr int NDS = NT * (L + 1);
    constexpr int M = NDS / 2;
    const int tid = threadIdx.x;
    const int N = g.Norb;
    const double* cph = scr + 64;
    const double* ctdi = scr + 64 + 4 * N;
    const double* ds = g.ds[L];
    const int* dsi = g.dsi[L];
    {
#pragma clang fp contract(off)
        const double w0 = PI2 * h.fhet;
        for (int i = tid; i < NDS; i += kLayerThreads)
        {
            const double* e5 = ds + 5 * (size_t)i;
            const double t = e5[0];
            double kdotr = 0.0;
            for (int d = 0; d < 3; d++) kdotr += h.k[d] * e5[1 + d];
            const double tsc = t - kdotr;
            const int j = even_index(tsc, g.ot0, g.ods, N);
            const double phase_ds = spline_at(cph, N, j, tsc - g.ot[j]);
            const double phase_het = w0 * t;
            const int si = dsi[i];
            const double sdx = e5[4];
            double x[3];
            for (int ch = 0; ch < 3; ch++)
            {
                const double* ca = ctdi + 4 * (size_t)N * (2 * ch);
                const double* cp = ctdi + 4 * (size_t)N * (2 * ch + 1);
                const double Amp = spline_at(ca, N, si, sdx);
                double Phase = spline_at(cp, N, si, sdx);
                Phase += phase_ds;
                Phase -= phase_het;
                double s, c;
                sincos_red(Phase, &s, &c);
                x[ch] = Amp * c;
            }
            U[i] = make_double2(x[0], x[1]);
            if (i & 1) V[i >> 1].y = x[2];
            else V[i >> 1].x = x[2];
        }
    }
    __syncthreads();
    // forward FFTs: U (NDS) and V (NDS/2)
    const bool three = (NDS % 3) == 0;
    const double2* twU = three ? g.tw3 : g.tw;
    const int tsU = (three ? kTw3N : kTwN) / NDS;
    lds_fft<NDS, 1, kLayerThreads>(U, twU, tsU);
    lds_fft<M, 1, kLayerThreads>(V, twU, tsU * 2);
    // pixel list (ucb/src/glass_ucb_waveform.c:1189-1212): candidates in
    // (time i, layer j) order, kept when kmin <= k < kmax
    int* lst = a.list + (size_t)e * a.vcap;
    {
        constexpr int NC = NT * L;
        constexpr int PER = (NC + kLayerThreads - 1) / kLayerThreads;
        const int c0 = tid * PER;
        int mine = 0;
        for (int c = c0; c < c0 + PER && c < NC; c++)
        {
            const int k = c / L + (h.jmin + c % L) * NT;
            mine += (k >= g.kmin && k < g.kmax);
        }
        int pos = block_exclusive_scan<kLayerThreads>(mine, cnt);
        if (tid == 0) a.nwave[e] = cnt[kLayerThreads / 64];
        for (int c = c0; c < c0 + PER && c < NC; c++)
        {
            const int k = c / L + (h.jmin + c % L) * NT;
            if (k >= g.kmin && k < g.kmax) lst[pos++] = k - g.kmin;
        }
    }
    __syncthreads();
    const int Nw = cnt[kLayerThreads / 64];
    const double* win = g.win[L];
    const double norm = sqrt((double)g.NF) / (L + 1.0);
    double* vals = a.values + (size_t)e * 3 * a.vcap;
    for (int ch = 0; ch < 3; ch++)
    {
        // windowed bands of the L layers, conjugated so the forward FFT
        // gives the conjugate of the inverse transform
        for (int it = tid; it < L * NT; it += kLayerThreads)
        {
            const int off = it / NT, jj = it % NT;
            const int j = jj - NT / 2;
            const int m = h.jmin + off;
            const int center = (off + 1) * NT / 2;
            const int idx = center + j;
            double2 v = make_double2(0.0, 0.0);
            if (jj > 0 && idx >= 0 && idx <= NDS / 2)
            {
                double wt = win[j < 0 ? -j : j];
                if (j == 0 && (m == 0 || m == g.NF)) wt /= 2.0;
                const double2 S = spectrum<NDS>(ch, idx, U, V, twU, tsU);
                v = make_double2(wt * S.x, -(wt * S.y));
            }
            B[off * NT + jj] = v;
        }
        __syncthreads();
        lds_fft<NT, L, kLayerThreads>(B, g.tw, kTwN / NT);
        for (int it = tid; it < L * NT; it += kLayerThreads)
        {
            const int n = it / L, off = it % L;
            const int m = h.jmin + off;
            const double2 r = B[off * NT + n];
            const double re = r.x / NT, im = -r.y / NT;
            double val;
            if (m == 0 || m == g.NF) val = (n % 2 == 0) ? re * M_SQRT2 * norm : 0.0;
            else
            {
                const int sgn = (m % 2 == 0) ? 1 : -1;
                val = ((n + m) % 2 == 0) ? re * norm : sgn * im * norm;
            }
            const int pos = n * L + off;
            if (pos < Nw) vals[ch * a.vcap + pos] = val;
        }
        __syncthreads();
    }
    if (!a.data) return;
    __threadfence_block();
    __syncthreads();
    // wavelet_nwip terms over the pixel list: snr^2 = sum h^T C h; the
    // logL change of adding the template to a zero model with the
    // reference's edge rule (gaussian_log_likelhood_wavelet,
    // utils/src/glass_model.c:805-855)
    double hh = 0.0, dl = 0.0;
    for (int t = tid; t < Nw; t += kLayerThreads)
    {
        const int k = lst[t];
        const double hv[3] = {vals[t], vals[a.vcap + t], vals[2 * a.vcap + t]};
        double C[3][3];
        C[0][0] = a.invC[k];
        C[1][1] = a.invC[(size_t)a.npix + k];
        C[2][2] = a.invC[2 * (size_t)a.npix + k];
        C[0][1] = C[1][0] = a.invC[3 * (size_t)a.npix + k];
        C[0][2] = C[2][0] = a.invC[4 * (size_t)a.npix + k];
        C[1][2] = C[2][1] = a.invC[5 * (size_t)a.npix + k];
        const double d[3] = {a.data[k], a.data[(size_t)a.npix + k], a.data[2 * (size_t)a.npix + k]};
        double th = 0.0, tdh = 0.0, tdd = 0.0;
        for (int p = 0; p < 3; p++)
            for (int q = 0; q < 3; q++)
            {
                th += hv[p] * hv[q] * C[p][q];
                tdh += d[p] * hv[q] * C[p][q];
                tdd += d[p] * d[q] * C[p][q];
            }
        hh += th;
        const int i = (k + g.kmin) % g.NT;
        if (k >= 0 && k < a.npix) dl += (i < g.imin || i > g.imax) ? 0.5 * tdd : tdh - 0.5 * th;
    }
    hh = block_sum<kLayerThreads>(hh);
    dl = block_sum<kLayerThreads>(dl);
    if (tid == 0)
    {
        a.snr[e] = sqrt(hh);
        a.dlogL[e] = dl;
    }
}

template <int NT>
__global__ void __launch_bounds__(kLayerThreads) layer_kernel(Geo g, int n, LayerArgs a,
                                                              const double* __restrict__ scratch)
{
    const int e = blockIdx.x;
    if (e >= n) return;
    extern __shared__ double2 lds2[];
    const double* scr = scratch + scratch_doubles(g.Norb) * (size_t)e;
    __shared__ Hdr H;
    if (threadIdx.x == 0) H = *(const Hdr*)scr;
    __syncthreads();
    const Hdr h = H;
    // LDS: U (4 NT) | V (2 NT) | B (3 NT); list counts in static LDS
    __shared__ int cnt[kLayerThreads + 1];
    double2* U = lds2;
    double2* V = U + 4 * NT;
    double2* B = V + 2 * NT;
    if (h.L == 1) layer_body<NT, 1>(g, h, e, a, U, V, B, cnt, scr);
    else if (h.L == 2) layer_body<NT, 2>(g, h, e, a, U, V, B, cnt, scr);
    else layer_body<NT, 3>(g, h, e, a, U, V, B, cnt, scr);
}

template <int NT>
constexpr size_t layer_lds_bytes()
{
    return sizeof(double2) * 9 * NT;
}

// ---------------------------------------------- K2 for NT of 1024..4096 --
//
// The heterodyned series of NT (L+1) <= 16384 samples and its spectrum do
// not fit one workgroup's LDS: they live in a global scratch slice per
// workgroup (U = X + iY and V = Z packed as before, each with a partner
// buffer) and are transformed by ping-pong Stockham stages through L2, with
// the same butterflies, twiddle values and stage order as the LDS path.
// The windowed band of each layer (NT points, 64 KB at NT = 4096) is
// inverse-transformed in LDS, one layer at a time.  Workgroups stride over
// the templates.

constexpr int kBigLayerThreads = 512;

// double2 per scratch slice: U and partner (NDS each) and V and partner
// (NDS / 2 each), NDS = NT (L + 1) at the largest L (3, or a forced count)
__host__ __device__ inline size_t big_layer_slice(int NT, int Lmax = kMaxLayers)
{
    return 3 * (size_t)NT * (size_t)((Lmax > kMaxLayers ? Lmax : kMaxLayers) + 1);
}

template <int NT, int L>
__device__ void layer_body_big(const Geo& g, const Hdr& h, int e, const LayerArgs& a, double2* G0, double2* B,
                               int* cnt, const double* __restrict__ scr)
{
    constexpr int NDS = NT * (L + 1);
    constexpr int M = NDS / 2;
    constexpr int NTHR = kBigLayerThreads;
    constexpr int SLOT = L > kMaxLayers ? kForcedSlot : L;
    const int tid = threadIdx.x;
    const int N = g.Norb;
    const double* cph = scr + 64;
    const double* ctdi = scr + 64 + 4 * N;
    const double* ds = g.ds[SLOT];
    const int* dsi = g.dsi[SLOT];
    double2* Ua = G0;
    double2* Ub = G0 + NDS;
    double2* Va = G0 + 2 * NDS;
    double2* Vb = Va + M;
    {
#pragma clang fp contract(off)
        const double w0 = PI2 * h.fhet;
        for (int i = tid; i < NDS; i += NTHR)
        {
            const double* e5 = ds + 5 * (size_t)i;
            const double t = e5[0];
            double kdotr = 0.0;
            for (int d = 0; d < 3; d++) kdotr += h.k[d] * e5[1 + d];
            const double tsc = t - kdotr;
            const int j = even_index(tsc, g.ot0, g.ods, N);
            const double phase_ds = spline_at(cph, N, j, tsc - g.ot[j]);
            const double phase_het = w0 * t;
            const int si = dsi[i];
            const double sdx = e5[4];
            double x[3];
            for (int ch = 0; ch < 3; ch++)
            {
                const double* ca = ctdi + 4 * (size_t)N * (2 * ch);
                const double* cp = ctdi + 4 * (size_t)N * (2 * ch + 1);
                const double Amp = spline_at(ca, N, si, sdx);
                double Phase = spline_at(cp, N, si, sdx);
                Phase += phase_ds;
                Phase -= phase_het;
                double sn, c;
                sincos_red(Phase, &sn, &c);
                x[ch] = Amp * c;
            }
            Ua[i] = make_double2(x[0], x[1]);
            if (i & 1) Va[i >> 1].y = x[2];
            else Va[i >> 1].x = x[2];
        }
    }
    __threadfence_block();
    __syncthreads();
    const bool three = (NDS % 3) == 0;
    const double2* twU = three ? g.tw3 : (NDS > kTw16N ? g.tw32 : g.tw16);
    const int tsU = (three ? kTw3N : (NDS > kTw16N ? kTw32N : kTw16N)) / NDS;
    const double2* U = pingpong_fft<NDS, NTHR>(Ua, Ub, twU, tsU);
    const double2* V = pingpong_fft<M, NTHR>(Va, Vb, twU, tsU * 2);
    int* lst = a.list + (size_t)e * a.vcap;
    {
        constexpr int NC = NT * L;
        constexpr int PER = (NC + NTHR - 1) / NTHR;
        const int c0 = tid * PER;
        int mine = 0;
        for (int c = c0; c < c0 + PER && c < NC; c++)
        {
            const int k = c / L + (h.jmin + c % L) * NT;
            mine += (k >= g.kmin && k < g.kmax);
        }
        int pos = block_exclusive_scan<NTHR>(mine, cnt);
        if (tid == 0) a.nwave[e] = cnt[NTHR / 64];
        for (int c = c0; c < c0 + PER && c < NC; c++)
        {
            const int k = c / L + (h.jmin + c % L) * NT;
            if (k >= g.kmin && k < g.kmax) lst[pos++] = k - g.kmin;
        }
    }
    __syncthreads();
    const int Nw = cnt[NTHR / 64];
    const double* win = g.win[SLOT];
    const double norm = sqrt((double)g.NF) / (L + 1.0);
    double* vals = a.values + (size_t)e * 3 * a.vcap;
    for (int ch = 0; ch < 3; ch++)
        for (int off = 0; off < L; off++)
        {
            const int m = h.jmin + off;
            const int center = (off + 1) * NT / 2;
            for (int jj = tid; jj < NT; jj += NTHR)
            {
                const int j = jj - NT / 2;
                const int idx = center + j;
                double2 v = make_double2(0.0, 0.0);
                if (jj > 0 && idx >= 0 && idx <= NDS / 2)
                {
                    double wt = win[j < 0 ? -j : j];
                    if (j == 0 && (m == 0 || m == g.NF)) wt /= 2.0;
                    const double2 S = spectrum<NDS>(ch, idx, U, V, twU, tsU);
                    v = make_double2(wt * S.x, -(wt * S.y));
                }
                B[jj] = v;
            }
            __syncthreads();
            lds_fft<NT, 1, NTHR>(B, g.tw, kTwN / NT);
            for (int nn = tid; nn < NT; nn += NTHR)
            {
                const double2 r = B[nn];
                const double re = r.x / NT, im = -r.y / NT;
                double val;
                if (m == 0 || m == g.NF) val = (nn % 2 == 0) ? re * M_SQRT2 * norm : 0.0;
                else
                {
                    const int sgn = (m % 2 == 0) ? 1 : -1;
                    val = ((nn + m) % 2 == 0) ? re * norm : sgn * im * norm;
                }
                const int pos = nn * L + off;
                if (pos < Nw) vals[ch * a.vcap + pos] = val;
            }
            __syncthreads();
        }
    if (a.data)
    {
        __threadfence_block();
        __syncthreads();
        double hh = 0.0, dl = 0.0;
        for (int t = tid; t < Nw; t += NTHR)
        {
            const int k = lst[t];
            const double hv[3] = {vals[t], vals[a.vcap + t], vals[2 * a.vcap + t]};
            double C[3][3];
            C[0][0] = a.invC[k];
            C[1][1] = a.invC[(size_t)a.npix + k];
            C[2][2] = a.invC[2 * (size_t)a.npix + k];
            C[0][1] = C[1][0] = a.invC[3 * (size_t)a.npix + k];
            C[0][2] = C[2][0] = a.invC[4 * (size_t)a.npix + k];
            C[1][2] = C[2][1] = a.invC[5 * (size_t)a.npix + k];
            const double d[3] = {a.data[k], a.data[(size_t)a.npix + k], a.data[2 * (size_t)a.npix + k]};
            double th = 0.0, tdh = 0.0, tdd = 0.0;
            for (int p = 0; p < 3; p++)
                for (int q = 0; q < 3; q++)
                {
                    th += hv[p] * hv[q] * C[p][q];
                    tdh += d[p] * hv[q] * C[p][q];
                    tdd += d[p] * d[q] * C[p][q];
                }
            hh += th;
            const int i = (k + g.kmin) % g.NT;
            if (k >= 0 && k < a.npix) dl += (i < g.imin || i > g.imax) ? 0.5 * tdd : tdh - 0.5 * th;
        }
        hh = block_sum<NTHR>(hh);
        dl = block_sum<NTHR>(dl);
        if (tid == 0)
        {
            a.snr[e] = sqrt(hh);
            a.dlogL[e] = dl;
        }
    }
}

template <int NT>
__global__ void __launch_bounds__(kBigLayerThreads) layer_kernel_big(Geo g, int n, LayerArgs a,
                                                                     const double* __restrict__ scratch)
{
    extern __shared__ double2 ldsb[];  // B: one layer's band, NT complex
    __shared__ Hdr H;
    __shared__ int cnt[kBigLayerThreads + 1];
    double2* G0 = a.big + a.bigslice * (size_t)blockIdx.x;
    for (int e = blockIdx.x; e < n; e += gridDim.x)
    {
        const double* scr = scratch + scratch_doubles(g.Norb) * (size_t)e;
        __syncthreads();  // the previous template is done with H, cnt and the slice
        if (threadIdx.x == 0) H = *(const Hdr*)scr;
        __syncthreads();
        const Hdr h = H;
        if (h.L == 1) layer_body_big<NT, 1>(g, h, e, a, G0, ldsb, cnt, scr);
        else if (h.L == 2) layer_body_big<NT, 2>(g, h, e, a, G0, ldsb, cnt, scr);
        else if (h.L == 3) layer_body_big<NT, 3>(g, h, e, a, G0, ldsb, cnt, scr);
        else if constexpr (NT <= 512)
        {
            // forced layer counts (forced_layers_ok)
            if (h.L == 7) layer_body_big<NT, 7>(g, h, e, a, G0, ldsb, cnt, scr);
            else if (h.L == 15) layer_body_big<NT, 15>(g, h, e, a, G0, ldsb, cnt, scr);
            else if (h.L == 31) layer_body_big<NT, 31>(g, h, e, a, G0, ldsb, cnt, scr);
            else if (h.L == 63) layer_body_big<NT, 63>(g, h, e, a, G0, ldsb, cnt, scr);
        }
    }
}

// ----------------------------------------------------------- launching --

static int launch_templates(const GeoHost& G, int n, const double* d_params, double* d_scratch,
                            const LayerArgs& a, hipStream_t s)
{
    const Geo& g = G.g;
    if (n <= 0) return GLASS_HIP_OK;
    hipLaunchKernelGGL(hdr_kernel, dim3((n + 255) / 256), dim3(256), 0, s, g, n, d_params, d_scratch);
    hipLaunchKernelGGL(resp_kernel, dim3(n), dim3(kRespThreads), resp_lds_bytes(g.Norb), s, g, n, d_params,
                       d_scratch);
    GH_CHECK(hipGetLastError());
    if (g.NT >= 1024 || g.Lf > 0)
    {
        if (!a.big || a.nbig <= 0 || a.bigslice < big_layer_slice(g.NT, g.Lf) || a.vcap < g.NT * (g.Lf > 3 ? g.Lf : 3))
        {
            set_error("wavelet templates: NT = %d (forced layers %d) needs the long-series scratch", g.NT, g.Lf);
            return GLASS_HIP_EINVAL;
        }
        const int grid = n < a.nbig ? n : a.nbig;
        const size_t lds = sizeof(double2) * (size_t)g.NT;
        if (g.NT == 256)
            hipLaunchKernelGGL(layer_kernel_big<256>, dim3(grid), dim3(kBigLayerThreads), lds, s, g, n, a,
                               d_scratch);
        else if (g.NT == 512)
            hipLaunchKernelGGL(layer_kernel_big<512>, dim3(grid), dim3(kBigLayerThreads), lds, s, g, n, a,
                               d_scratch);
        else if (g.NT == 1024)
            hipLaunchKernelGGL(layer_kernel_big<1024>, dim3(grid), dim3(kBigLayerThreads), lds, s, g, n, a,
                               d_scratch);
        else if (g.NT == 2048)
            hipLaunchKernelGGL(layer_kernel_big<2048>, dim3(grid), dim3(kBigLayerThreads), lds, s, g, n, a,
                               d_scratch);
        else
            hipLaunchKernelGGL(layer_kernel_big<4096>, dim3(grid), dim3(kBigLayerThreads), lds, s, g, n, a,
                               d_scratch);
    }
    else if (g.NT == 512)
        hipLaunchKernelGGL(layer_kernel<512>, dim3(n), dim3(kLayerThreads), layer_lds_bytes<512>(), s, g, n, a,
                           d_scratch);
    else
        hipLaunchKernelGGL(layer_kernel<256>, dim3(n), dim3(kLayerThreads), layer_lds_bytes<256>(), s, g, n, a,
                           d_scratch);
    GH_CHECK(hipGetLastError());
    return GLASS_HIP_OK;
}

// The calling thread's long-series scratch for n templates of grid NT
// (nothing below NT = 1024).  Sets the default pixel capacity 3 NT.
static int big_scratch(ThreadCtx* c, int NT, int n, LayerArgs& a)
{
    a.vcap = 3 * NT;
    if (NT < 1024) return GLASS_HIP_OK;
    a.bigslice = big_layer_slice(NT);
    if (int rc = c->dscr.ensure(sizeof(double2) * a.bigslice * (size_t)n)) return rc;
    a.big = (double2*)c->dscr.p;
    a.nbig = n;
    return GLASS_HIP_OK;
}

static int set_lds_limits()
{
    static std::once_flag once;
    static int rc = GLASS_HIP_OK;
    std::call_once(once, [] {
        if (hipFuncSetAttribute((const void*)resp_kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)resp_lds_bytes(kMaxKnots)) != hipSuccess ||
            hipFuncSetAttribute((const void*)layer_kernel<512>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)layer_lds_bytes<512>()) != hipSuccess ||
            hipFuncSetAttribute((const void*)layer_kernel_big<4096>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)(sizeof(double2) * 4096)) != hipSuccess)
        {
            set_error("hipFuncSetAttribute (dynamic LDS) failed");
            rc = GLASS_HIP_ERUNTIME;
        }
    });
    return rc;
}

// ------------------------------------------------ K3: wavelet Fisher --
//
// ucb_fisher_wavelet (ucb/src/glass_ucb_waveform.c:524-696) before the
// eigensolve, for sources whose 16 perturbed templates (slot 16 s + 2 i:
// params[i] + eps/2, slot + 1: params[i] - eps/2) are already in the
// template buffers.  One workgroup per source: the layer band the 16 pixel
// lists span becomes a dense local window of NT x (band) pixels in scratch;
// per parameter the minus template is scattered into a zeroed window and
// the derivative (h+ - h-)/eps written at the plus template's pixels with
// time index strictly inside (imin, imax); the 36 Fisher elements are then
// reduced over the window with the nine invC terms per pixel.

constexpr int kFishThreads = 256;
constexpr int kFishMaxBand = 6;  // layers the 16 templates of a source may span

struct FishArgs
{
    const int* nwave;      // [16 n]
    const int* list;       // [16 n][3 NT]
    const double* values;  // [16 n][3][3 NT]
    const double* invC;    // plane p of channel pair (a,b) = pmap[3a+b], element k - koff
    long long pstride;     // doubles between planes
    int koff, klen;        // data index of element 0; elements per plane
    int pmap[9];
    int npix;
    double* D;             // scratch [n][8][3][kFishMaxBand NT]
    double* M;             // scratch [n][3][kFishMaxBand NT]
    double* fisher;        // [n][64]
};

__device__ inline double fisher_terms(const double (&u)[3], const double (&v)[3], const double (&C)[9])
{
    // XX, YY, ZZ, XY, XZ, YZ, YX, ZX, ZY (glass_ucb_waveform.c:608-617)
    double f = u[0] * v[0] * C[0];
    f += u[1] * v[1] * C[4];
    f += u[2] * v[2] * C[8];
    f += u[0] * v[1] * C[1];
    f += u[0] * v[2] * C[2];
    f += u[1] * v[2] * C[5];
    f += u[1] * v[0] * C[3];
    f += u[2] * v[0] * C[6];
    f += u[2] * v[1] * C[7];
    return f;
}

__global__ void __launch_bounds__(kFishThreads) fisher_kernel(Geo g, int n, FishArgs a)
{
    const int s = blockIdx.x;
    if (s >= n) return;
    const int NT = g.NT, tid = threadIdx.x, cap = 3 * NT;
    const double invstep = 1. / 1.0e-6;
    __shared__ int band[2];
    __shared__ double red[36][kFishThreads / 64];
    if (tid == 0)
    {
        band[0] = 0x7fffffff;
        band[1] = -1;
    }
    __syncthreads();
    int jlo = 0x7fffffff, jhi = -1;
    for (int t = 0; t < 16; t++)
    {
        const int e = 16 * s + t;
        const int nw = a.nwave[e];
        const int* L = a.list + (size_t)e * cap;
        for (int m = tid; m < nw; m += kFishThreads)
        {
            const int j = (L[m] + g.kmin) / NT;
            jlo = min(jlo, j);
            jhi = max(jhi, j);
        }
    }
    atomicMin(&band[0], jlo);
    atomicMax(&band[1], jhi);
    __syncthreads();
    jlo = band[0];
    jhi = band[1];
    double* F = a.fisher + (size_t)s * 64;
    if (jhi < jlo || jhi - jlo + 1 > kFishMaxBand)
    {
        // no pixels: a zero matrix; a band too wide: NaN marks the failure
        const double v = jhi < jlo ? 0.0 : __builtin_nan("");
        for (int q = tid; q < 64; q += kFishThreads) F[q] = v;
        return;
    }
    const int W = (jhi - jlo + 1) * NT;
    double* D = a.D + (size_t)s * 24 * kFishMaxBand * NT;
    double* M = a.M + (size_t)s * 3 * kFishMaxBand * NT;
    for (int q = tid; q < 24 * W; q += kFishThreads) D[q] = 0.0;
    const int base = jlo * NT;
    for (int i = 0; i < 8; i++)
    {
        const int ep = 16 * s + 2 * i, em = ep + 1;
        for (int q = tid; q < 3 * W; q += kFishThreads) M[q] = 0.0;
        __syncthreads();
        {
            const int nw = a.nwave[em];
            const int* L = a.list + (size_t)em * cap;
            const double* V = a.values + (size_t)em * 3 * cap;
            for (int m = tid; m < nw; m += kFishThreads)
            {
                const int q = L[m] + g.kmin - base;
                for (int ch = 0; ch < 3; ch++) M[ch * W + q] = V[ch * cap + m];
            }
        }
        __syncthreads();
        {
            const int nw = a.nwave[ep];
            const int* L = a.list + (size_t)ep * cap;
            const double* V = a.values + (size_t)ep * 3 * cap;
            for (int m = tid; m < nw; m += kFishThreads)
            {
                const int k = L[m];
                const int it = (k + g.kmin) % NT;
                if (it > g.imin && it < g.imax && k >= 0 && k < a.npix)
                {
                    const int q = k + g.kmin - base;
                    for (int ch = 0; ch < 3; ch++)
                        D[(i * 3 + ch) * W + q] = (V[ch * cap + m] - M[ch * W + q]) * invstep;
                }
            }
        }
        __syncthreads();
    }
    double acc[36];
#pragma unroll
    for (int p = 0; p < 36; p++) acc[p] = 0.0;
    bool miss = false;
    for (int q = tid; q < W; q += kFishThreads)
    {
        const int k = base + q - g.kmin;
        if (k < 0 || k >= a.npix) continue;
        double d[8][3];
        bool any = false;
#pragma unroll
        for (int i = 0; i < 8; i++)
#pragma unroll
            for (int ch = 0; ch < 3; ch++)
            {
                d[i][ch] = D[(i * 3 + ch) * W + q];
                any |= d[i][ch] != 0.0;
            }
        if (!any) continue;
        const int ki = k - a.koff;
        if (ki < 0 || ki >= a.klen)
        {
            miss = true;
            continue;
        }
        double C[9];
#pragma unroll
        for (int p = 0; p < 9; p++) C[p] = a.invC[a.pmap[p] * a.pstride + ki];
        int idx = 0;
#pragma unroll
        for (int i = 0; i < 8; i++)
#pragma unroll
            for (int j = i; j < 8; j++) acc[idx++] += fisher_terms(d[i], d[j], C);
    }
    const int lane = tid & 63, wv = tid >> 6;
#pragma unroll
    for (int p = 0; p < 36; p++)
    {
        const double v = wave_sum(acc[p]);
        if (lane == 0) red[p][wv] = v;
    }
    __shared__ int missed;
    if (tid == 0) missed = 0;
    __syncthreads();
    if (miss) atomicOr(&missed, 1);
    __syncthreads();
    if (tid < 36)
    {
        double f = 0.0;
        for (int w = 0; w < kFishThreads / 64; w++) f += red[tid][w];
        if (f != f) f = 10.0;
        if (missed) f = __builtin_nan("");
        // element tid of the upper triangle, row-major
        int i = 0, r = tid;
        while (r >= 8 - i)
        {
            r -= 8 - i;
            i++;
        }
        const int j = i + r;
        F[i * 8 + j] = f;
        F[j * 8 + i] = f;
    }
}

// The 16 perturbed parameter vectors of each of n sources (slot 16 s + 2 i
// and + 1), as ucb_fisher_wavelet forms them (glass_ucb_waveform.c:556-576).
__global__ void fisher_params_kernel(int n, const double* __restrict__ params, double* __restrict__ out)
{
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= 16 * n) return;
    const int s = q / 16, t = q % 16, i = t / 2, sgn = t % 2;
    const double epsilon = 1.0e-6;
    double* o = out + (size_t)q * 8;
    for (int k = 0; k < 8; k++) o[k] = params[(size_t)s * 8 + k];
    if (sgn == 0) o[i] += epsilon / 2.;
    else o[i] -= epsilon / 2.;
    if (i == 1 || i == 4)
    {
        if (sgn == 0 && o[i] > 1.0) o[i] = 1.0;
        if (sgn == 1 && o[i] < -1.0) o[i] = -1.0;
    }
}

// -------------------------------------------------- geometry cache ------

struct GeoKey
{
    std::vector<double> blob;
    bool operator<(const GeoKey& o) const { return blob < o.blob; }
};

static std::mutex g_geo_mu;
static std::map<std::pair<int, GeoKey>, GeoHost*> g_geo;

static int geo_get(const glass_hip_wdm* w, int Norb, const double* ot, const double* ox, const double* oy,
                   const double* oz, double Tobs, double t0, const GeoHost** out, int Lf = 0)
{
    int rc = check_grid(w, Norb);
    if (rc) return rc;
    if (Lf != 0 && !forced_layers_ok(w->NT, Lf))
        return set_error("wavelet templates: forced layer count %d unsupported at NT = %d (L + 1 a power of two "
                         "from 8 to 64 with NT (L + 1) <= %d, NT = 256 or 512)",
                         Lf, w->NT, kMaxForcedSamples),
               GLASS_HIP_EINVAL;
    int dev = 0;
    GH_CHECK(hipGetDevice(&dev));
    GeoKey key;
    key.blob = {(double)w->NF, (double)w->NT, w->dt, w->df, w->A, (double)w->kmin, (double)w->kmax,
                (double)w->imin, (double)w->imax, Tobs, t0, (double)Norb, (double)Lf};
    key.blob.insert(key.blob.end(), ot, ot + Norb);
    // pos rows: spacecraft i coordinate d at 3i+d
    std::vector<double> pos(9 * (size_t)Norb);
    for (int i = 0; i < 3; i++)
    {
        std::memcpy(pos.data() + (3 * i + 0) * (size_t)Norb, ox + i * (size_t)Norb, sizeof(double) * Norb);
        std::memcpy(pos.data() + (3 * i + 1) * (size_t)Norb, oy + i * (size_t)Norb, sizeof(double) * Norb);
        std::memcpy(pos.data() + (3 * i + 2) * (size_t)Norb, oz + i * (size_t)Norb, sizeof(double) * Norb);
    }
    key.blob.insert(key.blob.end(), pos.begin(), pos.end());
    std::lock_guard<std::mutex> lk(g_geo_mu);
    auto it = g_geo.find({dev, key});
    if (it != g_geo.end())
    {
        *out = it->second;
        return GLASS_HIP_OK;
    }
    if ((rc = set_lds_limits())) return rc;
    GeoHost* G = new GeoHost();
    if ((rc = geo_build(*G, *w, Norb, ot, pos.data(), Tobs, t0, Lf)))
    {
        delete G;
        return rc;
    }
    g_geo[{dev, key}] = G;
    *out = G;
    return GLASS_HIP_OK;
}

}  // namespace wdm
}  // namespace glass

using namespace glass;
using namespace glass::wdm;

// ------------------------------------------------------------- drop-ins --

extern "C" int glass_hip_interpolated_analytic_orbit(const glass_hip_orbit* orbit, double Tobs, double t0, int* Norb,
                                                     double* t, double* x, double* y, double* z)
{
#pragma clang fp contract(off)
    if (!orbit || !Norb || Tobs <= 0) return set_error("interpolated_analytic_orbit: bad arguments"), GLASS_HIP_EINVAL;
    // initialize_interpolated_analytic_orbits (utils/src/glass_lisa.c:271-380)
    const double YEAR = 3.15581497632e7;
    int N = (int)(200 * Tobs / YEAR);
    if (N < 20) N = 20;
    *Norb = N;
    if (!t) return GLASS_HIP_OK;
    if (!x || !y || !z) return set_error("interpolated_analytic_orbit: NULL position array"), GLASS_HIP_EINVAL;
    const double buffer = Tobs / (double)(N - 1);
    const double cadence = (Tobs + 2.0 * buffer) / (double)(N - 1);
    for (int n = 0; n < N; n++)
    {
        const double tn = t0 + n * cadence - buffer;
        double p9[9];
        analytic_positions(*orbit, tn, p9);
        t[n] = tn;
        for (int i = 0; i < 3; i++)
        {
            x[i * (size_t)N + n] = p9[i];
            y[i * (size_t)N + n] = p9[3 + i];
            z[i * (size_t)N + n] = p9[6 + i];
        }
    }
    return GLASS_HIP_OK;
}

extern "C" int glass_hip_ucb_waveform_wavelet(const glass_hip_wdm* w, int Norb, const double* orbit_t,
                                              const double* orbit_x, const double* orbit_y,
                                              const double* orbit_z, double Tobs, double t0,
                                              const double* params, int* wavelet_list, int* Nwavelet, double* X,
                                              double* Y, double* Z)
{
    if (!orbit_t || !orbit_x || !orbit_y || !orbit_z || !params || !wavelet_list || !Nwavelet || !X || !Y || !Z)
        return set_error("ucb_waveform_wavelet: NULL argument"), GLASS_HIP_EINVAL;
    ThreadCtx* c;
    int rc = ctx_get(&c);
    if (rc) return rc;
    const GeoHost* G;
    if ((rc = geo_get(w, Norb, orbit_t, orbit_x, orbit_y, orbit_z, Tobs, t0, &G))) return rc;
    const int NT = w->NT;
    // device: params 8 | scratch | list 3NT ints | values 9NT | nwave
    const size_t nscr = scratch_doubles(Norb);
    const size_t bytes = sizeof(double) * (8 + nscr + 9 * NT) + sizeof(int) * (3 * NT + 4);
    if ((rc = c->dbuf.ensure(bytes)) || (rc = c->hbuf.ensure(bytes))) return rc;
    double* dp = (double*)c->dbuf.p;
    double* dscr = dp + 8;
    double* dval = dscr + nscr;
    int* dlist = (int*)(dval + 9 * NT);
    int* dnw = dlist + 3 * NT;
    double* hp = (double*)c->hbuf.p;
    std::memcpy(hp, params, sizeof(double) * 8);
    GH_CHECK(hipMemcpyAsync(dp, hp, sizeof(double) * 8, hipMemcpyHostToDevice, c->stream));
    LayerArgs a{};
    a.nwave = dnw;
    a.list = dlist;
    a.values = dval;
    if ((rc = big_scratch(c, NT, 1, a))) return rc;
    if ((rc = launch_templates(*G, 1, dp, dscr, a, c->stream))) return rc;
    GH_CHECK(hipMemcpyAsync(hp, dval, sizeof(double) * 9 * NT + sizeof(int) * (3 * NT + 1), hipMemcpyDeviceToHost,
                            c->stream));
    GH_CHECK(hipStreamSynchronize(c->stream));
    const double* hv = hp;
    const int* hl = (const int*)(hp + 9 * NT);
    const int nw = hl[3 * NT];
    *Nwavelet = nw;
    // insert the pixels at their list positions (the caller pre-zeroes, as
    // the reference's callers do)
    for (int t = 0; t < nw; t++)
    {
        wavelet_list[t] = hl[t];
        X[hl[t]] = hv[t];
        Y[hl[t]] = hv[3 * NT + t];
        Z[hl[t]] = hv[6 * NT + t];
    }
    return GLASS_HIP_OK;
}

// ---------------------------------------------------------------- batch --

struct glass_hip_wdm_batch
{
    int device = 0;
    const GeoHost* G = nullptr;
    int npix = 0, max_templates = 0;
    double* data = nullptr;   // 3 x npix
    double* invC = nullptr;   // 6 x npix
    double* scratch = nullptr;
    int* list = nullptr;
    double* values = nullptr;
    int* nwave = nullptr;
    // wavelet Fisher: perturbed parameters [max_templates x 8], derivative
    // and minus-template windows [max_templates / 16 sources] (allocated on
    // first use)
    double* fparams = nullptr;
    double* fD = nullptr;
    double* fM = nullptr;
    // long-series scratch of the NT >= 1024 (or forced layer count) layer
    // kernel: nbig slices of bigslice double2
    double2* big = nullptr;
    int nbig = 0;
    size_t bigslice = 0;
    int vcap = 0;  // pixels per template and channel
    int layers = 0;
};

static int wdm_batch_create(glass_hip_wdm_batch** plan, const glass_hip_wdm* w, int Norb, const double* orbit_t,
                            const double* orbit_x, const double* orbit_y, const double* orbit_z, double Tobs,
                            double t0, const double* data, const double* invC, int npix, int max_templates,
                            int layers)
{
    if (!plan || !orbit_t || !orbit_x || !orbit_y || !orbit_z || !data || !invC)
        return set_error("wdm_batch_create: NULL argument"), GLASS_HIP_EINVAL;
    if (npix <= 0 || max_templates <= 0)
        return set_error("wdm_batch_create: npix=%d max_templates=%d", npix, max_templates), GLASS_HIP_EINVAL;
    ThreadCtx* c;
    int rc = ctx_get(&c);
    if (rc) return rc;
    auto* P = new glass_hip_wdm_batch();
    P->device = c->device;
    auto fail = [&](int r) {
        glass_hip_wdm_batch_destroy(P);
        return r;
    };
    if ((rc = geo_get(w, Norb, orbit_t, orbit_x, orbit_y, orbit_z, Tobs, t0, &P->G, layers))) return fail(rc);
    if (w->kmax - w->kmin != npix)
    {
        set_error("wdm_batch_create: npix = %d but kmax - kmin = %d", npix, w->kmax - w->kmin);
        return fail(GLASS_HIP_EINVAL);
    }
    P->npix = npix;
    P->max_templates = max_templates;
    P->layers = layers;
    const int NT = w->NT;
    P->vcap = NT * (layers > kMaxLayers ? layers : kMaxLayers);
    if (hipMalloc(&P->data, sizeof(double) * 3 * (size_t)npix) != hipSuccess ||
        hipMalloc(&P->invC, sizeof(double) * 6 * (size_t)npix) != hipSuccess ||
        hipMalloc(&P->scratch, sizeof(double) * scratch_doubles(Norb) * (size_t)max_templates) != hipSuccess ||
        hipMalloc(&P->list, sizeof(int) * (size_t)P->vcap * max_templates) != hipSuccess ||
        hipMalloc(&P->values, sizeof(double) * 3 * (size_t)P->vcap * max_templates) != hipSuccess ||
        hipMalloc(&P->nwave, sizeof(int) * (size_t)max_templates) != hipSuccess)
    {
        set_error("wdm_batch_create: device allocation failed");
        return fail(GLASS_HIP_ERUNTIME);
    }
    if (NT >= 1024 || layers > 0)
    {
        // the layer kernel strides over the templates with two workgroups
        // per CU, each with its own long-series slice
        int ncu = 0;
        GH_CHECK(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, P->device));
        P->nbig = std::min(max_templates, 2 * std::max(ncu, 1));
        P->bigslice = big_layer_slice(NT, layers);
        if (hipMalloc(&P->big, sizeof(double2) * P->bigslice * (size_t)P->nbig) != hipSuccess)
        {
            set_error("wdm_batch_create: device allocation failed");
            return fail(GLASS_HIP_ERUNTIME);
        }
    }
    if (hipMemcpy(P->data, data, sizeof(double) * 3 * (size_t)npix, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(P->invC, invC, sizeof(double) * 6 * (size_t)npix, hipMemcpyHostToDevice) != hipSuccess)
    {
        set_error("wdm_batch_create: upload failed");
        return fail(GLASS_HIP_ERUNTIME);
    }
    *plan = P;
    return GLASS_HIP_OK;
}

extern "C" int glass_hip_wdm_batch_create(glass_hip_wdm_batch** plan, const glass_hip_wdm* w, int Norb,
                                          const double* orbit_t, const double* orbit_x, const double* orbit_y,
                                          const double* orbit_z, double Tobs, double t0, const double* data,
                                          const double* invC, int npix, int max_templates)
{
    return wdm_batch_create(plan, w, Norb, orbit_t, orbit_x, orbit_y, orbit_z, Tobs, t0, data, invC, npix,
                            max_templates, 0);
}

extern "C" int glass_hip_wdm_batch_create_layers(glass_hip_wdm_batch** plan, const glass_hip_wdm* w, int Norb,
                                                 const double* orbit_t, const double* orbit_x,
                                                 const double* orbit_y, const double* orbit_z, double Tobs,
                                                 double t0, const double* data, const double* invC, int npix,
                                                 int max_templates, int layers)
{
    if (layers <= 0)
        return set_error("wdm_batch_create_layers: layer count %d (use glass_hip_wdm_batch_create for the "
                         "reference's own)",
                         layers),
               GLASS_HIP_EINVAL;
    return wdm_batch_create(plan, w, Norb, orbit_t, orbit_x, orbit_y, orbit_z, Tobs, t0, data, invC, npix,
                            max_templates, layers);
}

extern "C" int glass_hip_wdm_batch_capacity(const glass_hip_wdm_batch* P, int* vcap)
{
    if (!P || !vcap) return set_error("wdm_batch_capacity: NULL argument"), GLASS_HIP_EINVAL;
    *vcap = P->vcap;
    return GLASS_HIP_OK;
}

extern "C" int glass_hip_wdm_batch_destroy(glass_hip_wdm_batch* P)
{
    if (!P) return GLASS_HIP_OK;
    (void)hipSetDevice(P->device);
    for (void* p : {(void*)P->data, (void*)P->invC, (void*)P->scratch, (void*)P->list, (void*)P->values,
                    (void*)P->nwave, (void*)P->fparams, (void*)P->fD, (void*)P->fM, (void*)P->big})
        if (p) (void)hipFree(p);
    delete P;
    return GLASS_HIP_OK;
}

extern "C" int glass_hip_wdm_batch_buffers(glass_hip_wdm_batch* P, int** d_nwave, int** d_list,
                                           double** d_values)
{
    if (!P) return set_error("wdm_batch_buffers: NULL plan"), GLASS_HIP_EINVAL;
    if (d_nwave) *d_nwave = P->nwave;
    if (d_list) *d_list = P->list;
    if (d_values) *d_values = P->values;
    return GLASS_HIP_OK;
}

extern "C" int glass_hip_wdm_batch_eval(glass_hip_wdm_batch* P, int n, const double* d_params, double* d_dlogL,
                                        double* d_snr, void* stream)
{
    if (!P || !d_params || !d_dlogL || !d_snr) return set_error("wdm_batch_eval: NULL argument"), GLASS_HIP_EINVAL;
    if (n < 0 || n > P->max_templates)
        return set_error("wdm_batch_eval: n = %d outside 0..%d", n, P->max_templates), GLASS_HIP_EINVAL;
    GH_CHECK(hipSetDevice(P->device));
    LayerArgs a{};
    a.data = P->data;
    a.invC = P->invC;
    a.npix = P->npix;
    a.nwave = P->nwave;
    a.list = P->list;
    a.values = P->values;
    a.dlogL = d_dlogL;
    a.snr = d_snr;
    a.big = P->big;
    a.nbig = P->nbig;
    a.bigslice = P->bigslice;
    a.vcap = P->vcap;
    return launch_templates(*P->G, n, d_params, P->scratch, a, (hipStream_t)stream);
}

extern "C" int glass_hip_wdm_batch_templates(glass_hip_wdm_batch* P, int n, int* nwave, int* list, double* values)
{
    if (!P || !nwave || !list || !values) return set_error("wdm_batch_templates: NULL argument"), GLASS_HIP_EINVAL;
    if (n < 0 || n > P->max_templates)
        return set_error("wdm_batch_templates: n = %d outside 0..%d", n, P->max_templates), GLASS_HIP_EINVAL;
    GH_CHECK(hipSetDevice(P->device));
    GH_CHECK(hipDeviceSynchronize());
    const size_t vcap = P->vcap;
    GH_CHECK(hipMemcpy(nwave, P->nwave, sizeof(int) * n, hipMemcpyDeviceToHost));
    GH_CHECK(hipMemcpy(list, P->list, sizeof(int) * vcap * n, hipMemcpyDeviceToHost));
    GH_CHECK(hipMemcpy(values, P->values, sizeof(double) * 3 * vcap * n, hipMemcpyDeviceToHost));
    return GLASS_HIP_OK;
}

extern "C" int glass_hip_wdm_batch_scratch(glass_hip_wdm_batch* P, int e, double* out, int nout)
{
    if (!P || !out) return set_error("wdm_batch_scratch: NULL argument"), GLASS_HIP_EINVAL;
    const size_t ns = scratch_doubles(P->G->g.Norb);
    if (e < 0 || e >= P->max_templates || nout < (int)ns)
        return set_error("wdm_batch_scratch: template %d / buffer %d (need %zu)", e, nout, ns), GLASS_HIP_EINVAL;
    GH_CHECK(hipSetDevice(P->device));
    GH_CHECK(hipDeviceSynchronize());
    GH_CHECK(hipMemcpy(out, P->scratch + ns * (size_t)e, sizeof(double) * ns, hipMemcpyDeviceToHost));
    return GLASS_HIP_OK;
}

// ------------------------------------------------------ wavelet Fisher --

extern "C" int glass_hip_ucb_fisher_wavelet(const glass_hip_wdm* w, int Norb, const double* orbit_t,
                                            const double* orbit_x, const double* orbit_y, const double* orbit_z,
                                            double Tobs, double t0, const double* params, const double* invC, int N,
                                            double* fisher)
{
    if (!w || !orbit_t || !orbit_x || !orbit_y || !orbit_z || !params || !invC || !fisher)
        return set_error("ucb_fisher_wavelet: NULL argument"), GLASS_HIP_EINVAL;
    if (N != w->kmax - w->kmin)
        return set_error("ucb_fisher_wavelet: N = %d but kmax - kmin = %d", N, w->kmax - w->kmin), GLASS_HIP_EINVAL;
    ThreadCtx* c;
    int rc = ctx_get(&c);
    if (rc) return rc;
    const GeoHost* G;
    if ((rc = geo_get(w, Norb, orbit_t, orbit_x, orbit_y, orbit_z, Tobs, t0, &G))) return rc;
    const int NT = w->NT;
    // the 16 perturbed parameter vectors (glass_ucb_waveform.c:556-576) and
    // the layer band their templates cover (ucb_wavelet_layers, the rule the
    // template kernels apply)
    double p16[16 * 8];
    int jlo = 1 << 30, jhi = -(1 << 30);
    for (int q = 0; q < 16; q++)
    {
        const int i = q / 2, sgn = q % 2;
        const double epsilon = 1.0e-6;
        double* o = p16 + 8 * q;
        std::memcpy(o, params, sizeof(double) * 8);
        if (sgn == 0) o[i] += epsilon / 2.;
        else o[i] -= epsilon / 2.;
        if (i == 1 || i == 4)
        {
            if (sgn == 0 && o[i] > 1.0) o[i] = 1.0;
            if (sgn == 1 && o[i] < -1.0) o[i] = -1.0;
        }
        int js, jw;
        wavelet_layers(o[0], o[7], Tobs, w->df, w->A, &js, &jw);
        jlo = std::min(jlo, js);
        jhi = std::max(jhi, js + jw - 1);
    }
    const int nb = jhi - jlo + 1;
    if (nb < 1 || nb > kFishMaxBand)
        return set_error("ucb_fisher_wavelet: templates span %d layers (at most %d)", nb, kFishMaxBand),
               GLASS_HIP_EINVAL;
    const int klen = nb * NT, koff = jlo * NT - w->kmin;
    // device: params 128 | scratch | values | invC band 9 klen | D | M | fisher 64 | list | nwave
    const size_t nscr = scratch_doubles(Norb);
    const size_t nD = 24 * (size_t)kFishMaxBand * NT, nM = 3 * (size_t)kFishMaxBand * NT;
    const size_t nd = 128 + 16 * nscr + 16 * 9 * (size_t)NT + 9 * (size_t)klen + nD + nM + 64;
    const size_t bytes = sizeof(double) * nd + sizeof(int) * (16 * 3 * (size_t)NT + 16);
    const size_t hbytes = sizeof(double) * (128 + 9 * (size_t)klen + 64);
    if ((rc = c->dbuf.ensure(bytes)) || (rc = c->hbuf.ensure(hbytes))) return rc;
    double* dp = (double*)c->dbuf.p;
    double* dscr = dp + 128;
    double* dval = dscr + 16 * nscr;
    double* dinv = dval + 16 * 9 * (size_t)NT;
    double* dD = dinv + 9 * (size_t)klen;
    double* dM = dD + nD;
    double* dF = dM + nM;
    int* dlist = (int*)(dF + 64);
    int* dnw = dlist + 16 * 3 * NT;
    double* hp = (double*)c->hbuf.p;
    std::memcpy(hp, p16, sizeof(p16));
    double* hinv = hp + 128;
    for (int p = 0; p < 9; p++)
        for (int q = 0; q < klen; q++)
        {
            const long long k = (long long)koff + q;
            hinv[(size_t)p * klen + q] = (k >= 0 && k < N) ? invC[(size_t)p * N + k] : 0.0;
        }
    GH_CHECK(hipMemcpyAsync(dp, hp, sizeof(double) * 128, hipMemcpyHostToDevice, c->stream));
    GH_CHECK(hipMemcpyAsync(dinv, hinv, sizeof(double) * 9 * klen, hipMemcpyHostToDevice, c->stream));
    LayerArgs a{};
    a.nwave = dnw;
    a.list = dlist;
    a.values = dval;
    if ((rc = big_scratch(c, NT, 16, a))) return rc;
    if ((rc = launch_templates(*G, 16, dp, dscr, a, c->stream))) return rc;
    FishArgs fa{};
    fa.nwave = dnw;
    fa.list = dlist;
    fa.values = dval;
    fa.invC = dinv;
    fa.pstride = klen;
    fa.koff = koff;
    fa.klen = klen;
    for (int p = 0; p < 9; p++) fa.pmap[p] = p;
    fa.npix = N;
    fa.D = dD;
    fa.M = dM;
    fa.fisher = dF;
    hipLaunchKernelGGL(fisher_kernel, dim3(1), dim3(kFishThreads), 0, c->stream, G->g, 1, fa);
    GH_CHECK(hipGetLastError());
    double* hF = hinv + 9 * (size_t)klen;
    GH_CHECK(hipMemcpyAsync(hF, dF, sizeof(double) * 64, hipMemcpyDeviceToHost, c->stream));
    GH_CHECK(hipStreamSynchronize(c->stream));
    for (int q = 0; q < 64; q++)
        if (hF[q] != hF[q]) return set_error("ucb_fisher_wavelet: template pixels outside the invC band"),
                                   GLASS_HIP_ERUNTIME;
    std::memcpy(fisher, hF, sizeof(double) * 64);
    return GLASS_HIP_OK;
}

extern "C" int glass_hip_wdm_batch_fisher(glass_hip_wdm_batch* P, int n, const double* d_params, double* d_fisher,
                                          void* stream)
{
    if (!P || !d_params || !d_fisher) return set_error("wdm_batch_fisher: NULL argument"), GLASS_HIP_EINVAL;
    if (n < 0 || 16LL * n > P->max_templates)
        return set_error("wdm_batch_fisher: n = %d sources need %lld template slots of %d", n, 16LL * n,
                         P->max_templates),
               GLASS_HIP_EINVAL;
    if (P->layers > 0)
        return set_error("wdm_batch_fisher: the plan forces %d layers per template; ucb_fisher_wavelet takes the "
                         "reference's own layer count",
                         P->layers),
               GLASS_HIP_EINVAL;
    if (n == 0) return GLASS_HIP_OK;
    GH_CHECK(hipSetDevice(P->device));
    const int NT = P->G->g.NT;
    const size_t nsrc = (size_t)P->max_templates / 16;
    if (!P->fparams)
    {
        GH_CHECK(hipMalloc(&P->fparams, sizeof(double) * 8 * (size_t)P->max_templates));
        GH_CHECK(hipMalloc(&P->fD, sizeof(double) * 24 * (size_t)kFishMaxBand * NT * nsrc));
        GH_CHECK(hipMalloc(&P->fM, sizeof(double) * 3 * (size_t)kFishMaxBand * NT * nsrc));
    }
    hipStream_t s = (hipStream_t)stream;
    hipLaunchKernelGGL(fisher_params_kernel, dim3((16 * n + 255) / 256), dim3(256), 0, s, n, d_params, P->fparams);
    GH_CHECK(hipGetLastError());
    LayerArgs a{};
    a.nwave = P->nwave;
    a.list = P->list;
    a.values = P->values;
    a.big = P->big;
    a.nbig = P->nbig;
    a.bigslice = P->bigslice;
    a.vcap = P->vcap;
    int rc;
    if ((rc = launch_templates(*P->G, 16 * n, P->fparams, P->scratch, a, s))) return rc;
    FishArgs fa{};
    fa.nwave = P->nwave;
    fa.list = P->list;
    fa.values = P->values;
    fa.invC = P->invC;
    fa.pstride = P->npix;
    fa.koff = 0;
    fa.klen = P->npix;
    const int pm[9] = {0, 3, 4, 3, 1, 5, 4, 5, 2};  // XX YY ZZ XY XZ YZ planes, symmetric
    for (int p = 0; p < 9; p++) fa.pmap[p] = pm[p];
    fa.npix = P->npix;
    fa.D = P->fD;
    fa.M = P->fM;
    fa.fisher = d_fisher;
    hipLaunchKernelGGL(fisher_kernel, dim3(n), dim3(kFishThreads), 0, s, P->G->g, n, fa);
    GH_CHECK(hipGetLastError());
    return GLASS_HIP_OK;
}
