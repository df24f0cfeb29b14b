// ucb_sampler.hip -- the UCB sampler's per-step work on the device, over
// the chains of a batched plan (fd_batch.hip).
//
// One step of every slot is four launches: ucb_propose_kernel (one thread
// per slot: the reference's draws in its order, the proposal, the two prior
// evaluations), the plan's batched eval of the proposals whose prior is
// finite, ucb_accept_kernel (snr cut, Hastings ratio, the accept draw; the
// accepted slots are appended to a list), and the plan's in-place accept of
// that list, sized on the device.  The reference (per chain, sequential):
//   ucb_mcmc                      ucb/src/glass_ucb_sampler.c:75-173
//   step loop (+ the RJ coin)     apps/src/ucb_mcmc.c:239-258
//   draw_from_uniform_prior       utils/src/glass_sampler.c:96-107
//   draw_from_fisher              ucb/src/glass_ucb_proposal.c:348-391
//   fm_shift, psi_phi_jump        ucb/src/glass_ucb_proposal.c:577-613
//   check_range, evaluate_ucb_prior, evaluate_uniform_priors,
//   evalaute_sky_location_prior   ucb/src/glass_ucb_prior.c:400-646
//   ptmcmc, adapt_temperature_ladder  utils/src/glass_sampler.c:19-94
// The F-statistic proposals read a per-window proposal density (the tensor
// of build_fstatistic_proposal, e.g. from glass_hip_fstat_proposal) set with
// glass_hip_ucb_sampler_set_fstat; their densities enter the Hastings ratio
// as in the reference, the other proposals' cancel exactly.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <vector>

#include "../../include/glass_hip_io.h"
#include "../../include/glass_hip_sampler.h"
#include "fd_batch_hooks.h"
#include "rng.h"
#include "runtime.h"

namespace glass {

namespace {

constexpr double kPi = 3.14159265358979323846;
constexpr double kTwoPi = 6.283185307179586;        // PI2 of the reference
constexpr double kYear = 3.15581497632e7;           // YEAR of the reference
constexpr int kWrapMax = 64;                        // periods stepped one at a time
constexpr int kFsMaxTries = 1 << 24;                // rejection tries of one F-statistic draw

struct SamplerDev
{
    // plan
    const int* c_group;
    const int* c_nsrc;
    double* sparams;
    const double* logl;
    int maxsrc;
    double T;
    // slots and chains
    unsigned* seed;
    double* temp;
    int* index;
    const double* lognorm;
    int* trials;
    int* accepts;
    int* swaps;
    int* pacc;
    int* iter;
    int nslot, nladder, ntemp;
    // priors [G][8][2], log prior volumes [G][8]
    const double* prior;
    const double* logpv;
    // eigen data [C][maxsrc][64], [C][maxsrc][8]
    const double* evec;
    const double* eval;
    // one step's proposals
    int* it_chain;
    int* it_src;
    double* it_params;
    double* it_logl;
    double* it_snr;
    double* lpx;
    double* lpy;
    int* prop;
    int* acc_list;
    int* nacc;
    double* lq;  // [S] log q(x|y) - log q(y|x) of the step's proposal
    // F-statistic proposal densities per window (null: none):
    // tensor [n_f][n_theta][n_phi], dims (n_f, n_theta, n_phi),
    // par (d_f, maxp, data qmin)
    const double* const* fs_ten;
    const int* fs_dims;
    const double* fs_par;
    int* fs_overrun;  // draws that gave up after kFsMaxTries rejections
    // options
    double w[9];
    int fix_sky, fix_freq, fix_fdot;
    double snr_min;
};

// while-loop periodic wrap of the reference, with the count of single
// steps bounded (a value kWrapMax periods or more away is reduced by fmod)
__device__ inline double wrap_up_down(double x, double lo, double hi, double period)
{
    int k = 0;
    while (x < lo && k < kWrapMax)
    {
        x += period;
        k++;
    }
    while (x > hi && k < kWrapMax)
    {
        x -= period;
        k++;
    }
    if (k == kWrapMax && (x < lo || x > hi))
    {
        x = lo + fmod(x - lo, period);
        if (x < lo) x += period;
    }
    return x;
}

// check_range (ucb/src/glass_ucb_prior.c:400-439); wraps in place
__device__ inline int check_range(double* p, const double* pr)
{
    for (int n = 0; n < 8; n++)
        if (p[n] != p[n]) return 1;
    if (p[0] < pr[0] || p[0] > pr[1]) return 1;
    if (p[1] < pr[2] || p[1] > pr[3]) return 1;
    if (p[2] < pr[4] || p[2] >= pr[5])
    {
        p[2] = atan2(sin(p[2]), cos(p[2]));
        if (p[2] < 0.0) p[2] += kTwoPi;
    }
    if (p[3] < pr[6] || p[3] > pr[7]) return 1;
    if (p[4] < pr[8] || p[4] > pr[9]) return 1;
    p[5] = wrap_up_down(p[5], pr[10], pr[11], kPi);
    p[6] = wrap_up_down(p[6], pr[12], pr[13], kTwoPi);
    if (p[7] < pr[14] || p[7] > pr[15]) return 1;
    return 0;
}

// evalaute_sky_location_prior with galaxyFlag 0 (ucb/src/glass_ucb_prior.c:
// 590-645); wraps the longitude in place
__device__ inline double sky_prior(double* p, const double* pr, const double* lpv)
{
    double logP = 0.0;
    if (p[1] < pr[2] || p[1] > pr[3])
        return -INFINITY;
    else
        logP -= lpv[1];
    if (pr[4] > 0.0 || pr[5] < kTwoPi)
    {
        if (p[2] < pr[4] || p[2] > pr[5]) return -INFINITY;
    }
    else if (p[2] < pr[4] || p[2] >= pr[5])
    {
        p[2] = atan2(sin(p[2]), cos(p[2]));
        if (p[2] < 0.0) p[2] += kTwoPi;
    }
    logP -= lpv[2];
    return logP;
}

// evaluate_uniform_priors (:560-588); lpv[0] = log(prior[0][1] - prior[0][0])
__device__ inline double uniform_priors(double* p, const double* pr, const double* lpv)
{
    if (check_range(p, pr)) return -INFINITY;
    double logP = 0.0;
    logP -= lpv[0];
    logP -= lpv[3];
    logP -= lpv[4];
    logP -= lpv[5];
    logP -= lpv[6];
    logP -= lpv[7];
    return logP;
}

// evaluate_ucb_prior with the galaxy and update flags off (:536-558)
__device__ inline double ucb_prior(double* p, const double* pr, const double* lpv)
{
    if (check_range(p, pr)) return -INFINITY;
    double logP = 0.0;
    logP += sky_prior(p, pr, lpv);
    logP += uniform_priors(p, pr, lpv);
    return logP;
}

struct FsProp
{
    const double* ten;
    int nf, nt, nph;
    double d_f, maxp, qmin, d_theta, d_phi;
};

__device__ inline FsProp fs_prop(const SamplerDev& S, int g)
{
    FsProp f;
    f.ten = S.fs_ten[g];
    f.nf = S.fs_dims[3 * g];
    f.nt = S.fs_dims[3 * g + 1];
    f.nph = S.fs_dims[3 * g + 2];
    f.d_f = S.fs_par[3 * g];
    f.maxp = S.fs_par[3 * g + 1];
    f.qmin = S.fs_par[3 * g + 2];
    f.d_theta = 2. / (double)f.nt;  // setup_fstatistic_proposal, ucb/src/glass_ucb_proposal.c:939-940
    f.d_phi = kTwoPi / (double)f.nph;
    return f;
}

// evaluate_fstatistic_proposal (ucb/src/glass_ucb_proposal.c:1640-1673);
// wraps p in place through the priors it calls
__device__ inline double fstat_density(double* p, const double* pr, const double* lpv, const FsProp& f, double T)
{
    double logP = 0.0;
    logP += sky_prior(p, pr, lpv);
    logP += uniform_priors(p, pr, lpv);
    const int i = (int)floor((p[0] - f.qmin) / (f.d_f * T));
    const int j = (int)floor((p[1] - -1) / f.d_theta);
    const int k = (int)floor((p[2]) / f.d_phi);
    if (i < 0 || i >= f.nf) return -INFINITY;
    if (j < 0 || j >= f.nt) return -INFINITY;
    if (k < 0 || k >= f.nph) return -INFINITY;
    logP += log(f.ten[((size_t)i * f.nt + j) * f.nph + k]);
    return logP;
}

// the rejection sampling of draw_from_fstatistic / jump_from_fstatistic
// (ucb/src/glass_ucb_proposal.c:1526-1638); jump: the reference's range
// check (which a draw of U = 1 fails) ends the proposal with y[0..2] as
// they were.  The reference loops until a cell is accepted; here a draw
// gives up after kFsMaxTries tries (about maxp x cells / sum of the density
// tries are expected, and set_fstat refuses tables that make that
// unbounded), counts the overrun and leaves y[0..2] unchanged.
__device__ inline void fstat_sample(unsigned* seed, double* y, const FsProp& f, double T, bool jump,
                                    int* overrun)
{
    double i, j, k, q, costheta, phi, p, alpha;
    int tries = 0;
    do
    {
        if (++tries > kFsMaxTries)
        {
            atomicAdd(overrun, 1);
            return;
        }
        i = rand_U(seed) * f.nf;
        j = rand_U(seed) * f.nt;
        k = rand_U(seed) * f.nph;
        q = f.qmin + i * f.d_f * T;
        costheta = -1. + j * f.d_theta;
        phi = k * f.d_phi;
        int ii = (int)i, jj = (int)j, kk = (int)k;
        if (jump && (ii < 0 || ii > f.nf - 1 || jj < 0 || jj > f.nt - 1 || kk < 0 || kk > f.nph - 1)) return;
        // draw_from_fstatistic has no such check; U = 1 would index one past
        ii = ii > f.nf - 1 ? f.nf - 1 : ii;
        jj = jj > f.nt - 1 ? f.nt - 1 : jj;
        kk = kk > f.nph - 1 ? f.nph - 1 : kk;
        p = f.ten[((size_t)ii * f.nt + jj) * f.nph + kk];
        alpha = rand_U(seed) * f.maxp;
    } while (p < alpha);
    y[0] = q;
    y[1] = costheta;
    y[2] = phi;
}

__global__ void ucb_propose_kernel(SamplerDev S)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s == 0) *S.nacc = 0;  // the accept list of this step starts empty
    if (s >= S.nslot) return;
    unsigned seed = S.seed[s];
    const int st = S.index[s];
    const int g = S.c_group[st];
    const int nlive = S.c_nsrc[st];
    const double* pr = S.prior + (size_t)g * 16;
    const double* lpv = S.logpv + (size_t)g * 8;
    (void)rand_U(&seed);  // the step loop's RJ coin (flags->rj off)
    int n = (int)(rand_U(&seed) * (double)nlive);
    if (n >= nlive) n = nlive - 1;  // U = 1 exactly
    int np;
    double draw;
    do
    {
        np = (int)floor(9.0 * rand_U(&seed));
        if (np > 8) np = 8;  // U = 1 exactly; the reference reads past its table
        draw = rand_U(&seed);
    } while (S.w[np] <= draw);
    S.trials[s * 9 + np] += 1;
    double* xs = S.sparams + ((size_t)st * S.maxsrc + n) * 8;
    double x[8], y[8];
    for (int k = 0; k < 8; k++) x[k] = y[k] = xs[k];
    if (np == 0)
    {
        for (int k = 0; k < 8; k++) y[k] = pr[2 * k] + rand_U(&seed) * (pr[2 * k + 1] - pr[2 * k]);
    }
    else if (np == 1 || np == 2)
    {
        const FsProp f = fs_prop(S, g);
        if (np == 1)
            for (int k = 0; k < 8; k++) y[k] = pr[2 * k] + rand_U(&seed) * (pr[2 * k + 1] - pr[2 * k]);
        else
            (void)(rand_U(&seed) < -0.5);  // the fm-shift branch of jump_from_fstatistic, never taken
        fstat_sample(&seed, y, f, S.T, np == 2, S.fs_overrun);
    }
    else if (np == 4)
    {
        const double* ev = S.evec + ((size_t)st * S.maxsrc + n) * 64;
        const double* la = S.eval + ((size_t)st * S.maxsrc + n) * 8;
        double amp[8], jump[8];
        for (int i = 0; i < 8; i++)
        {
            amp[i] = rand_N(&seed) / sqrt(la[i]);
            jump[i] = 0.0;
        }
        int i = (int)(rand_U(&seed) * 8.0);
        if (i > 7) i = 7;
        for (int j = 0; j < 8; j++) jump[j] += amp[i] * ev[j * 8 + i];
        for (int k = 0; k < 8; k++)
            if (jump[k] != jump[k]) jump[k] = 0.01 * x[k];
        for (int k = 0; k < 8; k++) y[k] = x[k] + jump[k];
        if (y[1] >= 1.) y[1] = x[1] - jump[1];
        if (y[1] <= -1.) y[1] = x[1] - jump[1];
        if (y[4] >= 1.) y[4] = x[4] - jump[4];
        if (y[4] <= -1.) y[4] = x[4] - jump[4];
    }
    else if (np == 5)
    {
        const double fm = S.T / kYear;
        int scale;
        do
            scale = (int)floor(-3. + 7. * rand_U(&seed));
        while (scale == 0);
        y[0] += scale * fm;
        y[2] = y[2] + -0.2 + rand_U(&seed) * 0.4;
    }
    else if (np == 6)
    {
        const double scale = rand_U(&seed) * kTwoPi;
        double d_phi = scale, d_psi = 0.5 * scale;
        if (rand_U(&seed) < 0.5) d_phi *= -1.;
        if (rand_U(&seed) < 0.5) d_psi *= -1.;
        y[5] += d_psi;
        y[6] += d_phi;
    }
    if (S.fix_sky)
    {
        y[1] = x[1];
        y[2] = x[2];
    }
    if (S.fix_freq) y[0] = x[0];
    if (S.fix_fdot) y[7] = x[7];
    // proposal densities, y first then x (:125-126); only the F-statistic
    // and uniform-prior proposals have one, and the latter cancels exactly
    double lqyx = 0.0, lqxy = 0.0;
    if (np == 0)
    {
        for (int k = 0; k < 8; k++) lqyx -= lpv[k];
        for (int k = 0; k < 8; k++) lqxy -= lpv[k];
    }
    else if (np == 1 || np == 2)
    {
        const FsProp f = fs_prop(S, g);
        lqyx = fstat_density(y, pr, lpv, f, S.T);
        lqxy = fstat_density(x, pr, lpv, f, S.T);
    }
    const double lpx = ucb_prior(x, pr, lpv);
    const double lpy = ucb_prior(y, pr, lpv);
    for (int k = 0; k < 8; k++) xs[k] = x[k];  // the prior's in-place wraps of the current state
    S.it_chain[s] = st;
    S.it_src[s] = lpy > -INFINITY ? n : -1;
    for (int k = 0; k < 8; k++) S.it_params[(size_t)s * 8 + k] = y[k];
    S.lpx[s] = lpx;
    S.lpy[s] = lpy;
    S.lq[s] = lqxy - lqyx;
    S.prop[s] = np;
    S.seed[s] = seed;
}

__global__ void ucb_accept_kernel(SamplerDev S)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= S.nslot || S.it_src[s] < 0) return;  // prior -inf: no likelihood and no accept draw
    unsigned seed = S.seed[s];
    const int st = S.it_chain[s];
    double lpy = S.lpy[s];
    if (S.it_snr[s] < S.snr_min) lpy = -INFINITY;
    double logH = 0.0;
    logH += (S.it_logl[s] - S.logl[st]) / S.temp[s];
    logH += lpy - S.lpx[s];
    logH += S.lq[s];  // log q(x|y) - log q(y|x)
    const double loga = log(rand_U(&seed));
    if (isfinite(logH) && logH > loga)
    {
        S.accepts[s * 9 + S.prop[s]] += 1;
        S.acc_list[atomicAdd(S.nacc, 1)] = s;
    }
    S.seed[s] = seed;
}

// ptmcmc (+ adapt_temperature_ladder) of every ladder, one thread each
__global__ void ucb_ptswap_kernel(SamplerDev S, int adapt)
{
    const int l = blockIdx.x * blockDim.x + threadIdx.x;
    if (l >= S.nladder) return;
    const int o = l * S.ntemp, NC = S.ntemp;
    for (int b = NC - 1; b > 0; b--)
    {
        const int a = b - 1;
        S.pacc[o + a] = 0;
        const int olda = S.index[o + a], oldb = S.index[o + b];
        const double heat1 = S.temp[o + a], heat2 = S.temp[o + b];
        const double logL1 = S.logl[olda] + S.lognorm[olda];
        const double logL2 = S.logl[oldb] + S.lognorm[oldb];
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
                S.pacc[o + a] = 1;
                S.swaps[o + a] += 1;
            }
        }
        S.seed[o + a] = seed;
    }
    if (!adapt || NC < 3) return;
    const double nu = 10, t0 = 10000.;
    const int mcmc = *S.iter;
    double* Tl = S.temp + o;
    double Sx[64];
    for (int ic = 1; ic < NC - 1 && ic < 64; ic++) Sx[ic] = log(Tl[ic] - Tl[ic - 1]);
    for (int ic = 1; ic < NC - 1 && ic < 64; ic++)
    {
        const double A0 = S.pacc[o + ic - 1], A1 = S.pacc[o + ic];
        Sx[ic] += (A0 - A1) * (t0 / ((double)mcmc + t0)) / nu;
        Tl[ic] = Tl[ic - 1] + exp(Sx[ic]);
        if (Tl[ic] / Tl[ic - 1] < 1.1) Tl[ic] = Tl[ic - 1] * 1.1;
    }
}

__global__ void iter_kernel(int* iter)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) *iter += 1;
}

// eigen data of live source k -> its (chain, source) slot
__global__ void scatter_eigen_kernel(int n, const int* __restrict__ ch, const int* __restrict__ sr, int maxsrc,
                                     const double* __restrict__ ev, const double* __restrict__ la,
                                     double* __restrict__ evec, double* __restrict__ eval)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n * 72) return;
    const int k = i / 72, e = i % 72;
    const size_t slot = (size_t)ch[k] * maxsrc + sr[k];
    if (e < 64)
        evec[slot * 64 + e] = ev[(size_t)k * 64 + e];
    else
        eval[slot * 8 + e - 64] = la[(size_t)k * 8 + e - 64];
}

}  // namespace

}  // namespace glass

using namespace glass;

struct glass_hip_ucb_sampler
{
    glass_hip_fd_batch* plan = nullptr;
    FdBatchView v{};
    int nladder = 0, ntemp = 0, nslot = 0;
    std::vector<void*> allocs;
    SamplerDev D{};
    // F-statistic proposal densities per window
    std::vector<double*> fs_buf;
    const double** d_fs_ten = nullptr;
    int* d_fs_dims = nullptr;
    double* d_fs_par = nullptr;
    // Fisher update scratch over the live sources
    double *fL = nullptr, *fgrad = nullptr, *fF = nullptr, *fvec = nullptr, *fval = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    int g_steps = -1, g_adapt = -1;
    // each step's proposal templates, kept by the eval for the accept
    double2* stage = nullptr;
};

template <typename T>
static int salloc(glass_hip_ucb_sampler* S, T** p, size_t n)
{
    void* q = nullptr;
    if (hipMalloc(&q, sizeof(T) * (n ? n : 1)) != hipSuccess)
    {
        set_error("ucb_sampler: hipMalloc of %zu bytes failed", sizeof(T) * (n ? n : 1));
        return GLASS_HIP_ERUNTIME;
    }
    S->allocs.push_back(q);
    *p = (T*)q;
    return GLASS_HIP_OK;
}

static int drop_graph(glass_hip_ucb_sampler* S)
{
    if (S->gexec) GH_CHECK(hipGraphExecDestroy(S->gexec));
    if (S->graph) GH_CHECK(hipGraphDestroy(S->graph));
    S->gexec = nullptr;
    S->graph = nullptr;
    S->g_steps = S->g_adapt = -1;
    return GLASS_HIP_OK;
}

extern "C" int glass_hip_ucb_sampler_destroy(glass_hip_ucb_sampler* S)
{
    if (!S) return GLASS_HIP_OK;
    (void)hipSetDevice(S->v.device);
    if (S->gexec) (void)hipGraphExecDestroy(S->gexec);
    if (S->graph) (void)hipGraphDestroy(S->graph);
    (void)hipDeviceSynchronize();
    for (void* p : S->allocs) (void)hipFree(p);
    for (double* p : S->fs_buf)
        if (p) (void)hipFree(p);
    delete S;
    return GLASS_HIP_OK;
}

extern "C" int glass_hip_ucb_sampler_create(glass_hip_ucb_sampler** out, glass_hip_fd_batch* plan, int nladder,
                                            int ntemp, const double* prior, const glass_hip_ucb_sampler_opts* opts,
                                            void* stream)
{
    if (!out || !plan || !prior || !opts || nladder <= 0 || ntemp <= 0 || ntemp > 64)
    {
        set_error("ucb_sampler_create: bad arguments (ladders %d, temperatures %d; at most 64 temperatures)",
                  nladder, ntemp);
        return GLASS_HIP_EINVAL;
    }
    *out = nullptr;
    FdBatchView v = fd_batch_view(plan);
    if ((long long)nladder * ntemp != v.C)
    {
        set_error("ucb_sampler_create: %d ladders x %d temperatures != %d chains of the plan", nladder, ntemp, v.C);
        return GLASS_HIP_EINVAL;
    }
    double wsum = 0.0;
    for (int i = 0; i < 9; i++)
    {
        const bool built = i == 0 || i == 1 || i == 2 || i == 4 || i == 5 || i == 6;
        if (!(opts->weight[i] >= 0.0) || (!built && opts->weight[i] != 0.0))
        {
            set_error("ucb_sampler_create: proposal %d has weight %g (only 0 uniform prior, 1/2 F-statistic, "
                      "4 Fisher, 5 fm shift, 6 psi-phi jump are built)", i, opts->weight[i]);
            return GLASS_HIP_EINVAL;
        }
        wsum += opts->weight[i];
    }
    if (!(wsum > 0.0))
    {
        set_error("ucb_sampler_create: every proposal weight is zero");
        return GLASS_HIP_EINVAL;
    }
    GH_CHECK(hipSetDevice(v.device));
    hipStream_t st = (hipStream_t)stream;
    std::vector<int> nsrc(v.C);
    GH_CHECK(hipMemcpyAsync(nsrc.data(), v.c_nsrc, sizeof(int) * v.C, hipMemcpyDeviceToHost, st));
    GH_CHECK(hipStreamSynchronize(st));
    for (int c = 0; c < v.C; c++)
        if (nsrc[c] < 1)
        {
            set_error("ucb_sampler_create: chain %d has no live source", c);
            return GLASS_HIP_EINVAL;
        }
    std::vector<double> lpv((size_t)v.G * 8);
    for (int g = 0; g < v.G; g++)
        for (int k = 0; k < 8; k++)
        {
            const double lo = prior[(size_t)g * 16 + 2 * k], hi = prior[(size_t)g * 16 + 2 * k + 1];
            if (!(hi > lo))
            {
                set_error("ucb_sampler_create: window %d parameter %d prior [%g, %g] is empty", g, k, lo, hi);
                return GLASS_HIP_EINVAL;
            }
            lpv[(size_t)g * 8 + k] = log(hi - lo);  // model->logPriorVolume, glass_ucb_prior.c:396
        }
    auto* S = new glass_hip_ucb_sampler();
    S->plan = plan;
    S->v = v;
    S->nladder = nladder;
    S->ntemp = ntemp;
    S->nslot = v.C;
    SamplerDev& D = S->D;
    const int N = v.C, M = v.ninit;
    int rc = 0;
    double *d_prior = nullptr, *d_lpv = nullptr, *d_evec = nullptr, *d_eval = nullptr, *d_norm = nullptr;
    if ((rc = salloc(S, &D.seed, N)) || (rc = salloc(S, &D.temp, N)) || (rc = salloc(S, &D.index, N)) ||
        (rc = salloc(S, &d_norm, N)) || (rc = salloc(S, &D.trials, 9 * (size_t)N)) ||
        (rc = salloc(S, &D.accepts, 9 * (size_t)N)) || (rc = salloc(S, &D.swaps, N)) ||
        (rc = salloc(S, &D.pacc, N)) || (rc = salloc(S, &D.iter, 1)) || (rc = salloc(S, &d_prior, 16 * (size_t)v.G)) ||
        (rc = salloc(S, &d_lpv, 8 * (size_t)v.G)) || (rc = salloc(S, &d_evec, 64 * (size_t)N * v.maxsrc)) ||
        (rc = salloc(S, &d_eval, 8 * (size_t)N * v.maxsrc)) || (rc = salloc(S, &D.it_chain, N)) ||
        (rc = salloc(S, &D.it_src, N)) || (rc = salloc(S, &D.it_params, 8 * (size_t)N)) ||
        (rc = salloc(S, &D.it_logl, N)) || (rc = salloc(S, &D.it_snr, N)) || (rc = salloc(S, &D.lpx, N)) ||
        (rc = salloc(S, &D.lpy, N)) || (rc = salloc(S, &D.prop, N)) || (rc = salloc(S, &D.acc_list, N)) ||
        (rc = salloc(S, &D.nacc, 1)) || (rc = salloc(S, &D.lq, N)) || (rc = salloc(S, &D.fs_overrun, 1)) || (rc = salloc(S, &S->d_fs_ten, v.G)) ||
        (rc = salloc(S, &S->d_fs_dims, 3 * (size_t)v.G)) || (rc = salloc(S, &S->d_fs_par, 3 * (size_t)v.G)) ||
        (rc = salloc(S, &S->fL, M)) || (rc = salloc(S, &S->fgrad, 8 * (size_t)M)) ||
        (rc = salloc(S, &S->fF, 64 * (size_t)M)) || (rc = salloc(S, &S->fvec, 64 * (size_t)M)) ||
        (rc = salloc(S, &S->fval, 8 * (size_t)M)) ||
        (rc = salloc(S, &S->stage, fd_batch_stage_bytes(plan, N) / sizeof(double2))))
    {
        glass_hip_ucb_sampler_destroy(S);
        return rc;
    }
    D.c_group = v.c_group;
    D.c_nsrc = v.c_nsrc;
    D.sparams = v.sparams;
    D.logl = v.logl;
    D.maxsrc = v.maxsrc;
    D.T = v.T;
    D.lognorm = d_norm;
    D.nslot = N;
    D.nladder = nladder;
    D.ntemp = ntemp;
    D.prior = d_prior;
    D.logpv = d_lpv;
    D.evec = d_evec;
    D.eval = d_eval;
    for (int i = 0; i < 9; i++) D.w[i] = opts->weight[i];
    D.fs_ten = S->d_fs_ten;
    D.fs_dims = S->d_fs_dims;
    D.fs_par = S->d_fs_par;
    S->fs_buf.assign(v.G, nullptr);
    D.fix_sky = opts->fix_sky;
    D.fix_freq = opts->fix_freq;
    D.fix_fdot = opts->fix_fdot;
    D.snr_min = opts->snr_min;
    // identity eigen data (unit eigenvalues) until the first Fisher update
    std::vector<double> ev(64 * (size_t)N * v.maxsrc, 0.0), la(8 * (size_t)N * v.maxsrc, 1.0);
    for (size_t m = 0; m < (size_t)N * v.maxsrc; m++)
        for (int i = 0; i < 8; i++) ev[m * 64 + i * 8 + i] = 1.0;
    std::vector<double> zero(N, 0.0);
    std::vector<const double*> nullt(v.G, nullptr);
    std::vector<int> zdims(3 * (size_t)v.G, 0);
    std::vector<double> zpar(3 * (size_t)v.G, 0.0);
    if (hipMemcpyAsync(S->d_fs_ten, nullt.data(), sizeof(double*) * v.G, hipMemcpyHostToDevice, st) ||
        hipMemcpyAsync(S->d_fs_dims, zdims.data(), sizeof(int) * 3 * v.G, hipMemcpyHostToDevice, st) ||
        hipMemcpyAsync(S->d_fs_par, zpar.data(), sizeof(double) * 3 * v.G, hipMemcpyHostToDevice, st) ||
        hipMemcpyAsync(d_prior, prior, sizeof(double) * 16 * v.G, hipMemcpyHostToDevice, st) ||
        hipMemcpyAsync(d_lpv, lpv.data(), sizeof(double) * 8 * v.G, hipMemcpyHostToDevice, st) ||
        hipMemcpyAsync(d_evec, ev.data(), sizeof(double) * ev.size(), hipMemcpyHostToDevice, st) ||
        hipMemcpyAsync(d_eval, la.data(), sizeof(double) * la.size(), hipMemcpyHostToDevice, st) ||
        hipMemcpyAsync(d_norm, zero.data(), sizeof(double) * N, hipMemcpyHostToDevice, st) ||
        hipStreamSynchronize(st))
    {
        set_error("ucb_sampler_create: upload failed");
        glass_hip_ucb_sampler_destroy(S);
        return GLASS_HIP_ERUNTIME;
    }
    *out = S;
    return GLASS_HIP_OK;
}

extern "C" int glass_hip_ucb_sampler_set_state(glass_hip_ucb_sampler* S, const unsigned* seeds,
                                               const double* temperature, const double* logLnorm, int adapt_iter0,
                                               void* stream)
{
    if (!S || !seeds || !temperature)
    {
        set_error("ucb_sampler_set_state: bad arguments");
        return GLASS_HIP_EINVAL;
    }
    GH_CHECK(hipSetDevice(S->v.device));
    hipStream_t st = (hipStream_t)stream;
    const int N = S->nslot;
    std::vector<int> idx(N), zero(9 * (size_t)N, 0);
    for (int i = 0; i < N; i++) idx[i] = i;
    std::vector<double> norm(N, 0.0);
    if (logLnorm) memcpy(norm.data(), logLnorm, sizeof(double) * N);
    SamplerDev& D = S->D;
    GH_CHECK(hipMemcpyAsync(D.seed, seeds, sizeof(unsigned) * N, hipMemcpyHostToDevice, st));
    GH_CHECK(hipMemcpyAsync(D.temp, temperature, sizeof(double) * N, hipMemcpyHostToDevice, st));
    GH_CHECK(hipMemcpyAsync((void*)D.lognorm, norm.data(), sizeof(double) * N, hipMemcpyHostToDevice, st));
    GH_CHECK(hipMemcpyAsync(D.index, idx.data(), sizeof(int) * N, hipMemcpyHostToDevice, st));
    GH_CHECK(hipMemcpyAsync(D.trials, zero.data(), sizeof(int) * 9 * N, hipMemcpyHostToDevice, st));
    GH_CHECK(hipMemcpyAsync(D.accepts, zero.data(), sizeof(int) * 9 * N, hipMemcpyHostToDevice, st));
    GH_CHECK(hipMemcpyAsync(D.swaps, zero.data(), sizeof(int) * N, hipMemcpyHostToDevice, st));
    GH_CHECK(hipMemcpyAsync(D.pacc, zero.data(), sizeof(int) * N, hipMemcpyHostToDevice, st));
    GH_CHECK(hipMemcpyAsync(D.iter, &adapt_iter0, sizeof(int), hipMemcpyHostToDevice, st));
    GH_CHECK(hipMemcpyAsync(D.fs_overrun, zero.data(), sizeof(int), hipMemcpyHostToDevice, st));
    GH_CHECK(hipStreamSynchronize(st));
    return GLASS_HIP_OK;
}

// A proposal density the device's rejection sampler can draw from: finite,
// non-negative cells, a positive maximum, and maxp no larger than it (the
// reference sets maxp to the table's maximum, ucb/src/glass_ucb_proposal.c:1172-1177;
// a larger maxp only adds rejections, an all-zero table never accepts).
static int fstat_table_check(const double* t, size_t n, double maxp)
{
    double mx = 0.0;
    for (size_t i = 0; i < n; i++)
    {
        if (!(t[i] >= 0.0) || !std::isfinite(t[i]))
        {
            set_error("ucb_sampler_set_fstat: density cell %zu is %g (must be finite and >= 0)", i, t[i]);
            return GLASS_HIP_EINVAL;
        }
        mx = t[i] > mx ? t[i] : mx;
    }
    if (!(mx > 0.0) || !(maxp <= mx * (1.0 + 1e-12)))
    {
        set_error("ucb_sampler_set_fstat: maxp %g must be > 0 and at most the density's maximum %g", maxp, mx);
        return GLASS_HIP_EINVAL;
    }
    return GLASS_HIP_OK;
}

extern "C" int glass_hip_ucb_sampler_set_fstat(glass_hip_ucb_sampler* S, int g, const double* tensor, int n_f,
                                               double d_f, int n_theta, int n_phi, double maxp, int data_qmin,
                                               void* stream)
{
    if (!tensor || n_f <= 0 || n_theta <= 0 || n_phi <= 0 || !(d_f > 0) || !(maxp > 0))
    {
        set_error("ucb_sampler_set_fstat: bad arguments (window %d)", g);
        return GLASS_HIP_EINVAL;
    }
    // the table is checked first, on the host
    const size_t n = (size_t)n_f * n_theta * n_phi;
    if (int rc = fstat_table_check(tensor, n, maxp)) return rc;
    if (!S || g < 0 || g >= S->v.G)
    {
        set_error("ucb_sampler_set_fstat: bad arguments (window %d)", g);
        return GLASS_HIP_EINVAL;
    }
    GH_CHECK(hipSetDevice(S->v.device));
    hipStream_t st = (hipStream_t)stream;
    if (S->graph && drop_graph(S)) return GLASS_HIP_ERUNTIME;
    GH_CHECK(hipStreamSynchronize(st));
    if (S->fs_buf[g]) GH_CHECK(hipFree(S->fs_buf[g]));
    S->fs_buf[g] = nullptr;
    GH_CHECK(hipMalloc(&S->fs_buf[g], sizeof(double) * n));
    const int dims[3] = {n_f, n_theta, n_phi};
    const double par[3] = {d_f, maxp, (double)data_qmin};
    const double* ptr = S->fs_buf[g];
    GH_CHECK(hipMemcpyAsync(S->fs_buf[g], tensor, sizeof(double) * n, hipMemcpyHostToDevice, st));
    GH_CHECK(hipMemcpyAsync(S->d_fs_ten + g, &ptr, sizeof(ptr), hipMemcpyHostToDevice, st));
    GH_CHECK(hipMemcpyAsync(S->d_fs_dims + 3 * g, dims, sizeof(dims), hipMemcpyHostToDevice, st));
    GH_CHECK(hipMemcpyAsync(S->d_fs_par + 3 * g, par, sizeof(par), hipMemcpyHostToDevice, st));
    GH_CHECK(hipStreamSynchronize(st));
    return GLASS_HIP_OK;
}

static int check_fstat(glass_hip_ucb_sampler* S)
{
    if (S->D.w[1] == 0.0 && S->D.w[2] == 0.0) return GLASS_HIP_OK;
    for (int g = 0; g < S->v.G; g++)
        if (!S->fs_buf[g])
        {
            set_error("ucb_sampler: F-statistic proposals weighted but window %d has no proposal density "
                      "(glass_hip_ucb_sampler_set_fstat)", g);
            return GLASS_HIP_EINVAL;
        }
    return GLASS_HIP_OK;
}

static int enqueue_fisher(glass_hip_ucb_sampler* S, hipStream_t st)
{
    const FdBatchView& v = S->v;
    if (int rc = glass_hip_fd_batch_derivs(S->plan, v.ninit, v.i_chain, v.i_src, S->fL, S->fgrad, S->fF, st))
        return rc;
    if (int rc = fisher_eigen_launch(v.ninit, S->fF, S->fvec, S->fval, nullptr, st)) return rc;
    const int tot = v.ninit * 72;
    scatter_eigen_kernel<<<(tot + 255) / 256, 256, 0, st>>>(v.ninit, v.i_chain, v.i_src, v.maxsrc, S->fvec, S->fval,
                                                            (double*)S->D.evec, (double*)S->D.eval);
    GH_CHECK(hipGetLastError());
    return GLASS_HIP_OK;
}

static int enqueue_steps(glass_hip_ucb_sampler* S, int nsteps, hipStream_t st)
{
    if (int rc = check_fstat(S)) return rc;
    // one thread per slot; one-wave workgroups spread the slots' long
    // serial proposal over as many CUs as there are waves
    const int N = S->nslot, g = (N + 63) / 64;
    SamplerDev& D = S->D;
    for (int k = 0; k < nsteps; k++)
    {
        ucb_propose_kernel<<<g, 64, 0, st>>>(D);
        GH_CHECK(hipGetLastError());
        if (int rc = fd_batch_eval_dev(S->plan, N, D.it_chain, D.it_src, D.it_params, D.it_logl, D.it_snr, st,
                                       S->stage))
            return rc;
        ucb_accept_kernel<<<g, 64, 0, st>>>(D);
        GH_CHECK(hipGetLastError());
        if (int rc = fd_batch_accept_staged_dev(S->plan, N, D.nacc, D.acc_list, D.it_chain, D.it_src, D.it_params,
                                                D.it_logl, S->stage, st))
            return rc;
    }
    return GLASS_HIP_OK;
}

static int enqueue_swap(glass_hip_ucb_sampler* S, int adapt, hipStream_t st)
{
    ucb_ptswap_kernel<<<(S->nladder + 63) / 64, 64, 0, st>>>(S->D, adapt ? 1 : 0);
    if (adapt) iter_kernel<<<1, 64, 0, st>>>(S->D.iter);
    GH_CHECK(hipGetLastError());
    return GLASS_HIP_OK;
}

extern "C" int glass_hip_ucb_sampler_fisher(glass_hip_ucb_sampler* S, void* stream)
{
    if (!S)
    {
        set_error("ucb_sampler_fisher: null sampler");
        return GLASS_HIP_EINVAL;
    }
    GH_CHECK(hipSetDevice(S->v.device));
    return enqueue_fisher(S, (hipStream_t)stream);
}

extern "C" int glass_hip_ucb_sampler_set_eigen(glass_hip_ucb_sampler* S, const double* evec, const double* eval,
                                               void* stream)
{
    if (!S || !evec || !eval)
    {
        set_error("ucb_sampler_set_eigen: bad arguments");
        return GLASS_HIP_EINVAL;
    }
    GH_CHECK(hipSetDevice(S->v.device));
    const size_t m = (size_t)S->nslot * S->v.maxsrc;
    hipStream_t st = (hipStream_t)stream;
    GH_CHECK(hipMemcpyAsync((void*)S->D.evec, evec, sizeof(double) * 64 * m, hipMemcpyHostToDevice, st));
    GH_CHECK(hipMemcpyAsync((void*)S->D.eval, eval, sizeof(double) * 8 * m, hipMemcpyHostToDevice, st));
    GH_CHECK(hipStreamSynchronize(st));
    return GLASS_HIP_OK;
}

extern "C" int glass_hip_ucb_sampler_get_eigen(glass_hip_ucb_sampler* S, double* evec, double* eval, void* stream)
{
    if (!S)
    {
        set_error("ucb_sampler_get_eigen: null sampler");
        return GLASS_HIP_EINVAL;
    }
    GH_CHECK(hipSetDevice(S->v.device));
    const size_t m = (size_t)S->nslot * S->v.maxsrc;
    hipStream_t st = (hipStream_t)stream;
    if (evec) GH_CHECK(hipMemcpyAsync(evec, S->D.evec, sizeof(double) * 64 * m, hipMemcpyDeviceToHost, st));
    if (eval) GH_CHECK(hipMemcpyAsync(eval, S->D.eval, sizeof(double) * 8 * m, hipMemcpyDeviceToHost, st));
    GH_CHECK(hipStreamSynchronize(st));
    return GLASS_HIP_OK;
}

extern "C" int glass_hip_ucb_sampler_steps(glass_hip_ucb_sampler* S, int nsteps, void* stream)
{
    if (!S || nsteps < 0)
    {
        set_error("ucb_sampler_steps: bad arguments");
        return GLASS_HIP_EINVAL;
    }
    GH_CHECK(hipSetDevice(S->v.device));
    return enqueue_steps(S, nsteps, (hipStream_t)stream);
}

extern "C" int glass_hip_ucb_sampler_ptswap(glass_hip_ucb_sampler* S, int adapt, void* stream)
{
    if (!S)
    {
        set_error("ucb_sampler_ptswap: null sampler");
        return GLASS_HIP_EINVAL;
    }
    GH_CHECK(hipSetDevice(S->v.device));
    return enqueue_swap(S, adapt, (hipStream_t)stream);
}

static int enqueue_sweep(glass_hip_ucb_sampler* S, int nsteps, int adapt, hipStream_t st)
{
    if (int rc = enqueue_steps(S, nsteps, st)) return rc;
    if (int rc = enqueue_fisher(S, st)) return rc;
    return enqueue_swap(S, adapt, st);
}

extern "C" int glass_hip_ucb_sampler_sweep(glass_hip_ucb_sampler* S, int nsteps, int adapt, int graph, void* stream)
{
    if (!S || nsteps < 0 || nsteps > 4096)
    {
        set_error("ucb_sampler_sweep: bad arguments (steps %d)", nsteps);
        return GLASS_HIP_EINVAL;
    }
    GH_CHECK(hipSetDevice(S->v.device));
    hipStream_t s = (hipStream_t)stream;
    if (!graph) return enqueue_sweep(S, nsteps, adapt, s);
    if (!S->gexec || S->g_steps != nsteps || S->g_adapt != (adapt ? 1 : 0))
    {
        if (int rc = drop_graph(S)) return rc;
        hipStream_t cs;
        GH_CHECK(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking));
        GH_CHECK(hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal));
        int rc = enqueue_sweep(S, nsteps, adapt, cs);
        hipGraph_t gr = nullptr;
        hipError_t ec = hipStreamEndCapture(cs, &gr);
        (void)hipStreamDestroy(cs);
        if (rc) return rc;
        if (ec != hipSuccess)
        {
            set_error("ucb_sampler_sweep: graph capture failed: %s", hipGetErrorString(ec));
            return GLASS_HIP_ERUNTIME;
        }
        S->graph = gr;
        GH_CHECK(hipGraphInstantiate(&S->gexec, gr, nullptr, nullptr, 0));
        S->g_steps = nsteps;
        S->g_adapt = adapt ? 1 : 0;
    }
    GH_CHECK(hipGraphLaunch(S->gexec, s));
    return GLASS_HIP_OK;
}

extern "C" int glass_hip_ucb_sampler_get_state(glass_hip_ucb_sampler* S, unsigned* seeds, int* index,
                                               double* temperature, int* trials, int* accepts, int* swaps,
                                               void* stream)
{
    if (!S)
    {
        set_error("ucb_sampler_get_state: null sampler");
        return GLASS_HIP_EINVAL;
    }
    GH_CHECK(hipSetDevice(S->v.device));
    hipStream_t st = (hipStream_t)stream;
    const int N = S->nslot;
    const SamplerDev& D = S->D;
    if (seeds) GH_CHECK(hipMemcpyAsync(seeds, D.seed, sizeof(unsigned) * N, hipMemcpyDeviceToHost, st));
    if (index) GH_CHECK(hipMemcpyAsync(index, D.index, sizeof(int) * N, hipMemcpyDeviceToHost, st));
    if (temperature) GH_CHECK(hipMemcpyAsync(temperature, D.temp, sizeof(double) * N, hipMemcpyDeviceToHost, st));
    if (trials) GH_CHECK(hipMemcpyAsync(trials, D.trials, sizeof(int) * 9 * N, hipMemcpyDeviceToHost, st));
    if (accepts) GH_CHECK(hipMemcpyAsync(accepts, D.accepts, sizeof(int) * 9 * N, hipMemcpyDeviceToHost, st));
    if (swaps) GH_CHECK(hipMemcpyAsync(swaps, D.swaps, sizeof(int) * N, hipMemcpyDeviceToHost, st));
    int overrun = 0;
    GH_CHECK(hipMemcpyAsync(&overrun, D.fs_overrun, sizeof(int), hipMemcpyDeviceToHost, st));
    GH_CHECK(hipStreamSynchronize(st));
    if (overrun)
    {
        set_error("ucb_sampler: %d F-statistic draws gave up after %d rejections (proposal left unchanged)",
                  overrun, kFsMaxTries);
        return GLASS_HIP_ERUNTIME;
    }
    return GLASS_HIP_OK;
}

extern "C" int glass_hip_ucb_sampler_save_chain_states(glass_hip_ucb_sampler* S, const char* const* dirs, int step,
                                                       const double* logLmax, const double* logLnorm, double t0,
                                                       const double* eta, int neta, void* stream)
{
    if (!S || !dirs || neta < 0 || (neta && !eta))
    {
        set_error("ucb_sampler_save_chain_states: bad arguments");
        return GLASS_HIP_EINVAL;
    }
    for (int l = 0; l < S->nladder; l++)
        if (!dirs[l])
        {
            set_error("ucb_sampler_save_chain_states: no directory for ladder %d", l);
            return GLASS_HIP_EINVAL;
        }
    GH_CHECK(hipSetDevice(S->v.device));
    hipStream_t st = (hipStream_t)stream;
    const FdBatchView& v = S->v;
    std::vector<int> index(S->nslot), nsrc(v.C);
    GH_CHECK(hipMemcpyAsync(index.data(), S->D.index, sizeof(int) * S->nslot, hipMemcpyDeviceToHost, st));
    GH_CHECK(hipMemcpyAsync(nsrc.data(), v.c_nsrc, sizeof(int) * v.C, hipMemcpyDeviceToHost, st));
    GH_CHECK(hipStreamSynchronize(st));
    std::vector<double> logl(v.C), sp((size_t)v.C * v.maxsrc * 8);
    if (int rc = glass_hip_fd_batch_state(S->plan, logl.data(), sp.data())) return rc;
    std::vector<char> path(4096);
    // save_chain_state (ucb/src/glass_ucb_io.c:418-446): file ic of a run
    // holds model[chain->index[ic]], the chain now at temperature ic
    for (int l = 0; l < S->nladder; l++)
        for (int t = 0; t < S->ntemp; t++)
        {
            const int c = index[(size_t)l * S->ntemp + t];
            if (c < 0 || c >= v.C)
            {
                set_error("ucb_sampler_save_chain_states: slot %d maps to chain %d", l * S->ntemp + t, c);
                return GLASS_HIP_ERUNTIME;
            }
            if (snprintf(path.data(), path.size(), "%s/chain_state_%i.dat", dirs[l], t) >= (int)path.size())
            {
                set_error("ucb_sampler_save_chain_states: directory name too long");
                return GLASS_HIP_EINVAL;
            }
            if (int rc = glass_hip_write_chain_state_file(path.data(), logLmax ? logLmax[l] : 0.0, step, nsrc[c],
                                                          logl[c], logLnorm ? logLnorm[c] : 0.0, t0, eta, neta,
                                                          sp.data() + (size_t)c * v.maxsrc * 8, v.T))
                return rc;
        }
    return GLASS_HIP_OK;
}
