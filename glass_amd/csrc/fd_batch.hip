// fd_batch.hip -- batched, device-resident UCB waveform + delta-logL
// (include/glass_hip_batch.h).
//
// One evaluation = one MCMC proposal of the reference's ucb_mcmc step
// (ucb/src/glass_ucb_sampler.c:139-155): regenerate one source's template
// and evaluate the chain's likelihood and the template's snr.  The sweep is
// a fixed launch sequence with no host synchronisation:
//   1. zero the BW-class counters (a kernel, not a memset: graph-safe)
//   2. align_kernel: ucb_alignment per proposal on the device; each block
//      histograms its proposals into the 8 BW classes (32..4096) in LDS and
//      reserves its slots with one atomic per class
//   3. per BW class, an eval_kernel<N, NTHR, LOOP> over all proposals,
//      launched with the class's tile (threads per workgroup; one
//      workgroup per proposal slot, or a multiple of the resident capacity
//      striding over the class list), chosen by glass_hip_fd_batch_autotune
//      from the compiled candidates: each workgroup takes an entry of its
//      class list, builds the template in LDS (slow series, 6 FFTs, TDI)
//      and reduces (r|dh) - (dh|dh)/2 over the union of the old and new
//      supports only, reading the chain residual, the group's invC and the
//      source's stored template with coalesced loads.
// The full-window likelihood is therefore never re-read per proposal: the
// algorithmic HBM traffic per proposal is ~144 bytes per template bin, and
// proposals of one chain find their window in L2.
#include <cmath>
#include <cstring>
#include <vector>

#include "../../include/glass_hip_batch.h"
#include "fd_batch_hooks.h"
#include "fd_derivs.h"
#include "fd_waveform.h"
#include "reduce.h"
#include "runtime.h"

namespace glass {

int parse_format(const char* format);

constexpr int kMaxSrcDefault = 8;   // source slots per chain of glass_hip_fd_batch_create
constexpr int kMaxSrcLimit = 64;    // largest max_src of glass_hip_fd_batch_create_ex
constexpr int kNClass = 8;     // BW = 32 << k, k = 0..7 (32..4096)
constexpr int kPosN = 4096;    // positions table samples (largest BW)

struct BatchDev
{
    // groups
    const int* g_qmin;
    const int* g_W;
    const long long* g_off;  // bins before group g
    const double2* data;     // [3][W_g] per group
    const double* invC;      // [6][W_g] per group
    // chains
    const int* c_group;
    const int* c_nsrc;
    const long long* c_roff;  // bins before chain c's residual
    double2* resid;           // [3][W] per chain
    double* logl;             // [C]
    double* sparams;          // [C][8][8]
    int* s_bw;                // [C*8]
    int* s_imin;              // [C*8]
    // templates: chain c's live sources s < c_nsrc[c] hold [3][cap] complex
    // at tmpl + c_toff[c] + 3 s cap, cap = g_cap[group] (the largest BW the
    // window allows: the largest power of two <= W)
    double2* tmpl;
    const int* g_cap;
    const long long* c_toff;
    int bwcap;                // largest g_cap
    int maxsrc;               // source slots per chain in sparams / s_bw / s_imin
    UcbConst* pc;             // per-item template constants, written by align_kernel
    // constants
    double T, t0, L, fstar;
    const double* pos;      // positions table, kPosN samples
    const double* pos_soa;  // structure-of-arrays position blocks per series length
    const double2* tw;
    int fmt;
};

// Positions for fd_link_bins at N bins: the structure-of-arrays block of N
// for the split shape (coalesced coordinate loads), else the record table.
template <int N, int NTHR>
__device__ __forceinline__ const double* batch_pos(const BatchDev& B)
{
    return fd_split(N, NTHR) ? B.pos_soa + pos_soa_offset(N) : B.pos;
}

// Work items: item e refers to entry q = sel ? sel[e] : e of the
// (chain, src, params[8]) arrays.
struct Items
{
    const int* chain;
    const int* src;
    const double* params;
    const int* sel;
    // optional device-side item count (items e >= *dn are left out); lets
    // graph-captured callers size a pass on the device
    const int* dn = nullptr;
    // optional template staging of an eval: item e's X, Y, Z bins go to
    // stg[e * 3 bwcap + ch * N + m] (N its template bins)
    double2* stg = nullptr;
};

__device__ inline int item_q(const Items& it, int e) { return it.sel ? it.sel[e] : e; }
__device__ inline const double* item_params(const Items& it, int e) { return it.params + (size_t)item_q(it, e) * 8; }
__device__ inline int item_chain(const Items& it, int e) { return it.chain[item_q(it, e)]; }
__device__ inline int item_src(const Items& it, int e) { return it.src[item_q(it, e)]; }

__device__ inline int chain_cap(const BatchDev& B, int c) { return B.g_cap[B.c_group[c]]; }
__device__ inline double2* src_tmpl(const BatchDev& B, int c, int s)
{
    return B.tmpl + B.c_toff[c] + (size_t)s * 3 * chain_cap(B, c);
}

__global__ void zero_int_kernel(int* p, int n)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

// ucb_alignment on the device (ucb/src/glass_ucb_waveform.c:727-736) and
// bucketing by BW class.  Class slots are reserved per block (an LDS
// histogram, then one global atomic per class per block) so the class
// counters are not hammered by every item.  Items whose BW falls outside
// the classes are counted in *err and left out.
__global__ __launch_bounds__(256) void align_kernel(int n, Items it, BatchDev B, int* __restrict__ bw_out,
                                                    int* __restrict__ imin_out, int* __restrict__ cls_count,
                                                    int* __restrict__ cls_list, int cls_stride, int* __restrict__ err,
                                                    double* __restrict__ logl_out = nullptr,
                                                    double* __restrict__ snr_out = nullptr)
{
    __shared__ int hist[kNClass + 1], base[kNClass + 1];
    if (threadIdx.x <= kNClass) hist[threadIdx.x] = 0;
    __syncthreads();
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    int k = -1, slot = 0;
    // items past the device count, or with a negative source index (a
    // proposal already rejected), are left out of every class list
    if (e < n && (!it.dn || e < *it.dn) && item_src(it, e) >= 0)
    {
        const double* p = item_params(it, e);
        int c = item_chain(it, e);
        int g = B.c_group[c];
        int W = B.g_W[g];
        double f0 = p[0] / B.T;
        double costh = p[1], amp = exp(p[3]);
        double dfdt = p[7] / (B.T * B.T);
        int BW = 2 * ucb_bandwidth(B.L, B.fstar, f0, dfdt, costh, amp, B.T, W);
        int qmin = (int)(f0 * B.T) - BW / 2;
        bw_out[e] = BW;
        imin_out[e] = qmin - B.g_qmin[g];
        // the template constants of the class kernels, once per item
        ucb_setup(p, 8, B.T, B.t0, B.L, B.fstar, B.pc[e]);
        k = 31 - __clz(BW) - 5;  // log2(BW) - 5
        if (BW > B.g_cap[g] || k < 0 || k >= kNClass || (BW & (BW - 1)))
        {
            k = kNClass;
            // no class kernel evaluates this item: a likelihood of -inf
            // (and snr 0) keeps a sampler from ever accepting it
            if (logl_out) logl_out[e] = -INFINITY;
            if (snr_out) snr_out[e] = 0.0;
        }
        slot = atomicAdd(&hist[k], 1);
    }
    __syncthreads();
    if (threadIdx.x < kNClass && hist[threadIdx.x]) base[threadIdx.x] = atomicAdd(&cls_count[threadIdx.x], hist[threadIdx.x]);
    if (threadIdx.x == kNClass && hist[kNClass]) atomicAdd(err, hist[kNClass]);
    __syncthreads();
    if (k >= 0 && k < kNClass) cls_list[k * cls_stride + base[k] + slot] = e;
}

// Work distribution of the class kernels: one workgroup per proposal slot
// (grid = proposal count; workgroups beyond the class count exit at once,
// and the body has no loop, so nothing is hoisted across templates and the
// register footprint stays that of one template), or a grid sized to a
// multiple of the resident capacity striding over the class list.

// symmetric 3x3 weighted bilinear form of complex 3-vectors:
// 2 * sum_ab Re(u_a conj v_b) w_ab, w planes XX,YY,ZZ,XY,XZ,YZ
__device__ inline double bform(const double2* u, const double2* v, const double* w)
{
    auto re = [](double2 a, double2 b) { return a.x * b.x + a.y * b.y; };
    double d = w[0] * re(u[0], v[0]) + w[1] * re(u[1], v[1]) + w[2] * re(u[2], v[2]);
    double o = w[3] * (re(u[0], v[1]) + re(u[1], v[0])) + w[4] * (re(u[0], v[2]) + re(u[2], v[0])) +
               w[5] * (re(u[1], v[2]) + re(u[2], v[1]));
    return 2.0 * (d + o);
}

__device__ inline void load_w(const double* __restrict__ iC, int W, int j, double* w)
{
#pragma unroll
    for (int q = 0; q < 6; q++) w[q] = iC[(size_t)q * W + j];
}

// Fused waveform + delta-logL + snr of class-list entry w.
template <int N, int NTHR>
__device__ __forceinline__ void eval_item(int w, double2* lds, const int* __restrict__ list, Items it,
                                          BatchDev B, const int* __restrict__ p_imin,
                                          double* __restrict__ logl_out, double* __restrict__ snr_out)
{
    {
        const int e = list[w];
        const int c = item_chain(it, e), s = item_src(it, e);
        const int g = B.c_group[c];
        const int W = B.g_W[g];
        // the constants straight into scalar registers (uniform loads):
        // no LDS copy and no barriers around it, and the slow series reads
        // them as scalar operands (3% faster at 512 bins than through LDS)
        UcbConst kc = B.pc[e];
        make_uniform(kc);
        const UcbConst& k = kc;

        const int so = c * B.maxsrc + s;
        const int imin_n = p_imin[e];
        const int imin_o = B.s_imin[so], bw_o = B.s_bw[so];
        const int cap = B.g_cap[g];
        const double2* ho = src_tmpl(B, c, s);
        const double2* R = B.resid + (size_t)B.c_roff[c] * 3;
        const double* iC = B.invC + (size_t)B.g_off[g] * 6;
        double acc = 0.0, s2 = 0.0;
        // the old template's values at bin m, loaded before the bin's
        // transform (their latency, an HBM read per source, overlaps the FFT)
        double2 hpre[3];
        auto pre = [&](int m) {
            const int j = imin_n + m, mo = j - imin_o;
            const bool inold = j >= 0 && j < W && mo >= 0 && mo < bw_o;
#pragma unroll
            for (int ch = 0; ch < 3; ch++) hpre[ch] = inold ? ho[ch * cap + mo] : make_double2(0.0, 0.0);
        };
        // new-template bins (and the old template where they overlap)
        // no tail barrier: the reduction's barrier below separates this
        // item's last LDS reads from the next item's writes
        fd_link_bins<N, NTHR, fd_split(N, NTHR), fd_split(N, NTHR), false>(
            k, lds, opaque(batch_pos<N, NTHR>(B)), kPosN / N, opaque(B.tw), kTwN / N, pre,
            [&](int m, const double2* dl) {
                double2 v[5];
                fd_tdi<N>(k, m, B.fmt, dl, v);
                if (it.stg)
#pragma unroll
                    for (int ch = 0; ch < 3; ch++) it.stg[(size_t)e * 3 * B.bwcap + ch * N + m] = v[ch];
                double wq[6];
                load_w(iC, W, m, wq);  // snr: template bin m against window bin m
                s2 += bform(v, v, wq);
                const int j = imin_n + m;
                if (j >= 0 && j < W)
                {
                    double2 d[3], r[3];
#pragma unroll
                    for (int ch = 0; ch < 3; ch++)
                    {
                        const double2 h0 = hpre[ch];  // zero outside the old template
                        d[ch] = make_double2(v[ch].x - h0.x, v[ch].y - h0.y);
                        r[ch] = R[(size_t)ch * W + j];
                    }
                    load_w(iC, W, j, wq);
                    acc += bform(r, d, wq) - 0.5 * bform(d, d, wq);
                }
            });
        // old-template bins outside the new support: dh = -h_old
        for (int mo = threadIdx.x; mo < bw_o; mo += NTHR)
        {
            const int j = imin_o + mo;
            if (j < 0 || j >= W || (j >= imin_n && j < imin_n + N)) continue;
            double2 d[3], r[3];
            double wq[6];
#pragma unroll
            for (int ch = 0; ch < 3; ch++)
            {
                double2 h0 = ho[ch * cap + mo];
                d[ch] = make_double2(-h0.x, -h0.y);
                r[ch] = R[(size_t)ch * W + j];
            }
            load_w(iC, W, j, wq);
            acc += bform(r, d, wq) - 0.5 * bform(d, d, wq);
        }
        // both sums over the workgroup behind one barrier (block_sum's
        // order: wave sums, then the waves' partials in turn); thread 0
        // writes the logL, thread 1 the snr.  The partials are next written
        // by the following item's reduction, many barriers later.
        constexpr int NW = (NTHR < 64 ? 64 : NTHR) / 64;
        __shared__ double part[NW][2];
        const double wa = wave_sum(acc), ws = wave_sum(s2);
        if ((threadIdx.x & 63) == 0)
        {
            part[threadIdx.x >> 6][0] = wa;
            part[threadIdx.x >> 6][1] = ws;
        }
        __syncthreads();
        if (threadIdx.x < 2)
        {
            double r = 0.0;
            for (int i = 0; i < NW; i++) r += part[i][threadIdx.x];
            if (threadIdx.x == 0)
                logl_out[e] = B.logl[c] + r;
            else
                snr_out[e] = sqrt(r);
        }
    }
}

// The proposals of one BW class.  LOOP = false: one workgroup per proposal
// slot (grid = proposal count; workgroups past the class count exit at
// once), so the body has no loop and nothing is hoisted across templates.
// LOOP = true: any grid, each workgroup strides over the class list.
// MINW: waves per SIMD the register allocation must allow (1: the
// compiler's choice; more trades register spills for resident waves).
template <int N, int NTHR, bool LOOP, int MINW>
__global__ __launch_bounds__(NTHR) __attribute__((amdgpu_waves_per_eu(MINW, 8))) void eval_kernel(const int* __restrict__ list, const int* __restrict__ count,
                                                    Items it, BatchDev B, const int* __restrict__ p_imin,
                                                    double* __restrict__ logl_out, double* __restrict__ snr_out)
{
    __shared__ double2 lds[FdShape<N, NTHR, fd_split(N, NTHR)>::LDS];
    const int cnt = *count;
    if constexpr (LOOP)
    {
        for (int w = blockIdx.x; w < cnt; w += gridDim.x)
            eval_item<N, NTHR>(w, lds, list, it, B, p_imin, logl_out, snr_out);
    }
    else
    {
        if (blockIdx.x < cnt) eval_item<N, NTHR>(blockIdx.x, lds, list, it, B, p_imin, logl_out, snr_out);
    }
}

// Compiled launch shapes of the class kernels (tile candidates), default
// first: (class k, template bins N, workgroup threads, waves per SIMD).
// From 128 to 1024 bins the default is the split shape (N/2 threads, two
// bins each, half the LDS of one sample per thread): at 126 VGPRs it keeps
// 4 waves per SIMD, so BW 1024 fits two workgroups per CU instead of one and
// BW 512 four 256-thread workgroups instead of three.  The N-thread shapes
// (one sample per thread, ~94 VGPRs) and their 5- and 6-wave variants
// (spilling some registers for more waves in flight across the FFT
// barriers) remain candidates.  The decimated classes (N >= 2048) keep
// their accumulators in registers only at 512 (and 1024 for N = 2048)
// threads.
#define GH_EVAL_TILES(X)                                                                                      \
    X(0, 32, 64, 1) X(1, 64, 64, 1) X(1, 64, 64, 6) X(2, 128, 64, 1) X(2, 128, 128, 1) X(2, 128, 128, 6)     \
    X(3, 256, 128, 1) X(3, 256, 128, 4) X(3, 256, 256, 1) X(3, 256, 256, 5) X(3, 256, 256, 6)                \
    X(4, 512, 256, 1) X(4, 512, 512, 1) X(4, 512, 512, 6) X(5, 1024, 512, 4) X(5, 1024, 512, 1)              \
    X(5, 1024, 1024, 1) X(5, 1024, 256, 1)                                                                   \
    X(6, 2048, 512, 1) X(6, 2048, 1024, 1) X(7, 4096, 512, 1)

// Template of each listed item into its source's slot.  LOOP: the grid
// strides over the class list (large rebuilds, where one workgroup per
// item would mostly launch workgroups that exit at once).
template <int N, int NTHR>
__device__ __forceinline__ void tmpl_item(int w, double2* lds, const int* __restrict__ list, Items it, BatchDev B)
{
    const int e = list[w];
    UcbConst kc = B.pc[e];  // into scalar registers, as in eval_item
    make_uniform(kc);
    const UcbConst& k = kc;
    const int c = item_chain(it, e), cap = chain_cap(B, c);
    double2* o = src_tmpl(B, c, item_src(it, e));
    fd_link_bins<N, NTHR, fd_split(N, NTHR), fd_split(N, NTHR)>(k, lds, opaque(batch_pos<N, NTHR>(B)), kPosN / N,
                                                                opaque(B.tw), kTwN / N,
                                             [&](int m, const double2* dl) {
                                                 double2 v[5];
                                                 fd_tdi<N>(k, m, B.fmt, dl, v);
#pragma unroll
                                                 for (int ch = 0; ch < 3; ch++) o[ch * cap + m] = v[ch];
                                             });
}

template <int N, int NTHR, bool LOOP>
__global__ __launch_bounds__(NTHR) void tmpl_kernel(const int* __restrict__ list, const int* __restrict__ count,
                                                    Items it, BatchDev B)
{
    __shared__ double2 lds[FdShape<N, NTHR, fd_split(N, NTHR)>::LDS];
    const int cnt = *count;
    if constexpr (LOOP)
    {
        for (int w = blockIdx.x; w < cnt; w += gridDim.x) tmpl_item<N, NTHR>(w, lds, list, it, B);
    }
    else
    {
        if ((int)blockIdx.x < cnt) tmpl_item<N, NTHR>(blockIdx.x, lds, list, it, B);
    }
}

// r_c = d_g - sum_s h_cs (add_ucb_model, ucb/src/glass_ucb_model.c:192-232)
// One workgroup per chain (grid = chain count), striding over its window.
__global__ __launch_bounds__(256) void resid_kernel(BatchDev B)
{
    const int c = blockIdx.x;
    const int g = B.c_group[c], W = B.g_W[g], cap = B.g_cap[g];
    const double2* D = B.data + (size_t)B.g_off[g] * 3;
    double2* R = B.resid + (size_t)B.c_roff[c] * 3;
    for (int j = threadIdx.x; j < W; j += blockDim.x)
        for (int ch = 0; ch < 3; ch++)
        {
            double2 v = D[(size_t)ch * W + j];
            for (int s = 0; s < B.c_nsrc[c]; s++)
            {
                const int so = c * B.maxsrc + s;
                const int m = j - B.s_imin[so];
                if (m >= 0 && m < B.s_bw[so])
                {
                    const double2 h = src_tmpl(B, c, s)[ch * cap + m];
                    v.x -= h.x;
                    v.y -= h.y;
                }
            }
            R[(size_t)ch * W + j] = v;
        }
}

// logL_c = -1/2 (r_c | r_c) over the whole window
__global__ __launch_bounds__(256) void chain_logl_kernel(BatchDev B)
{
    const int c = blockIdx.x;
    const int g = B.c_group[c], W = B.g_W[g];
    const double2* R = B.resid + (size_t)B.c_roff[c] * 3;
    const double* iC = B.invC + (size_t)B.g_off[g] * 6;
    double acc = 0.0;
    for (int j = threadIdx.x; j < W; j += 256)
    {
        double2 r[3];
        double w[6];
        for (int ch = 0; ch < 3; ch++) r[ch] = R[(size_t)ch * W + j];
        load_w(iC, W, j, w);
        acc += bform(r, r, w);
    }
    acc = block_sum<256>(acc);
    if (threadIdx.x == 0) B.logl[c] = -0.5 * acc;
}

// Accept, in place: (1) residual += h_old over the old support, (2) the new
// template overwrites the source's slot, (3) residual -= h_new over the new
// support and the source takes its new parameters.  Per bin the residual
// sees r + h_old - h_new in that order.  One workgroup per accepted
// proposal (grid = nacc).
__global__ __launch_bounds__(256) void accept_add_old_kernel(int nacc, Items it, BatchDev B)
{
    const int a = blockIdx.x;
    if (a >= nacc || (it.dn && a >= *it.dn)) return;
    const int c = item_chain(it, a), s = item_src(it, a), so = c * B.maxsrc + s;
    const int g = B.c_group[c], W = B.g_W[g], cap = B.g_cap[g];
    const int io = B.s_imin[so], bo = B.s_bw[so];
    double2* R = B.resid + (size_t)B.c_roff[c] * 3;
    const double2* ho = src_tmpl(B, c, s);
    for (int j = max(0, io) + threadIdx.x; j < min(W, io + bo); j += blockDim.x)
        for (int ch = 0; ch < 3; ch++)
        {
            double2 v = R[(size_t)ch * W + j];
            v.x += ho[ch * cap + j - io].x;
            v.y += ho[ch * cap + j - io].y;
            R[(size_t)ch * W + j] = v;
        }
}

__global__ __launch_bounds__(256) void accept_sub_new_kernel(int nacc, Items it, BatchDev B,
                                                             const int* __restrict__ bw_new,
                                                             const int* __restrict__ imin_new,
                                                             const double* __restrict__ logl_new)
{
    const int a = blockIdx.x;
    if (a >= nacc || (it.dn && a >= *it.dn)) return;
    const int c = item_chain(it, a), s = item_src(it, a), so = c * B.maxsrc + s;
    const int g = B.c_group[c], W = B.g_W[g], cap = B.g_cap[g];
    const int in = imin_new[a], bn = bw_new[a];
    double2* R = B.resid + (size_t)B.c_roff[c] * 3;
    const double2* hn = src_tmpl(B, c, s);
    for (int j = max(0, in) + threadIdx.x; j < min(W, in + bn); j += blockDim.x)
        for (int ch = 0; ch < 3; ch++)
        {
            double2 v = R[(size_t)ch * W + j];
            v.x -= hn[ch * cap + j - in].x;
            v.y -= hn[ch * cap + j - in].y;
            R[(size_t)ch * W + j] = v;
        }
    if (threadIdx.x < 8) B.sparams[(size_t)so * 8 + threadIdx.x] = item_params(it, a)[threadIdx.x];
    if (threadIdx.x == 0)
    {
        B.s_bw[so] = bn;
        B.s_imin[so] = in;
        B.logl[c] = logl_new[item_q(it, a)];
    }
}

}  // namespace glass

using namespace glass;

// Exact derivatives of chain sources at their current parameters: per-item
// arguments of the derivative kernels (fd_derivs.h) and BW-class lists
// (BW 32..4096, classes 0..7).  Items outside those (another BW, a chain or
// source index out of range) get NaN outputs.
constexpr int kNDClass = 8;

struct DerivItems
{
    int *pidx, *imin, *W, *tcap;
    long long *roff, *ioff, *toff;
    int* cnt;   // [kNDClass]
    int* list;  // [kNDClass][cap]
};

__global__ __launch_bounds__(256) void derivs_prep_kernel(int n, const int* __restrict__ chain,
                                                          const int* __restrict__ src, int nchain, BatchDev B,
                                                          DerivItems D, int cap, double* __restrict__ outL,
                                                          double* __restrict__ grad, double* __restrict__ fisher)
{
    __shared__ int hist[kNDClass + 1], base[kNDClass + 1];
    if (threadIdx.x <= kNDClass) hist[threadIdx.x] = 0;
    __syncthreads();
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    int k = -1, slot = 0;
    if (e < n)
    {
        const int c = chain[e], s = src[e];
        k = kNDClass;
        if (c >= 0 && c < nchain && s >= 0 && s < B.c_nsrc[c])
        {
            const int so = c * B.maxsrc + s, g = B.c_group[c];
            D.pidx[e] = so;
            D.imin[e] = B.s_imin[so];
            D.W[e] = B.g_W[g];
            D.roff[e] = B.c_roff[c] * 3;
            D.ioff[e] = B.g_off[g] * 6;
            D.toff[e] = B.c_toff[c] + (long long)s * 3 * B.g_cap[g];
            D.tcap[e] = B.g_cap[g];
            const int bw = B.s_bw[so];
            if (bw >= 32 && bw <= 4096 && (bw & (bw - 1)) == 0) k = 31 - __clz(bw) - 5;
        }
        slot = atomicAdd(&hist[k], 1);
    }
    __syncthreads();
    if (threadIdx.x < kNDClass && hist[threadIdx.x]) base[threadIdx.x] = atomicAdd(&D.cnt[threadIdx.x], hist[threadIdx.x]);
    __syncthreads();
    if (k >= 0 && k < kNDClass) D.list[k * cap + base[k] + slot] = e;
    if (k == kNDClass)
    {
        const double nan = __builtin_nan("");
        outL[e] = nan;
        for (int q = 0; q < 8; q++) grad[(size_t)e * 8 + q] = nan;
        for (int q = 0; q < 64; q++) fisher[(size_t)e * 64 + q] = nan;
    }
}

struct glass_hip_fd_batch
{
    int device = 0;
    glass_hip_orbit orbit{};
    int fmt = TDI_SANGRIA;
    double T = 0, t0 = 0;
    int G = 0, C = 0, maxprop = 0, ninit = 0;
    int Wmax = 0;
    std::vector<void*> allocs;
    BatchDev B{};
    // proposal buffers
    int *p_chain = nullptr, *p_src = nullptr;
    double *p_params = nullptr, *p_logl = nullptr, *p_snr = nullptr;
    int *p_bw = nullptr, *p_imin = nullptr;
    int* cls = nullptr;       // [kNClass] counts + [1] error
    int* cls_list = nullptr;  // [kNClass][cap]
    int cls_cap = 0;
    // init items: (chain, src) of every live source, its params gathered
    int *i_chain = nullptr, *i_src = nullptr, *i_slot = nullptr;
    double* i_params = nullptr;
    size_t tmpl_bytes = 0;    // template slots
    int maxsrc = kMaxSrcDefault;
    size_t alloc_bytes = 0;   // device memory the plan holds
    double2* d2048 = nullptr; // BW-2048 derivative scratch, one slice per CU
    double2* d4096 = nullptr; // BW-4096 derivative scratch, one slice per CU
    DerivItems D{};
    int ncu = 0;
    int tgrid[kNClass] = {0};
    glass_hip_tile tile[kNClass] = {};  // launch shape of each class's eval kernel
    // graph cache
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    int g_nprop = -1;
    // class kernels run concurrently: classes k with k % 4 = 1, 2, 3 go to
    // side streams forked from and joined back into the caller's stream
    // (so graph capture records them as parallel branches)
    hipStream_t side[3] = {};
    hipEvent_t evf = nullptr, evj[3] = {};
    bool concurrent = true;
    // profiling (eager evals only): events around each class kernel
    bool prof = false;
    hipEvent_t ev[kNClass][2] = {};
    double prof_ms[kNClass] = {0};
    int prof_launches[kNClass] = {0};
    int prof_count[kNClass] = {0};
    const void* g_key[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
    // sweep graph (eval + derivatives)
    hipGraph_t sgraph = nullptr;
    hipGraphExec_t sgexec = nullptr;
    long long s_key[12] = {};
};

template <typename T>
static int dalloc(glass_hip_fd_batch* P, T** p, size_t n)
{
    void* q = nullptr;
    GH_CHECK(hipMalloc(&q, sizeof(T) * (n ? n : 1)));
    P->allocs.push_back(q);
    P->alloc_bytes += sizeof(T) * (n ? n : 1);
    *p = (T*)q;
    return GLASS_HIP_OK;
}

#define GH_NTHR_(N) (::glass::fd_batch_threads(N))
#define GH_NTHR(N) GH_NTHR_(N)

// Workgroups of a class kernel resident per CU, as bounded by LDS (160 KB)
// and by the waves per SIMD the register allocation allows (4 by default,
// `waves` when forced higher).
static int resident_per_cu(int N, int nthr, int waves)
{
    const int ns = fd_split(N, nthr) ? N / 2 : (N <= 1024 ? N : GH_DECIM_NS);
    const int lds = (int)sizeof(double2) * 6 * ns + 512;
    int per = (160 * 1024) / lds;
    const int wv = 4 * (waves > 4 ? waves : 4) / ((nthr + 63) / 64);
    if (wv < per) per = wv;
    return per < 1 ? 1 : per;
}

// Waves per SIMD of the default tiles (the first candidate of each class):
// 4 at 1024 bins, where the compiler's choice lands a few VGPRs above the
// 128 that four waves allow.
constexpr int batch_default_waves(int N) { return N == 1024 ? 4 : 1; }

static int setup_grids(glass_hip_fd_batch* P)
{
    GH_CHECK(hipDeviceGetAttribute(&P->ncu, hipDeviceAttributeMultiprocessorCount, P->device));
    for (int k = 0; k < kNClass; k++)
    {
        const int N = 32 << k;
        P->tgrid[k] = P->ncu * resident_per_cu(N, fd_batch_threads(N), batch_default_waves(N));
        P->tile[k] = glass_hip_tile{fd_batch_threads(N), 0, batch_default_waves(N)};
    }
    return GLASS_HIP_OK;
}

// Grid multipliers of the tile candidates: 0 = one workgroup per proposal
// slot, m > 0 = m x (resident workgroups per CU) x CUs, striding.
constexpr int kGmul[] = {0, 1, 2, 4};
constexpr int kNGmul = sizeof(kGmul) / sizeof(kGmul[0]);

static bool tile_compiled(int k, int nthr, int waves)
{
#define GH_X(kk, N, T, MW) \
    if (k == kk && nthr == T && waves == MW) return true;
    GH_EVAL_TILES(GH_X)
#undef GH_X
    return false;
}

static int tile_grid(const glass_hip_fd_batch* P, int k, glass_hip_tile t, int n)
{
    if (t.gmul <= 0) return n;
    const long long g = (long long)t.gmul * P->ncu * resident_per_cu(32 << k, t.nthr, t.waves);
    return (int)(g < n ? g : n);
}

// The class-k eval kernel under tile t over the n proposal slots.
static int launch_eval(glass_hip_fd_batch* P, int k, glass_hip_tile t, int n, const int* list, const int* cnt,
                       const Items& it, double* logl_out, double* snr_out, hipStream_t s)
{
    const int g = tile_grid(P, k, t, n);
#define GH_X(kk, N, T, MW)                                                                                       \
    if (k == kk && t.nthr == T && t.waves == MW)                                                                 \
    {                                                                                                            \
        if (t.gmul > 0)                                                                                          \
            eval_kernel<N, T, true, MW><<<g, T, 0, s>>>(list, cnt, it, P->B, P->p_imin, logl_out, snr_out);      \
        else                                                                                                     \
            eval_kernel<N, T, false, MW><<<g, T, 0, s>>>(list, cnt, it, P->B, P->p_imin, logl_out, snr_out);     \
        return GLASS_HIP_OK;                                                                                     \
    }
    GH_EVAL_TILES(GH_X)
#undef GH_X
    set_error("batch: no compiled tile for class %d with %d threads, %d waves", k, t.nthr, t.waves);
    return GLASS_HIP_EINVAL;
}

// alignment + per-class launches for items; mode 0: eval, 1: templates
static int launch_items(glass_hip_fd_batch* P, int n, const Items& it, int mode, double* logl_out, double* snr_out,
                        hipStream_t s)
{
    if (n <= 0) return GLASS_HIP_OK;
    if (n > P->cls_cap)
    {
        set_error("batch: %d items exceed capacity %d", n, P->cls_cap);
        return GLASS_HIP_EINVAL;
    }
    zero_int_kernel<<<1, 64, 0, s>>>(P->cls, kNClass + 1);
    align_kernel<<<(n + 255) / 256, 256, 0, s>>>(n, it, P->B, P->p_bw, P->p_imin, P->cls, P->cls_list, P->cls_cap,
                                                 P->cls + kNClass, mode == 0 ? logl_out : nullptr,
                                                 mode == 0 ? snr_out : nullptr);
    // the per-class timings of the profiling mode need the classes in turn
    const bool par = P->concurrent && !(P->prof && mode == 0);
    if (par)
    {
        GH_CHECK(hipEventRecord(P->evf, s));
        for (int j = 0; j < 3; j++) GH_CHECK(hipStreamWaitEvent(P->side[j], P->evf, 0));
    }
    int lrc = GLASS_HIP_OK;
    for (int k = kNClass - 1; k >= 0; k--)
    {
        const int* list = P->cls_list + (size_t)k * P->cls_cap;
        const int* cnt = P->cls + k;
        const bool pr = P->prof && mode == 0;
        const hipStream_t sk = par && (k & 3) ? P->side[(k & 3) - 1] : s;
        if (pr) GH_CHECK(hipEventRecord(P->ev[k][0], sk));
        if ((32 << k) > P->B.bwcap)
        {
            // align_kernel routes BW above every window's cap to the error
            // count, so this class is empty; its one-workgroup-per-item grid
            // would only launch empty workgroups
        }
        else if (mode == 0)
        {
            if ((lrc = launch_eval(P, k, P->tile[k], n, list, cnt, it, logl_out, snr_out, sk))) break;
        }
        else
        {
            // one workgroup per item, or a striding grid of 4x the resident
            // capacity when the items are many more
            const bool loop = n > 4 * P->tgrid[k];
            const int g = loop ? 4 * P->tgrid[k] : n;
#define GH_L(kk, N)                                                                               \
    case kk:                                                                                      \
        if (loop)                                                                                 \
            tmpl_kernel<N, GH_NTHR(N), true><<<g, GH_NTHR(N), 0, sk>>>(list, cnt, it, P->B);      \
        else                                                                                      \
            tmpl_kernel<N, GH_NTHR(N), false><<<g, GH_NTHR(N), 0, sk>>>(list, cnt, it, P->B);     \
        break;
            switch (k)
            {
                GH_L(0, 32) GH_L(1, 64) GH_L(2, 128) GH_L(3, 256) GH_L(4, 512) GH_L(5, 1024) GH_L(6, 2048)
                GH_L(7, 4096)
            }
#undef GH_L
        }
        if (pr) GH_CHECK(hipEventRecord(P->ev[k][1], sk));
    }
    // joined also after a failed launch, so a capture is never left open
    if (par)
        for (int j = 0; j < 3; j++)
        {
            GH_CHECK(hipEventRecord(P->evj[j], P->side[j]));
            GH_CHECK(hipStreamWaitEvent(s, P->evj[j], 0));
        }
    if (lrc) return lrc;
    GH_CHECK(hipGetLastError());
    if (P->prof && mode == 0)
    {
        // synchronous in profiling mode: accumulate the class kernels' times
        int cnt[kNClass + 1];
        GH_CHECK(hipMemcpyAsync(cnt, P->cls, sizeof(int) * (kNClass + 1), hipMemcpyDeviceToHost, s));
        GH_CHECK(hipStreamSynchronize(s));
        for (int k = 0; k < kNClass; k++)
        {
            float ms = 0.f;
            GH_CHECK(hipEventElapsedTime(&ms, P->ev[k][0], P->ev[k][1]));
            P->prof_ms[k] += ms;
            P->prof_launches[k] += 1;
            P->prof_count[k] = cnt[k];
        }
    }
    return GLASS_HIP_OK;
}

// slot alignment from an alignment pass over the init items
__global__ void slot_align_kernel(int n, const int* __restrict__ slot, const int* __restrict__ bw,
                                  const int* __restrict__ imin, int* __restrict__ s_bw, int* __restrict__ s_imin)
{
    int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n) return;
    s_bw[slot[e]] = bw[e];
    s_imin[slot[e]] = imin[e];
}

__global__ void gather_params_kernel(int n, const int* __restrict__ slot, const double* __restrict__ sparams,
                                     double* __restrict__ out)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n * 8) return;
    out[i] = sparams[(size_t)slot[i / 8] * 8 + (i % 8)];
}

static int rebuild(glass_hip_fd_batch* P, hipStream_t s)
{
    const int n = P->ninit;
    gather_params_kernel<<<(n * 8 + 255) / 256, 256, 0, s>>>(n, P->i_slot, P->B.sparams, P->i_params);
    Items it{P->i_chain, P->i_src, P->i_params, nullptr};
    int rc;
    if ((rc = launch_items(P, n, it, 1, nullptr, nullptr, s))) return rc;
    slot_align_kernel<<<(n + 255) / 256, 256, 0, s>>>(n, P->i_slot, P->p_bw, P->p_imin, P->B.s_bw, P->B.s_imin);
    resid_kernel<<<P->C, 256, 0, s>>>(P->B);
    chain_logl_kernel<<<P->C, 256, 0, s>>>(P->B);
    GH_CHECK(hipGetLastError());
    return GLASS_HIP_OK;
}

static int check_err(glass_hip_fd_batch* P, hipStream_t s)
{
    int h = 0;
    GH_CHECK(hipMemcpyAsync(&h, P->cls + kNClass, sizeof(int), hipMemcpyDeviceToHost, s));
    GH_CHECK(hipStreamSynchronize(s));
    if (h)
    {
        set_error("batch: %d items with BW outside 32..%d", h, P->B.bwcap);
        return GLASS_HIP_EINVAL;
    }
    return GLASS_HIP_OK;
}

static int drop_graph(glass_hip_fd_batch* P);

extern "C" int glass_hip_fd_batch_create_ex(glass_hip_fd_batch** plan, const glass_hip_orbit* orbit,
                                            const char* format, double T, double t0, int ngroup,
                                            const int* group_qmin, const int* group_W, const double* data,
                                            const double* invC, int nchain, const int* chain_group,
                                            const int* chain_nsrc, int max_src, const double* src_params,
                                            int max_prop);

extern "C" int glass_hip_fd_batch_create(glass_hip_fd_batch** plan, const glass_hip_orbit* orbit, const char* format,
                                         double T, double t0, int ngroup, const int* group_qmin, const int* group_W,
                                         const double* data, const double* invC, int nchain, const int* chain_group,
                                         const int* chain_nsrc, const double* src_params, int max_prop)
{
    return glass_hip_fd_batch_create_ex(plan, orbit, format, T, t0, ngroup, group_qmin, group_W, data, invC, nchain,
                                        chain_group, chain_nsrc, kMaxSrcDefault, src_params, max_prop);
}

extern "C" int glass_hip_fd_batch_create_ex(glass_hip_fd_batch** plan, const glass_hip_orbit* orbit,
                                            const char* format, double T, double t0, int ngroup,
                                            const int* group_qmin, const int* group_W, const double* data,
                                            const double* invC, int nchain, const int* chain_group,
                                            const int* chain_nsrc, int max_src, const double* src_params,
                                            int max_prop)
{
    if (!plan || !orbit || ngroup <= 0 || nchain <= 0 || max_prop < 0 || !group_qmin || !group_W || !data ||
        !invC || !chain_group || !chain_nsrc || !src_params || max_src < 1 || max_src > kMaxSrcLimit)
    {
        set_error("fd_batch_create: bad arguments (max_src %d: 1..%d)", max_src, kMaxSrcLimit);
        return GLASS_HIP_EINVAL;
    }
    int fmt = parse_format(format);
    if (fmt < 0)
    {
        set_error("Unsupported data format %s", format ? format : "(null)");
        return GLASS_HIP_EINVAL;
    }
    std::vector<long long> goff(ngroup + 1, 0), roff(nchain + 1, 0);
    int Wmax = 0;
    for (int g = 0; g < ngroup; g++)
    {
        // ucb_bandwidth allows 2 * 2^floor(log2(W/2)) bins: at most the
        // 4096-bin template slots (and positions table) for W < 8192
        if (group_W[g] < 32 || group_W[g] >= 2 * kPosN)
        {
            set_error("fd_batch_create: window %d has W=%d outside 32..%d", g, group_W[g], 2 * kPosN - 1);
            return GLASS_HIP_EINVAL;
        }
        goff[g + 1] = goff[g] + group_W[g];
        if (group_W[g] > Wmax) Wmax = group_W[g];
    }
    std::vector<int> ic, is, islot;
    for (int c = 0; c < nchain; c++)
    {
        int g = chain_group[c];
        if (g < 0 || g >= ngroup || chain_nsrc[c] < 0 || chain_nsrc[c] > max_src)
        {
            set_error("fd_batch_create: chain %d has group %d / nsrc %d", c, g, chain_nsrc[c]);
            return GLASS_HIP_EINVAL;
        }
        roff[c + 1] = roff[c] + group_W[g];
        for (int s = 0; s < chain_nsrc[c]; s++)
        {
            ic.push_back(c);
            is.push_back(s);
            islot.push_back(c * max_src + s);
        }
    }
    ThreadCtx* cx;
    int rc;
    if ((rc = ctx_get(&cx))) return rc;
    glass_hip_fd_batch* P = new glass_hip_fd_batch();
    P->orbit = *orbit;
    P->fmt = fmt;
    P->T = T;
    P->t0 = t0;
    P->G = ngroup;
    P->C = nchain;
    P->maxprop = max_prop;
    P->ninit = (int)ic.size();
    P->Wmax = Wmax;
    P->cls_cap = std::max(max_prop, std::max(P->ninit, nchain));
    GH_CHECK(hipGetDevice(&P->device));
    // template slot capacity per window: the largest BW the window allows,
    // 2 * 2^floor(log2(W/2)) <= W (ucb_bandwidth clamps to it), at most 4096
    std::vector<int> gcap(ngroup);
    P->B.bwcap = 32;
    for (int g = 0; g < ngroup; g++)
    {
        int cp = 32;
        while (cp * 2 <= group_W[g] && cp < kPosN) cp *= 2;
        gcap[g] = cp;
        P->B.bwcap = std::max(P->B.bwcap, cp);
    }
    std::vector<long long> toff(nchain + 1, 0);
    for (int c = 0; c < nchain; c++)
        toff[c + 1] = toff[c] + (long long)chain_nsrc[c] * 3 * gcap[chain_group[c]];
    P->tmpl_bytes = sizeof(double2) * (size_t)toff[nchain];
    P->maxsrc = max_src;
    P->B.maxsrc = max_src;
    auto fail = [&](int r) {
        glass_hip_fd_batch_destroy(P);
        return r;
    };
    int *g_qmin, *g_W, *c_group, *c_nsrc, *g_cap;
    long long *g_off, *c_roff, *c_toff;
    double2 *ddata, *resid, *tmpl;
    double *dinvC, *logl, *sparams;
    int *s_bw, *s_imin;
    const size_t nslot = (size_t)nchain * max_src;
    if ((rc = dalloc(P, &g_qmin, ngroup)) || (rc = dalloc(P, &g_W, ngroup)) || (rc = dalloc(P, &g_off, ngroup)) ||
        (rc = dalloc(P, &ddata, 3 * (size_t)goff[ngroup])) || (rc = dalloc(P, &dinvC, 6 * (size_t)goff[ngroup])) ||
        (rc = dalloc(P, &c_group, nchain)) || (rc = dalloc(P, &c_nsrc, nchain)) || (rc = dalloc(P, &c_roff, nchain)) ||
        (rc = dalloc(P, &resid, 3 * (size_t)roff[nchain])) || (rc = dalloc(P, &logl, nchain)) ||
        (rc = dalloc(P, &sparams, nslot * 8)) || (rc = dalloc(P, &s_bw, nslot)) || (rc = dalloc(P, &s_imin, nslot)) ||
        (rc = dalloc(P, &tmpl, (size_t)toff[nchain])) || (rc = dalloc(P, &g_cap, ngroup)) ||
        (rc = dalloc(P, &c_toff, nchain)) ||
        (rc = dalloc(P, &P->p_chain, P->cls_cap)) ||
        (rc = dalloc(P, &P->p_src, P->cls_cap)) || (rc = dalloc(P, &P->p_params, (size_t)P->cls_cap * 8)) ||
        (rc = dalloc(P, &P->p_logl, P->cls_cap)) || (rc = dalloc(P, &P->p_snr, P->cls_cap)) ||
        (rc = dalloc(P, &P->p_bw, P->cls_cap)) || (rc = dalloc(P, &P->p_imin, P->cls_cap)) ||
        (rc = dalloc(P, &P->cls, kNClass + 1)) || (rc = dalloc(P, &P->cls_list, (size_t)kNClass * P->cls_cap)) ||
        (rc = dalloc(P, &P->B.pc, P->cls_cap)) ||
        (rc = dalloc(P, &P->i_chain, P->ninit)) || (rc = dalloc(P, &P->i_src, P->ninit)) ||
        (rc = dalloc(P, &P->i_slot, P->ninit)) || (rc = dalloc(P, &P->i_params, (size_t)P->ninit * 8)) ||
        (rc = dalloc(P, &P->D.pidx, P->cls_cap)) || (rc = dalloc(P, &P->D.imin, P->cls_cap)) ||
        (rc = dalloc(P, &P->D.W, P->cls_cap)) || (rc = dalloc(P, &P->D.tcap, P->cls_cap)) ||
        (rc = dalloc(P, &P->D.roff, P->cls_cap)) || (rc = dalloc(P, &P->D.ioff, P->cls_cap)) ||
        (rc = dalloc(P, &P->D.toff, P->cls_cap)) || (rc = dalloc(P, &P->D.cnt, kNDClass)) ||
        (rc = dalloc(P, &P->D.list, (size_t)kNDClass * P->cls_cap)))
        return fail(rc);
    auto up = [&](void* d, const void* h, size_t b) { return hipMemcpy(d, h, b, hipMemcpyHostToDevice); };
    if (up(g_qmin, group_qmin, sizeof(int) * ngroup) || up(g_W, group_W, sizeof(int) * ngroup) ||
        up(g_off, goff.data(), sizeof(long long) * ngroup) || up(ddata, data, 48 * (size_t)goff[ngroup]) ||
        up(dinvC, invC, 48 * (size_t)goff[ngroup]) || up(c_group, chain_group, sizeof(int) * nchain) ||
        up(c_nsrc, chain_nsrc, sizeof(int) * nchain) || up(c_roff, roff.data(), sizeof(long long) * nchain) ||
        up(sparams, src_params, sizeof(double) * nslot * 8) || up(g_cap, gcap.data(), sizeof(int) * ngroup) ||
        up(c_toff, toff.data(), sizeof(long long) * nchain) ||
        (P->ninit && (up(P->i_chain, ic.data(), sizeof(int) * P->ninit) ||
                      up(P->i_src, is.data(), sizeof(int) * P->ninit) ||
                      up(P->i_slot, islot.data(), sizeof(int) * P->ninit))))
    {
        set_error("fd_batch_create: upload failed");
        return fail(GLASS_HIP_ERUNTIME);
    }
    const double *pos, *pos_soa;
    const double2* tw;
    static_assert(kPosN == kPosSoaMax, "one position grid for both tables");
    if ((rc = positions(*orbit, T, t0, kPosN, &pos)) || (rc = positions_soa(*orbit, T, t0, &pos_soa)) ||
        (rc = twiddles(&tw)))
        return fail(rc);
    BatchDev& B = P->B;
    B.g_qmin = g_qmin;
    B.g_W = g_W;
    B.g_off = g_off;
    B.data = ddata;
    B.invC = dinvC;
    B.c_group = c_group;
    B.c_nsrc = c_nsrc;
    B.c_roff = c_roff;
    B.resid = resid;
    B.logl = logl;
    B.sparams = sparams;
    B.s_bw = s_bw;
    B.s_imin = s_imin;
    B.tmpl = tmpl;
    B.g_cap = g_cap;
    B.c_toff = c_toff;
    B.T = T;
    B.t0 = t0;
    B.L = orbit->L;
    B.fstar = orbit->fstar;
    B.pos = pos;
    B.pos_soa = pos_soa;
    B.tw = tw;
    B.fmt = fmt;
    if ((rc = setup_grids(P))) return fail(rc);
    for (int j = 0; j < 3; j++)
        if (hipStreamCreateWithFlags(&P->side[j], hipStreamNonBlocking) != hipSuccess ||
            hipEventCreateWithFlags(&P->evj[j], hipEventDisableTiming) != hipSuccess)
        {
            set_error("fd_batch_create: side streams");
            return fail(GLASS_HIP_ERUNTIME);
        }
    if (hipEventCreateWithFlags(&P->evf, hipEventDisableTiming) != hipSuccess)
    {
        set_error("fd_batch_create: fork event");
        return fail(GLASS_HIP_ERUNTIME);
    }
    if (P->B.bwcap >= 2048 && (rc = dalloc(P, &P->d2048, (size_t)P->ncu * kDeriv2048Scratch))) return fail(rc);
    if (P->B.bwcap >= 4096 && (rc = dalloc(P, &P->d4096, (size_t)P->ncu * kDeriv4096Scratch))) return fail(rc);
    if ((rc = rebuild(P, cx->stream)) || (rc = check_err(P, cx->stream))) return fail(rc);
    *plan = P;
    return GLASS_HIP_OK;
}

extern "C" int glass_hip_fd_batch_destroy(glass_hip_fd_batch* P)
{
    if (!P) return GLASS_HIP_OK;
    (void)hipSetDevice(P->device);
    if (P->gexec) (void)hipGraphExecDestroy(P->gexec);
    for (int k = 0; k < kNClass; k++)
        for (int j = 0; j < 2; j++)
            if (P->ev[k][j]) (void)hipEventDestroy(P->ev[k][j]);
    if (P->graph) (void)hipGraphDestroy(P->graph);
    if (P->sgexec) (void)hipGraphExecDestroy(P->sgexec);
    if (P->sgraph) (void)hipGraphDestroy(P->sgraph);
    (void)hipDeviceSynchronize();
    for (int j = 0; j < 3; j++)
    {
        if (P->side[j]) (void)hipStreamDestroy(P->side[j]);
        if (P->evj[j]) (void)hipEventDestroy(P->evj[j]);
    }
    if (P->evf) (void)hipEventDestroy(P->evf);
    for (void* p : P->allocs) (void)hipFree(p);
    delete P;
    return GLASS_HIP_OK;
}

extern "C" int glass_hip_fd_batch_buffers(glass_hip_fd_batch* P, int** d_chain, int** d_src, double** d_params,
                                          double** d_logL, double** d_snr)
{
    if (!P)
    {
        set_error("fd_batch_buffers: null plan");
        return GLASS_HIP_EINVAL;
    }
    if (d_chain) *d_chain = P->p_chain;
    if (d_src) *d_src = P->p_src;
    if (d_params) *d_params = P->p_params;
    if (d_logL) *d_logL = P->p_logl;
    if (d_snr) *d_snr = P->p_snr;
    return GLASS_HIP_OK;
}

extern "C" int glass_hip_fd_batch_eval(glass_hip_fd_batch* P, int nprop, const int* d_chain, const int* d_src,
                                       const double* d_params, double* d_logL, double* d_snr, void* stream)
{
    if (!P || nprop < 0 || nprop > P->maxprop || (nprop && (!d_chain || !d_src || !d_params || !d_logL || !d_snr)))
    {
        set_error("fd_batch_eval: bad arguments (nprop %d, max %d)", nprop, P ? P->maxprop : 0);
        return GLASS_HIP_EINVAL;
    }
    GH_CHECK(hipSetDevice(P->device));
    Items it{d_chain, d_src, d_params, nullptr};
    return launch_items(P, nprop, it, 0, d_logL, d_snr, (hipStream_t)stream);
}

extern "C" int glass_hip_fd_batch_eval_graph(glass_hip_fd_batch* P, int nprop, const int* d_chain, const int* d_src,
                                             const double* d_params, double* d_logL, double* d_snr, void* stream)
{
    if (!P)
    {
        set_error("fd_batch_eval_graph: null plan");
        return GLASS_HIP_EINVAL;
    }
    GH_CHECK(hipSetDevice(P->device));
    hipStream_t s = (hipStream_t)stream;
    const void* key[5] = {d_chain, d_src, d_params, d_logL, d_snr};
    bool same = P->gexec && P->g_nprop == nprop;
    for (int i = 0; i < 5 && same; i++) same = key[i] == P->g_key[i];
    if (!same)
    {
        if (int rc = drop_graph(P)) return rc;
        // capture on a private stream so the caller's stream may be the
        // legacy default stream
        hipStream_t cs;
        GH_CHECK(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking));
        GH_CHECK(hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal));
        int rc = glass_hip_fd_batch_eval(P, nprop, d_chain, d_src, d_params, d_logL, d_snr, cs);
        hipGraph_t g = nullptr;
        hipError_t ec = hipStreamEndCapture(cs, &g);
        (void)hipStreamDestroy(cs);
        if (rc) return rc;
        if (ec != hipSuccess)
        {
            set_error("graph capture failed: %s", hipGetErrorString(ec));
            return GLASS_HIP_ERUNTIME;
        }
        P->graph = g;
        GH_CHECK(hipGraphInstantiate(&P->gexec, g, nullptr, nullptr, 0));
        P->g_nprop = nprop;
        for (int i = 0; i < 5; i++) P->g_key[i] = key[i];
    }
    GH_CHECK(hipGraphLaunch(P->gexec, s));
    return GLASS_HIP_OK;
}

extern "C" int glass_hip_fd_batch_accept(glass_hip_fd_batch* P, int nacc, const int* d_idx, const int* d_chain,
                                         const int* d_src, const double* d_params, const double* d_logL, void* stream)
{
    if (!P || nacc < 0 || nacc > P->C || (nacc && (!d_idx || !d_chain || !d_src || !d_params || !d_logL)))
    {
        set_error("fd_batch_accept: bad arguments");
        return GLASS_HIP_EINVAL;
    }
    if (nacc == 0) return GLASS_HIP_OK;
    GH_CHECK(hipSetDevice(P->device));
    hipStream_t s = (hipStream_t)stream;
    Items it{d_chain, d_src, d_params, d_idx};
    int rc;
    // new templates into the scratch slots 0..nacc (item a -> slot a)
    accept_add_old_kernel<<<nacc, 256, 0, s>>>(nacc, it, P->B);
    if ((rc = launch_items(P, nacc, it, 1, nullptr, nullptr, s))) return rc;
    accept_sub_new_kernel<<<nacc, 256, 0, s>>>(nacc, it, P->B, P->p_bw, P->p_imin, d_logL);
    GH_CHECK(hipGetLastError());
    return GLASS_HIP_OK;
}

// Accept with templates staged by the eval (Items::stg of that eval): per
// accepted item, residual += h_old over the old support, the staged
// template into the source's slot, residual -= h_new over the new support
// (per bin r + h_old - h_new, as the three-pass accept), new parameters.
// bw_e / imin_e / logl_new are indexed by the eval's item (item_q).  At most
// one accepted item per chain.  One workgroup per accepted item.
__global__ __launch_bounds__(256) void accept_staged_kernel(int nacc, Items it, BatchDev B,
                                                            const double2* __restrict__ stg,
                                                            const int* __restrict__ bw_e,
                                                            const int* __restrict__ imin_e,
                                                            const double* __restrict__ logl_new)
{
    const int a = blockIdx.x;
    if (a >= nacc || (it.dn && a >= *it.dn)) return;
    const int q = item_q(it, a);
    const int c = it.chain[q], s = it.src[q], so = c * B.maxsrc + s;
    const int g = B.c_group[c], W = B.g_W[g], cap = B.g_cap[g];
    const int io = B.s_imin[so], bo = B.s_bw[so];
    const int in = imin_e[q], bn = bw_e[q];
    // an item the alignment left out of every class (BW above the window's
    // slot) has no staged template: never written into the slot
    if (bn > cap || bn > B.bwcap) return;
    double2* R = B.resid + (size_t)B.c_roff[c] * 3;
    double2* h = src_tmpl(B, c, s);
    const double2* hs = stg + (size_t)q * 3 * B.bwcap;
    // one pass over the union of both supports: (r + h_old) - h_new per bin
    const int j0 = max(0, min(io, in)), j1 = min(W, max(io + bo, in + bn));
    for (int j = j0 + threadIdx.x; j < j1; j += blockDim.x)
    {
        const bool o = j >= io && j < io + bo, nw = j >= in && j < in + bn;
        if (!o && !nw) continue;
#pragma unroll
        for (int ch = 0; ch < 3; ch++)
        {
            double2 v = R[(size_t)ch * W + j];
            if (o)
            {
                v.x += h[ch * cap + j - io].x;
                v.y += h[ch * cap + j - io].y;
            }
            if (nw)
            {
                v.x -= hs[ch * bn + j - in].x;
                v.y -= hs[ch * bn + j - in].y;
            }
            R[(size_t)ch * W + j] = v;
        }
    }
    __syncthreads();  // the old template is read before its slot is overwritten
    for (int m = threadIdx.x; m < bn; m += blockDim.x)
#pragma unroll
        for (int ch = 0; ch < 3; ch++) h[ch * cap + m] = hs[ch * bn + m];
    if (threadIdx.x < 8) B.sparams[(size_t)so * 8 + threadIdx.x] = it.params[(size_t)q * 8 + threadIdx.x];
    if (threadIdx.x == 0)
    {
        B.s_bw[so] = bn;
        B.s_imin[so] = in;
        B.logl[c] = logl_new[q];
    }
}

// ------------------------------------------- hooks for the device sampler ---

namespace glass {

FdBatchView fd_batch_view(glass_hip_fd_batch* P)
{
    FdBatchView v{};
    v.device = P->device;
    v.C = P->C;
    v.G = P->G;
    v.maxsrc = P->maxsrc;
    v.ninit = P->ninit;
    v.cap = P->cls_cap;
    v.T = P->T;
    v.c_group = P->B.c_group;
    v.c_nsrc = P->B.c_nsrc;
    v.sparams = P->B.sparams;
    v.logl = P->B.logl;
    v.i_chain = P->i_chain;
    v.i_src = P->i_src;
    return v;
}

int fd_batch_eval_dev(glass_hip_fd_batch* P, int n, const int* chain, const int* src, const double* params,
                      double* logl, double* snr, hipStream_t s, double2* stage)
{
    Items it{chain, src, params, nullptr};
    it.stg = stage;
    return launch_items(P, n, it, 0, logl, snr, s);
}

size_t fd_batch_stage_bytes(glass_hip_fd_batch* P, int n) { return sizeof(double2) * (size_t)n * 3 * P->B.bwcap; }

int fd_batch_accept_staged_dev(glass_hip_fd_batch* P, int nmax, const int* d_n, const int* d_idx, const int* chain,
                               const int* src, const double* params, const double* logl, const double2* stage,
                               hipStream_t s)
{
    if (nmax <= 0) return GLASS_HIP_OK;
    if (nmax > P->C || nmax > P->cls_cap || !stage)
    {
        set_error("batch accept: %d items exceed the plan (%d chains, capacity %d) or no staged templates", nmax,
                  P->C, P->cls_cap);
        return GLASS_HIP_EINVAL;
    }
    Items it{chain, src, params, d_idx, d_n};
    accept_staged_kernel<<<nmax, 256, 0, s>>>(nmax, it, P->B, stage, P->p_bw, P->p_imin, logl);
    GH_CHECK(hipGetLastError());
    return GLASS_HIP_OK;
}

}  // namespace glass

extern "C" int glass_hip_fd_batch_refresh(glass_hip_fd_batch* P, void* stream)
{
    if (!P)
    {
        set_error("fd_batch_refresh: null plan");
        return GLASS_HIP_EINVAL;
    }
    GH_CHECK(hipSetDevice(P->device));
    return rebuild(P, (hipStream_t)stream);
}

extern "C" int glass_hip_fd_batch_set_concurrent(glass_hip_fd_batch* P, int on)
{
    if (!P)
    {
        set_error("fd_batch_set_concurrent: null plan");
        return GLASS_HIP_EINVAL;
    }
    if (int rc = drop_graph(P)) return rc;
    P->concurrent = on != 0;
    return GLASS_HIP_OK;
}

extern "C" int glass_hip_fd_batch_profile(glass_hip_fd_batch* P, int on)
{
    if (!P)
    {
        set_error("fd_batch_profile: null plan");
        return GLASS_HIP_EINVAL;
    }
    GH_CHECK(hipSetDevice(P->device));
    if (on && !P->ev[0][0])
        for (int k = 0; k < kNClass; k++)
            for (int j = 0; j < 2; j++) GH_CHECK(hipEventCreate(&P->ev[k][j]));
    P->prof = on != 0;
    for (int k = 0; k < kNClass; k++)
    {
        P->prof_ms[k] = 0;
        P->prof_launches[k] = 0;
        P->prof_count[k] = 0;
    }
    return GLASS_HIP_OK;
}

extern "C" int glass_hip_fd_batch_profile_read(glass_hip_fd_batch* P, double* ms, int* launches, int* count)
{
    if (!P || !ms || !launches || !count)
    {
        set_error("fd_batch_profile_read: bad arguments");
        return GLASS_HIP_EINVAL;
    }
    for (int k = 0; k < kNClass; k++)
    {
        ms[k] = P->prof_ms[k];
        launches[k] = P->prof_launches[k];
        count[k] = P->prof_count[k];
    }
    return GLASS_HIP_OK;
}

extern "C" int glass_hip_fd_batch_state(glass_hip_fd_batch* P, double* logL, double* src_params)
{
    if (!P || !logL)
    {
        set_error("fd_batch_state: bad arguments");
        return GLASS_HIP_EINVAL;
    }
    GH_CHECK(hipSetDevice(P->device));
    GH_CHECK(hipDeviceSynchronize());
    int rc;
    if ((rc = check_err(P, nullptr))) return rc;
    GH_CHECK(hipMemcpy(logL, P->B.logl, sizeof(double) * P->C, hipMemcpyDeviceToHost));
    if (src_params)
        GH_CHECK(hipMemcpy(src_params, P->B.sparams, sizeof(double) * P->C * P->maxsrc * 8, hipMemcpyDeviceToHost));
    return GLASS_HIP_OK;
}

// ------------------------------------------------------------- tiles ---

static int drop_graph(glass_hip_fd_batch* P)
{
    if (P->gexec) GH_CHECK(hipGraphExecDestroy(P->gexec));
    if (P->graph) GH_CHECK(hipGraphDestroy(P->graph));
    if (P->sgexec) GH_CHECK(hipGraphExecDestroy(P->sgexec));
    if (P->sgraph) GH_CHECK(hipGraphDestroy(P->sgraph));
    P->gexec = nullptr;
    P->graph = nullptr;
    P->sgexec = nullptr;
    P->sgraph = nullptr;
    P->g_nprop = -1;
    return GLASS_HIP_OK;
}

extern "C" int glass_hip_fd_batch_tile_candidates(int k, glass_hip_tile* out, int max, int* n)
{
    if (k < 0 || k >= kNClass || !n || (max > 0 && !out))
    {
        set_error("fd_batch_tile_candidates: bad arguments");
        return GLASS_HIP_EINVAL;
    }
    int c = 0;
#define GH_X(kk, N, T, MW)                                          \
    if (k == kk)                                                    \
        for (int m = 0; m < kNGmul; m++, c++)                       \
            if (c < max) out[c] = glass_hip_tile{T, kGmul[m], MW};
    GH_EVAL_TILES(GH_X)
#undef GH_X
    *n = c;
    return GLASS_HIP_OK;
}

extern "C" int glass_hip_fd_batch_get_tiles(glass_hip_fd_batch* P, glass_hip_tile* tiles)
{
    if (!P || !tiles)
    {
        set_error("fd_batch_get_tiles: bad arguments");
        return GLASS_HIP_EINVAL;
    }
    for (int k = 0; k < kNClass; k++) tiles[k] = P->tile[k];
    return GLASS_HIP_OK;
}

extern "C" int glass_hip_fd_batch_set_tiles(glass_hip_fd_batch* P, const glass_hip_tile* tiles)
{
    if (!P || !tiles)
    {
        set_error("fd_batch_set_tiles: bad arguments");
        return GLASS_HIP_EINVAL;
    }
    for (int k = 0; k < kNClass; k++)
        if (!tile_compiled(k, tiles[k].nthr, tiles[k].waves) || tiles[k].gmul < 0 || tiles[k].gmul > 64)
        {
            set_error("fd_batch_set_tiles: class %d (BW %d): no compiled tile with %d threads, %d waves, grid x%d",
                      k, 32 << k, tiles[k].nthr, tiles[k].waves, tiles[k].gmul);
            return GLASS_HIP_EINVAL;
        }
    GH_CHECK(hipSetDevice(P->device));
    if (int rc = drop_graph(P)) return rc;
    for (int k = 0; k < kNClass; k++) P->tile[k] = tiles[k];
    return GLASS_HIP_OK;
}

extern "C" int glass_hip_fd_batch_autotune(glass_hip_fd_batch* P, int nprop, const int* d_chain, const int* d_src,
                                           const double* d_params, int reps, glass_hip_tile* best,
                                           double* best_ms, void* stream)
{
    if (!P || nprop <= 0 || nprop > P->maxprop || !d_chain || !d_src || !d_params || reps < 1)
    {
        set_error("fd_batch_autotune: bad arguments (nprop %d, max %d, reps %d)", nprop, P ? P->maxprop : 0, reps);
        return GLASS_HIP_EINVAL;
    }
    GH_CHECK(hipSetDevice(P->device));
    hipStream_t s = (hipStream_t)stream;
    Items it{d_chain, d_src, d_params, nullptr};
    // class lists of these proposals, as an eval builds them
    zero_int_kernel<<<1, 64, 0, s>>>(P->cls, kNClass + 1);
    align_kernel<<<(nprop + 255) / 256, 256, 0, s>>>(nprop, it, P->B, P->p_bw, P->p_imin, P->cls, P->cls_list,
                                                     P->cls_cap, P->cls + kNClass);
    GH_CHECK(hipGetLastError());
    int cnt[kNClass + 1];
    GH_CHECK(hipMemcpyAsync(cnt, P->cls, sizeof(cnt), hipMemcpyDeviceToHost, s));
    GH_CHECK(hipStreamSynchronize(s));
    if (cnt[kNClass])
    {
        set_error("fd_batch_autotune: %d proposals with BW outside 32..%d", cnt[kNClass], P->B.bwcap);
        return GLASS_HIP_EINVAL;
    }
    if (int rc = drop_graph(P)) return rc;
    hipEvent_t e0, e1;
    GH_CHECK(hipEventCreate(&e0));
    GH_CHECK(hipEventCreate(&e1));
    int rc = GLASS_HIP_OK;
    for (int k = 0; k < kNClass && !rc; k++)
    {
        double bms = 0.0;
        if (cnt[k] > 0)
        {
            glass_hip_tile cand[64];
            int nc = 0;
            glass_hip_fd_batch_tile_candidates(k, cand, 64, &nc);
            const int* list = P->cls_list + (size_t)k * P->cls_cap;
            const int* cntp = P->cls + k;
            bms = 1e300;
            for (int c = 0; c < nc && c < 64 && !rc; c++)
            {
                // one untimed launch, then reps timed ones; results go to
                // the plan's own logL/snr buffers
                if ((rc = launch_eval(P, k, cand[c], nprop, list, cntp, it, P->p_logl, P->p_snr, s))) break;
                if (hipEventRecord(e0, s) != hipSuccess) rc = GLASS_HIP_ERUNTIME;
                for (int r = 0; r < reps && !rc; r++)
                    rc = launch_eval(P, k, cand[c], nprop, list, cntp, it, P->p_logl, P->p_snr, s);
                float ms = 0.f;
                if (!rc && (hipEventRecord(e1, s) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
                            hipEventElapsedTime(&ms, e0, e1) != hipSuccess))
                {
                    set_error("fd_batch_autotune: timing failed");
                    rc = GLASS_HIP_ERUNTIME;
                }
                if (!rc && ms / reps < bms)
                {
                    bms = ms / reps;
                    P->tile[k] = cand[c];
                }
            }
        }
        if (best) best[k] = P->tile[k];
        if (best_ms) best_ms[k] = bms;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (!rc && hipGetLastError() != hipSuccess) rc = GLASS_HIP_ERUNTIME;
    return rc;
}

// ------------------------------------------------------- derivatives ---

extern "C" int glass_hip_fd_batch_derivs(glass_hip_fd_batch* P, int n, const int* d_chain, const int* d_src,
                                         double* d_logL, double* d_grad, double* d_fisher, void* stream)
{
    if (!P || n < 0 || n > P->cls_cap || (n && (!d_chain || !d_src || !d_logL || !d_grad || !d_fisher)))
    {
        set_error("fd_batch_derivs: bad arguments (n %d, capacity %d)", n, P ? P->cls_cap : 0);
        return GLASS_HIP_EINVAL;
    }
    if (n == 0) return GLASS_HIP_OK;
    GH_CHECK(hipSetDevice(P->device));
    hipStream_t s = (hipStream_t)stream;
    zero_int_kernel<<<1, 64, 0, s>>>(P->D.cnt, kNDClass);
    derivs_prep_kernel<<<(n + 255) / 256, 256, 0, s>>>(n, d_chain, d_src, P->C, P->B, P->D, P->cls_cap, d_logL,
                                                       d_grad, d_fisher);
    GH_CHECK(hipGetLastError());
    const BatchDev& B = P->B;
    DerivArgs A{};
    A.params = B.sparams;
    A.pidx = P->D.pidx;
    A.imin = P->D.imin;
    A.iW = P->D.W;
    A.resid = B.resid;
    A.iroff = P->D.roff;
    A.invC = B.invC;
    A.iioff = P->D.ioff;
    A.tadd = B.tmpl;
    A.taddoff = P->D.toff;
    A.taddcap = P->D.tcap;
    A.mode = 0;
    A.T = B.T;
    A.t0 = B.t0;
    A.L = B.L;
    A.fstar = B.fstar;
    A.pos = B.pos;
    A.tw = B.tw;
    A.fmt = B.fmt;
    A.outL = d_logL;
    A.grad = d_grad;
    A.fisher = d_fisher;
    // the classes on the plan's side streams as in launch_items (k % 4)
    const bool par = P->concurrent;
    if (par)
    {
        GH_CHECK(hipEventRecord(P->evf, s));
        for (int j = 0; j < 3; j++) GH_CHECK(hipStreamWaitEvent(P->side[j], P->evf, 0));
    }
    int lrc = GLASS_HIP_OK;
    for (int k = kNDClass - 1; k >= 0 && !lrc; k--)
    {
        const hipStream_t sk = par && (k & 3) ? P->side[(k & 3) - 1] : s;
        A.list = P->D.list + (size_t)k * P->cls_cap;
        A.count = P->D.cnt + k;
        if (32 << k == 2048)
        {
            // BW 2048 only occurs when some window allows it (then d2048 exists)
            if (P->d2048) lrc = launch_derivs_2048(A, std::min(n, P->ncu), P->d2048, sk);
        }
        else if (32 << k == 4096)
        {
            if (P->d4096) lrc = launch_derivs_4096(A, std::min(n, P->ncu), P->d4096, sk);
        }
        else
            lrc = launch_derivs(32 << k, A, n, sk);
    }
    if (par)
        for (int j = 0; j < 3; j++)
        {
            GH_CHECK(hipEventRecord(P->evj[j], P->side[j]));
            GH_CHECK(hipStreamWaitEvent(s, P->evj[j], 0));
        }
    return lrc;
}

extern "C" int glass_hip_fd_batch_sweep_graph(glass_hip_fd_batch* P, int nprop, const int* d_chain, const int* d_src,
                                              const double* d_params, double* d_logL, double* d_snr, int nder,
                                              const int* d_dchain, const int* d_dsrc, double* d_dlogL,
                                              double* d_grad, double* d_fisher, void* stream)
{
    if (!P)
    {
        set_error("fd_batch_sweep_graph: null plan");
        return GLASS_HIP_EINVAL;
    }
    GH_CHECK(hipSetDevice(P->device));
    const long long key[12] = {nprop, (long long)d_chain, (long long)d_src, (long long)d_params,
                               (long long)d_logL, (long long)d_snr, nder, (long long)d_dchain,
                               (long long)d_dsrc, (long long)d_dlogL, (long long)d_grad, (long long)d_fisher};
    bool same = P->sgexec != nullptr;
    for (int i = 0; i < 12 && same; i++) same = key[i] == P->s_key[i];
    if (!same)
    {
        if (P->sgexec) GH_CHECK(hipGraphExecDestroy(P->sgexec));
        if (P->sgraph) GH_CHECK(hipGraphDestroy(P->sgraph));
        P->sgexec = nullptr;
        P->sgraph = nullptr;
        hipStream_t cs;
        GH_CHECK(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking));
        GH_CHECK(hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal));
        int rc = glass_hip_fd_batch_eval(P, nprop, d_chain, d_src, d_params, d_logL, d_snr, cs);
        if (!rc) rc = glass_hip_fd_batch_derivs(P, nder, d_dchain, d_dsrc, d_dlogL, d_grad, d_fisher, cs);
        hipGraph_t g = nullptr;
        hipError_t ec = hipStreamEndCapture(cs, &g);
        (void)hipStreamDestroy(cs);
        if (rc)
        {
            if (g) (void)hipGraphDestroy(g);
            return rc;
        }
        if (ec != hipSuccess)
        {
            set_error("sweep graph capture failed: %s", hipGetErrorString(ec));
            return GLASS_HIP_ERUNTIME;
        }
        P->sgraph = g;
        GH_CHECK(hipGraphInstantiate(&P->sgexec, g, nullptr, nullptr, 0));
        for (int i = 0; i < 12; i++) P->s_key[i] = key[i];
    }
    GH_CHECK(hipGraphLaunch(P->sgexec, (hipStream_t)stream));
    return GLASS_HIP_OK;
}

extern "C" int glass_hip_fd_batch_memory(glass_hip_fd_batch* P, long long* template_bytes, long long* total_bytes)
{
    if (!P)
    {
        set_error("fd_batch_memory: null plan");
        return GLASS_HIP_EINVAL;
    }
    if (template_bytes) *template_bytes = (long long)P->tmpl_bytes;
    if (total_bytes) *total_bytes = (long long)P->alloc_bytes;
    return GLASS_HIP_OK;
}
