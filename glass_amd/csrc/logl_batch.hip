// logl_batch.hip -- many independent Fourier-domain Gaussian likelihoods in
// one launch (include/glass_hip_batch.h, glass_hip_logl_batch_*).
//
// Item e is one call of gaussian_log_likelihood (utils/src/glass_model.c:
// 726-774) with Nchannel 3 over a window of W_e bins:
//   logL_e = -1/2 chi^2,  chi^2 = XX + YY + ZZ + 2 (XY + XZ + YZ),
// each term a fourier_nwip (utils/src/glass_math.c:362-380, factor 2) of
// the residual r = d - m.  Per bin the kernel reads the 3 data and 3 model
// channels (complex, 96 B) and the 6 distinct inverse-covariance planes
// (48 B): 144 B/bin, 36 FP64 flops/bin -- an HBM stream.
//
// Windows are uneven (512..4096 bins in a tempered sampler), so the plan
// cuts every item into chunks of at most kChunk bins, one workgroup each;
// partials land in plan memory and a second kernel sums each item's chunks
// in chunk order.  Results are deterministic run to run.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <mutex>
#include <vector>

#include "../../include/glass_hip_batch.h"
#include "reduce.h"
#include "runtime.h"

namespace glass {

constexpr int kLThr = 256;
constexpr int kChunk = 2048;                 // bins per workgroup
constexpr int kPerThr = kChunk / kLThr;      // bins per thread

struct LoglItem
{
    long long doff, moff, ioff;  // bins before the item in data / model / invC
    int W, c0, c1, pad;          // window bins, chunk range [c0, c1)
};

// Chunk w: item, first bin, bin count.
struct LoglChunk
{
    int item, b0, nb, pad;
};

__device__ inline double bin_chi2(const double2* __restrict__ D, const double2* __restrict__ M,
                                  const double* __restrict__ I, size_t W, int n)
{
    double2 x = D[n], y = D[W + n], z = D[2 * W + n];
    const double2 a = M[n], b = M[W + n], c = M[2 * W + n];
    const double ixx = I[n], iyy = I[W + n], izz = I[2 * W + n];
    const double ixy = I[3 * W + n], ixz = I[4 * W + n], iyz = I[5 * W + n];
    x.x -= a.x; x.y -= a.y;
    y.x -= b.x; y.y -= b.y;
    z.x -= c.x; z.y -= c.y;
    const double diag = (x.x * x.x + x.y * x.y) * ixx + (y.x * y.x + y.y * y.y) * iyy + (z.x * z.x + z.y * z.y) * izz;
    const double off = (x.x * y.x + x.y * y.y) * ixy + (x.x * z.x + x.y * z.y) * ixz + (y.x * z.x + y.y * z.y) * iyz;
    return diag + 2.0 * off;
}

__global__ __launch_bounds__(kLThr) void logl_chunk_kernel(const LoglChunk* __restrict__ chunks,
                                                           const LoglItem* __restrict__ items,
                                                           const double2* __restrict__ data,
                                                           const double2* __restrict__ model,
                                                           const double* __restrict__ invC, double* __restrict__ part)
{
    const LoglChunk ck = chunks[blockIdx.x];
    const LoglItem it = items[ck.item];
    const size_t W = (size_t)it.W;
    const double2* D = data + 3 * it.doff + ck.b0;
    const double2* M = model + 3 * it.moff + ck.b0;
    const double* I = invC + 6 * it.ioff + ck.b0;
    double acc = 0.0;
#pragma unroll
    for (int k = 0; k < kPerThr; k++)
    {
        const int n = threadIdx.x + k * kLThr;
        if (n < ck.nb) acc += bin_chi2(D, M, I, W, n);
    }
    acc = block_sum<kLThr>(acc);
    if (threadIdx.x == 0) part[blockIdx.x] = acc;
}

__global__ void logl_finish_kernel(int n, const LoglItem* __restrict__ items, const double* __restrict__ part,
                                   double* __restrict__ logl)
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n) return;
    const LoglItem it = items[e];
    double s = 0.0;
    for (int c = it.c0; c < it.c1; c++) s += part[c];
    logl[e] = -s;  // -chi^2/2 with chi^2 = 2 * sum
}

// ---- allocation-free form: a fixed grid of cpi chunks per item --------
//
// Item e owns chunks e*cpi .. e*cpi + cpi - 1 of kChunk bins each (cpi =
// ceil(max_W / kChunk)); chunk k covers bins [k kChunk, (k+1) kChunk) of the
// item's window, as the plan's chunk table does, so the partial sums and
// their order are the plan's.  The item table is read from the caller's
// device arrays; the partials live in the file-scope array below.

constexpr long long kNoAllocParts = 1 << 20;  // partials: 8 MiB per device
__device__ __align__(256) double scratch[kNoAllocParts];

__global__ __launch_bounds__(kLThr) void logl_chunk_grid_kernel(int cpi, int max_W, const int* __restrict__ Wv,
                                                                const long long* __restrict__ doff,
                                                                const long long* __restrict__ moff,
                                                                const long long* __restrict__ ioff,
                                                                const double2* __restrict__ data,
                                                                const double2* __restrict__ model,
                                                                const double* __restrict__ invC,
                                                                double* __restrict__ part)
{
    const int e = blockIdx.x / cpi, k = blockIdx.x - e * cpi;
    const int Wi = Wv[e];
    const int b0 = k * kChunk;
    const int nb = (Wi > 0 && Wi <= max_W) ? min(kChunk, Wi - b0) : 0;
    if (nb <= 0)
    {
        if (threadIdx.x == 0) part[blockIdx.x] = 0.0;
        return;
    }
    const size_t W = (size_t)Wi;
    const double2* D = data + 3 * doff[e] + b0;
    const double2* M = model + 3 * moff[e] + b0;
    const double* I = invC + 6 * ioff[e] + b0;
    double acc = 0.0;
#pragma unroll
    for (int j = 0; j < kPerThr; j++)
    {
        const int n = threadIdx.x + j * kLThr;
        if (n < nb) acc += bin_chi2(D, M, I, W, n);
    }
    acc = block_sum<kLThr>(acc);
    if (threadIdx.x == 0) part[blockIdx.x] = acc;
}

__global__ void logl_finish_grid_kernel(int n, int cpi, int max_W, const int* __restrict__ Wv,
                                        const double* __restrict__ part, double* __restrict__ logl)
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n) return;
    const int Wi = Wv[e];
    if (Wi <= 0 || Wi > max_W)
    {
        logl[e] = NAN;  // a window the call's max_W does not cover
        return;
    }
    const int nck = (Wi + kChunk - 1) / kChunk;
    double s = 0.0;
    for (int c = 0; c < nck; c++) s += part[(size_t)e * cpi + c];
    logl[e] = -s;
}

}  // namespace glass

using namespace glass;

struct glass_hip_logl_batch
{
    int n = 0, nchunk = 0;
    int device = 0;
    LoglItem* items = nullptr;
    LoglChunk* chunks = nullptr;
    double* part = nullptr;
};

static int upload(glass_hip_logl_batch* P, const std::vector<LoglItem>& items, const std::vector<LoglChunk>& chunks)
{
    GH_CHECK(hipGetDevice(&P->device));
    GH_CHECK(hipMalloc((void**)&P->items, sizeof(LoglItem) * P->n));
    GH_CHECK(hipMalloc((void**)&P->chunks, sizeof(LoglChunk) * P->nchunk));
    GH_CHECK(hipMalloc((void**)&P->part, sizeof(double) * P->nchunk));
    GH_CHECK(hipMemcpy(P->items, items.data(), sizeof(LoglItem) * P->n, hipMemcpyHostToDevice));
    GH_CHECK(hipMemcpy(P->chunks, chunks.data(), sizeof(LoglChunk) * P->nchunk, hipMemcpyHostToDevice));
    return GLASS_HIP_OK;
}

extern "C" int glass_hip_logl_batch_create(glass_hip_logl_batch** plan, int n, const int* W,
                                           const long long* data_off, const long long* model_off,
                                           const long long* invC_off)
{
    if (!plan || n <= 0 || !W || !data_off || !model_off || !invC_off)
    {
        set_error("logl_batch_create: bad arguments");
        return GLASS_HIP_EINVAL;
    }
    *plan = nullptr;
    std::vector<LoglItem> items(n);
    std::vector<LoglChunk> chunks;
    for (int e = 0; e < n; e++)
    {
        if (W[e] <= 0 || data_off[e] < 0 || model_off[e] < 0 || invC_off[e] < 0)
        {
            set_error("logl_batch_create: item %d: W %d, offsets %lld %lld %lld", e, W[e], data_off[e],
                      model_off[e], invC_off[e]);
            return GLASS_HIP_EINVAL;
        }
        LoglItem& it = items[e];
        it.doff = data_off[e];
        it.moff = model_off[e];
        it.ioff = invC_off[e];
        it.W = W[e];
        it.c0 = (int)chunks.size();
        for (int b0 = 0; b0 < W[e]; b0 += kChunk) chunks.push_back({e, b0, std::min(kChunk, W[e] - b0), 0});
        it.c1 = (int)chunks.size();
        it.pad = 0;
    }
    auto* P = new glass_hip_logl_batch;
    P->n = n;
    P->nchunk = (int)chunks.size();
    if (int rc = upload(P, items, chunks))
    {
        glass_hip_logl_batch_destroy(P);
        return rc;
    }
    *plan = P;
    return GLASS_HIP_OK;
}

extern "C" int glass_hip_logl_batch_destroy(glass_hip_logl_batch* P)
{
    if (!P) return GLASS_HIP_OK;
    (void)hipFree(P->items);
    (void)hipFree(P->chunks);
    (void)hipFree(P->part);
    delete P;
    return GLASS_HIP_OK;
}

extern "C" int glass_hip_logl_batch_eval(glass_hip_logl_batch* P, const double* d_data, const double* d_model,
                                         const double* d_invC, double* d_logL, void* stream)
{
    if (!P || !d_data || !d_model || !d_invC || !d_logL)
    {
        set_error("logl_batch_eval: bad arguments");
        return GLASS_HIP_EINVAL;
    }
    hipStream_t s = (hipStream_t)stream;
    logl_chunk_kernel<<<P->nchunk, kLThr, 0, s>>>(P->chunks, P->items, (const double2*)d_data,
                                                  (const double2*)d_model, d_invC, P->part);
    logl_finish_kernel<<<(P->n + 255) / 256, 256, 0, s>>>(P->n, P->items, P->part, d_logL);
    GH_CHECK(hipGetLastError());
    return GLASS_HIP_OK;
}

// Device address of `scratch` on each device, looked up once per device
// with that device current (host code never passes the symbol itself).
static std::atomic<double*> g_scratch_ptr[64];
static std::mutex g_scratch_mu;

static int scratch_ptr(double** out)
{
    int dev = 0;
    GH_CHECK(hipGetDevice(&dev));
    if (dev < 0 || dev >= 64)
    {
        set_error("logl_batch_eval_noalloc: device %d beyond the 64 this library tracks", dev);
        return GLASS_HIP_ENODEV;
    }
    double* p = g_scratch_ptr[dev].load(std::memory_order_acquire);
    if (!p)
    {
        std::lock_guard<std::mutex> lk(g_scratch_mu);
        p = g_scratch_ptr[dev].load(std::memory_order_relaxed);
        if (!p)
        {
            void* q = nullptr;
            GH_CHECK(hipGetSymbolAddress(&q, HIP_SYMBOL(glass::scratch)));
            p = (double*)q;
            g_scratch_ptr[dev].store(p, std::memory_order_release);
        }
    }
    *out = p;
    return GLASS_HIP_OK;
}

extern "C" int glass_hip_logl_batch_eval_noalloc(int n, const int* d_W, const long long* d_data_off,
                                                 const long long* d_model_off, const long long* d_invC_off,
                                                 int max_W, const double* d_data, const double* d_model,
                                                 const double* d_invC, double* d_logL, void* stream)
{
    if (n <= 0 || max_W <= 0 || !d_W || !d_data_off || !d_model_off || !d_invC_off || !d_data || !d_model ||
        !d_invC || !d_logL)
    {
        set_error("logl_batch_eval_noalloc: bad arguments (n %d, max_W %d)", n, max_W);
        return GLASS_HIP_EINVAL;
    }
    const int cpi = (max_W + kChunk - 1) / kChunk;
    if ((long long)n * cpi > kNoAllocParts)
    {
        set_error("logl_batch_eval_noalloc: %d items x %d chunks of %d bins exceed the %lld partials of the "
                  "static scratch", n, cpi, kChunk, kNoAllocParts);
        return GLASS_HIP_EINVAL;
    }
    double* part = nullptr;
    if (int rc = scratch_ptr(&part)) return rc;
    hipStream_t s = (hipStream_t)stream;
    logl_chunk_grid_kernel<<<n * cpi, kLThr, 0, s>>>(cpi, max_W, d_W, d_data_off, d_model_off, d_invC_off,
                                                     (const double2*)d_data, (const double2*)d_model, d_invC, part);
    logl_finish_grid_kernel<<<(n + 255) / 256, 256, 0, s>>>(n, cpi, max_W, d_W, part, d_logL);
    GH_CHECK(hipGetLastError());
    return GLASS_HIP_OK;
}
