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
#include <cmath>
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
