// fd_derivs.h -- arguments of the exact-derivative kernels (fd_derivs.hip),
// shared with the batched plan (fd_batch.hip), which launches them over its
// own chains' sources.
#pragma once
#include <hip/hip_runtime.h>

namespace glass {

struct DerivArgs
{
    // item e: params at params + 8 (pidx ? pidx[e] : e); template window
    // offset imin[e] in a window of W bins (iW[e] when iW != NULL); residual
    // at resid + e rstride/2 complex (resid + iroff[e] when iroff != NULL),
    // invC at invC + e istride (invC + iioff[e]); NULL resid = zero residual
    const double* params;
    const int* pidx;
    const int* imin;
    int W;
    const int* iW;
    const double2* resid;
    long long rstride;  // doubles
    const long long* iroff;
    const double* invC;
    long long istride;  // doubles
    const long long* iioff;
    // optional stored template of the item's source ([3][taddcap[e]] complex
    // at tadd + taddoff[e], bin m at window bin imin + m), added back to the
    // residual so that it excludes the source
    const double2* tadd;
    const long long* taddoff;
    const int* taddcap;
    // optional class list: workgroup w takes item list[w] while w < *count
    // (without a list the grid-striding kernels take items w < nitems)
    const int* list;
    const int* count;
    int nitems;
    int mode;           // 0: window-aligned weights; 1: template-local (reference)
    double T, t0, L, fstar;
    const double* pos;
    const double2* tw;
    int fmt;
    double* outL;
    double* grad;
    double* fisher;
};

// deriv_kernel<N> over grid workgroups (N a power of two, 32..1024).
int launch_derivs(int N, const DerivArgs& A, int grid, hipStream_t s);

// The BW 2048 and 4096 kernels: grid workgroups striding over A.list (or
// over items 0..A.nitems-1 without a list), workgroup w with a scratch
// slice of deriv_scratch(N) complex at hs + w slice.
constexpr int kDeriv2048Threads = 512;
constexpr size_t kDeriv2048Scratch = 18 * 2048;
constexpr size_t kDeriv4096Scratch = 24 * 4096;
constexpr size_t deriv_scratch(int N) { return N == 2048 ? kDeriv2048Scratch : (N == 4096 ? kDeriv4096Scratch : 0); }
int launch_derivs_2048(const DerivArgs& A, int grid, double2* hs, hipStream_t s);
int launch_derivs_4096(const DerivArgs& A, int grid, double2* hs, hipStream_t s);

}  // namespace glass
