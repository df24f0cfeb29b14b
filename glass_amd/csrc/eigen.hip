// eigen.hip -- eigen-decomposition of many symmetric 8x8 matrices (the
// Fisher matrices of the UCB parameters), replacing the LAPACKE_dgeev call
// of matrix_eigenstuff (utils/src/glass_math.c:475-499).
//
// Method: cyclic Jacobi with the symmetric 2x2 Schur rotation (Golub & Van
// Loan, Matrix Computations, section 8.5), rows p < q in order, sweeps until
// the off-diagonal mass falls below eps^2 of the Frobenius norm (at most
// kMaxSweep).  One thread per matrix; the matrix (upper triangle) and the
// rotation product live in LDS with the thread index as the fastest stride,
// so the 64 lanes of a wave touch 64 consecutive words.  Jacobi keeps the
// small eigenvalues of a graded matrix to high relative accuracy, which the
// Fisher matrices (eigenvalues over many decades) need.
//
// Output convention (what draw_from_fisher reads, ucb/src/glass_ucb_proposal.c:
// 356-364): evec[i][j] = component i of eigenvector j, eval[j]; eigenvalues
// ascending; each eigenvector's largest-magnitude component positive.  dgeev
// fixes neither order nor sign for a symmetric matrix, so the reference's
// values agree with these up to that permutation and sign.
//
// Floating-point contraction is off in the kernel so the arithmetic matches
// the C restatement of the same method operation by operation.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>

#include "../../include/glass_hip_sampler.h"
#include "fd_batch_hooks.h"
#include "runtime.h"

namespace glass {

namespace {

constexpr int kEigThr = 64;
constexpr int kMaxSweep = 40;
constexpr int kNT = 36;  // upper-triangle entries of 8x8

__device__ __host__ constexpr int tri(int p, int q) { return p * 8 - p * (p - 1) / 2 + (q - p); }

__global__ __launch_bounds__(kEigThr) void jacobi8_kernel(int n, const double* __restrict__ F,
                                                          double* __restrict__ evec, double* __restrict__ eval,
                                                          int* __restrict__ sweeps)
{
#pragma clang fp contract(off)
    __shared__ double sa[kNT][kEigThr];
    __shared__ double sv[64][kEigThr];
    const int t = threadIdx.x;
    const int m = blockIdx.x * kEigThr + t;
    if (m >= n) return;  // no barriers below: every thread works alone
    const double* f = F + (size_t)m * 64;
    auto A = [&](int p, int q) -> double& { return p <= q ? sa[tri(p, q)][t] : sa[tri(q, p)][t]; };
    auto V = [&](int p, int q) -> double& { return sv[p * 8 + q][t]; };
    for (int p = 0; p < 8; p++)
        for (int q = p; q < 8; q++) A(p, q) = f[p * 8 + q];
    for (int p = 0; p < 8; p++)
        for (int q = 0; q < 8; q++) V(p, q) = p == q ? 1.0 : 0.0;
    double nrm = 0.0;
    for (int p = 0; p < 8; p++)
        for (int q = 0; q < 8; q++) nrm += A(p, q) * A(p, q);
    const double tol = 1.2e-32 * nrm;  // (2^-53)^2 ||A||_F^2
    int sw = 0;
    for (; sw < kMaxSweep; sw++)
    {
        double off = 0.0;
        for (int p = 0; p < 7; p++)
            for (int q = p + 1; q < 8; q++) off += A(p, q) * A(p, q);
        if (!(off > tol)) break;
        for (int p = 0; p < 7; p++)
            for (int q = p + 1; q < 8; q++)
            {
                const double apq = A(p, q);
                if (apq == 0.0) continue;
                const double app = A(p, p), aqq = A(q, q);
                const double theta = (aqq - app) / (2.0 * apq);
                const double tt = (theta >= 0.0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(1.0 + theta * theta));
                const double c = 1.0 / sqrt(1.0 + tt * tt), s = tt * c;
                A(p, p) = app - tt * apq;
                A(q, q) = aqq + tt * apq;
                A(p, q) = 0.0;
                for (int r = 0; r < 8; r++)
                {
                    if (r == p || r == q) continue;
                    const double arp = A(r, p), arq = A(r, q);
                    A(r, p) = c * arp - s * arq;
                    A(r, q) = s * arp + c * arq;
                }
                for (int r = 0; r < 8; r++)
                {
                    const double vrp = V(r, p), vrq = V(r, q);
                    V(r, p) = c * vrp - s * vrq;
                    V(r, q) = s * vrp + c * vrq;
                }
            }
    }
    // ascending order (insertion sort of the indices), sign convention
    int ord[8];
    double lam[8];
    for (int j = 0; j < 8; j++)
    {
        ord[j] = j;
        lam[j] = A(j, j);
    }
    for (int i = 1; i < 8; i++)
    {
        int k = ord[i];
        double v = lam[i];
        int j = i - 1;
        while (j >= 0 && lam[j] > v)
        {
            lam[j + 1] = lam[j];
            ord[j + 1] = ord[j];
            j--;
        }
        lam[j + 1] = v;
        ord[j + 1] = k;
    }
    double* ev = evec + (size_t)m * 64;
    for (int j = 0; j < 8; j++)
    {
        const int c = ord[j];
        int imax = 0;
        double amax = -1.0;
        for (int i = 0; i < 8; i++)
            if (fabs(V(i, c)) > amax)
            {
                amax = fabs(V(i, c));
                imax = i;
            }
        const double sg = V(imax, c) < 0.0 ? -1.0 : 1.0;
        for (int i = 0; i < 8; i++) ev[i * 8 + j] = sg * V(i, c);
        eval[(size_t)m * 8 + j] = lam[j];
    }
    if (sweeps) sweeps[m] = sw;
}

}  // namespace

int fisher_eigen_launch(int n, const double* F, double* evec, double* eval, int* sweeps, hipStream_t s)
{
    if (n <= 0) return GLASS_HIP_OK;
    jacobi8_kernel<<<(n + kEigThr - 1) / kEigThr, kEigThr, 0, s>>>(n, F, evec, eval, sweeps);
    GH_CHECK(hipGetLastError());
    return GLASS_HIP_OK;
}

}  // namespace glass

using namespace glass;

extern "C" int glass_hip_fisher_eigen(int n, const double* d_F, double* d_evec, double* d_eval, int* d_sweeps,
                                      void* stream)
{
    if (n < 0 || (n && (!d_F || !d_evec || !d_eval)))
    {
        set_error("fisher_eigen: bad arguments (n %d)", n);
        return GLASS_HIP_EINVAL;
    }
    return fisher_eigen_launch(n, d_F, d_evec, d_eval, d_sweeps, (hipStream_t)stream);
}

extern "C" int glass_hip_matrix_eigenstuff(const double* matrix, double* evectors, double* evalues, int N)
{
    if (N != 8 || !matrix || !evectors || !evalues)
    {
        set_error("matrix_eigenstuff: only N = 8 (the UCB parameter count) is supported, got %d", N);
        return GLASS_HIP_EINVAL;
    }
    ThreadCtx* cx;
    int rc;
    if ((rc = ctx_get(&cx))) return rc;
    if ((rc = cx->dbuf.ensure(sizeof(double) * 136 + 256))) return rc;
    if ((rc = cx->hbuf.ensure(sizeof(double) * 136 + 256))) return rc;
    double* d = (double*)cx->dbuf.p;
    memcpy(cx->hbuf.p, matrix, sizeof(double) * 64);
    GH_CHECK(hipMemcpyAsync(d, cx->hbuf.p, sizeof(double) * 64, hipMemcpyHostToDevice, cx->stream));
    if ((rc = fisher_eigen_launch(1, d, d + 64, d + 128, nullptr, cx->stream))) return rc;
    GH_CHECK(hipMemcpyAsync(cx->hbuf.p, d + 64, sizeof(double) * 72, hipMemcpyDeviceToHost, cx->stream));
    GH_CHECK(hipStreamSynchronize(cx->stream));
    memcpy(evectors, cx->hbuf.p, sizeof(double) * 64);
    memcpy(evalues, (double*)cx->hbuf.p + 64, sizeof(double) * 8);
    return GLASS_HIP_OK;
}
