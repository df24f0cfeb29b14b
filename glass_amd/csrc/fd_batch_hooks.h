// fd_batch_hooks.h -- what the device sampler (ucb_sampler.hip) uses of the
// batched plan (fd_batch.hip) and of the eigensolver (eigen.hip).
#pragma once
#include <hip/hip_runtime.h>

#include "../../include/glass_hip_batch.h"

namespace glass {

struct FdBatchView
{
    int device, C, G, maxsrc, ninit, cap;
    double T;
    const int* c_group;  // [C]
    const int* c_nsrc;   // [C]
    double* sparams;     // [C][maxsrc][8]
    double* logl;        // [C]
    const int* i_chain;  // [ninit] (chain, source) of every live source
    const int* i_src;
};

FdBatchView fd_batch_view(glass_hip_fd_batch* P);

// eval of n items (chain, src, params[8]); items with src < 0 are skipped
// and their outputs left as they were.  stage (optional, of
// fd_batch_stage_bytes(P, n)): the items' templates are kept there for
// fd_batch_accept_staged_dev.
int fd_batch_eval_dev(glass_hip_fd_batch* P, int n, const int* chain, const int* src, const double* params,
                      double* logl, double* snr, hipStream_t s, double2* stage = nullptr);

size_t fd_batch_stage_bytes(glass_hip_fd_batch* P, int n);

// accept the items d_idx[0 .. *d_n) (entries of chain/src/params/logl of
// an eval with stage), at most nmax of them, one per chain, taking their
// templates from that stage (no template rebuild)
int fd_batch_accept_staged_dev(glass_hip_fd_batch* P, int nmax, const int* d_n, const int* d_idx, const int* chain,
                               const int* src, const double* params, const double* logl, const double2* stage,
                               hipStream_t s);

// eigen-decomposition of n symmetric 8x8 matrices (eigen.hip)
int fisher_eigen_launch(int n, const double* F, double* evec, double* eval, int* sweeps, hipStream_t s);

}  // namespace glass
