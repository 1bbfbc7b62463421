// spmv.hpp -- launch interface of the ABHSF SpMV / SpMM kernels.
#pragma once

#include "common.hpp"

namespace abhsf {

struct SpmvLaunch {
    int bcol_bytes;  // 2 or 4
    int val_bytes;   // 8 (f64) or 4 (f32, compensated accumulation)
};

// Largest power-of-two rhs tile whose s x R accumulator fits the LDS budget.
int max_rhs_tile(int s, bool comp);

// Y = alpha*A*X + beta*Y, X n x nrhs (ldx), Y m x nrhs (ldy), row-major, device.
void spmm(const SpmvLaunch &L, const DevMat &A, int nrhs, double alpha, const double *X, int64_t ldx, double beta,
          double *Y, int64_t ldy, hipStream_t st);

}  // namespace abhsf
