// Entry points of the MFMA-accumulated float32 kernel sums (ops_mfma.hip).
#pragma once

#include "ops.hpp"
#include "pair_mfma.hpp"

namespace lddmm {

int pair_engine();
int mfma_applyK(const lddmm_kernel* kd, const double* y, int n, const double* x, int m, int dim, const double* a, int k,
                double* out, hipStream_t st);
int mfma_diffKT_plain(const lddmm_kernel* kd, const double* y, int n, const double* x, int m, int dim,
                      const double* p, const double* a, int k, double* out, hipStream_t st);

}  // namespace lddmm
