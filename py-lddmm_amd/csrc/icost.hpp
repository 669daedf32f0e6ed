// Internal (regularity) costs of a velocity field on a triangulated surface
// (py-lddmm base/surfaceDistances.py:631-872), used by the single-surface
// entry point and by the shooting loops with internal costs.
#pragma once

#include "common.hpp"

namespace lddmm {

int launch_fill(double* p, size_t count, double v, hipStream_t st);

struct IcList {
    int n;
    int kinds[LDDMM_MAX_ICOSTS];
    double w[LDDMM_MAX_ICOSTS];
};

int ic_check(const lddmm_internal_cost* ic, IcList* out);

// Workspace doubles needed by ic_eval for nf faces.
size_t ic_scratch_doubles(int64_t nf, bool grads);

// With x [nv,3] the vertices and phi [nv,3] the field:
//   energy (device f64[1], may be null) = escale * sum_k w_k C_k;
//   grad_phi / grad_x (may be null) = gscale * sum_k w_k dC_k/dphi, dC_k/dx.
// `scratch` holds ic_scratch_doubles(nf, grads) doubles.  Deterministic.
int ic_eval(const IcList& L, int nv, int nf, const int32_t* faces, const int32_t* adj_ptr, const int32_t* adj_idx,
            const double* x, const double* phi, double escale, double* energy, double gscale, double* grad_phi,
            double* grad_x, double* scratch, hipStream_t st);

}  // namespace lddmm
