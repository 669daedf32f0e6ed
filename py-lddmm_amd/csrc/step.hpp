// One launch per shooting step for small point sets (the registration sizes):
// each workgroup owns kStepTgt targets and sums over ALL sources, so the step's
// epilogue (x_{t+1} = x_t + dt v and the energy dot, or p_t and the gradient)
// runs in the same kernel and no chunk partials go through HBM.
//
// Workgroup: kStepThreads lanes (16 wave64).  Lane tid takes target
// tid % kStepTgt and source slice tid / kStepTgt (kStepSlices slices): in each
// LDS tile of kStepTile source records it applies the op's pair function to
// records slice, slice + kStepSlices, ...  A wave holds 4 slices of the 16
// targets, so each LDS read has 4 distinct addresses, each broadcast to 16
// lanes.  Tile sums are folded into a (hi, lo) two-sum accumulator in float32
// (compensated form) or added in float64; the slices of a target are then
// combined in float64, first across the wave's 4 slices by lane shuffles,
// then across the 16 waves in wave order through LDS (a fixed order: the
// result does not depend on scheduling).
//
// For larger sets the chunked kernels (pair.hpp) are faster: each LDS record
// (one wave-uniform read) feeds two targets per lane there, against one pair
// per four-address read here.  Measured on MI355X (forward and adjoint steps,
// float32 and float64): equal or faster up to 4,000 points, slower by 6,000
// (1.5x at 10,000); kStepMaxN is where the shooting loops switch.
#pragma once

#include "ops.hpp"

namespace lddmm {

constexpr int kStepTgt = 16;                         // targets per workgroup
constexpr int kStepThreads = 1024;                   // lanes per workgroup
constexpr int kStepSlices = kStepThreads / kStepTgt; // source slices (64)
constexpr int kStepTile = 512;                       // source records per LDS tile
constexpr int kStepWaves = kStepThreads / 64;
constexpr int64_t kStepMaxN = 4096;                  // largest n served by the one-launch steps

// Per-target sums over all m sources: on return, lanes tid < kStepTgt hold the
// NCH channel sums of target blockIdx.x * kStepTgt + tid in v[] (float64).
template <typename T, class OP>
__device__ __forceinline__ void step_sums(const typename OP::template Args<T>& args, int n, int m, T* lds,
                                          double (*red)[OP::NCH][kStepTgt], double* v) {
    constexpr int NCH = OP::NCH;
    constexpr int REC = OP::REC;
    const int tid = threadIdx.x;
    const int tg = tid % kStepTgt;
    const int slice = tid / kStepTgt;
    const int i = blockIdx.x * kStepTgt + tg;

    typename OP::template Tgt<T> tgt;
    OP::load_target(args, i < n ? i : n - 1, tgt);

    T hi[NCH], lo[NCH];
#pragma unroll
    for (int c = 0; c < NCH; ++c) { hi[c] = T(0); lo[c] = T(0); }

    typename OP::Raw raw;
    if (tid < kStepTile) OP::load_raw(args, tid < m ? tid : -1, raw);
    for (int jt = 0; jt < m; jt += kStepTile) {
        __syncthreads();                                 // previous tile consumed
        if (tid < kStepTile) OP::store_rec(args, raw, &lds[tid * REC]);
        __syncthreads();
        if (tid < kStepTile) {
            const int jn = jt + kStepTile + tid;         // prefetch the next tile
            OP::load_raw(args, jn < m ? jn : -1, raw);
        }
        // records past m are zero-weight padding; every slice runs the same count
        const int cnt = m - jt < kStepTile ? m - jt : kStepTile;
        const int tn = (cnt + kStepSlices - 1) / kStepSlices * kStepSlices;
        T acc[NCH];
#pragma unroll
        for (int c = 0; c < NCH; ++c) acc[c] = T(0);
#pragma unroll 2
        for (int r = slice; r < tn; r += kStepSlices) OP::pair(args, tgt, &lds[r * REC], acc);
#pragma unroll
        for (int c = 0; c < NCH; ++c) {
            if constexpr (kCompensated<T>) two_sum_acc(hi[c], lo[c], acc[c]);
            else hi[c] += acc[c];
        }
    }

    // slices of one target: across the wave (lanes tg, tg + 16, tg + 32, tg + 48) ...
    const int wave = tid / 64;
#pragma unroll
    for (int c = 0; c < NCH; ++c) {
        double s = kCompensated<T> ? double(hi[c]) + double(lo[c]) : double(hi[c]);
        s += __shfl_xor(s, 16, 64);
        s += __shfl_xor(s, 32, 64);
        if ((tid & 63) < kStepTgt) red[wave][c][tg] = s;
    }
    __syncthreads();
    // ... then across the waves, in wave order
    if (tid < kStepTgt) {
#pragma unroll
        for (int c = 0; c < NCH; ++c) {
            double s = 0.0;
            for (int w = 0; w < kStepWaves; ++w) s += red[w][c][tid];
            v[c] = s;
        }
    }
}

// Sum over the first kStepTgt lanes (one wave), lane 0 holds the result.
__device__ __forceinline__ double step_lane_sum(double d) {
#pragma unroll
    for (int o = kStepTgt / 2; o > 0; o >>= 1) d += __shfl_xor(d, o, 64);
    return d;
}

// Forward Euler step: x_{t+1} = x_t + dt K(x_t) a_t; dots[b] = dscale sum_i a_i.v_i
// over the workgroup's targets (the energy terms, summed later in a fixed order).
template <typename T, int KW, bool MS>
__global__ void __launch_bounds__(kStepThreads)
k_step_shoot(const typename OpApplyK<KW, MS>::template Args<T> args, int n, double dt,
             const double* __restrict__ x, const double* __restrict__ a, double* __restrict__ xn,
             double* __restrict__ dots, double dscale) {
    using OP = OpApplyK<KW, MS>;
    __shared__ __attribute__((aligned(16))) T lds[kStepTile * OP::REC];
    __shared__ double red[kStepWaves][OP::NCH][kStepTgt];
    double v[OP::NCH];
    step_sums<T, OP>(args, n, n, lds, red, v);
    const int tid = threadIdx.x;
    if (tid < 64) {
        const int i = blockIdx.x * kStepTgt + tid;
        double d = 0.0;
        if (tid < kStepTgt && i < n) {
#pragma unroll
            for (int c = 0; c < KW; ++c) {
                const size_t o = size_t(i) * KW + c;
                xn[o] = x[o] + dt * v[c];
                d += a[o] * v[c];
            }
        }
        d = step_lane_sum(d);
        if (tid == 0 && dots) dots[blockIdx.x] = dscale * d;
    }
}

// Adjoint step with the Euclidean gradient (OpAdjoint's two sums):
// p_t = p_{t+1} + dt fdiff (lddmm covector sum), grad_t = -dt (kernel sum of q).
template <typename T, int KW, bool MS>
__global__ void __launch_bounds__(kStepThreads)
k_step_adjoint(const typename OpAdjoint<KW, MS>::template Args<T> args, int n, double dt, double fdiff,
               const double* __restrict__ pn, double* __restrict__ pc, double* __restrict__ g) {
    using OP = OpAdjoint<KW, MS>;
    __shared__ __attribute__((aligned(16))) T lds[kStepTile * OP::REC];
    __shared__ double red[kStepWaves][OP::NCH][kStepTgt];
    double v[OP::NCH];
    step_sums<T, OP>(args, n, n, lds, red, v);
    const int tid = threadIdx.x;
    const int i = blockIdx.x * kStepTgt + tid;
    if (tid < kStepTgt && i < n) {
#pragma unroll
        for (int c = 0; c < KW; ++c) {
            const size_t o = size_t(i) * KW + c;
            pc[o] = pn[o] + dt * v[c] * fdiff;
            g[o] = -dt * v[3 + c];
        }
    }
}

}  // namespace lddmm
