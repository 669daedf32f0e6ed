// One launch per shooting step for small point sets (the registration sizes):
// each workgroup owns `tpw` targets and sums over ALL sources, so the step's
// epilogue (x_{t+1} = x_t + dt v and the energy dot, or p_t and the gradient)
// runs in the same kernel and no chunk partials go through HBM.
//
// tpw = ceil(n / CUs) (at most kStepTgtMax): the workgroups cover the CUs
// once, each CU summing about n / CUs targets (2,562 points on 256 CUs: 11
// targets, 233 workgroups; a power-of-two count would leave CUs idle or give
// some CUs two workgroups).  Workgroup: kStepThreads lanes (16 wave64).  Lane
// tid takes target tid % tpw and source slice tid / tpw (nsl =
// kStepThreads / tpw slices; lanes past tpw * nsl only stage sources): in
// each LDS tile of kStepTile source records it applies the op's pair
// function to records slice, slice + nsl, ...  Tile sums are folded into a
// (hi, lo) two-sum accumulator in float32 (compensated form) or added in
// float64; the slices of a target are then combined in float64 through LDS
// (the tile buffer, reused): kStepGroups lanes per target each add every
// kStepGroups-th slice in slice order, then one lane per target adds the group
// sums in group order (a fixed order: the result does not depend on
// scheduling).
//
// For larger sets the chunked kernels (pair.hpp) are faster: each LDS record
// (one wave-uniform read) feeds two targets per lane there.  Measured on
// MI355X (forward and adjoint steps, float32 and float64): equal or faster up
// to 4,000 points, slower by 6,000 (1.5x at 10,000); kStepMaxN is where the
// shooting loops switch.
#pragma once

#include "ops.hpp"

#include <type_traits>

namespace lddmm {

constexpr int kStepTgtMax = 16;                      // largest targets per workgroup
constexpr int kStepThreads = 1024;                   // lanes per workgroup
constexpr int kStepTile = 512;                       // source records per LDS tile
constexpr int kStepGroups = 32;                      // partial sums per target in the final combine
constexpr int64_t kStepMaxN = 4096;                  // largest n served by the one-launch steps

// Targets per workgroup for n points on num_cu CUs.
inline int step_targets(int n, int num_cu) {
    int t = num_cu > 0 ? (n + num_cu - 1) / num_cu : kStepTgtMax;
    return t < 1 ? 1 : (t > kStepTgtMax ? kStepTgtMax : t);
}

template <typename T, class OP>
constexpr size_t step_smem_bytes() {
    constexpr size_t tile = size_t(kStepTile) * OP::REC * sizeof(T);
    constexpr size_t red = size_t(kStepThreads) * OP::NCH * sizeof(double);
    return tile > red ? tile : red;
}

// Per-target sums over all m sources: on return, lanes tid < tpw hold the NCH
// channel sums of target blockIdx.x * tpw + tid in v[] (float64).  smem holds
// step_smem_bytes<T, OP>() bytes, 16-byte aligned.  LEAN runs the pair loop
// without unrolling (the float64 multi-scale adjoint: its unrolled loop needs
// more than the 128 registers a 1024-lane workgroup has).
template <typename T, class OP, bool LEAN = false>
__device__ __forceinline__ void step_sums(const typename OP::template Args<T>& args, int n, int m, int tpw,
                                          char* smem, double* v) {
    constexpr int NCH = OP::NCH;
    constexpr int REC = OP::REC;
    T* lds = reinterpret_cast<T*>(smem);
    const int tid = threadIdx.x;
    const int nsl = kStepThreads / tpw;
    const int tg = tid % tpw;
    const int slice = tid / tpw;
    const bool active = slice < nsl;
    const int i = blockIdx.x * tpw + tg;

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
        const int cnt = m - jt < kStepTile ? m - jt : kStepTile;
        if constexpr (LEAN && !kCompensated<T>) {
            // float64, lean: the pairs add straight into the running sums
            if (active) {
#pragma unroll 1
                for (int r = slice; r < cnt; r += nsl) OP::pair(args, tgt, &lds[r * REC], hi);
            }
            continue;
        }
        T acc[NCH];
#pragma unroll
        for (int c = 0; c < NCH; ++c) acc[c] = T(0);
        if (active) {
            if constexpr (LEAN) {
#pragma unroll 1
                for (int r = slice; r < cnt; r += nsl) OP::pair(args, tgt, &lds[r * REC], acc);
            } else {
#pragma unroll 2
                for (int r = slice; r < cnt; r += nsl) OP::pair(args, tgt, &lds[r * REC], acc);
            }
        }
#pragma unroll
        for (int c = 0; c < NCH; ++c) {
            if constexpr (kCompensated<T>) two_sum_acc(hi[c], lo[c], acc[c]);
            else hi[c] += acc[c];
        }
    }

    // combine the slices of each target: red[(slice * NCH + c) * tpw + tg]
    double* red = reinterpret_cast<double*>(smem);
    __syncthreads();                                     // last tile consumed
    if (active) {
#pragma unroll
        for (int c = 0; c < NCH; ++c)
            red[(slice * NCH + c) * tpw + tg] = kCompensated<T> ? double(hi[c]) + double(lo[c]) : double(hi[c]);
    }
    __syncthreads();
    // group sums: lane (g, tg) for g < kStepGroups adds slices g, g + kStepGroups, ...
    double gs[NCH];
    const int gtg = tid % tpw;
    const int g = tid / tpw;
    const bool gl = g < kStepGroups;
#pragma unroll
    for (int c = 0; c < NCH; ++c) gs[c] = 0.0;
    if (gl) {
        for (int sl = g; sl < nsl; sl += kStepGroups) {
#pragma unroll
            for (int c = 0; c < NCH; ++c) gs[c] += red[(sl * NCH + c) * tpw + gtg];
        }
    }
    __syncthreads();                                     // every slice sum read
    if (gl) {
#pragma unroll
        for (int c = 0; c < NCH; ++c) red[(g * NCH + c) * tpw + gtg] = gs[c];
    }
    __syncthreads();
    if (tid < tpw) {
#pragma unroll
        for (int c = 0; c < NCH; ++c) {
            double s = 0.0;
            for (int q = 0; q < kStepGroups; ++q) s += red[(q * NCH + c) * tpw + tid];
            v[c] = s;
        }
    }
}

// Sum of d over the first 64 lanes (one wave; lanes past the workgroup's
// targets hold 0), lane 0 holds the result.
__device__ __forceinline__ double step_lane_sum(double d) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) d += __shfl_xor(d, o, 64);
    return d;
}

// Forward Euler step: x_{t+1} = x_t + dt K(x_t) a_t; dots[b] = dscale sum_i a_i.v_i
// over the workgroup's targets (the energy terms, summed later in a fixed order).
template <typename T, int KW, bool MS>
__global__ void __launch_bounds__(kStepThreads)
k_step_shoot(const typename OpApplyK<KW, MS>::template Args<T> args, int n, int tpw, double dt,
             const double* __restrict__ x, const double* __restrict__ a, double* __restrict__ xn,
             double* __restrict__ dots, double dscale) {
    using OP = OpApplyK<KW, MS>;
    __shared__ __attribute__((aligned(16))) char smem[step_smem_bytes<T, OP>()];
    const int tid = threadIdx.x;
    const int i = blockIdx.x * tpw + tid;
    const bool own = tid < tpw && i < n;
    // the epilogue's inputs, loaded before the sums so that their latency
    // overlaps the pair work
    double xi[KW], ai[KW];
#pragma unroll
    for (int c = 0; c < KW; ++c) {
        xi[c] = own ? x[size_t(i) * KW + c] : 0.0;
        ai[c] = own ? a[size_t(i) * KW + c] : 0.0;
    }
    double v[OP::NCH];
    step_sums<T, OP>(args, n, n, tpw, smem, v);
    if (tid < 64) {
        double d = 0.0;
        if (own) {
#pragma unroll
            for (int c = 0; c < KW; ++c) {
                xn[size_t(i) * KW + c] = xi[c] + dt * v[c];
                d += ai[c] * v[c];
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
k_step_adjoint(const typename OpAdjoint<KW, MS>::template Args<T> args, int n, int tpw, double dt,
               double fdiff, const double* __restrict__ pn, double* __restrict__ pc, double* __restrict__ g) {
    using OP = OpAdjoint<KW, MS>;
    __shared__ __attribute__((aligned(16))) char smem[step_smem_bytes<T, OP>()];
    const int tid = threadIdx.x;
    const int i = blockIdx.x * tpw + tid;
    const bool own = tid < tpw && i < n;
    // p_{t+1} loaded before the sums (latency overlap), except in the
    // multi-scale float64 form, which has no registers to spare for it
    constexpr bool kLean = sizeof(T) == 8 && MS;
    double pi[KW];
#pragma unroll
    for (int c = 0; c < KW; ++c) pi[c] = !kLean && own ? pn[size_t(i) * KW + c] : 0.0;
    double v[OP::NCH];
    step_sums<T, OP, kLean>(args, n, n, tpw, smem, v);
    if (own) {
#pragma unroll
        for (int c = 0; c < KW; ++c) {
            const size_t o = size_t(i) * KW + c;
            pc[o] = (!kLean ? pi[c] : pn[o]) + dt * v[c] * fdiff;
            g[o] = -dt * v[3 + c];
        }
    }
}

// Geodesic (EPDiff) forward step, one Gaussian scale (OpEPDiffFwd's two sums):
// x_{k+1} = x_k + dt K a_k, a_{k+1} = a_k - dt fdiff DiffKT(x_k, a_k, a_k).
template <typename T, int KW>
__global__ void __launch_bounds__(kStepThreads)
k_step_epdiff_fwd(const typename OpEPDiffFwd<KW>::template Args<T> args, int n, int tpw, double dt, double fdiff,
                  const double* __restrict__ x, const double* __restrict__ a, double* __restrict__ xn,
                  double* __restrict__ an) {
    using OP = OpEPDiffFwd<KW>;
    __shared__ __attribute__((aligned(16))) char smem[step_smem_bytes<T, OP>()];
    const int tid = threadIdx.x;
    const int i = blockIdx.x * tpw + tid;
    const bool own = tid < tpw && i < n;
    double xi[KW], ai[KW];
#pragma unroll
    for (int c = 0; c < KW; ++c) {
        xi[c] = own ? x[size_t(i) * KW + c] : 0.0;
        ai[c] = own ? a[size_t(i) * KW + c] : 0.0;
    }
    double v[OP::NCH];
    step_sums<T, OP>(args, n, n, tpw, smem, v);
    if (own) {
#pragma unroll
        for (int c = 0; c < KW; ++c) {
            const size_t o = size_t(i) * KW + c;
            xn[o] = xi[c] + dt * v[c] * 1.0;
            an[o] = ai[c] + (-dt) * v[KW + c] * fdiff;
        }
    }
}

}  // namespace lddmm
