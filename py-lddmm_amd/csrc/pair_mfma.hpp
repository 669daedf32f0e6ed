// Float32 pair sums whose accumulation runs on the matrix cores.
//
// For ops of the form  out_i = sum_j g(|u_i - u_j|^2) * B_j  with a per-source
// row B_j of at most 16 (or 32) columns, the Gaussian values are computed on
// the VALU directly in the A-operand layout of v_mfma_f32_16x16x4f32 and
// multiplied by the source columns on the MFMA unit:
//   A[i][k] = g(target i, source k)  (lane l: i = l % 16, k = l / 16)
//   B[k][c] = column c of source k    (lane l: k = l / 16, c = l % 16)
//   D[i][c] += sum_k A[i][k] B[k][c]  (lane l: i = 4*(l/16) + r, c = l % 16)
// so each pair costs 3 sub + 3 FMA + one v_exp_f32 on the VALU and the
// weight FMAs move to the MFMA pipe, which runs beside the VALU.
// applyDiffKT's per-pair coefficient c_ij (u_i - u_j) is rewritten as
//   sum_j g c_ij (u_i - u_j) = u_i (p_i . sum_j g a_j) - (sum_j g u_j a_j^T) p_i
// i.e. columns [a_j | u_j (x) a_j], recombined per target in float64.
//
// Workgroup: 256 lanes = 4 waves; each wave owns G groups of 16 targets.
// Sources are staged through LDS in tiles of kTile: coordinates (float4) and
// the column row (NCP floats).  Tile sums are folded into (hi, lo) two-sum
// accumulators; chunk partials use the Partials layout [chunk][col][n].
#pragma once

#include "pair.hpp"

namespace lddmm {

typedef float v4f __attribute__((ext_vector_type(4)));

template <class OP, int G>
__global__ void __launch_bounds__(kBlock)
k_pair_mfma(const typename OP::Args args, int n, int m, int chunk, Partials<float> part) {
    constexpr int NCP = OP::NCP;      // padded columns: 16 or 32
    constexpr int NB = NCP / 16;      // MFMA column blocks
    __shared__ __attribute__((aligned(16))) float lds_u[kTile * 4];
    __shared__ __attribute__((aligned(16))) float lds_b[kTile * NCP];

    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;
    const int li = lane & 15;          // target row within a group / B column
    const int lk = lane >> 4;          // source within a quad / D row block
    const int tbase = blockIdx.x * (4 * 16 * G) + wave * 16 * G;

    float tu[G][3];
#pragma unroll
    for (int g = 0; g < G; ++g) {
        int i = tbase + 16 * g + li;
        OP::load_target(args, i < n ? i : n - 1, tu[g]);
    }
    v4f acc[G][NB];
    float hi[G][NB][4], lo[G][NB][4];
#pragma unroll
    for (int g = 0; g < G; ++g)
#pragma unroll
        for (int b = 0; b < NB; ++b) {
            acc[g][b] = v4f{0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int r = 0; r < 4; ++r) { hi[g][b][r] = 0.f; lo[g][b][r] = 0.f; }
        }

    int j0, j1;
    OP::Chunks::range(args, blockIdx.y, chunk, m, j0, j1);

    typename OP::Raw raw;
    OP::load_raw(args, j0 + tid < j1 ? j0 + tid : -1, raw);

    for (int jt = j0; jt < j1; jt += kTile) {
        __syncthreads();
        OP::store_rec(args, raw, &lds_u[tid * 4], &lds_b[tid * NCP]);
        __syncthreads();
        {
            int jn = jt + kTile + tid;
            OP::load_raw(args, jn < j1 ? jn : -1, raw);
        }
#pragma unroll 2
        for (int q = 0; q < kTile / 4; ++q) {
            const int j = 4 * q + lk;
            const float4 xj = *reinterpret_cast<const float4*>(&lds_u[j * 4]);
            float bcol[NB];
#pragma unroll
            for (int b = 0; b < NB; ++b) bcol[b] = lds_b[j * NCP + 16 * b + li];
#pragma unroll
            for (int g = 0; g < G; ++g) {
                float dx = tu[g][0] - xj.x, dy = tu[g][1] - xj.y, dz = tu[g][2] - xj.z;
                float r2 = fmaf(dz, dz, fmaf(dy, dy, dx * dx));
                float K = OP::gauss(args, r2);
#pragma unroll
                for (int b = 0; b < NB; ++b)
                    acc[g][b] = __builtin_amdgcn_mfma_f32_16x16x4f32(K, bcol[b], acc[g][b], 0, 0, 0);
            }
        }
#pragma unroll
        for (int g = 0; g < G; ++g)
#pragma unroll
            for (int b = 0; b < NB; ++b) {
#pragma unroll
                for (int r = 0; r < 4; ++r) two_sum_acc(hi[g][b][r], lo[g][b][r], acc[g][b][r]);
                acc[g][b] = v4f{0.f, 0.f, 0.f, 0.f};
            }
    }

#pragma unroll
    for (int g = 0; g < G; ++g)
#pragma unroll
        for (int b = 0; b < NB; ++b) {
            const int c = 16 * b + li;
            if (c >= OP::NCH) continue;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int i = tbase + 16 * g + 4 * lk + r;
                if (i < n) {
                    size_t o = (size_t(blockIdx.y) * OP::NCH + c) * size_t(n) + i;
                    part.hi[o] = hi[g][b][r];
                    part.lo[o] = lo[g][b][r];
                }
            }
        }
}

template <class OP, int G>
int plan_mfma(int n, int m, PairPlan* plan) {
    DeviceInfo di;
    LDDMM_TRY(device_info(&di));
    int occ = blocks_per_cu(reinterpret_cast<const void*>(&k_pair_mfma<OP, G>), kBlock, 0);
    // plan_pair's target block is kBlock * R lanes; here it is 64 * G targets
    *plan = plan_pair(n, m, (64 * G) / kBlock > 0 ? (64 * G) / kBlock : 1, di.num_cu, occ, 64);
    plan->bx = (n + 64 * G - 1) / (64 * G);
    return LDDMM_OK;
}

template <class OP, int G>
int run_mfma(const typename OP::Args& args, int n, int m, const PairPlan& plan, Partials<float> part,
             hipStream_t st) {
    dim3 grid(plan.bx, plan.nchunk);
    const bool prof = profiling();
    if (prof) LDDMM_TRY(profile_begin(st));
    hipLaunchKernelGGL((k_pair_mfma<OP, G>), grid, dim3(kBlock), 0, st, args, n, m, plan.chunk, part);
    LDDMM_CHECK_LAUNCH();
    if (prof) LDDMM_TRY(profile_end(st));
    return LDDMM_OK;
}

// ---------------------------------------------------------------------------
// Ops
// ---------------------------------------------------------------------------
template <bool MS>
__device__ __forceinline__ float mfma_gauss(const ScaleSet<float>& s, const float* w, float r2) {
    if constexpr (!MS) {
        return __builtin_amdgcn_exp2f(-r2);
    } else {
        float g = 0.f;
        for (int q = 0; q < s.ns; ++q) g += w[q] * __builtin_amdgcn_exp2f(-s.beta[q] * r2);
        return g;
    }
}

// applyK with k <= 16 weight columns: B_j = a_j
template <int NC, bool MS>
struct MApplyK {
    static constexpr int NCH = NC;
    static constexpr int NCP = 16;
    using Chunks = ContiguousChunks;
    struct Args {
        const double *y, *x, *a, *c0;
        int dim, lda, col0, kc;   // kc <= NC columns are read, the rest are zero
        ScaleSet<float> sc;
    };
    struct Raw { double x0, x1, x2, a[NC]; };
    __device__ static void load_target(const Args& A, int i, float* u) {
        for (int d = 0; d < 3; ++d) u[d] = coord<float>(A.y, i, A.dim, d, A.c0, A.sc.cs);
    }
    __device__ static void load_raw(const Args& A, int j, Raw& r) {
        if (j < 0) {
            r.x0 = A.c0[0]; r.x1 = A.c0[1]; r.x2 = A.c0[2];
#pragma unroll
            for (int c = 0; c < NC; ++c) r.a[c] = 0.0;
            return;
        }
        const double* xp = A.x + size_t(j) * A.dim;
        r.x0 = xp[0];
        r.x1 = A.dim > 1 ? xp[1] : 0.0;
        r.x2 = A.dim > 2 ? xp[2] : 0.0;
        const double* ap = A.a + size_t(j) * A.lda + A.col0;
#pragma unroll
        for (int c = 0; c < NC; ++c) r.a[c] = c < A.kc ? ap[c] : 0.0;
    }
    __device__ static void store_rec(const Args& A, const Raw& r, float* u, float* b) {
        u[0] = float((r.x0 - A.c0[0]) * A.sc.cs);
        u[1] = float((r.x1 - A.c0[1]) * A.sc.cs);
        u[2] = float((r.x2 - A.c0[2]) * A.sc.cs);
        u[3] = 0.f;
#pragma unroll
        for (int c = 0; c < 16; ++c) b[c] = c < NC ? float(r.a[c < NC ? c : 0]) : 0.f;
    }
    __device__ static float gauss(const Args& A, float r2) { return mfma_gauss<MS>(A.sc, A.sc.wK, r2); }
};

// applyDiffKT plain with k <= 4: B_j = [a_j (k) | u_j[0] a_j | u_j[1] a_j | u_j[2] a_j]
template <int KW, bool MS>
struct MDiffKT {
    static constexpr int NCH = 4 * KW;
    static constexpr int NCP = 16;
    using Chunks = ContiguousChunks;
    struct Args {
        const double *y, *x, *a, *c0;
        int dim, lda, col0;
        ScaleSet<float> sc;
    };
    struct Raw { double x0, x1, x2, a[KW]; };
    __device__ static void load_target(const Args& A, int i, float* u) {
        for (int d = 0; d < 3; ++d) u[d] = coord<float>(A.y, i, A.dim, d, A.c0, A.sc.cs);
    }
    __device__ static void load_raw(const Args& A, int j, Raw& r) {
        if (j < 0) {
            r.x0 = A.c0[0]; r.x1 = A.c0[1]; r.x2 = A.c0[2];
#pragma unroll
            for (int c = 0; c < KW; ++c) r.a[c] = 0.0;
            return;
        }
        const double* xp = A.x + size_t(j) * A.dim;
        r.x0 = xp[0];
        r.x1 = A.dim > 1 ? xp[1] : 0.0;
        r.x2 = A.dim > 2 ? xp[2] : 0.0;
        const double* ap = A.a + size_t(j) * A.lda + A.col0;
#pragma unroll
        for (int c = 0; c < KW; ++c) r.a[c] = ap[c];
    }
    __device__ static void store_rec(const Args& A, const Raw& r, float* u, float* b) {
        // centred, scaled coordinates in float64 before rounding, as for the targets
        double ud[3] = {(r.x0 - A.c0[0]) * A.sc.cs, (r.x1 - A.c0[1]) * A.sc.cs, (r.x2 - A.c0[2]) * A.sc.cs};
        u[0] = float(ud[0]); u[1] = float(ud[1]); u[2] = float(ud[2]); u[3] = 0.f;
#pragma unroll
        for (int c = 0; c < 16; ++c) b[c] = 0.f;
#pragma unroll
        for (int c = 0; c < KW; ++c) {
            b[c] = float(r.a[c]);
#pragma unroll
            for (int d = 0; d < 3; ++d) b[KW + d * KW + c] = float(ud[d] * r.a[c]);
        }
    }
    __device__ static float gauss(const Args& A, float r2) { return mfma_gauss<MS>(A.sc, A.sc.wD, r2); }
};

// out_i = f * (u_i (p_i . S_a) - S_ua p_i),  S_a = sum g a_j, S_ua[d][c] = sum g u_j[d] a_j[c]
__global__ void k_reduce_diffkt_mfma(Partials<float> part, int nchunk, int KW, int n, const double* y,
                                     const double* p, int dim, int lda, int col0, const double* c0, double cs,
                                     double f, double* out, int accumulate);

}  // namespace lddmm
