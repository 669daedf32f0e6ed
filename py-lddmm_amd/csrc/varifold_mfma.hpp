// Float32 varifold energy sums with the cross terms on the matrix cores
// (single-scale Gaussian, opt-in through lddmm_set_varifold_engine).
//
// For a block of 16 target faces i and 16 source faces j, one
// v_mfma_f32_16x16x4f32 gives the 16x16 block of cosines c_ij = sqrt(d) n_i.n_j:
//   A[i][k] = sqrt(d) n_i[k]   (lane l: i = l % 16, k = l / 16; k = 3 is 0)
//   B[k][j] = n_j[k]           (lane l: k = l / 16, j = l % 16; k = 3 is 0)
//   D[i][j] = c_ij             (lane l: i = 4 (l / 16) + r, r < 4, j = l % 16)
// so each lane evaluates the 4 pairs (4 (l/16) + r, l % 16) of the block and the
// three-term dot of every pair moves to the MFMA pipe.  With R2 a second MFMA
// gives the squared distances in the same layout, relative to an origin o at
// the group's first target (so the cancellation in |a|^2 + |b|^2 - 2 a.b stays
// small for the nearby pairs whose Gaussian is not negligible):
//   A2[i][k] = (u_i - o)[k] for k < 3, 1 for k = 3
//   B2[k][j] = -2 (u_j - o)[k] for k < 3, |u_j - o|^2 for k = 3
//   r_ij^2 = |u_i - o|^2 + D2[i][j]
// The rest of each pair (Gaussian, 1 + c^2, weight) is the VALU op's
// (OpVarifoldE): acc_i += exp2(-r^2) a_j (1 + c^2).  Tile sums are folded into
// (hi, lo) two-sum accumulators; at the end the 16 lanes holding one row's
// columns are combined in float64 by shuffles in a fixed order and one lane
// writes the row's chunk partial ([chunk][n] as hi, lo), read by the VALU
// path's epilogue (k_vf_energy_rows).
#pragma once

#include "pair_mfma.hpp"
#include "varifold.hpp"

namespace lddmm {

template <int G, bool R2>
__global__ void __launch_bounds__(kBlock)
k_vf_energy_mfma(const typename OpVarifoldE<false>::template Args<float> A, int n, int m, int chunk,
                 Partials<float> part) {
    __shared__ __attribute__((aligned(16))) float lds[kTile * 8];
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;
    const int li = lane & 15;          // A row / B and D column
    const int lk = lane >> 4;          // A and B k index / D row quad
    const int tbase = blockIdx.x * (4 * 16 * G) + wave * 16 * G;

    float an[G];                       // A operand: sqrt(d) n_i[lk] of target li
    float au[G];                       // R2: A2 operand
    float tu[G][4][3];                 // centres of the D rows 4 lk + r
    float q[G][4];                     // R2: |u_i - o|^2 of the D rows
    float o[G][3];                     // R2: group origin
#pragma unroll
    for (int g = 0; g < G; ++g) {
        VfTgt<float> t;
        const int ia = tbase + 16 * g + li;
        vf_target(A, A.sc.cs, ia < n ? ia : n - 1, t);
        an[g] = lk == 0 ? t.n0 : lk == 1 ? t.n1 : lk == 2 ? t.n2 : 0.f;
        if constexpr (R2) {
            VfTgt<float> t0;
            const int i0 = tbase + 16 * g;
            vf_target(A, A.sc.cs, i0 < n ? i0 : n - 1, t0);
            o[g][0] = t0.u0; o[g][1] = t0.u1; o[g][2] = t0.u2;
            au[g] = lk == 0 ? t.u0 - o[g][0] : lk == 1 ? t.u1 - o[g][1] : lk == 2 ? t.u2 - o[g][2] : 1.f;
        }
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            VfTgt<float> tr;
            const int ir = tbase + 16 * g + 4 * lk + r;
            vf_target(A, A.sc.cs, ir < n ? ir : n - 1, tr);
            tu[g][r][0] = tr.u0; tu[g][r][1] = tr.u1; tu[g][r][2] = tr.u2;
            if constexpr (R2) {
                const float ex = tr.u0 - o[g][0], ey = tr.u1 - o[g][1], ez = tr.u2 - o[g][2];
                q[g][r] = fmaf(ez, ez, fmaf(ey, ey, ex * ex));
            }
        }
    }
    float hi[G][4], lo[G][4];
#pragma unroll
    for (int g = 0; g < G; ++g)
#pragma unroll
        for (int r = 0; r < 4; ++r) { hi[g][r] = 0.f; lo[g][r] = 0.f; }

    int j0, j1;
    OpVarifoldE<false>::Chunks::range(A, blockIdx.y, chunk, m, j0, j1);

    VarifoldRaw raw;
    vf_load_raw<float>(A, j0 + tid < j1 ? j0 + tid : -1, raw);
    for (int jt = j0; jt < j1; jt += kTile) {
        __syncthreads();                             // previous tile consumed
        vf_store(A, A.sc.cs, raw, &lds[tid * 8]);    // records past j1 carry zero weight
        __syncthreads();
        {
            const int jn = jt + kTile + tid;         // prefetch the next tile
            vf_load_raw<float>(A, jn < j1 ? jn : -1, raw);
        }
        float acc[G][4];
#pragma unroll
        for (int g = 0; g < G; ++g)
#pragma unroll
            for (int r = 0; r < 4; ++r) acc[g][r] = 0.f;
#pragma unroll 2
        for (int b = 0; b < kTile / 16; ++b) {
            const float* rec = &lds[(16 * b + li) * 8];
            const float4 p = *reinterpret_cast<const float4*>(rec);       // centre j, weight a_j
            const float bn = rec[4 + lk];                                   // n_j[lk]; rec[7] = 0
#pragma unroll
            for (int g = 0; g < G; ++g) {
                const v4f cs = __builtin_amdgcn_mfma_f32_16x16x4f32(an[g], bn, v4f{0.f, 0.f, 0.f, 0.f}, 0, 0, 0);
                v4f d2;
                if constexpr (R2) {
                    const float ex = p.x - o[g][0], ey = p.y - o[g][1], ez = p.z - o[g][2];
                    const float bq = lk == 0 ? -2.f * ex : lk == 1 ? -2.f * ey : lk == 2 ? -2.f * ez
                                                                                  : fmaf(ez, ez, fmaf(ey, ey, ex * ex));
                    d2 = __builtin_amdgcn_mfma_f32_16x16x4f32(au[g], bq, v4f{0.f, 0.f, 0.f, 0.f}, 0, 0, 0);
                }
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    float r2;
                    if constexpr (R2) {
                        r2 = fmaxf(q[g][r] + d2[r], 0.f);
                    } else {
                        const float dx = tu[g][r][0] - p.x, dy = tu[g][r][1] - p.y, dz = tu[g][r][2] - p.z;
                        r2 = fmaf(dz, dz, fmaf(dy, dy, dx * dx));
                    }
                    const float K = __builtin_amdgcn_exp2f(-r2);
                    const float c = cs[r];
                    acc[g][r] = fmaf(K * p.w, fmaf(c, c, 1.f), acc[g][r]);
                }
            }
        }
#pragma unroll
        for (int g = 0; g < G; ++g)
#pragma unroll
            for (int r = 0; r < 4; ++r) two_sum_acc(hi[g][r], lo[g][r], acc[g][r]);
    }

    // row 4 lk + r of group g: its 16 columns sit in lanes 16 lk .. 16 lk + 15
#pragma unroll
    for (int g = 0; g < G; ++g)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            double s = double(hi[g][r]) + double(lo[g][r]);
#pragma unroll
            for (int off = 1; off < 16; off <<= 1) s += __shfl_xor(s, off, 64);
            const int i = tbase + 16 * g + 4 * lk + r;
            if (li == r && i < n) {
                const size_t oi = size_t(blockIdx.y) * size_t(n) + i;
                const float h = float(s);
                part.hi[oi] = h;
                part.lo[oi] = float(s - double(h));
            }
        }
}

template <int G, bool R2>
int plan_vf_energy_mfma(int n, int m, PairPlan* plan) {
    DeviceInfo di;
    LDDMM_TRY(device_info(&di));
    int occ = blocks_per_cu(reinterpret_cast<const void*>(&k_vf_energy_mfma<G, R2>), kBlock, 0);
    *plan = plan_pair(n, m, (64 * G) / kBlock > 0 ? (64 * G) / kBlock : 1, di.num_cu, occ, 64);
    plan->bx = (n + 64 * G - 1) / (64 * G);
    return LDDMM_OK;
}

template <int G, bool R2>
int run_vf_energy_mfma(const typename OpVarifoldE<false>::template Args<float>& A, int n, int m,
                       const PairPlan& plan, Partials<float> part, hipStream_t st) {
    dim3 grid(plan.bx, plan.nchunk);
    const bool prof = profiling();
    if (prof) LDDMM_TRY(profile_begin(st));
    hipLaunchKernelGGL((k_vf_energy_mfma<G, R2>), grid, dim3(kBlock), 0, st, A, n, m, plan.chunk, part);
    LDDMM_CHECK_LAUNCH();
    if (prof) LDDMM_TRY(profile_end(st));
    return LDDMM_OK;
}

}  // namespace lddmm
