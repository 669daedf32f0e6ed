// Device-resident landmark shooting: forward Euler flow with its energy, the
// exact discrete adjoint (lddmm covector), EPDiff geodesic shooting and its
// adjoint.  State stays in HBM across the T steps; each step is a handful of
// launches on the caller's stream (capturable into a hipGraph by the caller).
//   forward : base/pointEvolution.py:88-164, energy base/surfaceMatching.py:533-548
//   adjoint : base/surfaceMatching.py:728-908, base/pointEvolution.py:167-236
//   EPDiff  : base/pointEvolution.py:385-472 (forward), :489-618 (adjoint)
#include "ops.hpp"
#include "icost.hpp"
#include "step.hpp"

#include <algorithm>

namespace lddmm {

namespace {

// x_next = x + dt v; dots[block] = scale * sum_i a_i . v_i   (v = summed chunk partials)
// 64 rows x 4 chunk lanes per workgroup (lanes combined in lane order).
template <typename T, int KW>
__global__ void __launch_bounds__(256) k_shoot_epi(Partials<T> part, int nchunk, int n, const double* __restrict__ x,
                                                   const double* __restrict__ a, double dt, double* __restrict__ xn,
                                                   double* __restrict__ dots, double dscale) {
    __shared__ double red[KW][kEpiLanes][kEpiRows];
    __shared__ double dred[kEpiRows];
    const int fl = threadIdx.x % kEpiRows, lane = threadIdx.x / kEpiRows;
    const int i = blockIdx.x * kEpiRows + fl;
#pragma unroll
    for (int c = 0; c < KW; ++c) red[c][lane][fl] = i < n ? sum_lane(part, 0, nchunk, KW, i, c, lane) : 0.0;
    __syncthreads();
    if (lane == 0) {
        double d = 0.0;
        if (i < n) {
#pragma unroll
            for (int c = 0; c < KW; ++c) {
                const double v = ((red[c][0][fl] + red[c][1][fl]) + red[c][2][fl]) + red[c][3][fl];
                const size_t o = size_t(i) * KW + c;
                xn[o] = x[o] + dt * v;
                d += a[o] * v;
            }
        }
        dred[fl] = d;
    }
    __syncthreads();
    for (int s = kEpiRows / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s) dred[threadIdx.x] += dred[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0 && dots) dots[blockIdx.x] = dscale * dred[0];
}

// out = alpha * a + beta * b (elementwise)
__global__ void k_axpby(size_t count, double alpha, const double* a, double beta, const double* b, double* out) {
    size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < count) out[i] = (a ? alpha * a[i] : 0.0) + (b ? beta * b[i] : 0.0);
}

__global__ void __launch_bounds__(1024) k_sum_doubles(const double* rows, int n, double* out) {
    __shared__ double red[1024];
    double s = 0.0;
    for (int i = threadIdx.x; i < n; i += 1024) s += rows[i];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int o = 512; o > 0; o >>= 1) {
        if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = red[0];
}

// ---------------------------------------------------------------------------
// Affine part of the kernel (Kernel.applyK, base/kernelFunctions.py:476-505;
// Kernel.applyDiffKT, :601-625).  With xx = x - c, M(u) = sum_j xx_j u_j^T
// (d x d, M[p][q] = sum_j xx_j[p] u_j[q]) and S(u) = sum_j u_j:
//   velocity:  z_i += (x_i - c) w1 f(M(a)) + w2 S(a)
//   lddmm covector: zpx_i += p_i w1 f(M(a))^T + a_i w1 (f(M(p))^T - 2 r f(M(a))^T)
// f = identity ('affine') or the antisymmetric part (M - M^T)/2
// ('euclidean': the reference sums <E, X> E over an orthonormal basis of
// antisymmetric matrices, which is that projection).  The energy term
// sum_i a_i.z_i(affine) = w1 <f(M(a)), M(a)> + w2 |S(a)|^2.
// Moments are reduced deterministically: a fixed grid for a given n writes
// per-workgroup partials, one workgroup adds them in order.
// ---------------------------------------------------------------------------
constexpr int kAffGrid = 256;
constexpr int kAffVals = 24;   // two weight arrays x (9 moments + 3 sums)

struct AffineArgs {
    int kind;                  // LDDMM_AFFINE_FULL / LDDMM_AFFINE_EUCLIDEAN
    double w1, w2, c[3];
};

AffineArgs affine_args(const lddmm_kernel* kd) {
    return {kd->affine, kd->w1, kd->w2, {kd->center[0], kd->center[1], kd->center[2]}};
}

__global__ void __launch_bounds__(256)
k_aff_partial(int n, int dim, const double* __restrict__ x, AffineArgs A, const double* __restrict__ u0,
              const double* __restrict__ u1, double* __restrict__ part) {
    __shared__ double red[kAffVals][256];
    double acc[kAffVals];
#pragma unroll
    for (int v = 0; v < kAffVals; ++v) acc[v] = 0.0;
    for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) {
        double xx[3], uu[3];
#pragma unroll
        for (int p = 0; p < 3; ++p) xx[p] = p < dim ? x[size_t(i) * dim + p] - A.c[p] : 0.0;
#pragma unroll
        for (int w = 0; w < 2; ++w) {
            const double* u = w ? u1 : u0;
            if (!u) continue;
#pragma unroll
            for (int q = 0; q < 3; ++q) uu[q] = q < dim ? u[size_t(i) * dim + q] : 0.0;
#pragma unroll
            for (int p = 0; p < 3; ++p)
#pragma unroll
                for (int q = 0; q < 3; ++q) acc[w * 12 + p * 3 + q] = fma(xx[p], uu[q], acc[w * 12 + p * 3 + q]);
#pragma unroll
            for (int q = 0; q < 3; ++q) acc[w * 12 + 9 + q] += uu[q];
        }
    }
#pragma unroll
    for (int v = 0; v < kAffVals; ++v) red[v][threadIdx.x] = acc[v];
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o)
            for (int v = 0; v < kAffVals; ++v) red[v][threadIdx.x] += red[v][threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x < kAffVals) part[size_t(blockIdx.x) * kAffVals + threadIdx.x] = red[threadIdx.x][0];
}

// f(M) (or its transpose) into B[9].
__device__ __forceinline__ void aff_f(const double* M, int kind, bool transpose, double* B) {
    for (int p = 0; p < 3; ++p)
        for (int q = 0; q < 3; ++q) {
            double v = kind == LDDMM_AFFINE_EUCLIDEAN ? 0.5 * (M[p * 3 + q] - M[q * 3 + p]) : M[p * 3 + q];
            B[transpose ? q * 3 + p : p * 3 + q] = v;
        }
}

// mom[24] = ordered sum of the partials; *energy_out (if set) = escale *
// (w1 <f(M0), M0> + w2 |S0|^2).
__global__ void __launch_bounds__(64)
k_aff_final(const double* __restrict__ part, int nblocks, AffineArgs A, double* __restrict__ mom,
            double* __restrict__ energy_out, double escale) {
    __shared__ double sh[kAffVals];
    const int v = threadIdx.x;
    if (v < kAffVals) {
        double s = 0.0;
        for (int b = 0; b < nblocks; ++b) s += part[size_t(b) * kAffVals + v];
        sh[v] = s;
        mom[v] = s;
    }
    __syncthreads();
    if (v == 0 && energy_out) {
        double B[9];
        aff_f(sh, A.kind, false, B);
        double e = 0.0;
        for (int k = 0; k < 9; ++k) e = fma(B[k], sh[k], e);
        double t = 0.0;
        for (int q = 0; q < 3; ++q) t = fma(sh[9 + q], sh[9 + q], t);
        *energy_out = escale * (A.w1 * e + A.w2 * t);
    }
}

// out_i += scale * (r1_i B1 + r2_i B2 + t): mode 0 (velocity): r1 = x_i - c,
// B1 = w1 f(M0), t = w2 S0; mode 1 (lddmm covector): r1 = p_i, B1 = w1 f(M0)^T,
// r2 = a_i, B2 = w1 (f(M1)^T - 2 r f(M0)^T).
__global__ void __launch_bounds__(256)
k_aff_apply(int n, int dim, int mode, AffineArgs A, double r, const double* __restrict__ mom,
            const double* __restrict__ x, const double* __restrict__ r1, const double* __restrict__ r2,
            double scale, double* __restrict__ out) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    double B1[9], B2[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0}, t[3] = {0, 0, 0}, u1[3] = {0, 0, 0}, u2[3] = {0, 0, 0};
    if (mode == 0) {
        aff_f(mom, A.kind, false, B1);
        for (int q = 0; q < 3; ++q) t[q] = A.w2 * mom[9 + q];
        for (int p = 0; p < dim; ++p) u1[p] = x[size_t(i) * dim + p] - A.c[p];
    } else {
        double F1[9];
        aff_f(mom, A.kind, true, B1);
        aff_f(mom + 12, A.kind, true, F1);
        for (int k = 0; k < 9; ++k) B2[k] = A.w1 * (F1[k] - 2.0 * r * B1[k]);
        for (int p = 0; p < dim; ++p) {
            u1[p] = r1[size_t(i) * dim + p];
            u2[p] = r2[size_t(i) * dim + p];
        }
    }
    for (int k = 0; k < 9; ++k) B1[k] *= A.w1;
    for (int q = 0; q < dim; ++q) {
        double z = t[q];
        for (int p = 0; p < 3; ++p) z = fma(u1[p], B1[p * 3 + q], fma(u2[p], B2[p * 3 + q], z));
        out[size_t(i) * dim + q] += scale * z;
    }
}

// Workspace of the affine part: [kAffGrid * kAffVals partials][kAffVals moments].
constexpr size_t kAffBytes = size_t(kAffGrid + 1) * kAffVals * sizeof(double);

int aff_moments(const AffineArgs& A, int n, int dim, const double* x, const double* u0, const double* u1,
                double* ws, double* energy_out, double escale, hipStream_t st) {
    const int grid = std::max(1, std::min(kAffGrid, (n + 255) / 256));
    hipLaunchKernelGGL(k_aff_partial, dim3(unsigned(grid)), dim3(256), 0, st, n, dim, x, A, u0, u1, ws);
    LDDMM_CHECK_LAUNCH();
    hipLaunchKernelGGL(k_aff_final, dim3(1), dim3(64), 0, st, ws, grid, A, ws + size_t(kAffGrid) * kAffVals,
                       energy_out, escale);
    LDDMM_CHECK_LAUNCH();
    return LDDMM_OK;
}

int aff_apply(const AffineArgs& A, int n, int dim, int mode, double r, const double* ws, const double* x,
              const double* r1, const double* r2, double scale, double* out, hipStream_t st) {
    hipLaunchKernelGGL(k_aff_apply, dim3(unsigned((n + 255) / 256)), dim3(256), 0, st, n, dim, mode, A, r,
                       ws + size_t(kAffGrid) * kAffVals, x, r1, r2, scale, out);
    LDDMM_CHECK_LAUNCH();
    return LDDMM_OK;
}

// Affine part of a geodesic (EPDiff) step: Kernel.applyK, applyDiffKT,
// applyDDiffK11and12 and applyDiffK1and2 each add theirs
// (base/kernelFunctions.py:476-505, :549-561, :601-625, :665-674).  With
// M(u) = sum_j xx_j u_j^T, S(u) = sum_j u_j, N = sum_j a_j pa_j^T and f as
// above (a row vector r times a matrix B: (r B)[q] = sum_p r[p] B[p][q]):
//   forward:  x_{k+1} += dt (xx_i w1 f(M(a)) + w2 S(a))
//             a_{k+1} -= dt w1 a_i f(M(a))^T
//   adjoint:  px_k += dt w1 (px_i f(M(a))^T + a_i f(M(px))^T - a_i f(N))
//             pa_k += dt (w1 xx_i f(M(px)) + w2 S(px) - w1 (pa_i f(M(a)) + xx_i f(N)^T))
// ('euclidean': each of the reference's sums <E, X> E over the orthonormal
// antisymmetric basis is the projection f(X)).  Moments: a fixed grid for a
// given n, per-lane sums, in-wave shuffles, waves and then workgroups added
// in order (deterministic).
constexpr int kEpdAffVals = 33;   // M(a) 9, S(a) 3, M(px) 9, S(px) 3, N 9

__global__ void __launch_bounds__(256)
k_epd_aff_partial(int n, int dim, const double* __restrict__ x, AffineArgs A, const double* __restrict__ a,
                  const double* __restrict__ px, const double* __restrict__ pa, double* __restrict__ part) {
    __shared__ double red[4][kEpdAffVals];
    double acc[kEpdAffVals];
#pragma unroll
    for (int v = 0; v < kEpdAffVals; ++v) acc[v] = 0.0;
    for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) {
        double xx[3], aa[3], pp[3], qq[3];
#pragma unroll
        for (int p = 0; p < 3; ++p) {
            const bool in = p < dim;
            const size_t o = size_t(i) * dim + p;
            xx[p] = in ? x[o] - A.c[p] : 0.0;
            aa[p] = in ? a[o] : 0.0;
            pp[p] = in && px ? px[o] : 0.0;
            qq[p] = in && pa ? pa[o] : 0.0;
        }
#pragma unroll
        for (int p = 0; p < 3; ++p)
#pragma unroll
            for (int q = 0; q < 3; ++q) {
                acc[p * 3 + q] = fma(xx[p], aa[q], acc[p * 3 + q]);
                acc[12 + p * 3 + q] = fma(xx[p], pp[q], acc[12 + p * 3 + q]);
                acc[24 + p * 3 + q] = fma(aa[p], qq[q], acc[24 + p * 3 + q]);
            }
#pragma unroll
        for (int q = 0; q < 3; ++q) {
            acc[9 + q] += aa[q];
            acc[21 + q] += pp[q];
        }
    }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int v = 0; v < kEpdAffVals; ++v) {
        double s = acc[v];
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
        if (lane == 0) red[wave][v] = s;
    }
    __syncthreads();
    if (threadIdx.x < kEpdAffVals) {
        const int v = threadIdx.x;
        part[size_t(blockIdx.x) * kEpdAffVals + v] = ((red[0][v] + red[1][v]) + red[2][v]) + red[3][v];
    }
}

__global__ void __launch_bounds__(64)
k_epd_aff_final(const double* __restrict__ part, int nblocks, double* __restrict__ mom) {
    const int v = threadIdx.x;
    if (v >= kEpdAffVals) return;
    double s = 0.0;
    for (int b = 0; b < nblocks; ++b) s += part[size_t(b) * kEpdAffVals + v];
    mom[v] = s;
}

// mode 0: forward step (x, a -> out1 = x_{k+1}, out2 = a_{k+1}); mode 1:
// adjoint step (x, a, px = px_{k+1}, pa = pa_{k+1} -> out1 = px_k, out2 = pa_k).
__global__ void __launch_bounds__(256)
k_epd_aff_apply(int n, int dim, int mode, AffineArgs A, double dt, const double* __restrict__ mom,
                const double* __restrict__ x, const double* __restrict__ a, const double* __restrict__ px,
                const double* __restrict__ pa, double* __restrict__ out1, double* __restrict__ out2) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    double Fa[9], Fp[9], Fn[9], xx[3] = {0, 0, 0}, aa[3] = {0, 0, 0}, pp[3] = {0, 0, 0}, qq[3] = {0, 0, 0};
    aff_f(mom, A.kind, false, Fa);
    for (int p = 0; p < dim; ++p) {
        const size_t o = size_t(i) * dim + p;
        xx[p] = x[o] - A.c[p];
        aa[p] = a[o];
        if (mode == 1) {
            pp[p] = px[o];
            qq[p] = pa[o];
        }
    }
    if (mode == 0) {
        for (int q = 0; q < dim; ++q) {
            double zx = A.w2 * mom[9 + q], za = 0.0;
            for (int p = 0; p < 3; ++p) {
                zx = fma(A.w1 * xx[p], Fa[p * 3 + q], zx);
                za = fma(A.w1 * aa[p], Fa[q * 3 + p], za);
            }
            out1[size_t(i) * dim + q] += dt * zx;
            out2[size_t(i) * dim + q] += -dt * za;
        }
        return;
    }
    aff_f(mom + 12, A.kind, false, Fp);
    aff_f(mom + 24, A.kind, false, Fn);
    for (int q = 0; q < dim; ++q) {
        double zpx = 0.0, zpa = A.w2 * mom[21 + q];
        for (int p = 0; p < 3; ++p) {
            zpx = fma(A.w1, pp[p] * Fa[q * 3 + p] + aa[p] * Fp[q * 3 + p] - aa[p] * Fn[p * 3 + q], zpx);
            zpa = fma(A.w1, xx[p] * Fp[p * 3 + q] - qq[p] * Fa[p * 3 + q] - xx[p] * Fn[q * 3 + p], zpa);
        }
        out1[size_t(i) * dim + q] += dt * zpx;
        out2[size_t(i) * dim + q] += dt * zpa;
    }
}

// Workspace of the geodesic affine part: [kAffGrid * kEpdAffVals partials][kEpdAffVals moments].
constexpr size_t kEpdAffBytes = size_t(kAffGrid + 1) * kEpdAffVals * sizeof(double);

int epd_aff_step(const AffineArgs& A, int n, int dim, int mode, double dt, const double* x, const double* a,
                 const double* px, const double* pa, double* out1, double* out2, double* ws, hipStream_t st) {
    const int grid = std::max(1, std::min(kAffGrid, (n + 255) / 256));
    double* mom = ws + size_t(kAffGrid) * kEpdAffVals;
    hipLaunchKernelGGL(k_epd_aff_partial, dim3(unsigned(grid)), dim3(256), 0, st, n, dim, x, A, a,
                       mode == 1 ? px : nullptr, mode == 1 ? pa : nullptr, ws);
    LDDMM_CHECK_LAUNCH();
    hipLaunchKernelGGL(k_epd_aff_final, dim3(1), dim3(64), 0, st, ws, grid, mom);
    LDDMM_CHECK_LAUNCH();
    hipLaunchKernelGGL(k_epd_aff_apply, dim3(unsigned((n + 255) / 256)), dim3(256), 0, st, n, dim, mode, A, dt, mom,
                       x, a, px, pa, out1, out2);
    LDDMM_CHECK_LAUNCH();
    return LDDMM_OK;
}

int axpby(size_t count, double alpha, const double* a, double beta, const double* b, double* out,
          hipStream_t st) {
    if (count == 0) return LDDMM_OK;
    hipLaunchKernelGGL(k_axpby, dim3(unsigned((count + 255) / 256)), dim3(256), 0, st, count, alpha, a, beta,
                       b, out);
    LDDMM_CHECK_LAUNCH();
    return LDDMM_OK;
}

int check_shoot(const lddmm_kernel* kd, int T, int64_t n, int dim, unsigned flags) {
    LDDMM_TRY(validate_kernel(kd, flags));
    LDDMM_REQUIRE(T >= 1, "T must be >= 1 (got %d)", T);
    LDDMM_REQUIRE(n >= 0 && n <= kMaxPoints, "bad n");
    LDDMM_REQUIRE(dim == 2 || dim == 3, "dim must be 2 or 3 (got %d)", dim);
    return LDDMM_OK;
}

ChanScale uniform_scale(double v) {
    ChanScale c{};
    for (int i = 0; i < 8; ++i) c.s[i] = v;
    return c;
}

template <typename T, class OP>
int reduce_job(const Job<T, OP>& j, Partials<T> part, int n, int dim, double scale, double* out, int accumulate,
               const double* base, double alpha, hipStream_t st) {
    return run_reduce(part, j.plan.nchunk, OP::NCH, dim, n, uniform_scale(scale), out, dim, 0, accumulate, base,
                      alpha, st);
}

// Internal costs on the template's triangulation (shooting with internalCost).
struct IcCtx {
    IcList L;
    const int32_t* faces;
    int nf;
    const int32_t* adj_ptr;
    const int32_t* adj_idx;
};

// x_next = x + dt v; dots[block] = dscale * sum_i a_i . v_i   (v given)
__global__ void __launch_bounds__(256)
k_shoot_epi_v(int64_t count, const double* __restrict__ x, const double* __restrict__ a,
              const double* __restrict__ v, double dt, double* __restrict__ xn, double* __restrict__ dots,
              double dscale) {
    __shared__ double red[256];
    const int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x;
    double d = 0.0;
    if (i < count) {
        xn[i] = x[i] + dt * v[i];
        d = a[i] * v[i];
    }
    red[threadIdx.x] = d;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) dots[blockIdx.x] = dscale * red[0];
}

// Reduce of the fused adjoint step (OpAdjoint partials): channel c < KW gives
// p_t = p_{t+1} + dt * fdiff * v_c, channel KW + c the gradient g = -dt * v.
// 64 rows x 4 chunk lanes per workgroup; c = blockIdx.y in [0, 2 KW).
template <typename T, int KW>
__global__ void __launch_bounds__(256) k_reduce_adjoint(Partials<T> part, int nchunk, int n, double dt, double fdiff,
                                                        const double* __restrict__ pn, double* __restrict__ pc,
                                                        double* __restrict__ g) {
    __shared__ double red[kEpiLanes][kEpiRows];
    const int fl = threadIdx.x % kEpiRows, lane = threadIdx.x / kEpiRows;
    const int i = blockIdx.x * kEpiRows + fl;
    const int c = blockIdx.y;
    const int ch = c < KW ? c : 3 + (c - KW);            // partial channel
    red[lane][fl] = i < n ? sum_lane(part, 0, nchunk, 3 + KW, i, ch, lane) : 0.0;
    __syncthreads();
    if (lane != 0 || i >= n) return;
    const double v = ((red[0][fl] + red[1][fl]) + red[2][fl]) + red[3][fl];
    if (c < KW) {
        const size_t o = size_t(i) * KW + c;
        pc[o] = pn[o] + dt * v * fdiff;          // the order of reduce_job(..., fdiff, pc, 0, pn, dt)
    } else {
        g[size_t(i) * KW + (c - KW)] = -dt * v;
    }
}

// Fused geodesic forward step (OpEPDiffFwd partials): c < KW gives
// x_{k+1} = x + dt K a, c >= KW gives a_{k+1} = a - dt fdiff DiffKT (the
// operation order of two reduce_job calls).
template <typename T, int KW>
__global__ void __launch_bounds__(256) k_reduce_epdiff_fwd(Partials<T> part, int nchunk, int n, double dt,
                                                           double fdiff, const double* __restrict__ x,
                                                           const double* __restrict__ a, double* __restrict__ xn,
                                                           double* __restrict__ an) {
    __shared__ double red[kEpiLanes][kEpiRows];
    const int fl = threadIdx.x % kEpiRows, lane = threadIdx.x / kEpiRows;
    const int i = blockIdx.x * kEpiRows + fl;
    const int c = blockIdx.y;
    red[lane][fl] = i < n ? sum_lane(part, 0, nchunk, KW + 3, i, c, lane) : 0.0;
    __syncthreads();
    if (lane != 0 || i >= n) return;
    const double v = ((red[0][fl] + red[1][fl]) + red[2][fl]) + red[3][fl];
    if (c < KW) {
        const size_t o = size_t(i) * KW + c;
        xn[o] = x[o] + dt * v * 1.0;
    } else {
        const size_t o = size_t(i) * KW + (c - KW);
        an[o] = a[o] + (-dt) * v * fdiff;
    }
}

// Fused geodesic adjoint step (OpEPDiffAdj partials): c < KW gives
// pxo = px + dt fdiff DiffKT_sym - dt fdd DDiffK, c >= KW gives
// pao = pa + dt K px - dt fd1 DiffK1and2 (the operation order of the four
// reduce_job calls of the unfused step).
template <typename T, int KW>
__global__ void __launch_bounds__(256) k_reduce_epdiff_adj(Partials<T> part, int nchunk, int n, double dt,
                                                           double fdiff, double fdd, double fd1,
                                                           const double* __restrict__ px,
                                                           const double* __restrict__ pa, double* __restrict__ pxo,
                                                           double* __restrict__ pao) {
    constexpr int NCH = 6 + 2 * KW;
    __shared__ double red[2][kEpiLanes][kEpiRows];
    const int fl = threadIdx.x % kEpiRows, lane = threadIdx.x / kEpiRows;
    const int i = blockIdx.x * kEpiRows + fl;
    const int c = blockIdx.y;
    const int cc = c < KW ? c : c - KW;
    const int ch1 = c < KW ? c : 6 + cc;                 // DiffKT_sym / K px
    const int ch2 = c < KW ? 3 + c : 6 + KW + cc;        // DDiffK / DiffK1and2
    red[0][lane][fl] = i < n ? sum_lane(part, 0, nchunk, NCH, i, ch1, lane) : 0.0;
    red[1][lane][fl] = i < n ? sum_lane(part, 0, nchunk, NCH, i, ch2, lane) : 0.0;
    __syncthreads();
    if (lane != 0 || i >= n) return;
    const double v1 = ((red[0][0][fl] + red[0][1][fl]) + red[0][2][fl]) + red[0][3][fl];
    const double v2 = ((red[1][0][fl] + red[1][1][fl]) + red[1][2][fl]) + red[1][3][fl];
    const size_t o = size_t(i) * KW + cc;
    if (c < KW) {
        const double u = px[o] + dt * v1 * fdiff;
        pxo[o] = u + (-dt) * v2 * fdd;
    } else {
        const double u = pa[o] + dt * v1 * 1.0;
        pao[o] = u + (-dt) * v2 * fd1;
    }
}

// ---------------------------------------------------------------------------
template <typename T, bool MS, int KW>
int shoot_forward_t(const lddmm_kernel* kd, const double* at, int T_, int n, double* xt, double r,
                    double* energy, const IcCtx* ic, hipStream_t st) {
    ScaleSet<T> sc = make_scales<T>(kd, nullptr);
    using OPK = OpApplyK<KW, MS>;
    Job<T, OPK> jk;
    LDDMM_TRY(plan_job(jk, n, n, !MS));
    const bool aff = kd->affine != LDDMM_AFFINE_NONE;
    // small sets: one launch per step (pair sums and epilogue in one workgroup)
    const bool small = !ic && !aff && int64_t(n) <= kStepMaxN;
    int tpw = kStepTgtMax;                                      // targets per k_step_shoot workgroup
    if (small) {
        DeviceInfo di;
        LDDMM_TRY(device_info(&di));
        tpw = step_targets(n, di.num_cu);
    }
    const int nb = small ? (n + tpw - 1) / tpw                  // blocks of k_step_shoot
                         : (n + kEpiRows - 1) / kEpiRows;     // blocks of k_shoot_epi
    const int nbv = int((int64_t(n) * KW + 255) / 256);       // blocks of k_shoot_epi_v
    const int ndot = ic ? nbv + 1 : nb + 1;                     // dot slots per step
    void* ws;
    const AffineArgs AA = affine_args(kd);
    const size_t icb = ic ? (size_t(n) * KW + ic_scratch_doubles(ic->nf, false)) * sizeof(double) + 512 : 0;
    LDDMM_TRY(workspace(256 + partial_bytes<T>(jk.plan, KW, n) + size_t(T_) * ndot * sizeof(double) +
                            (aff ? kAffBytes + 256 : 0) + icb + 512,
                        st, &ws));
    char* base = static_cast<char*>(ws);
    Carve cv;
    double* c0 = reinterpret_cast<double*>(base + cv.take(32));
    double* dots = reinterpret_cast<double*>(base + cv.take(size_t(T_) * ndot * sizeof(double)));
    double* aws = aff ? reinterpret_cast<double*>(base + cv.take(kAffBytes)) : nullptr;
    double* vbuf = ic ? reinterpret_cast<double*>(base + cv.take(size_t(n) * KW * sizeof(double))) : nullptr;
    double* icws = ic ? reinterpret_cast<double*>(base + cv.take(ic_scratch_doubles(ic->nf, false) *
                                                                 sizeof(double))) : nullptr;
    Partials<T> part = carve_partials<T>(base, cv, jk.plan, KW, n);
    const double dt = 1.0 / T_;
    const size_t stride = size_t(n) * KW;
    // one centre for the whole flow: any central point serves the centred
    // float32 coordinates, and the shape's centre moves little over one shoot
    LDDMM_TRY(launch_center(xt, n, xt, 0, KW, c0, st));
    for (int t = 0; t < T_; ++t) {
        const double* x = xt + t * stride;
        const double* a = at + t * stride;
        jk.args = {x, x, a, c0, KW, KW, 0, sc};
        if (small) {
            hipLaunchKernelGGL((k_step_shoot<T, KW, MS>), dim3(unsigned(nb)), dim3(kStepThreads), 0, st, jk.args, n,
                               tpw, dt, x, a, xt + (t + 1) * stride, dots + size_t(t) * nb, r * dt);
            LDDMM_CHECK_LAUNCH();
            continue;
        }
        LDDMM_TRY(run_job(jk, n, n, part, st));
        if (!ic) {
            hipLaunchKernelGGL((k_shoot_epi<T, KW>), dim3(unsigned(nb)), dim3(256), 0, st, part, jk.plan.nchunk, n,
                               x, a, dt, xt + (t + 1) * stride, dots + size_t(t) * nb, r * dt);
            LDDMM_CHECK_LAUNCH();
            if (aff) {
                LDDMM_TRY(aff_moments(AA, n, KW, x, a, nullptr, aws, dots + size_t(T_) * nb + t, r * dt, st));
                LDDMM_TRY(aff_apply(AA, n, KW, 0, 0.0, aws, x, nullptr, nullptr, dt, xt + (t + 1) * stride, st));
            }
        } else {
            // v_t = K(x_t) a_t (with the affine part) is needed by the cost: materialise it
            LDDMM_TRY(reduce_job(jk, part, n, KW, 1.0, vbuf, 0, nullptr, 1.0, st));
            if (aff) {
                LDDMM_TRY(aff_moments(AA, n, KW, x, a, nullptr, aws, nullptr, 0.0, st));
                LDDMM_TRY(aff_apply(AA, n, KW, 0, 0.0, aws, x, nullptr, nullptr, 1.0, vbuf, st));
            }
            hipLaunchKernelGGL(k_shoot_epi_v, dim3(unsigned(nbv)), dim3(256), 0, st, int64_t(n) * KW, x, a, vbuf, dt,
                               xt + (t + 1) * stride, dots + size_t(t) * nbv, r * dt);
            LDDMM_CHECK_LAUNCH();
            LDDMM_TRY(ic_eval(ic->L, n, ic->nf, ic->faces, ic->adj_ptr, ic->adj_idx, x, vbuf, dt,
                              dots + size_t(T_) * nbv + t, 0.0, nullptr, nullptr, icws, st));
        }
    }
    if (energy) {
        const int count = ic ? T_ * (nbv + 1) : T_ * (aff ? nb + 1 : nb);
        hipLaunchKernelGGL(k_sum_doubles, dim3(1), dim3(1024), 0, st, dots, count, energy);
        LDDMM_CHECK_LAUNCH();
    }
    return LDDMM_OK;
}

template <typename T, bool MS, int KW>
int shoot_adjoint_t(const lddmm_kernel* kd, const double* at, const double* xt, int T_, int n,
                    const double* px1, double r, int euclidean, double* pxt, double* grad, const IcCtx* ic,
                    hipStream_t st) {
    SingleScaleConsts ss;
    ScaleSet<T> sc = make_scales<T>(kd, &ss);
    const double fdiff = MS ? -1.0 : ss.diff;
    using OPD = OpDiffKT<KW, true, MS>;
    using OPK = OpApplyK<KW, MS>;
    using OPA = OpAdjoint<KW, MS>;
    Job<T, OPD> jd;
    Job<T, OPK> jk;
    Job<T, OPA> ja;
    LDDMM_TRY(plan_job(jd, n, n, !MS));
    LDDMM_TRY(plan_job(jk, n, n, !MS));
    const size_t stride = size_t(n) * KW;
    const bool aff = kd->affine != LDDMM_AFFINE_NONE;
    // the fused step (one pair launch + one reduce) covers the Euclidean gradient
    // without affine part or internal costs
    const bool fused = euclidean && !aff && !ic;
    if (fused) LDDMM_TRY(plan_job(ja, n, n, !MS));
    const AffineArgs AA = affine_args(kd);
    const size_t icb = ic ? (4 * stride + ic_scratch_doubles(ic->nf, true)) * sizeof(double) + 1280 : 0;
    size_t need = 256 + std::max({partial_bytes<T>(jd.plan, 3, n), partial_bytes<T>(jk.plan, KW, n),
                                  fused ? partial_bytes<T>(ja.plan, 3 + KW, n) : size_t(0)}) +
                  stride * sizeof(double) * (pxt ? 1 : 3) + (aff ? kAffBytes + 256 : 0) + icb + 1024;
    void* ws;
    LDDMM_TRY(workspace(need, st, &ws));
    char* base = static_cast<char*>(ws);
    Carve cv;
    double* c0 = reinterpret_cast<double*>(base + cv.take(32));
    double* aws = aff ? reinterpret_cast<double*>(base + cv.take(kAffBytes)) : nullptr;
    // internal costs: v_t, Lv, DLv, p_{t+1} - Lv and the cost's scratch
    double *vbuf = nullptr, *lv = nullptr, *dlv = nullptr, *pe = nullptr, *icws = nullptr;
    if (ic) {
        vbuf = reinterpret_cast<double*>(base + cv.take(stride * sizeof(double)));
        lv = reinterpret_cast<double*>(base + cv.take(stride * sizeof(double)));
        dlv = reinterpret_cast<double*>(base + cv.take(stride * sizeof(double)));
        pe = reinterpret_cast<double*>(base + cv.take(stride * sizeof(double)));
        icws = reinterpret_cast<double*>(base + cv.take(ic_scratch_doubles(ic->nf, true) * sizeof(double)));
    }
    double* dat = reinterpret_cast<double*>(base + cv.take(stride * sizeof(double)));
    double* pbuf[2] = {nullptr, nullptr};
    if (!pxt) {
        pbuf[0] = reinterpret_cast<double*>(base + cv.take(stride * sizeof(double)));
        pbuf[1] = reinterpret_cast<double*>(base + cv.take(stride * sizeof(double)));
    }
    size_t poff = cv.off;
    auto P = [&](int t) -> double* { return pxt ? pxt + size_t(t) * stride : pbuf[t & 1]; };
    LDDMM_HIP(hipMemcpyAsync(P(T_), px1, stride * sizeof(double), hipMemcpyDeviceToDevice, st));
    const double dt = 1.0 / T_;
    LDDMM_TRY(launch_center(xt, n, xt, 0, KW, c0, st));       // one centre (x_0) for all steps
    for (int t = T_ - 1; t >= 0; --t) {
        const double* x = xt + size_t(t) * stride;
        const double* a = at + size_t(t) * stride;
        double* pn = P(t + 1);
        double* pc = P(t);
        // p used by the covector sum: p_{t+1}, or p_{t+1} - Lv with internal costs
        const double* pu = pn;
        if (ic) {
            Carve c1; c1.off = poff;
            Partials<T> partV = carve_partials<T>(base, c1, jk.plan, KW, n);
            jk.args = {x, x, a, c0, KW, KW, 0, sc};
            LDDMM_TRY(run_job(jk, n, n, partV, st));
            LDDMM_TRY(reduce_job(jk, partV, n, KW, 1.0, vbuf, 0, nullptr, 1.0, st));
            if (aff) {
                LDDMM_TRY(aff_moments(AA, n, KW, x, a, nullptr, aws, nullptr, 0.0, st));
                LDDMM_TRY(aff_apply(AA, n, KW, 0, 0.0, aws, x, nullptr, nullptr, 1.0, vbuf, st));
            }
            LDDMM_TRY(ic_eval(ic->L, n, ic->nf, ic->faces, ic->adj_ptr, ic->adj_idx, x, vbuf, 0.0, nullptr, 1.0, lv,
                              dlv, icws, st));
            LDDMM_TRY(axpby(stride, 1.0, pn, -1.0, lv, pe, st));
            pu = pe;
        }
        if (fused) {
            ja.args = {x, x, pu, a, c0, KW, KW, 0, 2.0 * r, sc};
            if (int64_t(n) <= kStepMaxN) {
                // small sets: one launch per step
                DeviceInfo di;
                LDDMM_TRY(device_info(&di));
                const int tpw = step_targets(n, di.num_cu);
                hipLaunchKernelGGL((k_step_adjoint<T, KW, MS>), dim3(unsigned((n + tpw - 1) / tpw)),
                                   dim3(kStepThreads), 0, st, ja.args, n, tpw, dt, fdiff, pn, pc,
                                   grad + size_t(t) * stride);
                LDDMM_CHECK_LAUNCH();
                continue;
            }
            Carve ca; ca.off = poff;
            Partials<T> partA = carve_partials<T>(base, ca, ja.plan, 3 + KW, n);
            LDDMM_TRY(run_job(ja, n, n, partA, st));
            hipLaunchKernelGGL((k_reduce_adjoint<T, KW>), dim3(unsigned((n + kEpiRows - 1) / kEpiRows), 2 * KW),
                               dim3(256), 0, st, partA, ja.plan.nchunk, n, dt, fdiff, pn, pc,
                               grad + size_t(t) * stride);
            LDDMM_CHECK_LAUNCH();
            continue;
        }
        Carve c2; c2.off = poff;
        Partials<T> partD = carve_partials<T>(base, c2, jd.plan, 3, n);
        jd.args = {x, x, pu, a, c0, KW, KW, 0, 2.0 * r, sc};
        LDDMM_TRY(run_job(jd, n, n, partD, st));
        // dat = 2 r a_t - p_{t+1} (+ Lv)  (read before p_t overwrites the other buffer)
        LDDMM_TRY(axpby(stride, 2.0 * r, a, -1.0, pu, dat, st));
        LDDMM_TRY(reduce_job(jd, partD, n, KW, fdiff, pc, 0, pn, dt, st));
        if (ic) LDDMM_TRY(axpby(stride, 1.0, pc, -dt, dlv, pc, st));
        if (aff) {
            LDDMM_TRY(aff_moments(AA, n, KW, x, a, pu, aws, nullptr, 0.0, st));
            LDDMM_TRY(aff_apply(AA, n, KW, 1, r, aws, x, pu, a, dt, pc, st));
        }
        double* g = grad + size_t(t) * stride;
        if (euclidean) {
            Carve c3; c3.off = poff;
            Partials<T> partK = carve_partials<T>(base, c3, jk.plan, KW, n);
            jk.args = {x, x, dat, c0, KW, KW, 0, sc};
            LDDMM_TRY(run_job(jk, n, n, partK, st));
            LDDMM_TRY(reduce_job(jk, partK, n, KW, 1.0, g, 0, nullptr, dt, st));
            if (aff) {
                LDDMM_TRY(aff_moments(AA, n, KW, x, dat, nullptr, aws, nullptr, 0.0, st));
                LDDMM_TRY(aff_apply(AA, n, KW, 0, 0.0, aws, x, nullptr, nullptr, dt, g, st));
            }
        } else {
            LDDMM_TRY(axpby(stride, dt, dat, 0.0, nullptr, g, st));
        }
    }
    return LDDMM_OK;
}

// ---------------------------------------------------------------------------
template <typename T, bool MS, int KW>
int epdiff_forward_t(const lddmm_kernel* kd, int T_, int n, double* xt, double* at, hipStream_t st) {
    SingleScaleConsts ss;
    ScaleSet<T> sc = make_scales<T>(kd, &ss);
    const double fdiff = MS ? -1.0 : ss.diff;
    using OPK = OpApplyK<KW, MS>;
    using OPD = OpDiffKT<KW, false, MS>;
    using OPF = OpEPDiffFwd<KW>;
    Job<T, OPK> jk;
    Job<T, OPD> jd;
    Job<T, OPF> jf;
    LDDMM_TRY(plan_job(jk, n, n, !MS));
    LDDMM_TRY(plan_job(jd, n, n, !MS));
    if (!MS) LDDMM_TRY(plan_job(jf, n, n, true));
    const bool aff = kd->affine != LDDMM_AFFINE_NONE;
    const AffineArgs AA = affine_args(kd);
    void* ws;
    LDDMM_TRY(workspace(256 + (aff ? kEpdAffBytes + 256 : 0) +
                            std::max({partial_bytes<T>(jk.plan, KW, n), partial_bytes<T>(jd.plan, 3, n),
                                      MS ? size_t(0) : partial_bytes<T>(jf.plan, KW + 3, n)}) + 512,
                        st, &ws));
    char* base = static_cast<char*>(ws);
    Carve cv;
    double* c0 = reinterpret_cast<double*>(base + cv.take(32));
    double* aws = aff ? reinterpret_cast<double*>(base + cv.take(kEpdAffBytes)) : nullptr;
    size_t poff = cv.off;
    const size_t stride = size_t(n) * KW;
    const double dt = 1.0 / T_;
    int tpw = 0;                                              // > 0: one-launch steps
    if (!MS && int64_t(n) <= kStepMaxN) {
        DeviceInfo di;
        LDDMM_TRY(device_info(&di));
        tpw = step_targets(n, di.num_cu);
    }
    LDDMM_TRY(launch_center(xt, n, xt, 0, KW, c0, st));       // one centre (x_0) for all steps
    for (int k = 0; k < T_; ++k) {
        const double* x = xt + size_t(k) * stride;
        const double* a = at + size_t(k) * stride;
        double* xn = xt + size_t(k + 1) * stride;
        double* an = at + size_t(k + 1) * stride;
        if constexpr (!MS) {
            // one Gaussian scale: both sums of a step in one pass (OpEPDiffFwd);
            // small sets in one launch per step, larger ones chunked + one reduce
            jf.args = {x, a, c0, sc};
            if (tpw > 0) {
                hipLaunchKernelGGL((k_step_epdiff_fwd<T, KW>), dim3(unsigned((n + tpw - 1) / tpw)),
                                   dim3(kStepThreads), 0, st, jf.args, n, tpw, dt, fdiff, x, a, xn, an);
                LDDMM_CHECK_LAUNCH();
                if (aff) LDDMM_TRY(epd_aff_step(AA, n, KW, 0, dt, x, a, nullptr, nullptr, xn, an, aws, st));
                continue;
            }
            Carve c1; c1.off = poff;
            Partials<T> pf = carve_partials<T>(base, c1, jf.plan, KW + 3, n);
            LDDMM_TRY(run_job(jf, n, n, pf, st));
            hipLaunchKernelGGL((k_reduce_epdiff_fwd<T, KW>), dim3(unsigned((n + kEpiRows - 1) / kEpiRows), 2 * KW),
                               dim3(256), 0, st, pf, jf.plan.nchunk, n, dt, fdiff, x, a, xn, an);
            LDDMM_CHECK_LAUNCH();
        } else {
            Carve c1; c1.off = poff;
            Partials<T> pk = carve_partials<T>(base, c1, jk.plan, KW, n);
            jk.args = {x, x, a, c0, KW, KW, 0, sc};
            LDDMM_TRY(run_job(jk, n, n, pk, st));
            LDDMM_TRY(reduce_job(jk, pk, n, KW, 1.0, xn, 0, x, dt, st));
            Carve c2; c2.off = poff;
            Partials<T> pd = carve_partials<T>(base, c2, jd.plan, 3, n);
            jd.args = {x, x, a, a, c0, KW, KW, 0, 0.0, sc};
            LDDMM_TRY(run_job(jd, n, n, pd, st));
            LDDMM_TRY(reduce_job(jd, pd, n, KW, fdiff, an, 0, a, -dt, st));
        }
        if (aff) LDDMM_TRY(epd_aff_step(AA, n, KW, 0, dt, x, a, nullptr, nullptr, xn, an, aws, st));
    }
    return LDDMM_OK;
}

template <typename T, bool MS, int KW>
int epdiff_adjoint_t(const lddmm_kernel* kd, int T_, int n, const double* xt, const double* at, const double* px1,
                     const double* pa1, double* pxt, double* pat, double* da0, hipStream_t st) {
    SingleScaleConsts ss;
    ScaleSet<T> sc = make_scales<T>(kd, &ss);
    // Kernel.applyDiffK1and2 reaches applykdiff1and2, which always uses KP = -1
    lddmm_kernel kd1 = *kd;
    kd1.kp = -1.0;
    SingleScaleConsts ss1;
    ScaleSet<T> sc1 = make_scales<T>(&kd1, &ss1);
    const bool ms1 = MS;
    using OPDs = OpDiffKT<KW, true, MS>;
    using OPDD = OpDDiffK<KW, LDDMM_DDIFFK_11AND12, MS>;
    using OPK = OpApplyK<KW, MS>;
    using OPD12 = OpDiffK<KW, LDDMM_DIFFK_1AND2, MS>;
    Job<T, OPDs> j1;
    Job<T, OPDD> j2;
    Job<T, OPK> j3;
    Job<T, OPD12> j4;
    using OPA = OpEPDiffAdj<KW>;
    Job<T, OPA> ja;
    LDDMM_TRY(plan_job(j1, n, n, !MS));
    LDDMM_TRY(plan_job(j2, n, n, !MS));
    LDDMM_TRY(plan_job(j3, n, n, !MS));
    LDDMM_TRY(plan_job(j4, n, n, !MS));
    if (!MS) LDDMM_TRY(plan_job(ja, n, n, true));
    size_t pb = std::max(std::max(partial_bytes<T>(j1.plan, 3, n), partial_bytes<T>(j2.plan, 3, n)),
                         std::max(partial_bytes<T>(j3.plan, KW, n), partial_bytes<T>(j4.plan, KW, n)));
    if (!MS) pb = std::max(pb, partial_bytes<T>(ja.plan, 6 + 2 * KW, n));
    const bool aff = kd->affine != LDDMM_AFFINE_NONE;
    const AffineArgs AA = affine_args(kd);
    void* ws;
    LDDMM_TRY(workspace(256 + (aff ? kEpdAffBytes + 256 : 0) + pb + 512, st, &ws));
    char* base = static_cast<char*>(ws);
    Carve cv;
    double* c0 = reinterpret_cast<double*>(base + cv.take(32));
    double* aws = aff ? reinterpret_cast<double*>(base + cv.take(kEpdAffBytes)) : nullptr;
    size_t poff = cv.off;
    const size_t stride = size_t(n) * KW;
    const double dt = 1.0 / T_;
    LDDMM_HIP(hipMemcpyAsync(pxt + size_t(T_) * stride, px1, stride * sizeof(double), hipMemcpyDeviceToDevice, st));
    LDDMM_HIP(hipMemcpyAsync(pat + size_t(T_) * stride, pa1, stride * sizeof(double), hipMemcpyDeviceToDevice, st));
    const double fdiff = MS ? -1.0 : ss.diff;
    const double fdd = MS ? 1.0 : ss.ddiff;
    const double fd1 = ms1 ? -1.0 : ss1.diff;
    LDDMM_TRY(launch_center(xt, n, xt, 0, KW, c0, st));       // one centre (x_0) for all steps
    for (int t = T_ - 1; t >= 0; --t) {
        const double* x = xt + size_t(t) * stride;
        const double* a = at + size_t(t) * stride;
        const double* px = pxt + size_t(t + 1) * stride;
        const double* pa = pat + size_t(t + 1) * stride;
        double* pxo = pxt + size_t(t) * stride;
        double* pao = pat + size_t(t) * stride;
        Carve c; c.off = poff;
        if constexpr (!MS) {
            // one Gaussian scale: the four sums in one pass (OpEPDiffAdj) + one
            // reduce.  At registration sizes this is as fast as a one-launch
            // step: the op's 12-value source records make the step kernel's
            // per-lane LDS reads its limit.
            Partials<T> pa4 = carve_partials<T>(base, c, ja.plan, 6 + 2 * KW, n);
            ja.args = {x, a, px, pa, c0, T(ss.gamma), sc};
            LDDMM_TRY(run_job(ja, n, n, pa4, st));
            hipLaunchKernelGGL((k_reduce_epdiff_adj<T, KW>), dim3(unsigned((n + kEpiRows - 1) / kEpiRows), 2 * KW),
                               dim3(256), 0, st, pa4, ja.plan.nchunk, n, dt, fdiff, fdd, fd1, px, pa, pxo, pao);
            LDDMM_CHECK_LAUNCH();
            if (aff) LDDMM_TRY(epd_aff_step(AA, n, KW, 1, dt, x, a, px, pa, pxo, pao, aws, st));
            continue;
        }
        Partials<T> part = carve_partials<T>(base, c, j1.plan, 3, n);
        // zpx = DiffKT_sym(x, px, a) - DDiffK11and12(x, a, a, pa)
        j1.args = {x, x, px, a, c0, KW, KW, 0, 0.0, sc};
        LDDMM_TRY(run_job(j1, n, n, part, st));
        LDDMM_TRY(reduce_job(j1, part, n, KW, fdiff, pxo, 0, px, dt, st));
        c.off = poff;
        part = carve_partials<T>(base, c, j2.plan, 3, n);
        j2.args = {x, a, a, pa, c0, KW, KW, 0, T(ss.gamma), sc};
        LDDMM_TRY(run_job(j2, n, n, part, st));
        LDDMM_TRY(reduce_job(j2, part, n, KW, fdd, pxo, 1, nullptr, -dt, st));
        // zpa = K(x) px - DiffK1and2(x, pa, a)
        c.off = poff;
        part = carve_partials<T>(base, c, j3.plan, KW, n);
        j3.args = {x, x, px, c0, KW, KW, 0, sc};
        LDDMM_TRY(run_job(j3, n, n, part, st));
        LDDMM_TRY(reduce_job(j3, part, n, KW, 1.0, pao, 0, pa, dt, st));
        c.off = poff;
        part = carve_partials<T>(base, c, j4.plan, KW, n);
        j4.args = {x, pa, a, c0, KW, KW, 0, sc1};
        LDDMM_TRY(run_job(j4, n, n, part, st));
        LDDMM_TRY(reduce_job(j4, part, n, KW, fd1, pao, 1, nullptr, -dt, st));
        if (aff) LDDMM_TRY(epd_aff_step(AA, n, KW, 1, dt, x, a, px, pa, pxo, pao, aws, st));
    }
    // da0 = K(x0) a0 - pa_0
    LDDMM_TRY(axpby(stride, -1.0, pat, 0.0, nullptr, da0, st));
    Carve c; c.off = poff;
    Partials<T> part = carve_partials<T>(base, c, j3.plan, KW, n);
    j3.args = {xt, xt, at, c0, KW, KW, 0, sc};
    LDDMM_TRY(run_job(j3, n, n, part, st));
    LDDMM_TRY(reduce_job(j3, part, n, KW, 1.0, da0, 1, nullptr, 1.0, st));
    if (aff) {
        // the affine part of K(x0) a0
        LDDMM_TRY(aff_moments(AA, n, KW, xt, at, nullptr, aws, nullptr, 0.0, st));
        LDDMM_TRY(aff_apply(AA, n, KW, 0, 0.0, aws, xt, nullptr, nullptr, 1.0, da0, st));
    }
    return LDDMM_OK;
}

}  // namespace
}  // namespace lddmm

using namespace lddmm;

#define LDDMM_DISPATCH(FN, ...)                                                          \
    do {                                                                                 \
        const bool ms_ = general_path(kd);                                                 \
        const bool f32_ = kd->prec == LDDMM_PREC_F32C;                                   \
        if (dim == 3) {                                                                  \
            if (f32_) return ms_ ? FN<float, true, 3>(__VA_ARGS__) : FN<float, false, 3>(__VA_ARGS__);    \
            return ms_ ? FN<double, true, 3>(__VA_ARGS__) : FN<double, false, 3>(__VA_ARGS__);            \
        }                                                                                \
        if (f32_) return ms_ ? FN<float, true, 2>(__VA_ARGS__) : FN<float, false, 2>(__VA_ARGS__);        \
        return ms_ ? FN<double, true, 2>(__VA_ARGS__) : FN<double, false, 2>(__VA_ARGS__);                \
    } while (0)

extern "C" int lddmm_shoot_forward(const lddmm_kernel* kd, const double* at, int32_t T, int64_t n, int32_t dim,
                                   double* xt, double regweight, double* energy, void* stream) {
    LDDMM_TRY(check_shoot(kd, T, n, dim, kValAffine));
    hipStream_t st = static_cast<hipStream_t>(stream);
    if (n == 0) return energy ? launch_fill(energy, 1, 0.0, st) : LDDMM_OK;
    LDDMM_REQUIRE(at && xt, "null at / xt");
    LDDMM_DISPATCH(shoot_forward_t, kd, at, T, int(n), xt, regweight, energy, nullptr, st);
}

extern "C" int lddmm_shoot_adjoint(const lddmm_kernel* kd, const double* at, const double* xt, int32_t T,
                                   int64_t n, int32_t dim, const double* px1, double regweight, int32_t euclidean,
                                   double* pxt, double* grad, void* stream) {
    LDDMM_TRY(check_shoot(kd, T, n, dim, kValAffine));
    if (n == 0) return LDDMM_OK;
    LDDMM_REQUIRE(at && xt && px1 && grad, "null pointer");
    hipStream_t st = static_cast<hipStream_t>(stream);
    LDDMM_DISPATCH(shoot_adjoint_t, kd, at, xt, T, int(n), px1, regweight, euclidean, pxt, grad, nullptr, st);
}

namespace {
int ic_ctx(const lddmm_surface* s, const lddmm_internal_cost* ic, bool need_adj, IcCtx* out) {
    LDDMM_TRY(ic_check(ic, &out->L));
    LDDMM_REQUIRE(s != nullptr, "internal cost: null surface");
    LDDMM_REQUIRE(s->nv >= 0 && s->nf >= 0 && s->nf < (int64_t(1) << 30) && s->nv < (int64_t(1) << 31),
                  "internal cost: bad surface sizes");
    LDDMM_REQUIRE(s->nf == 0 || s->faces, "internal cost: null faces");
    LDDMM_REQUIRE(!need_adj || s->nv == 0 || (s->adj_ptr && s->adj_idx), "internal cost: adjacency required");
    out->faces = s->faces;
    out->nf = int(s->nf);
    out->adj_ptr = s->adj_ptr;
    out->adj_idx = s->adj_idx;
    return LDDMM_OK;
}
}  // namespace

extern "C" int lddmm_shoot_forward_ic(const lddmm_kernel* kd, const lddmm_surface* s, const lddmm_internal_cost* ic,
                                      const double* at, int32_t T, double* xt, double regweight, double* energy,
                                      void* stream) {
    IcCtx C;
    LDDMM_TRY(ic_ctx(s, ic, false, &C));
    const int64_t n = s->nv;
    const int dim = 3;
    LDDMM_TRY(check_shoot(kd, T, n, dim, kValAffine));
    hipStream_t st = static_cast<hipStream_t>(stream);
    if (n == 0) return energy ? launch_fill(energy, 1, 0.0, st) : LDDMM_OK;
    LDDMM_REQUIRE(at && xt, "null at / xt");
    LDDMM_DISPATCH(shoot_forward_t, kd, at, T, int(n), xt, regweight, energy, &C, st);
}

extern "C" int lddmm_shoot_adjoint_ic(const lddmm_kernel* kd, const lddmm_surface* s, const lddmm_internal_cost* ic,
                                      const double* at, const double* xt, int32_t T, const double* px1,
                                      double regweight, int32_t euclidean, double* pxt, double* grad, void* stream) {
    IcCtx C;
    LDDMM_TRY(ic_ctx(s, ic, true, &C));
    const int64_t n = s->nv;
    const int dim = 3;
    LDDMM_TRY(check_shoot(kd, T, n, dim, kValAffine));
    if (n == 0) return LDDMM_OK;
    LDDMM_REQUIRE(at && xt && px1 && grad, "null pointer");
    hipStream_t st = static_cast<hipStream_t>(stream);
    LDDMM_DISPATCH(shoot_adjoint_t, kd, at, xt, T, int(n), px1, regweight, euclidean, pxt, grad, &C, st);
}

extern "C" int lddmm_epdiff_forward(const lddmm_kernel* kd, int32_t T, int64_t n, int32_t dim, double* xt,
                                    double* at, void* stream) {
    LDDMM_TRY(check_shoot(kd, T, n, dim, kValAffine));
    if (n == 0) return LDDMM_OK;
    LDDMM_REQUIRE(xt && at, "null pointer");
    hipStream_t st = static_cast<hipStream_t>(stream);
    LDDMM_DISPATCH(epdiff_forward_t, kd, T, int(n), xt, at, st);
}

extern "C" int lddmm_epdiff_adjoint(const lddmm_kernel* kd, int32_t T, int64_t n, int32_t dim, const double* xt,
                                    const double* at, const double* px1, const double* pa1, double* pxt,
                                    double* pat, double* da0, void* stream) {
    LDDMM_TRY(check_shoot(kd, T, n, dim, kValAffine));
    if (n == 0) return LDDMM_OK;
    LDDMM_REQUIRE(xt && at && px1 && pa1 && pxt && pat && da0, "null pointer");
    hipStream_t st = static_cast<hipStream_t>(stream);
    LDDMM_DISPATCH(epdiff_adjoint_t, kd, T, int(n), xt, at, px1, pa1, pxt, pat, da0, st);
}
