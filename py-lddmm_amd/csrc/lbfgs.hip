// L-BFGS two-loop recursion on the device (py-lddmm base/bfgs.py:188-207).
//
// Each step of the recursion is one launch that applies the update of q
// elementwise and accumulates the next dot product on the updated values.
// A launch writes only its per-workgroup partial sums; the launch after it,
// which is needed anyway, starts by having every workgroup add those partials
// in one fixed tree order, multiply by rho and use the scalar (workgroup 0
// also stores it where later steps read it).  Kernel boundaries make the
// partials visible, so no fences, atomics or counters are needed, and the
// reduction order depends only on n.  The host issues 2m + 1 launches for m
// stored pairs and never waits.
#include "common.hpp"

#include <mutex>
#include <unordered_map>
#include <vector>

namespace lddmm {
namespace {

constexpr int kLbBlock = 256;
constexpr int kLbGrid = 256;        // fixed: the reduction order depends only on n
constexpr int kLbMaxPairs = 1024;
static_assert(kLbGrid == kLbBlock, "a workgroup reduces one partial per lane");

// Per-device state: two [kLbGrid] partial buffers (alternating launches),
// [2*kLbMaxPairs + 8] scalars.
constexpr int kLbcMaxPairs = 64;        // compact form: stored pairs per call
constexpr int kLbcMaxProbes = 2;
constexpr int kLbcMaxTargets = 2 * kLbcMaxPairs;
constexpr int kLbcMaxSlices = 512;
constexpr int kLbcGroup = 8;            // targets per multidot workgroup

struct LbState {
    double* partials[2] = {nullptr, nullptr};
    double* scal = nullptr;
    double* mdpart = nullptr;           // [kLbcMaxSlices][kLbcMaxProbes][kLbcMaxTargets]
    double* coef = nullptr;             // [2 * kLbcMaxPairs + 1]
};
std::mutex g_lb_mutex;
std::unordered_map<int, LbState> g_lb;

int lb_state(LbState* out) {
    int dev = -1;
    LDDMM_HIP(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lk(g_lb_mutex);
    auto it = g_lb.find(dev);
    if (it == g_lb.end()) {
        LbState s;
        for (auto& p : s.partials) LDDMM_HIP(hipMalloc(reinterpret_cast<void**>(&p), kLbGrid * sizeof(double)));
        LDDMM_HIP(hipMalloc(reinterpret_cast<void**>(&s.scal), (2 * kLbMaxPairs + 8) * sizeof(double)));
        LDDMM_HIP(hipMalloc(reinterpret_cast<void**>(&s.mdpart),
                            size_t(kLbcMaxSlices) * kLbcMaxProbes * kLbcMaxTargets * sizeof(double)));
        LDDMM_HIP(hipMalloc(reinterpret_cast<void**>(&s.coef), (2 * kLbcMaxPairs + 1) * sizeof(double)));
        it = g_lb.emplace(dev, s).first;
    }
    *out = it->second;
    return LDDMM_OK;
}

// Fixed-order sum of red[0 .. kLbBlock) into red[0] (all lanes call it):
// each wave adds its 64 values by a butterfly, then the wave sums are added
// in wave order.
__device__ __forceinline__ void lb_tree(double* red) {
    __syncthreads();
    double v = red[threadIdx.x];
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
    __syncthreads();
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        double s = 0.0;
#pragma unroll
        for (int w = 0; w < kLbBlock / 64; ++w) s += red[w];
        red[0] = s;
    }
    __syncthreads();
}

// Scalars the step uses: the previous launch's dot (prev = *prev_mul * sum of
// prev_part, stored to *prev_store by workgroup 0) and stored ones.
struct LbStep {
    const double* prev_part;    // partials of the previous launch (null: no previous dot)
    const double* prev_mul;     // rho of the previous dot (null: 1)
    double* prev_store;         // where workgroup 0 stores it (may be null)
    const double* u;            // q += coef * u (null: no update)
    const double* a;            // coef = sign * (A - B); A = *a, or prev when a_prev
    int a_prev;
    int b_prev;                 // B = prev when b_prev, else 0
    double sign;
    const double* sc;           // then q *= *sc (null: no scaling)
    const double* d;            // partial sums of <d, q> into part_out (null: no dot)
    double* part_out;
};

__global__ void __launch_bounds__(kLbBlock) k_lb_step(int64_t n, double* __restrict__ q, LbStep st) {
    __shared__ double red[kLbBlock];
    double prev = 0.0;
    if (st.prev_part) {
        red[threadIdx.x] = st.prev_part[threadIdx.x];
        lb_tree(red);
        prev = (st.prev_mul ? *st.prev_mul : 1.0) * red[0];
        if (blockIdx.x == 0 && threadIdx.x == 0 && st.prev_store) *st.prev_store = prev;
        __syncthreads();
    }
    double coef = 0.0;
    if (st.u) coef = st.sign * ((st.a_prev ? prev : *st.a) - (st.b_prev ? prev : 0.0));
    const double s = st.sc ? *st.sc : 1.0;
    double acc = 0.0;
    for (int64_t i = int64_t(blockIdx.x) * kLbBlock + threadIdx.x; i < n; i += int64_t(gridDim.x) * kLbBlock) {
        double qi = q[i];
        if (st.u) qi = fma(coef, st.u[i], qi);
        if (st.sc) qi *= s;
        if (st.u || st.sc) q[i] = qi;
        if (st.d) acc = fma(st.d[i], qi, acc);
    }
    if (!st.d) return;
    red[threadIdx.x] = acc;
    lb_tree(red);
    if (threadIdx.x == 0) st.part_out[blockIdx.x] = red[0];
}

// rho = 1 / <y, s>, c = <s, y> / <y, y> (c = 1 when below 1e-10), from the
// partials of the two dots (same tree order as the steps).
__global__ void __launch_bounds__(kLbBlock) k_lb_pair_scalars(const double* pys, const double* pyy, double* rho,
                                                              double* c) {
    __shared__ double red[kLbBlock];
    red[threadIdx.x] = pys[threadIdx.x];
    lb_tree(red);
    const double ys = red[0];
    __syncthreads();
    red[threadIdx.x] = pyy[threadIdx.x];
    lb_tree(red);
    const double yy = red[0];
    if (threadIdx.x == 0) {
        *rho = 1.0 / ys;
        double v = ys / yy;
        *c = v < 1e-10 ? 1.0 : v;
    }
}

// ---------------------------------------------------------------------------
// Compact form of the same inverse-Hessian product (Byrd, Nocedal and Schnabel,
// "Representations of quasi-Newton matrices and their use in limited memory
// methods", Math. Programming 63, 1994): with S = [s_0 .. s_{m-1}],
// Y = [y_0 .. y_{m-1}] (oldest first), R = upper triangle of S^T Y (diagonal
// included), D = diag(s_i . y_i) and H0 = gamma I,
//   H g = gamma g + S w - gamma Y u,   u = R^{-1} S^T g,
//   w = R^{-T} ((D + gamma Y^T Y) u - gamma Y^T g).
// Equal to the two-loop recursion in exact arithmetic; it needs the m x m
// Gram blocks (cached per stored pair by the caller), S^T g and Y^T g (one
// batched dot launch) and one pass over the vectors.
// ---------------------------------------------------------------------------
struct LbcPtrs {
    const double* v[kLbcMaxTargets];
};

// Pass 1 of the batched dots: workgroup (slice, group) adds, for up to two
// probes and the group's kLbcGroup targets, the products over its slice of
// the vectors; lanes are combined by a fixed tree.  part[(slice * np + p) * nt + t].
__global__ void __launch_bounds__(kLbBlock) k_lbc_dots(LbcPtrs probes, int np, LbcPtrs targets, int nt, int64_t n,
                                                       int64_t slice, double* __restrict__ part) {
    constexpr int kWaves = kLbBlock / 64;
    __shared__ double red[kWaves][kLbcMaxProbes * kLbcGroup];
    const int t0 = blockIdx.y * kLbcGroup;
    const int64_t e0 = int64_t(blockIdx.x) * slice, e1 = e0 + slice < n ? e0 + slice : n;
    double acc[kLbcMaxProbes][kLbcGroup];
#pragma unroll
    for (int p = 0; p < kLbcMaxProbes; ++p)
#pragma unroll
        for (int k = 0; k < kLbcGroup; ++k) acc[p][k] = 0.0;
    for (int64_t e = e0 + threadIdx.x; e < e1; e += kLbBlock) {
        const double a0 = probes.v[0][e];
        const double a1 = np > 1 ? probes.v[1][e] : 0.0;
#pragma unroll
        for (int k = 0; k < kLbcGroup; ++k) {
            if (t0 + k < nt) {
                const double b = targets.v[t0 + k][e];
                acc[0][k] = fma(a0, b, acc[0][k]);
                acc[1][k] = fma(a1, b, acc[1][k]);
            }
        }
    }
    // each wave adds its lanes by a butterfly, then the waves' sums are
    // added in wave order (a fixed order)
#pragma unroll
    for (int p = 0; p < kLbcMaxProbes; ++p)
#pragma unroll
        for (int k = 0; k < kLbcGroup; ++k) {
            double v = acc[p][k];
#pragma unroll
            for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
            if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6][p * kLbcGroup + k] = v;
        }
    __syncthreads();
    if (threadIdx.x < kLbcMaxProbes * kLbcGroup) {
        const int p = threadIdx.x / kLbcGroup, k = threadIdx.x % kLbcGroup;
        double v = 0.0;
#pragma unroll
        for (int w = 0; w < kWaves; ++w) v += red[w][threadIdx.x];
        if (p < np && t0 + k < nt) part[(size_t(blockIdx.x) * np + p) * nt + t0 + k] = v;
    }
}

// Pass 2: out[p * nt + t] = sum over slices (one workgroup per dot, fixed tree).
__global__ void __launch_bounds__(kLbBlock) k_lbc_dots_fold(const double* __restrict__ part, int nslice, int ndot,
                                                            double* __restrict__ out) {
    __shared__ double red[kLbBlock];
    const int d = blockIdx.x;
    double acc = 0.0;
    for (int sl = threadIdx.x; sl < nslice; sl += kLbBlock) acc += part[size_t(sl) * ndot + d];
    red[threadIdx.x] = acc;
    lb_tree(red);
    if (threadIdx.x == 0) out[d] = red[0];
}

// Column j of the cached Gram blocks: cols.v[j] holds 2 * len[j] dots of y_j
// with s and y of the pairs up to j (window order when j was stored, j last).
struct LbcCols {
    const double* v[kLbcMaxPairs];
    int len[kLbcMaxPairs];
};

// One wave: the m x m triangular solves; coef = [w (m), -gamma u (m), gamma].
__global__ void __launch_bounds__(64) k_lbc_solve(int m, LbcCols cols, const double* __restrict__ gdots,
                                                  double* __restrict__ coef) {
    __shared__ double R[kLbcMaxPairs][kLbcMaxPairs + 1];
    __shared__ double YY[kLbcMaxPairs][kLbcMaxPairs + 1];
    __shared__ double u[kLbcMaxPairs], b[kLbcMaxPairs], w[kLbcMaxPairs];
    const int t = threadIdx.x;
    if (t < m) {
        const int L = cols.len[t];
        const double* c = cols.v[t];
        for (int i = 0; i <= t; ++i) {
            const int o = L - 1 - (t - i);
            R[i][t] = c[o];
            YY[i][t] = c[L + o];
            YY[t][i] = c[L + o];
        }
        b[t] = gdots[t];                       // S^T g
    }
    __syncthreads();
    double gamma = R[m - 1][m - 1] / YY[m - 1][m - 1];
    if (gamma < 1e-10) gamma = 1.0;
    // u = R^{-1} S^T g (back substitution, column by column)
    for (int i = m - 1; i >= 0; --i) {
        if (t == 0) u[i] = b[i] / R[i][i];
        __syncthreads();
        if (t < i) b[t] -= R[t][i] * u[i];
        __syncthreads();
    }
    // b = (D + gamma Y^T Y) u - gamma Y^T g, then w = R^{-T} b (forward substitution)
    if (t < m) {
        double yu = 0.0;
        for (int j = 0; j < m; ++j) yu = fma(YY[t][j], u[j], yu);
        b[t] = R[t][t] * u[t] + gamma * yu - gamma * gdots[m + t];
    }
    __syncthreads();
    for (int i = 0; i < m; ++i) {
        if (t == 0) w[i] = b[i] / R[i][i];
        __syncthreads();
        if (t > i && t < m) b[t] -= R[i][t] * w[i];
        __syncthreads();
    }
    if (t < m) {
        coef[t] = w[t];
        coef[m + t] = -gamma * u[t];
    }
    if (t == 0) coef[2 * m] = gamma;
}

// q = gamma g + sum_i coef[i] s_i + sum_i coef[m + i] y_i  (q may alias g).
__global__ void __launch_bounds__(kLbBlock) k_lbc_combine(int64_t n, int m, const double* g, LbcPtrs sy,
                                                          const double* __restrict__ coef, double* q) {
    __shared__ double cf[kLbcMaxTargets + 1];
    for (int k = threadIdx.x; k <= 2 * m; k += kLbBlock) cf[k] = coef[k];
    __syncthreads();
    const double gamma = cf[2 * m];
    for (int64_t e = int64_t(blockIdx.x) * kLbBlock + threadIdx.x; e < n; e += int64_t(gridDim.x) * kLbBlock) {
        double v = gamma * g[e];
        for (int k = 0; k < 2 * m; ++k) v = fma(cf[k], sy.v[k][e], v);
        q[e] = v;
    }
}

int lb_launch(int64_t n, double* q, const LbStep& s, hipStream_t st) {
    // a step with a dot uses the fixed grid (its partials feed the next launch)
    int64_t blocks = (n + kLbBlock - 1) / kLbBlock;
    unsigned grid = unsigned(s.d ? kLbGrid : std::max<int64_t>(1, std::min<int64_t>(blocks, 4096)));
    hipLaunchKernelGGL(k_lb_step, dim3(grid), dim3(kLbBlock), 0, st, n, q, s);
    LDDMM_CHECK_LAUNCH();
    return LDDMM_OK;
}

}  // namespace
}  // namespace lddmm

using namespace lddmm;

extern "C" int lddmm_lbfgs_pair_scalars(const double* s, const double* y, int64_t n, double* rho, double* c,
                                        void* stream) {
    LDDMM_REQUIRE(n >= 1 && s && y && rho && c, "lbfgs_pair_scalars: bad arguments");
    LbState S;
    LDDMM_TRY(lb_state(&S));
    hipStream_t st = static_cast<hipStream_t>(stream);
    // <y, s> and <y, y>: q is not modified without u / sc, so y serves as q
    LbStep a{};
    a.d = s;
    a.part_out = S.partials[0];
    LDDMM_TRY(lb_launch(n, const_cast<double*>(y), a, st));
    LbStep b{};
    b.d = y;
    b.part_out = S.partials[1];
    LDDMM_TRY(lb_launch(n, const_cast<double*>(y), b, st));
    hipLaunchKernelGGL(k_lb_pair_scalars, dim3(1), dim3(kLbBlock), 0, st, S.partials[0], S.partials[1], rho, c);
    LDDMM_CHECK_LAUNCH();
    return LDDMM_OK;
}

extern "C" int lddmm_lbfgs_two_loop(const double* g, int64_t n, int32_t m, const double* const* s,
                                    const double* const* y, const double* const* rho, const double* c_newest,
                                    double* q, void* stream) {
    LDDMM_REQUIRE(n >= 1 && g && q, "lbfgs_two_loop: bad arguments");
    LDDMM_REQUIRE(m >= 0 && m <= kLbMaxPairs, "lbfgs_two_loop: m must be in [0, %d]", kLbMaxPairs);
    LDDMM_REQUIRE(m == 0 || (s && y && rho && c_newest), "lbfgs_two_loop: null history");
    for (int k = 0; k < m; ++k) LDDMM_REQUIRE(s[k] && y[k] && rho[k], "lbfgs_two_loop: null history entry %d", k);
    LbState S;
    LDDMM_TRY(lb_state(&S));
    hipStream_t st = static_cast<hipStream_t>(stream);
    if (q != g) LDDMM_HIP(hipMemcpyAsync(q, g, size_t(n) * sizeof(double), hipMemcpyDeviceToDevice, st));
    if (m == 0) return LDDMM_OK;
    double* alpha = S.scal;                 // alpha[k]
    double* beta = S.scal + kLbMaxPairs;    // beta[k] (kept for inspection)
    int buf = 0;                            // partial buffer the next dot writes
    const double* last = nullptr;           // partials of the previous launch's dot
    // first loop, newest to oldest: alpha_k = rho_k <s_k, q>; q -= alpha_k y_k
    for (int k = m - 1; k >= 0; --k) {
        LbStep p{};
        if (k < m - 1) {                    // q -= alpha_{k+1} y_{k+1}, alpha_{k+1} from the previous dot
            p.prev_part = last; p.prev_mul = rho[k + 1]; p.prev_store = alpha + k + 1;
            p.u = y[k + 1]; p.a_prev = 1; p.sign = -1.0;
        }
        p.d = s[k];
        p.part_out = S.partials[buf];
        LDDMM_TRY(lb_launch(n, q, p, st));
        last = S.partials[buf];
        buf ^= 1;
    }
    // q -= alpha_0 y_0; q *= c; beta_0 = rho_0 <y_0, q>
    {
        LbStep p{};
        p.prev_part = last; p.prev_mul = rho[0]; p.prev_store = alpha;
        p.u = y[0]; p.a_prev = 1; p.sign = -1.0;
        p.sc = c_newest;
        p.d = y[0];
        p.part_out = S.partials[buf];
        LDDMM_TRY(lb_launch(n, q, p, st));
        last = S.partials[buf];
        buf ^= 1;
    }
    // second loop, oldest to newest: q += (alpha_k - beta_k) s_k; beta_{k+1} = rho_{k+1} <y_{k+1}, q>
    for (int k = 0; k < m; ++k) {
        LbStep p{};
        p.prev_part = last; p.prev_mul = rho[k]; p.prev_store = beta + k;
        p.u = s[k]; p.a = alpha + k; p.b_prev = 1; p.sign = 1.0;
        if (k + 1 < m) {
            p.d = y[k + 1];
            p.part_out = S.partials[buf];
        }
        LDDMM_TRY(lb_launch(n, q, p, st));
        last = S.partials[buf];
        buf ^= 1;
    }
    return LDDMM_OK;
}

extern "C" int lddmm_lbfgs_multidot(const double* const* probes, int32_t np, const double* const* targets,
                                    int32_t nt, int64_t n, double* out, void* stream) {
    LDDMM_REQUIRE(n >= 1 && probes && targets && out, "lbfgs_multidot: bad arguments");
    LDDMM_REQUIRE(np >= 1 && np <= kLbcMaxProbes, "lbfgs_multidot: np must be in [1, %d]", kLbcMaxProbes);
    LDDMM_REQUIRE(nt >= 1 && nt <= kLbcMaxTargets, "lbfgs_multidot: nt must be in [1, %d]", kLbcMaxTargets);
    LbcPtrs P{}, T{};
    for (int p = 0; p < np; ++p) { LDDMM_REQUIRE(probes[p], "lbfgs_multidot: null probe"); P.v[p] = probes[p]; }
    for (int t = 0; t < nt; ++t) { LDDMM_REQUIRE(targets[t], "lbfgs_multidot: null target"); T.v[t] = targets[t]; }
    LbState S;
    LDDMM_TRY(lb_state(&S));
    hipStream_t st = static_cast<hipStream_t>(stream);
    // slices of whole workgroup passes; at most kLbcMaxSlices of them
    int64_t slice = std::max<int64_t>(4 * kLbBlock, (n + kLbcMaxSlices - 1) / kLbcMaxSlices);
    slice = (slice + kLbBlock - 1) / kLbBlock * kLbBlock;
    const int nslice = int((n + slice - 1) / slice);
    dim3 grid(unsigned(nslice), unsigned((nt + kLbcGroup - 1) / kLbcGroup));
    hipLaunchKernelGGL(k_lbc_dots, grid, dim3(kLbBlock), 0, st, P, int(np), T, int(nt), n, slice, S.mdpart);
    LDDMM_CHECK_LAUNCH();
    hipLaunchKernelGGL(k_lbc_dots_fold, dim3(unsigned(np * nt)), dim3(kLbBlock), 0, st, S.mdpart, nslice,
                       int(np * nt), out);
    LDDMM_CHECK_LAUNCH();
    return LDDMM_OK;
}

extern "C" int lddmm_lbfgs_compact(const double* g, int64_t n, int32_t m, const double* const* s,
                                   const double* const* y, const double* const* cols, const int32_t* col_len,
                                   const double* gdots, double* q, void* stream) {
    LDDMM_REQUIRE(n >= 1 && g && q && gdots, "lbfgs_compact: bad arguments");
    LDDMM_REQUIRE(m >= 1 && m <= kLbcMaxPairs, "lbfgs_compact: m must be in [1, %d]", kLbcMaxPairs);
    LDDMM_REQUIRE(s && y && cols && col_len, "lbfgs_compact: null history");
    LbcPtrs V{};
    LbcCols C{};
    for (int k = 0; k < m; ++k) {
        LDDMM_REQUIRE(s[k] && y[k] && cols[k], "lbfgs_compact: null history entry %d", k);
        LDDMM_REQUIRE(col_len[k] >= k + 1, "lbfgs_compact: column %d holds %d pairs, needs %d", k, col_len[k], k + 1);
        V.v[k] = s[k];
        V.v[m + k] = y[k];
        C.v[k] = cols[k];
        C.len[k] = col_len[k];
    }
    LbState S;
    LDDMM_TRY(lb_state(&S));
    hipStream_t st = static_cast<hipStream_t>(stream);
    hipLaunchKernelGGL(k_lbc_solve, dim3(1), dim3(64), 0, st, int(m), C, gdots, S.coef);
    LDDMM_CHECK_LAUNCH();
    const int64_t blocks = std::max<int64_t>(1, std::min<int64_t>((n + kLbBlock - 1) / kLbBlock, 4096));
    hipLaunchKernelGGL(k_lbc_combine, dim3(unsigned(blocks)), dim3(kLbBlock), 0, st, n, int(m), g, V, S.coef, q);
    LDDMM_CHECK_LAUNCH();
    return LDDMM_OK;
}
