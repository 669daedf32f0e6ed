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
struct LbState {
    double* partials[2] = {nullptr, nullptr};
    double* scal = nullptr;
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
        it = g_lb.emplace(dev, s).first;
    }
    *out = it->second;
    return LDDMM_OK;
}

// Fixed-order tree sum of red[0 .. kLbBlock) into red[0] (all lanes call it).
__device__ __forceinline__ void lb_tree(double* red) {
    __syncthreads();
    for (int k = kLbBlock / 2; k > 0; k >>= 1) {
        if (threadIdx.x < k) red[threadIdx.x] += red[threadIdx.x + k];
        __syncthreads();
    }
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
