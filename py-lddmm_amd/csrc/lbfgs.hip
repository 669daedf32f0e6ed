// L-BFGS two-loop recursion on the device (py-lddmm base/bfgs.py:188-207).
//
// Each step of the recursion is one launch that first applies the previous
// step's update to q elementwise and then accumulates the next dot product
// on the updated values; the dot product is reduced deterministically (each
// workgroup writes its partial sum, the last workgroup to finish adds the
// partials in a fixed tree order) and stored, already multiplied by its rho, in
// device memory, where the next launch reads it.  The host issues 2m + 1
// launches for m stored pairs and never waits.
#include "common.hpp"

#include <mutex>
#include <unordered_map>
#include <vector>

namespace lddmm {
namespace {

constexpr int kLbBlock = 256;
constexpr int kLbGrid = 256;        // fixed: the reduction order depends only on n
constexpr int kLbMaxPairs = 1024;
static_assert(kLbGrid <= kLbBlock, "the last workgroup reads one partial per lane");

// Per-device state: [kLbGrid] partials, [2*kLbMaxPairs + 8] scalars, counter.
struct LbState {
    double* partials = nullptr;
    double* scal = nullptr;
    unsigned* counter = nullptr;
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
        LDDMM_HIP(hipMalloc(reinterpret_cast<void**>(&s.partials), kLbGrid * sizeof(double)));
        LDDMM_HIP(hipMalloc(reinterpret_cast<void**>(&s.scal), (2 * kLbMaxPairs + 8) * sizeof(double)));
        LDDMM_HIP(hipMalloc(reinterpret_cast<void**>(&s.counter), 64));
        LDDMM_HIP(hipMemset(s.counter, 0, 64));        // once, before any use; the kernels reset it
        it = g_lb.emplace(dev, s).first;
    }
    *out = it->second;
    return LDDMM_OK;
}

// q = (q + sign * (a - b) * u) * sc  (each part optional), then
// *out = mul * <d, q>  (if d), with mul = *mul_ptr or 1.
__global__ void __launch_bounds__(kLbBlock)
k_lb_step(int64_t n, double* __restrict__ q, const double* __restrict__ u, const double* __restrict__ a,
          const double* __restrict__ b, double sign, const double* __restrict__ sc, const double* __restrict__ d,
          double* __restrict__ out, const double* __restrict__ mul_ptr, double* __restrict__ partials,
          unsigned* __restrict__ counter) {
    __shared__ double red[kLbBlock];
    __shared__ bool last;
    double coef = 0.0;
    if (u) coef = sign * (*a - (b ? *b : 0.0));
    const double s = sc ? *sc : 1.0;
    double acc = 0.0;
    for (int64_t i = int64_t(blockIdx.x) * kLbBlock + threadIdx.x; i < n; i += int64_t(gridDim.x) * kLbBlock) {
        double qi = q[i];
        if (u) qi = fma(coef, u[i], qi);
        if (sc) qi *= s;
        if (u || sc) q[i] = qi;
        if (d) acc = fma(d[i], qi, acc);
    }
    if (!d) return;
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int k = kLbBlock / 2; k > 0; k >>= 1) {
        if (threadIdx.x < k) red[threadIdx.x] += red[threadIdx.x + k];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        partials[blockIdx.x] = red[0];
        __threadfence();
        unsigned ticket = atomicAdd(counter, 1u);
        last = ticket == gridDim.x - 1;
    }
    __syncthreads();
    if (!last) return;
    // last workgroup: every lane reads one partial (gridDim.x <= kLbBlock), then a
    // fixed tree, so the order depends only on the grid
    __threadfence();
    red[threadIdx.x] = threadIdx.x < gridDim.x ? partials[threadIdx.x] : 0.0;
    __syncthreads();
    for (int k = kLbBlock / 2; k > 0; k >>= 1) {
        if (threadIdx.x < k) red[threadIdx.x] += red[threadIdx.x + k];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        *out = (mul_ptr ? *mul_ptr : 1.0) * red[0];
        atomicExch(counter, 0u);
    }
}

// rho = 1 / <y, s>, c = <s, y> / <y, y> (c = 1 when below 1e-10), from two dots.
__global__ void k_lb_pair_scalars(const double* ys, const double* yy, double* rho, double* c) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        *rho = 1.0 / *ys;
        double v = *ys / *yy;
        *c = v < 1e-10 ? 1.0 : v;
    }
}

int lb_launch(int64_t n, double* q, const double* u, const double* a, const double* b, double sign,
              const double* sc, const double* d, double* out, const double* mul, const LbState& S, hipStream_t st) {
    int64_t blocks = (n + kLbBlock - 1) / kLbBlock;
    unsigned grid = unsigned(d ? kLbGrid : std::max<int64_t>(1, std::min<int64_t>(blocks, 4096)));
    hipLaunchKernelGGL(k_lb_step, dim3(grid), dim3(kLbBlock), 0, st, n, q, u, a, b, sign, sc, d, out, mul,
                       S.partials, S.counter);
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
    double* ys = S.scal + 2 * kLbMaxPairs;
    double* yy = ys + 1;
    // <y, s> and <y, y>: q is not modified without u / sc, so y serves as q
    LDDMM_TRY(lb_launch(n, const_cast<double*>(y), nullptr, nullptr, nullptr, 0.0, nullptr, s, ys, nullptr, S, st));
    LDDMM_TRY(lb_launch(n, const_cast<double*>(y), nullptr, nullptr, nullptr, 0.0, nullptr, y, yy, nullptr, S, st));
    hipLaunchKernelGGL(k_lb_pair_scalars, dim3(1), dim3(64), 0, st, ys, yy, rho, c);
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
    double* alpha = S.scal;                 // alpha[k], beta[k]
    double* beta = S.scal + kLbMaxPairs;
    // first loop, newest to oldest: alpha_k = rho_k <s_k, q>; q -= alpha_k y_k
    for (int k = m - 1; k >= 0; --k) {
        const bool prev = k < m - 1;
        LDDMM_TRY(lb_launch(n, q, prev ? y[k + 1] : nullptr, prev ? alpha + k + 1 : nullptr, nullptr, -1.0, nullptr,
                            s[k], alpha + k, rho[k], S, st));
    }
    // q -= alpha_0 y_0; q *= c; beta_0 = rho_0 <y_0, q>
    LDDMM_TRY(lb_launch(n, q, y[0], alpha, nullptr, -1.0, c_newest, y[0], beta, rho[0], S, st));
    // second loop, oldest to newest: q += (alpha_k - beta_k) s_k; beta_{k+1} = rho_{k+1} <y_{k+1}, q>
    for (int k = 0; k < m; ++k) {
        const bool next = k + 1 < m;
        LDDMM_TRY(lb_launch(n, q, s[k], alpha + k, beta + k, 1.0, nullptr, next ? y[k + 1] : nullptr,
                            next ? beta + k + 1 : nullptr, next ? rho[k + 1] : nullptr, S, st));
    }
    return LDDMM_OK;
}
