// Conjugate gradients for K(x, x) a = b on the device: the reference's
// Kernel.solve (base/kernelFunctions.py:676-680), which runs linearcg
// (base/conjugateGradient.py:351-408) with K(x, x) as the operator.
//
// One iteration is three launches on the caller's stream: a single-workgroup
// kernel that forms the search direction, the kernel sum q = K(x, x) p
// (lddmm_applyK), and a single-workgroup kernel that takes the step and
// evaluates both stopping tests.  The CG scalars and a stop flag live in
// device memory; once the flag is set the two CG kernels return at once, so
// the iterate is the one at the stop.  Every reduction is one workgroup
// adding per-lane partials in a fixed order (deterministic).  The host checks
// the flag every kCgPoll iterations only, to skip products after an early stop.
//
// Two sequences:
// * LDDMM_CG_REFERENCE is linearcg as it executes: at the first iteration the
//   direction p is the residual array itself (p = r), so the in-place
//   residual update r -= alpha q also moves p.  The stopping test of that
//   iteration reads the updated residual, and the second direction is
//   r1 + beta r1 = (1 + beta) r1.  Later directions follow the textbook
//   recursion p = r + beta p.
// * LDDMM_CG_TEXTBOOK keeps p and r apart (Hestenes-Stiefel conjugate
//   gradients).
// Both: zero start, return zero when |b|^2 < 1e-10, stop when
// alpha mean|p| < 1e-10 or when the quadratic energy a.Ka - 2 b.a changes by
// less than 1e-20 of its previous value, or after iter_max products.
// linearcg's per-iteration residual |K a - b|^2 is an extra product used only
// for its log and is not computed.
#include "common.hpp"

namespace lddmm {

namespace {

constexpr int kCgThreads = 1024;
constexpr int kCgWaves = kCgThreads / 64;
constexpr int kCgPoll = 16;

// state layout (doubles): [0] mu, [1] previous energy, [2] stop flag, [3] iterations
enum { kMu = 0, kEner = 1, kStop = 2, kIter = 3, kStateLen = 8 };

// Sum of one value per lane over the workgroup, in a fixed order: lanes of a
// wave by a butterfly, then the waves' sums in wave order.  Every lane gets
// the result.  `lds` holds kCgWaves doubles; the call begins and ends with a
// barrier, so consecutive calls may reuse it.
__device__ double block_sum(double v, double* lds) {
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
    __syncthreads();
    if ((threadIdx.x & 63) == 0) lds[threadIdx.x >> 6] = v;
    __syncthreads();
    double s = 0.0;
#pragma unroll
    for (int w = 0; w < kCgWaves; ++w) s += lds[w];
    __syncthreads();
    return s;
}

// Direction of iteration `it`: mu_new = r.r; first iteration p = r (zero
// result when mu_new < 1e-10), later p = r + (mu_new / mu) p.
__global__ void __launch_bounds__(kCgThreads) k_cg_direction(const double* __restrict__ r, double* __restrict__ p,
                                                             int64_t len, double* __restrict__ state, int first) {
    __shared__ double lds[kCgWaves];
    if (state[kStop] != 0.0) return;
    double s = 0.0;
    for (int64_t i = threadIdx.x; i < len; i += kCgThreads) s += r[i] * r[i];
    const double mu_new = block_sum(s, lds);
    if (first) {
        if (mu_new < 1e-10) {
            if (threadIdx.x == 0) state[kStop] = 1.0;
            return;
        }
        for (int64_t i = threadIdx.x; i < len; i += kCgThreads) p[i] = r[i];
    } else {
        const double beta = mu_new / state[kMu];
        for (int64_t i = threadIdx.x; i < len; i += kCgThreads) p[i] = r[i] + beta * p[i];
    }
    __syncthreads();                 // every lane has read state[kMu]
    if (threadIdx.x == 0) state[kMu] = mu_new;
}

// Step of iteration `it` with q = K p: alpha = mu / p.q; a += alpha p,
// r -= alpha q, z += alpha q (z = K a - 2 b); then the stopping tests on
// alpha mean|p| and on the energy z.a.  `alias_first`: the reference
// sequence's first iteration, where p is the residual array, so p takes the
// updated residual and mean|p| is read from it.
__global__ void __launch_bounds__(kCgThreads) k_cg_step(double* __restrict__ a, double* __restrict__ r,
                                                        double* __restrict__ p, const double* __restrict__ q,
                                                        double* __restrict__ z, int64_t len,
                                                        double* __restrict__ state, int alias_first) {
    __shared__ double lds[kCgWaves];
    if (state[kStop] != 0.0) return;
    double s = 0.0;
    for (int64_t i = threadIdx.x; i < len; i += kCgThreads) s += p[i] * q[i];
    const double alpha = state[kMu] / block_sum(s, lds);
    double sp = 0.0, se = 0.0;
    for (int64_t i = threadIdx.x; i < len; i += kCgThreads) {
        const double pi = p[i], qi = q[i];
        const double ai = a[i] + alpha * pi;
        const double ri = r[i] - alpha * qi;
        const double zi = z[i] + alpha * qi;
        a[i] = ai;
        r[i] = ri;
        z[i] = zi;
        if (alias_first) {
            p[i] = ri;
            sp += fabs(ri);
        } else {
            sp += fabs(pi);
        }
        se += zi * ai;
    }
    const double mean_p = block_sum(sp, lds) / double(len);
    const double ener = block_sum(se, lds);
    if (threadIdx.x == 0) {
        const double prev = state[kEner];
        state[kIter] += 1.0;
        if (alpha * mean_p < 1e-10)
            state[kStop] = 1.0;
        else if (fabs(ener - prev) < 1e-20 * fabs(prev))
            state[kStop] = 1.0;
        state[kEner] = ener;
    }
}

// a = 0, r = b, z = -2 b, state zeroed.
__global__ void __launch_bounds__(256) k_cg_init(const double* __restrict__ b, double* __restrict__ a,
                                                 double* __restrict__ r, double* __restrict__ z, int64_t len,
                                                 double* __restrict__ state) {
    const int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x;
    if (i < len) {
        const double bi = b[i];
        a[i] = 0.0;
        r[i] = bi;
        z[i] = -2.0 * bi;
    }
    if (blockIdx.x == 0 && threadIdx.x < kStateLen) state[threadIdx.x] = 0.0;
}

}  // namespace

}  // namespace lddmm

extern "C" int lddmm_kernel_solve(const lddmm_kernel* kd, const double* x, int64_t n, int32_t dim,
                                  const double* b, int32_t k, int32_t iter_max, int32_t sequence, double* a,
                                  double* work, void* stream) {
    using namespace lddmm;
    LDDMM_TRY(validate_kernel(kd));
    LDDMM_REQUIRE(n >= 0 && n <= kMaxPoints, "kernel_solve: n outside [0, %lld]", (long long)kMaxPoints);
    LDDMM_REQUIRE(dim >= 1 && dim <= 3, "kernel_solve: dim %d outside [1, 3]", dim);
    LDDMM_REQUIRE(k >= 1, "kernel_solve: k must be >= 1");
    LDDMM_REQUIRE(iter_max >= 0, "kernel_solve: negative iter_max");
    LDDMM_REQUIRE(sequence == LDDMM_CG_REFERENCE || sequence == LDDMM_CG_TEXTBOOK,
                  "kernel_solve: unknown sequence %d", sequence);
    if (n == 0) return LDDMM_OK;
    LDDMM_REQUIRE(x && b && a && work, "kernel_solve: null pointer");
    hipStream_t st = static_cast<hipStream_t>(stream);
    const int64_t len = n * int64_t(k);
    double* r = work;
    double* p = work + len;
    double* q = work + 2 * len;
    double* z = work + 3 * len;
    double* state = work + 4 * len;
    const int64_t blocks = (len + 255) / 256;
    LDDMM_REQUIRE(blocks < (int64_t(1) << 31), "kernel_solve: too many values");
    hipLaunchKernelGGL(k_cg_init, dim3(unsigned(blocks)), dim3(256), 0, st, b, a, r, z, len, state);
    LDDMM_CHECK_LAUNCH();
    for (int it = 0; it < iter_max; ++it) {
        if (it > 0 && it % kCgPoll == 0) {
            double stop = 0.0;
            LDDMM_HIP(hipMemcpyAsync(&stop, state + kStop, sizeof(double), hipMemcpyDeviceToHost, st));
            LDDMM_HIP(hipStreamSynchronize(st));
            if (stop != 0.0) break;
        }
        hipLaunchKernelGGL(k_cg_direction, dim3(1), dim3(kCgThreads), 0, st, r, p, len, state, int(it == 0));
        LDDMM_CHECK_LAUNCH();
        LDDMM_TRY(lddmm_applyK(kd, x, n, x, n, dim, p, k, q, stream));
        hipLaunchKernelGGL(k_cg_step, dim3(1), dim3(kCgThreads), 0, st, a, r, p, q, z, len, state,
                           int(it == 0 && sequence == LDDMM_CG_REFERENCE));
        LDDMM_CHECK_LAUNCH();
    }
    return LDDMM_OK;
}
