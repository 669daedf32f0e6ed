// Device-resident landmark shooting: forward Euler flow with its energy, the
// exact discrete adjoint (lddmm covector), EPDiff geodesic shooting and its
// adjoint.  State stays in HBM across the T steps; each step is a handful of
// launches on the caller's stream (capturable into a hipGraph by the caller).
//   forward : base/pointEvolution.py:88-164, energy base/surfaceMatching.py:533-548
//   adjoint : base/surfaceMatching.py:728-908, base/pointEvolution.py:167-236
//   EPDiff  : base/pointEvolution.py:385-472 (forward), :489-618 (adjoint)
#include "ops.hpp"

#include <algorithm>

namespace lddmm {

namespace {

// x_next = x + dt v; dots[block] = scale * sum_i a_i . v_i   (v = summed chunk partials)
template <typename T, int KW>
__global__ void k_shoot_epi(Partials<T> part, int nchunk, int n, const double* __restrict__ x,
                            const double* __restrict__ a, double dt, double* __restrict__ xn,
                            double* __restrict__ dots, double dscale) {
    __shared__ double red[256];
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    double d = 0.0;
    if (i < n) {
        for (int c = 0; c < KW; ++c) {
            double v = sum_chunks(part, nchunk, KW, i, c, nullptr);
            size_t o = size_t(i) * KW + c;
            xn[o] = x[o] + dt * v;
            d += a[o] * v;
        }
    }
    red[threadIdx.x] = d;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0 && dots) dots[blockIdx.x] = dscale * red[0];
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

int axpby(size_t count, double alpha, const double* a, double beta, const double* b, double* out,
          hipStream_t st) {
    if (count == 0) return LDDMM_OK;
    hipLaunchKernelGGL(k_axpby, dim3(unsigned((count + 255) / 256)), dim3(256), 0, st, count, alpha, a, beta,
                       b, out);
    LDDMM_CHECK_LAUNCH();
    return LDDMM_OK;
}

int check_shoot(const lddmm_kernel* kd, int T, int64_t n, int dim) {
    LDDMM_TRY(validate_kernel(kd));
    LDDMM_REQUIRE(T >= 1, "T must be >= 1 (got %d)", T);
    LDDMM_REQUIRE(n >= 0 && n < (int64_t(1) << 31), "bad n");
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

// ---------------------------------------------------------------------------
template <typename T, bool MS, int KW>
int shoot_forward_t(const lddmm_kernel* kd, const double* at, int T_, int n, double* xt, double r,
                    double* energy, hipStream_t st) {
    ScaleSet<T> sc = make_scales<T>(kd, nullptr);
    using OPK = OpApplyK<KW, MS>;
    Job<T, OPK> jk;
    LDDMM_TRY(plan_job(jk, n, n, !MS));
    const int nb = (n + 255) / 256;
    void* ws;
    LDDMM_TRY(workspace(256 + partial_bytes<T>(jk.plan, KW, n) + size_t(T_) * nb * sizeof(double) + 512, st, &ws));
    char* base = static_cast<char*>(ws);
    Carve cv;
    double* c0 = reinterpret_cast<double*>(base + cv.take(32));
    double* dots = reinterpret_cast<double*>(base + cv.take(size_t(T_) * nb * sizeof(double)));
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
        LDDMM_TRY(run_job(jk, n, n, part, st));
        hipLaunchKernelGGL((k_shoot_epi<T, KW>), dim3(unsigned(nb)), dim3(256), 0, st, part, jk.plan.nchunk, n, x,
                           a, dt, xt + (t + 1) * stride, dots + size_t(t) * nb, r * dt);
        LDDMM_CHECK_LAUNCH();
    }
    if (energy) {
        hipLaunchKernelGGL(k_sum_doubles, dim3(1), dim3(1024), 0, st, dots, T_ * nb, energy);
        LDDMM_CHECK_LAUNCH();
    }
    return LDDMM_OK;
}

template <typename T, bool MS, int KW>
int shoot_adjoint_t(const lddmm_kernel* kd, const double* at, const double* xt, int T_, int n,
                    const double* px1, double r, int euclidean, double* pxt, double* grad, hipStream_t st) {
    SingleScaleConsts ss;
    ScaleSet<T> sc = make_scales<T>(kd, &ss);
    const double fdiff = MS ? -1.0 : ss.diff;
    using OPD = OpDiffKT<KW, true, MS>;
    using OPK = OpApplyK<KW, MS>;
    Job<T, OPD> jd;
    Job<T, OPK> jk;
    LDDMM_TRY(plan_job(jd, n, n, !MS));
    LDDMM_TRY(plan_job(jk, n, n, !MS));
    const size_t stride = size_t(n) * KW;
    size_t need = 256 + std::max(partial_bytes<T>(jd.plan, 3, n), partial_bytes<T>(jk.plan, KW, n)) +
                  stride * sizeof(double) * (pxt ? 1 : 3) + 1024;
    void* ws;
    LDDMM_TRY(workspace(need, st, &ws));
    char* base = static_cast<char*>(ws);
    Carve cv;
    double* c0 = reinterpret_cast<double*>(base + cv.take(32));
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
        Carve c2; c2.off = poff;
        Partials<T> partD = carve_partials<T>(base, c2, jd.plan, 3, n);
        jd.args = {x, x, pn, a, c0, KW, KW, 0, 2.0 * r, sc};
        LDDMM_TRY(run_job(jd, n, n, partD, st));
        // dat = 2 r a_t - p_{t+1}  (read before p_t overwrites the other buffer)
        LDDMM_TRY(axpby(stride, 2.0 * r, a, -1.0, pn, dat, st));
        LDDMM_TRY(reduce_job(jd, partD, n, KW, fdiff, pc, 0, pn, dt, st));
        double* g = grad + size_t(t) * stride;
        if (euclidean) {
            Carve c3; c3.off = poff;
            Partials<T> partK = carve_partials<T>(base, c3, jk.plan, KW, n);
            jk.args = {x, x, dat, c0, KW, KW, 0, sc};
            LDDMM_TRY(run_job(jk, n, n, partK, st));
            LDDMM_TRY(reduce_job(jk, partK, n, KW, 1.0, g, 0, nullptr, dt, st));
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
    Job<T, OPK> jk;
    Job<T, OPD> jd;
    LDDMM_TRY(plan_job(jk, n, n, !MS));
    LDDMM_TRY(plan_job(jd, n, n, !MS));
    void* ws;
    LDDMM_TRY(workspace(256 + std::max(partial_bytes<T>(jk.plan, KW, n), partial_bytes<T>(jd.plan, 3, n)) + 512,
                        st, &ws));
    char* base = static_cast<char*>(ws);
    Carve cv;
    double* c0 = reinterpret_cast<double*>(base + cv.take(32));
    size_t poff = cv.off;
    const size_t stride = size_t(n) * KW;
    const double dt = 1.0 / T_;
    LDDMM_TRY(launch_center(xt, n, xt, 0, KW, c0, st));       // one centre (x_0) for all steps
    for (int k = 0; k < T_; ++k) {
        const double* x = xt + size_t(k) * stride;
        const double* a = at + size_t(k) * stride;
        Carve c1; c1.off = poff;
        Partials<T> pk = carve_partials<T>(base, c1, jk.plan, KW, n);
        jk.args = {x, x, a, c0, KW, KW, 0, sc};
        LDDMM_TRY(run_job(jk, n, n, pk, st));
        LDDMM_TRY(reduce_job(jk, pk, n, KW, 1.0, xt + size_t(k + 1) * stride, 0, x, dt, st));
        Carve c2; c2.off = poff;
        Partials<T> pd = carve_partials<T>(base, c2, jd.plan, 3, n);
        jd.args = {x, x, a, a, c0, KW, KW, 0, 0.0, sc};
        LDDMM_TRY(run_job(jd, n, n, pd, st));
        LDDMM_TRY(reduce_job(jd, pd, n, KW, fdiff, at + size_t(k + 1) * stride, 0, a, -dt, st));
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
    LDDMM_TRY(plan_job(j1, n, n, !MS));
    LDDMM_TRY(plan_job(j2, n, n, !MS));
    LDDMM_TRY(plan_job(j3, n, n, !MS));
    LDDMM_TRY(plan_job(j4, n, n, !MS));
    size_t pb = std::max(std::max(partial_bytes<T>(j1.plan, 3, n), partial_bytes<T>(j2.plan, 3, n)),
                         std::max(partial_bytes<T>(j3.plan, KW, n), partial_bytes<T>(j4.plan, KW, n)));
    void* ws;
    LDDMM_TRY(workspace(256 + pb + 512, st, &ws));
    char* base = static_cast<char*>(ws);
    Carve cv;
    double* c0 = reinterpret_cast<double*>(base + cv.take(32));
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
    }
    // da0 = K(x0) a0 - pa_0
    LDDMM_TRY(axpby(stride, -1.0, pat, 0.0, nullptr, da0, st));
    Carve c; c.off = poff;
    Partials<T> part = carve_partials<T>(base, c, j3.plan, KW, n);
    j3.args = {xt, xt, at, c0, KW, KW, 0, sc};
    LDDMM_TRY(run_job(j3, n, n, part, st));
    LDDMM_TRY(reduce_job(j3, part, n, KW, 1.0, da0, 1, nullptr, 1.0, st));
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
    LDDMM_TRY(check_shoot(kd, T, n, dim));
    hipStream_t st = static_cast<hipStream_t>(stream);
    if (n == 0) return energy ? launch_fill(energy, 1, 0.0, st) : LDDMM_OK;
    LDDMM_REQUIRE(at && xt, "null at / xt");
    LDDMM_DISPATCH(shoot_forward_t, kd, at, T, int(n), xt, regweight, energy, st);
}

extern "C" int lddmm_shoot_adjoint(const lddmm_kernel* kd, const double* at, const double* xt, int32_t T,
                                   int64_t n, int32_t dim, const double* px1, double regweight, int32_t euclidean,
                                   double* pxt, double* grad, void* stream) {
    LDDMM_TRY(check_shoot(kd, T, n, dim));
    if (n == 0) return LDDMM_OK;
    LDDMM_REQUIRE(at && xt && px1 && grad, "null pointer");
    hipStream_t st = static_cast<hipStream_t>(stream);
    LDDMM_DISPATCH(shoot_adjoint_t, kd, at, xt, T, int(n), px1, regweight, euclidean, pxt, grad, st);
}

extern "C" int lddmm_epdiff_forward(const lddmm_kernel* kd, int32_t T, int64_t n, int32_t dim, double* xt,
                                    double* at, void* stream) {
    LDDMM_TRY(check_shoot(kd, T, n, dim));
    if (n == 0) return LDDMM_OK;
    LDDMM_REQUIRE(xt && at, "null pointer");
    hipStream_t st = static_cast<hipStream_t>(stream);
    LDDMM_DISPATCH(epdiff_forward_t, kd, T, int(n), xt, at, st);
}

extern "C" int lddmm_epdiff_adjoint(const lddmm_kernel* kd, int32_t T, int64_t n, int32_t dim, const double* xt,
                                    const double* at, const double* px1, const double* pa1, double* pxt,
                                    double* pat, double* da0, void* stream) {
    LDDMM_TRY(check_shoot(kd, T, n, dim));
    if (n == 0) return LDDMM_OK;
    LDDMM_REQUIRE(xt && at && px1 && pa1 && pxt && pat && da0, "null pointer");
    hipStream_t st = static_cast<hipStream_t>(stream);
    LDDMM_DISPATCH(epdiff_adjoint_t, kd, T, int(n), xt, at, px1, pa1, pxt, pat, da0, st);
}
