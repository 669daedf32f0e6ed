// Float32 applyK / applyDiffKT (plain) with MFMA accumulation (pair_mfma.hpp),
// selected by lddmm_set_pair_engine(LDDMM_ENGINE_MFMA) for LDDMM_PREC_F32C.
#include "ops_mfma.hpp"

#include <algorithm>
#include <atomic>

namespace lddmm {

namespace {
std::atomic<int> g_engine{LDDMM_ENGINE_VALU};
constexpr int kG = 4;   // 16-target groups per wave
}

int pair_engine() { return g_engine.load(); }

__global__ void k_reduce_diffkt_mfma(Partials<float> part, int nchunk, int KW, int n, const double* y,
                                     const double* p, int dim, int lda, int col0, const double* c0, double cs,
                                     double f, double* out, int accumulate) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int nch = 4 * KW;
    double S[16];
    for (int c = 0; c < nch; ++c) S[c] = sum_chunks(part, nchunk, nch, i, c, nullptr);
    double pa = 0.0;
    for (int c = 0; c < KW; ++c) pa += p[size_t(i) * lda + col0 + c] * S[c];
    for (int d = 0; d < dim; ++d) {
        double u = (y[size_t(i) * dim + d] - c0[d]) * cs;
        double v = u * pa;
        for (int c = 0; c < KW; ++c) v -= S[KW + d * KW + c] * p[size_t(i) * lda + col0 + c];
        size_t o = size_t(i) * dim + d;
        out[o] = (accumulate ? out[o] : 0.0) + f * v;
    }
}

template <class F>
static int with_nc(int nc, F&& f) {
    switch (nc) {
        case 1: return f(std::integral_constant<int, 1>{});
        case 2: return f(std::integral_constant<int, 2>{});
        case 3: return f(std::integral_constant<int, 3>{});
        case 4: return f(std::integral_constant<int, 4>{});
        default: return f(std::integral_constant<int, 16>{});
    }
}

template <bool MS>
int mfma_applyK_t(const lddmm_kernel* kd, const double* y, int n, const double* x, int m, int dim, const double* a,
                  int k, double* out, hipStream_t st) {
    ScaleSet<float> sc = make_scales<float>(kd, nullptr);
    const int ng = (k + 15) / 16;
    size_t need = 0;
    std::vector<PairPlan> plans(ng);
    for (int g = 0; g < ng; ++g) {
        int kc = std::min(16, k - 16 * g);
        int rc = with_nc(kc, [&](auto NCc) {
            constexpr int NC = decltype(NCc)::value;
            LDDMM_TRY((plan_mfma<MApplyK<NC, MS>, kG>(n, m, &plans[g])));
            need = std::max(need, partial_bytes<float>(plans[g], NC, n));
            return LDDMM_OK;
        });
        if (rc) return rc;
    }
    void* ws;
    LDDMM_TRY(workspace(need + 256, st, &ws));
    char* base = static_cast<char*>(ws);
    Carve cv;
    double* c0 = reinterpret_cast<double*>(base + cv.take(32));
    LDDMM_TRY(launch_center(y, n, x, m, dim, c0, st));
    size_t off0 = cv.off;
    for (int g = 0; g < ng; ++g) {
        int kc = std::min(16, k - 16 * g);
        int rc = with_nc(kc, [&](auto NCc) {
            constexpr int NC = decltype(NCc)::value;
            using OP = MApplyK<NC, MS>;
            typename OP::Args A{y, x, a, c0, dim, k, 16 * g, kc, sc};
            Carve c2; c2.off = off0;
            Partials<float> part = carve_partials<float>(base, c2, plans[g], NC, n);
            LDDMM_TRY((run_mfma<OP, kG>(A, n, m, plans[g], part, st)));
            ChanScale cs{};
            for (int c = 0; c < 8; ++c) cs.s[c] = 1.0;
            // k_reduce scales at most 8 channels: write the columns in runs of 8
            for (int c0_ = 0; c0_ < kc; c0_ += 8) {
                int nout = std::min(8, kc - c0_);
                Partials<float> pp = part;
                pp.hi += size_t(c0_) * n;
                pp.lo += size_t(c0_) * n;
                // channel stride inside a chunk stays NC
                LDDMM_TRY(run_reduce(pp, plans[g].nchunk, NC, nout, n, cs, out, k, 16 * g + c0_, 0, nullptr, 1.0, st));
            }
            return LDDMM_OK;
        });
        if (rc) return rc;
    }
    return LDDMM_OK;
}

template <bool MS>
int mfma_diffKT_t(const lddmm_kernel* kd, const double* y, int n, const double* x, int m, int dim, const double* p,
                  const double* a, int k, double* out, hipStream_t st) {
    SingleScaleConsts ss;
    ScaleSet<float> sc = make_scales<float>(kd, &ss);
    const double f = MS ? -1.0 : ss.diff;
    const int ng = (k + 3) / 4;
    size_t need = 0;
    std::vector<PairPlan> plans(ng);
    for (int g = 0; g < ng; ++g) {
        int kw = std::min(4, k - 4 * g);
        int rc = with_kw(kw, [&](auto KWc) {
            constexpr int KW = decltype(KWc)::value;
            LDDMM_TRY((plan_mfma<MDiffKT<KW, MS>, kG>(n, m, &plans[g])));
            need = std::max(need, partial_bytes<float>(plans[g], 4 * KW, n));
            return LDDMM_OK;
        });
        if (rc) return rc;
    }
    void* ws;
    LDDMM_TRY(workspace(need + 256, st, &ws));
    char* base = static_cast<char*>(ws);
    Carve cv;
    double* c0 = reinterpret_cast<double*>(base + cv.take(32));
    LDDMM_TRY(launch_center(y, n, x, m, dim, c0, st));
    size_t off0 = cv.off;
    for (int g = 0; g < ng; ++g) {
        int kw = std::min(4, k - 4 * g);
        int rc = with_kw(kw, [&](auto KWc) {
            constexpr int KW = decltype(KWc)::value;
            using OP = MDiffKT<KW, MS>;
            typename OP::Args A{y, x, a, c0, dim, k, 4 * g, sc};
            Carve c2; c2.off = off0;
            Partials<float> part = carve_partials<float>(base, c2, plans[g], 4 * KW, n);
            LDDMM_TRY((run_mfma<OP, kG>(A, n, m, plans[g], part, st)));
            hipLaunchKernelGGL(k_reduce_diffkt_mfma, dim3(unsigned((n + 255) / 256)), dim3(256), 0, st, part,
                               plans[g].nchunk, KW, n, y, p, dim, k, 4 * g, c0, double(sc.cs), f, out, g > 0 ? 1 : 0);
            LDDMM_CHECK_LAUNCH();
            return LDDMM_OK;
        });
        if (rc) return rc;
    }
    return LDDMM_OK;
}

int mfma_applyK(const lddmm_kernel* kd, const double* y, int n, const double* x, int m, int dim, const double* a, int k,
                double* out, hipStream_t st) {
    return general_path(kd) ? mfma_applyK_t<true>(kd, y, n, x, m, dim, a, k, out, st)
                              : mfma_applyK_t<false>(kd, y, n, x, m, dim, a, k, out, st);
}

int mfma_diffKT_plain(const lddmm_kernel* kd, const double* y, int n, const double* x, int m, int dim,
                      const double* p, const double* a, int k, double* out, hipStream_t st) {
    return general_path(kd) ? mfma_diffKT_t<true>(kd, y, n, x, m, dim, p, a, k, out, st)
                              : mfma_diffKT_t<false>(kd, y, n, x, m, dim, p, a, k, out, st);
}

}  // namespace lddmm

extern "C" int lddmm_set_pair_engine(int32_t engine) {
    if (engine != LDDMM_ENGINE_VALU && engine != LDDMM_ENGINE_MFMA) {
        lddmm::set_error("unknown pair engine %d", engine);
        return LDDMM_EINVAL;
    }
    lddmm::g_engine.store(engine);
    return LDDMM_OK;
}
