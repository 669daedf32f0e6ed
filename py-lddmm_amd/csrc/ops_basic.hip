// Gaussian kernel sums on point sets: applyK, applyDiffKT, applyktensor,
// applyDiffK, applyDDiffK (pair machinery in pair.hpp) and the dense-weight
// sums applykmat / applykdiffmat.
#include "ops_mfma.hpp"

#include <algorithm>
#include <mutex>
#include <vector>

namespace lddmm {

// ---------------------------------------------------------------------------
// Small shared kernels
// ---------------------------------------------------------------------------
__global__ void k_fill(double* p, size_t count, double v) {
    size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    size_t stride = size_t(gridDim.x) * blockDim.x;
    for (; i < count; i += stride) p[i] = v;
}

int launch_fill(double* p, size_t count, double v, hipStream_t st) {
    if (count == 0) return LDDMM_OK;
    size_t nb = std::min<size_t>((count + 255) / 256, 4096);
    hipLaunchKernelGGL(k_fill, dim3(unsigned(nb)), dim3(256), 0, st, p, count, v);
    LDDMM_CHECK_LAUNCH();
    return LDDMM_OK;
}

// Centre used to shift coordinates before the float32 conversion: the midpoint
// of the bounding box of a sample of each set (every point of a set of at most
// kCenterSample points; otherwise kCenterRuns evenly spaced runs of
// kCenterRun consecutive points, which a wave reads as contiguous memory).
// Any point inside the clouds' hull bounds the coordinate magnitudes equally
// well, so a sample keeps this pass at a few microseconds.
constexpr int kCenterSample = 2048;
constexpr int kCenterRun = 32;
constexpr int kCenterRuns = kCenterSample / kCenterRun;
constexpr int kCenterThreads = 256;
constexpr int kCenterPer = kCenterSample / kCenterThreads;   // samples per lane and set

// Bounding box of the sampled points of one set into (l, h), one coordinate
// per register (dim <= 3); a lane's kCenterPer loads are all issued before
// any is used.
__device__ __forceinline__ void center_scan(const double* p, int cnt, int dim, double (&l)[3], double (&h)[3]) {
    double v[kCenterPer][3];
    bool ok[kCenterPer];
#pragma unroll
    for (int k = 0; k < kCenterPer; ++k) {
        const int j = threadIdx.x + k * kCenterThreads;         // sample number
        int64_t i = j;
        if (cnt > kCenterSample)
            i = int64_t(j / kCenterRun) * (cnt - kCenterRun) / (kCenterRuns - 1) + j % kCenterRun;
        ok[k] = i < cnt;
#pragma unroll
        for (int d = 0; d < 3; ++d) v[k][d] = ok[k] && d < dim ? p[size_t(i) * dim + d] : l[d];
    }
#pragma unroll
    for (int k = 0; k < kCenterPer; ++k) {
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            if (ok[k] && d < dim) {
                l[d] = fmin(l[d], v[k][d]);
                h[d] = fmax(h[d], v[k][d]);
            }
        }
    }
}

__global__ void __launch_bounds__(kCenterThreads) k_center(const double* y, int n, const double* x, int m, int dim,
                                                           double* c0) {
    constexpr int W = kCenterThreads / 64;
    __shared__ double lo[3][W], hi[3][W];
    const int t = threadIdx.x;
    double l[3] = {1e300, 1e300, 1e300}, h[3] = {-1e300, -1e300, -1e300};
    center_scan(y, n, dim, l, h);
    center_scan(x, m, dim, l, h);
    // min / max are exact in any order: wave butterflies, then the W wave values
#pragma unroll
    for (int d = 0; d < 3; ++d) {
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) {
            l[d] = fmin(l[d], __shfl_xor(l[d], o, 64));
            h[d] = fmax(h[d], __shfl_xor(h[d], o, 64));
        }
    }
    if ((t & 63) == 0) {
#pragma unroll
        for (int d = 0; d < 3; ++d) { lo[d][t >> 6] = l[d]; hi[d][t >> 6] = h[d]; }
    }
    __syncthreads();
    if (t < 4) {
        double c = 0.0;
        if (t < dim && t < 3) {
            double a = lo[t][0], b = hi[t][0];
#pragma unroll
            for (int w = 1; w < W; ++w) { a = fmin(a, lo[t][w]); b = fmax(b, hi[t][w]); }
            // an empty or non-finite box centres at 0
            if (a <= b) c = 0.5 * (a + b);
        }
        if (!isfinite(c)) c = 0.0;
        c0[t] = c;
    }
}

int launch_center(const double* y, int n, const double* x, int m, int dim, double* c0, hipStream_t st) {
    hipLaunchKernelGGL(k_center, dim3(1), dim3(kCenterThreads), 0, st, y, n, x, m, dim, c0);
    LDDMM_CHECK_LAUNCH();
    return LDDMM_OK;
}

PairPlan plan_pair(int n, int m, int R, int num_cu, int occ, int quantum, int max_nchunk) {
    PairPlan p;
    // n <= kMaxPoints (checked by every entry point), so the block count and
    // the kernels' first-target index blockIdx.x * kBlock * R stay in int
    p.bx = int((int64_t(n) + kBlock * R - 1) / (kBlock * R));
    if (m <= 0 || n <= 0) { p.nchunk = 0; p.chunk = kTile; return p; }
    occ = std::max(1, occ);
    num_cu = std::max(1, num_cu);
    quantum = std::max(1, quantum);
    max_nchunk = std::max(1, max_nchunk);
    // Time model (VALU-bound pair kernels): each CU runs its blocks `occ` at a
    // time, so its time is its block count times one block's work (chunk plus a
    // fixed cost of about half a tile for the target loads and partial
    // stores); a round with fewer than min(4, occ) resident blocks per CU costs
    // as much as one with that many (too few waves to hide latency).  The
    // busiest CU sets the launch time, so blocks are spread evenly rather
    // than filling num_cu * occ slots with some CUs one block ahead.
    const long minres = std::min(4, occ);
    const long min_chunk = std::max(quantum, 64);
    const long smax = std::max(1L, (long(m) + min_chunk - 1) / min_chunk);
    const long p0 = (long(p.bx) + num_cu - 1) / num_cu;
    double best = 1e300;
    // for each block count P on the busiest CU, the most chunks that keep it
    for (long P = p0; P <= p0 + 3L * occ + 4; ++P) {
        const long S = std::max(1L, std::min(std::min(smax, long(max_nchunk)), P * num_cu / p.bx));
        const long chunk = ((m + S - 1) / S + quantum - 1) / quantum * quantum;
        const long se = (m + chunk - 1) / chunk;
        const long per_cu = (long(p.bx) * se + num_cu - 1) / num_cu;
        const long full = per_cu / occ, rest = per_cu % occ;
        const double t = double(full * occ + (rest ? std::max(rest, minres) : 0));
        const double cost = t * double(chunk + kTile / 2) + 0.25 * double(se);
        if (cost < best) { best = cost; p.nchunk = int(se); p.chunk = int(chunk); }
    }
    return p;
}

// ---------------------------------------------------------------------------
// Host drivers
// ---------------------------------------------------------------------------
namespace {

int check_common(const lddmm_kernel* kd, int64_t n, int64_t m, int dim) {
    LDDMM_TRY(validate_kernel(kd));
    LDDMM_REQUIRE(n >= 0 && m >= 0, "negative sizes");
    LDDMM_REQUIRE(n <= kMaxPoints && m <= kMaxPoints, "sizes above %lld points", (long long)kMaxPoints);
    LDDMM_REQUIRE(dim >= 1 && dim <= 3, "dim %d outside [1, 3]", dim);
    return LDDMM_OK;
}

}  // namespace

// Sizes per op are small enough to keep every temporary in one workspace
// request: [c0 (256 B)] [partials of the largest job].
template <typename T, bool MS>
int applyK_t(const lddmm_kernel* kd, const double* y, int n, const double* x, int m, int dim,
             const double* a, int k, double* out, hipStream_t st) {
    SingleScaleConsts ss;
    ScaleSet<T> sc = make_scales<T>(kd, &ss);
    const int ng = (k + 3) / 4;
    size_t need = 0;
    std::vector<PairPlan> plans(ng);
    std::vector<int> Rs(ng);
    for (int g = 0; g < ng; ++g) {
        int kw = std::min(4, k - 4 * g);
        int rc = with_kw(kw, [&](auto KWc) {
            constexpr int KW = decltype(KWc)::value;
            Job<T, OpApplyK<KW, MS>> j;
            int r = plan_job(j, n, m, !MS);
            plans[g] = j.plan; Rs[g] = j.R;
            need = std::max(need, partial_bytes<T>(j.plan, KW, n));
            return r;
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
            Job<T, OpApplyK<KW, MS>> j;
            j.args = {y, x, a, c0, dim, k, 4 * g, sc};
            j.plan = plans[g]; j.R = Rs[g];
            Carve c2; c2.off = off0;
            Partials<T> part = carve_partials<T>(base, c2, j.plan, KW, n);
            LDDMM_TRY(run_job(j, n, m, part, st));
            ChanScale cs{};
            for (int c = 0; c < 8; ++c) cs.s[c] = 1.0;
            return run_reduce(part, j.plan.nchunk, KW, KW, n, cs, out, k, 4 * g, 0, nullptr, 1.0, st);
        });
        if (rc) return rc;
    }
    return LDDMM_OK;
}

template <typename T, bool SYM, bool MS>
int diffKT_t(const lddmm_kernel* kd, const double* y, int n, const double* x, int m, int dim,
             const double* p, const double* a, int k, double r2w, double* out, hipStream_t st) {
    SingleScaleConsts ss;
    ScaleSet<T> sc = make_scales<T>(kd, &ss);
    const int ng = (k + 3) / 4;
    size_t need = 0;
    std::vector<PairPlan> plans(ng);
    std::vector<int> Rs(ng);
    for (int g = 0; g < ng; ++g) {
        int kw = std::min(4, k - 4 * g);
        int rc = with_kw(kw, [&](auto KWc) {
            constexpr int KW = decltype(KWc)::value;
            Job<T, OpDiffKT<KW, SYM, MS>> j;
            int r = plan_job(j, n, m, !MS);
            plans[g] = j.plan; Rs[g] = j.R;
            need = std::max(need, partial_bytes<T>(j.plan, 3, n));
            return r;
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
    const double f = MS ? -1.0 : ss.diff;
    for (int g = 0; g < ng; ++g) {
        int kw = std::min(4, k - 4 * g);
        int rc = with_kw(kw, [&](auto KWc) {
            constexpr int KW = decltype(KWc)::value;
            Job<T, OpDiffKT<KW, SYM, MS>> j;
            j.args = {y, x, p, a, c0, dim, k, 4 * g, r2w, sc};
            j.plan = plans[g]; j.R = Rs[g];
            Carve c2; c2.off = off0;
            Partials<T> part = carve_partials<T>(base, c2, j.plan, 3, n);
            LDDMM_TRY(run_job(j, n, m, part, st));
            ChanScale cs{};
            for (int c = 0; c < 8; ++c) cs.s[c] = f;
            // column groups add up: c_ij is a sum over weight columns
            return run_reduce(part, j.plan.nchunk, 3, dim, n, cs, out, dim, 0, g > 0 ? 1 : 0, nullptr, 1.0, st);
        });
        if (rc) return rc;
    }
    return LDDMM_OK;
}

template <typename T, bool MS>
int ktensor_t(const lddmm_kernel* kd, const double* y, int n, const double* x, int m, int dim,
              const double* ay, const double* ax, const double* by, const double* bx, int kb,
              double* out, hipStream_t st) {
    ScaleSet<T> sc = make_scales<T>(kd, nullptr);
    return with_kw(kb, [&](auto KBc) {
        constexpr int KB = decltype(KBc)::value;
        Job<T, OpKTensor<KB, MS>> j;
        LDDMM_TRY(plan_job(j, n, m, !MS));
        void* ws;
        LDDMM_TRY(workspace(partial_bytes<T>(j.plan, 1, n) + 256, st, &ws));
        char* base = static_cast<char*>(ws);
        Carve cv;
        double* c0 = reinterpret_cast<double*>(base + cv.take(32));
        LDDMM_TRY(launch_center(y, n, x, m, dim, c0, st));
        j.args = {y, x, ay, ax, by, bx, c0, dim, kb, sc};
        Partials<T> part = carve_partials<T>(base, cv, j.plan, 1, n);
        LDDMM_TRY(run_job(j, n, m, part, st));
        ChanScale cs{};
        cs.s[0] = 1.0;
        return run_reduce(part, j.plan.nchunk, 1, 1, n, cs, out, 1, 0, 0, nullptr, 1.0, st);
    });
}

template <typename T, int OPT, bool MS>
int diffK_t(const lddmm_kernel* kd, const double* x, int n, int dim, const double* a1,
            const double* a2, int k, double* out, hipStream_t st) {
    SingleScaleConsts ss;
    ScaleSet<T> sc = make_scales<T>(kd, &ss);
    const double sign = OPT == LDDMM_DIFFK_2 ? -1.0 : 1.0;
    const double f = sign * (MS ? -1.0 : ss.diff);
    const int ng = (k + 3) / 4;
    size_t need = 0;
    std::vector<PairPlan> plans(ng);
    std::vector<int> Rs(ng);
    for (int g = 0; g < ng; ++g) {
        int kw = std::min(4, k - 4 * g);
        int rc = with_kw(kw, [&](auto KWc) {
            constexpr int KW = decltype(KWc)::value;
            Job<T, OpDiffK<KW, OPT, MS>> j;
            int r = plan_job(j, n, n, !MS);
            plans[g] = j.plan; Rs[g] = j.R;
            need = std::max(need, partial_bytes<T>(j.plan, KW, n));
            return r;
        });
        if (rc) return rc;
    }
    void* ws;
    LDDMM_TRY(workspace(need + 256, st, &ws));
    char* base = static_cast<char*>(ws);
    Carve cv;
    double* c0 = reinterpret_cast<double*>(base + cv.take(32));
    LDDMM_TRY(launch_center(x, n, x, 0, dim, c0, st));
    size_t off0 = cv.off;
    for (int g = 0; g < ng; ++g) {
        int kw = std::min(4, k - 4 * g);
        int rc = with_kw(kw, [&](auto KWc) {
            constexpr int KW = decltype(KWc)::value;
            Job<T, OpDiffK<KW, OPT, MS>> j;
            j.args = {x, a1, a2, c0, dim, k, 4 * g, sc};
            j.plan = plans[g]; j.R = Rs[g];
            Carve c2; c2.off = off0;
            Partials<T> part = carve_partials<T>(base, c2, j.plan, KW, n);
            LDDMM_TRY(run_job(j, n, n, part, st));
            ChanScale cs{};
            for (int c = 0; c < 8; ++c) cs.s[c] = f;
            return run_reduce(part, j.plan.nchunk, KW, KW, n, cs, out, k, 4 * g, 0, nullptr, 1.0, st);
        });
        if (rc) return rc;
    }
    return LDDMM_OK;
}

template <typename T, int OPT, bool MS>
int ddiffK_t(const lddmm_kernel* kd, const double* x, int n, int dim, const double* a1,
             const double* a2, int k, const double* p, double* out, hipStream_t st) {
    SingleScaleConsts ss;
    ScaleSet<T> sc = make_scales<T>(kd, &ss);
    const double f = MS ? 1.0 : ss.ddiff;
    const int ng = (k + 3) / 4;
    size_t need = 0;
    std::vector<PairPlan> plans(ng);
    std::vector<int> Rs(ng);
    for (int g = 0; g < ng; ++g) {
        int kw = std::min(4, k - 4 * g);
        int rc = with_kw(kw, [&](auto KWc) {
            constexpr int KW = decltype(KWc)::value;
            Job<T, OpDDiffK<KW, OPT, MS>> j;
            int r = plan_job(j, n, n, !MS);
            plans[g] = j.plan; Rs[g] = j.R;
            need = std::max(need, partial_bytes<T>(j.plan, 3, n));
            return r;
        });
        if (rc) return rc;
    }
    void* ws;
    LDDMM_TRY(workspace(need + 256, st, &ws));
    char* base = static_cast<char*>(ws);
    Carve cv;
    double* c0 = reinterpret_cast<double*>(base + cv.take(32));
    LDDMM_TRY(launch_center(x, n, x, 0, dim, c0, st));
    size_t off0 = cv.off;
    for (int g = 0; g < ng; ++g) {
        int kw = std::min(4, k - 4 * g);
        int rc = with_kw(kw, [&](auto KWc) {
            constexpr int KW = decltype(KWc)::value;
            Job<T, OpDDiffK<KW, OPT, MS>> j;
            j.args = {x, a1, a2, p, c0, dim, k, 4 * g, T(ss.gamma), sc};
            j.plan = plans[g]; j.R = Rs[g];
            Carve c2; c2.off = off0;
            Partials<T> part = carve_partials<T>(base, c2, j.plan, 3, n);
            LDDMM_TRY(run_job(j, n, n, part, st));
            ChanScale cs{};
            for (int c = 0; c < 8; ++c) cs.s[c] = f;
            return run_reduce(part, j.plan.nchunk, 3, dim, n, cs, out, dim, 0, g > 0 ? 1 : 0, nullptr, 1.0, st);
        });
        if (rc) return rc;
    }
    return LDDMM_OK;
}

// ---------------------------------------------------------------------------
// applyK with no allocation: the centre and the chunk partials live in a
// fixed device array, `scratch`, instead of the grow-only workspace.  Host
// code never passes `scratch` itself to a kernel (in host code the name is a
// stand-in, not a device address): each device's address of the array is
// looked up once with hipGetSymbolAddress and kept in g_scratch_dev.
// ---------------------------------------------------------------------------
constexpr size_t kNoAllocScratchBytes = size_t(LDDMM_NOALLOC_SCRATCH_BYTES);
static __device__ __align__(256) char scratch[kNoAllocScratchBytes];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mutex;
char* g_scratch_dev[kMaxDevices] = {};

// This device's address of `scratch` (looked up on the first call per device).
int scratch_address(char** out) {
    int dev = 0;
    LDDMM_HIP(hipGetDevice(&dev));
    LDDMM_REQUIRE(dev >= 0 && dev < kMaxDevices, "device %d beyond the %d the library tracks", dev, kMaxDevices);
    std::lock_guard<std::mutex> g(g_scratch_mutex);
    if (!g_scratch_dev[dev]) {
        void* p = nullptr;
        LDDMM_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        LDDMM_REQUIRE(p != nullptr, "no device address for the static scratch array");
        g_scratch_dev[dev] = static_cast<char*>(p);
    }
    *out = g_scratch_dev[dev];
    return LDDMM_OK;
}

// Same launches as applyK_t, with each weight group's partials carved from
// `base` (cap bytes).  A group whose plan would not fit gets fewer, larger
// source chunks (same sums, another summation order); one that does not fit
// with a single chunk is rejected before anything is launched.
template <typename T, bool MS>
int applyK_fixed(const lddmm_kernel* kd, const double* y, int n, const double* x, int m, int dim,
                 const double* a, int k, double* out, char* base, size_t cap, hipStream_t st) {
    SingleScaleConsts ss;
    ScaleSet<T> sc = make_scales<T>(kd, &ss);
    Carve cv;
    double* c0 = reinterpret_cast<double*>(base + cv.take(32));
    const size_t off0 = cv.off;
    const size_t room = cap - off0;
    const int ng = (k + 3) / 4;
    // every group is planned and checked before the first launch
    for (int pass = 0; pass < 2; ++pass) {
        if (pass == 1) LDDMM_TRY(launch_center(y, n, x, m, dim, c0, st));
        for (int g = 0; g < ng; ++g) {
            const int kw = std::min(4, k - 4 * g);
            int rc = with_kw(kw, [&](auto KWc) {
                constexpr int KW = decltype(KWc)::value;
                Job<T, OpApplyK<KW, MS>> j;
                LDDMM_TRY(plan_job(j, n, m, !MS));
                if (partial_bytes<T>(j.plan, KW, n) > room) {
                    PairPlan one = j.plan;
                    one.nchunk = 1;
                    const size_t per = partial_bytes<T>(one, KW, n);
                    if (per > room) {
                        set_error("applyK_noalloc: %lld targets x %d weight columns need %zu bytes of partials, "
                                  "the static scratch holds %zu (n * min(k, 4) * 8 <= %zu)",
                                  (long long)n, KW, per + off0, cap, room);
                        return LDDMM_EINVAL;
                    }
                    LDDMM_TRY(plan_job(j, n, m, !MS, int(std::min<size_t>(room / per, 1 << 30))));
                }
                if (pass == 0) return LDDMM_OK;
                j.args = {y, x, a, c0, dim, k, 4 * g, sc};
                Carve c2;
                c2.off = off0;
                Partials<T> part = carve_partials<T>(base, c2, j.plan, KW, n);
                LDDMM_TRY(run_job(j, n, m, part, st));
                ChanScale cs{};
                for (int c = 0; c < 8; ++c) cs.s[c] = 1.0;
                return run_reduce(part, j.plan.nchunk, KW, KW, n, cs, out, k, 4 * g, 0, nullptr, 1.0, st);
            });
            if (rc) return rc;
        }
    }
    return LDDMM_OK;
}
}  // namespace

// ---------------------------------------------------------------------------
// Dense-weight sums (memory-bound on beta): one wave64 per target row, lanes
// stride over sources, float64 arithmetic, wave reduction at the end.
// ---------------------------------------------------------------------------
__device__ __forceinline__ double wave_sum(double v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

template <bool DIFF>
__global__ void __launch_bounds__(256)
k_dense(const double* y, int n, const double* x, int m, int dim, const double* beta, int kb,
        ScaleSet<double> sc, double* out) {
    const int lane = threadIdx.x & 63;
    const int i = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (i >= n) return;
    double yi[3] = {0, 0, 0};
    for (int d = 0; d < dim; ++d) yi[d] = y[size_t(i) * dim + d];
    double acc[4] = {0, 0, 0, 0};
    const int nc = DIFF ? dim : kb;
    for (int c0 = 0; c0 < nc; c0 += 4) {
        for (int c = 0; c < 4; ++c) acc[c] = 0.0;
        for (int j = lane; j < m; j += 64) {
            double dd[3] = {0, 0, 0};
            double r2 = 0.0;
            for (int d = 0; d < dim; ++d) {
                dd[d] = yi[d] - x[size_t(j) * dim + d];
                r2 += dd[d] * dd[d];
            }
            if constexpr (DIFF) {
                double g = gD<double, true>(sc, r2);
                double b = beta[size_t(i) * m + j] * g;
                for (int c = 0; c < 4 && c0 + c < dim; ++c) acc[c] -= b * dd[c0 + c];
            } else {
                double g = gK<double, true>(sc, r2);
                const double* bp = beta + (size_t(i) * m + j) * kb + c0;
                for (int c = 0; c < 4 && c0 + c < kb; ++c) acc[c] += g * bp[c];
            }
        }
        for (int c = 0; c < 4 && c0 + c < nc; ++c) {
            double v = wave_sum(acc[c]);
            if (lane == 0) out[size_t(i) * nc + c0 + c] = v;
        }
    }
}

}  // namespace lddmm

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
using namespace lddmm;

extern "C" int lddmm_applyK(const lddmm_kernel* kd, const double* y, int64_t n, const double* x,
                            int64_t m, int32_t dim, const double* a, int32_t k, double* out,
                            void* stream) {
    LDDMM_TRY(check_common(kd, n, m, dim));
    LDDMM_REQUIRE(k >= 1, "k must be >= 1");
    if (n == 0) return LDDMM_OK;
    LDDMM_REQUIRE(y && out, "null y/out");
    hipStream_t st = static_cast<hipStream_t>(stream);
    if (m == 0) return launch_fill(out, size_t(n) * k, 0.0, st);
    LDDMM_REQUIRE(x && a, "null x/a");
    const bool ms = general_path(kd);
    if (kd->prec == LDDMM_PREC_F32C && kd->kind == LDDMM_KERNEL_GAUSS && pair_engine() == LDDMM_ENGINE_MFMA)
        return mfma_applyK(kd, y, int(n), x, int(m), dim, a, k, out, st);
    if (kd->prec == LDDMM_PREC_F32C)
        return ms ? applyK_t<float, true>(kd, y, n, x, m, dim, a, k, out, st)
                  : applyK_t<float, false>(kd, y, n, x, m, dim, a, k, out, st);
    return ms ? applyK_t<double, true>(kd, y, n, x, m, dim, a, k, out, st)
              : applyK_t<double, false>(kd, y, n, x, m, dim, a, k, out, st);
}

extern "C" int lddmm_applyK_noalloc(const lddmm_kernel* kd, const double* y, int64_t n, const double* x,
                                    int64_t m, int32_t dim, const double* a, int32_t k, double* out,
                                    void* stream) {
    LDDMM_TRY(check_common(kd, n, m, dim));
    LDDMM_REQUIRE(k >= 1, "k must be >= 1");
    if (n == 0) return LDDMM_OK;
    LDDMM_REQUIRE(y && out, "null y/out");
    hipStream_t st = static_cast<hipStream_t>(stream);
    if (m == 0) return launch_fill(out, size_t(n) * k, 0.0, st);
    LDDMM_REQUIRE(x && a, "null x/a");
    char* base = nullptr;
    LDDMM_TRY(scratch_address(&base));
    ProfileOff no_events;            // the timing hooks would create HIP events
    const bool ms = general_path(kd);
    if (kd->prec == LDDMM_PREC_F32C)
        return ms ? applyK_fixed<float, true>(kd, y, n, x, m, dim, a, k, out, base, kNoAllocScratchBytes, st)
                  : applyK_fixed<float, false>(kd, y, n, x, m, dim, a, k, out, base, kNoAllocScratchBytes, st);
    return ms ? applyK_fixed<double, true>(kd, y, n, x, m, dim, a, k, out, base, kNoAllocScratchBytes, st)
              : applyK_fixed<double, false>(kd, y, n, x, m, dim, a, k, out, base, kNoAllocScratchBytes, st);
}

extern "C" int lddmm_applyDiffKT(const lddmm_kernel* kd, const double* y, int64_t n, const double* x,
                                 int64_t m, int32_t dim, const double* p, const double* a, int32_t k,
                                 int32_t mode, double regweight, double* out, void* stream) {
    LDDMM_TRY(check_common(kd, n, m, dim));
    LDDMM_REQUIRE(k >= 1, "k must be >= 1");
    LDDMM_REQUIRE(mode >= LDDMM_DIFFKT_PLAIN && mode <= LDDMM_DIFFKT_LDDMM, "bad mode %d", mode);
    LDDMM_REQUIRE(mode == LDDMM_DIFFKT_PLAIN || n == m, "sym/lddmm need n == m (got %lld, %lld)",
                  (long long)n, (long long)m);
    if (n == 0) return LDDMM_OK;
    LDDMM_REQUIRE(y && out && p, "null y/out/p");
    hipStream_t st = static_cast<hipStream_t>(stream);
    if (m == 0) return launch_fill(out, size_t(n) * dim, 0.0, st);
    LDDMM_REQUIRE(x && a, "null x/a");
    const bool ms = general_path(kd);
    const bool sym = mode != LDDMM_DIFFKT_PLAIN;
    const double r2w = mode == LDDMM_DIFFKT_LDDMM ? 2.0 * regweight : 0.0;
    if (kd->prec == LDDMM_PREC_F32C && !sym && kd->kind == LDDMM_KERNEL_GAUSS &&
        pair_engine() == LDDMM_ENGINE_MFMA)
        return mfma_diffKT_plain(kd, y, int(n), x, int(m), dim, p, a, k, out, st);
#define LDDMM_DKT(T, S, M) diffKT_t<T, S, M>(kd, y, n, x, m, dim, p, a, k, r2w, out, st)
    if (kd->prec == LDDMM_PREC_F32C) {
        if (sym) return ms ? LDDMM_DKT(float, true, true) : LDDMM_DKT(float, true, false);
        return ms ? LDDMM_DKT(float, false, true) : LDDMM_DKT(float, false, false);
    }
    if (sym) return ms ? LDDMM_DKT(double, true, true) : LDDMM_DKT(double, true, false);
    return ms ? LDDMM_DKT(double, false, true) : LDDMM_DKT(double, false, false);
#undef LDDMM_DKT
}

extern "C" int lddmm_applyDiv(const lddmm_kernel* kd, const double* y, int64_t n, const double* x, int64_t m,
                              int32_t dim, const double* a, double* out, void* stream) {
    LDDMM_TRY(check_common(kd, n, m, dim));
    if (n == 0) return LDDMM_OK;
    LDDMM_REQUIRE(y && out, "null y/out");
    hipStream_t st = static_cast<hipStream_t>(stream);
    if (m == 0) return launch_fill(out, size_t(n), 0.0, st);
    LDDMM_REQUIRE(x && a, "null x/a");
    auto run = [&](auto Tz, auto MSc) -> int {
        using T = decltype(Tz);
        constexpr bool MS = decltype(MSc)::value;
        SingleScaleConsts ss;
        ScaleSet<T> sc = make_scales<T>(kd, &ss);
        Job<T, OpDiv<MS>> j;
        LDDMM_TRY(plan_job(j, int(n), int(m), !MS));
        void* ws;
        LDDMM_TRY(workspace(partial_bytes<T>(j.plan, 1, int(n)) + 512, st, &ws));
        char* base = static_cast<char*>(ws);
        Carve cv;
        double* c0 = reinterpret_cast<double*>(base + cv.take(32));
        LDDMM_TRY(launch_center(y, int(n), x, int(m), dim, c0, st));
        Partials<T> part = carve_partials<T>(base, cv, j.plan, 1, int(n));
        j.args = {y, x, a, c0, dim, sc};
        LDDMM_TRY(run_job(j, int(n), int(m), part, st));
        ChanScale cs{};
        cs.s[0] = MS ? -1.0 : ss.diff;
        return run_reduce(part, j.plan.nchunk, 1, 1, int(n), cs, out, 1, 0, 0, nullptr, 1.0, st);
    };
    const bool ms = general_path(kd);
    if (kd->prec == LDDMM_PREC_F32C)
        return ms ? run(float{}, std::true_type{}) : run(float{}, std::false_type{});
    return ms ? run(double{}, std::true_type{}) : run(double{}, std::false_type{});
}

extern "C" int lddmm_applyktensor(const lddmm_kernel* kd, const double* y, int64_t n, const double* x,
                                  int64_t m, int32_t dim, const double* ay, const double* ax,
                                  const double* betay, const double* betax, int32_t kb, double* out,
                                  void* stream) {
    LDDMM_TRY(check_common(kd, n, m, dim));
    if (kb < 1 || kb > 4) {
        set_error("applyktensor supports 1 <= kb <= 4 (got %d)", kb);
        return LDDMM_EUNSUPPORTED;
    }
    if (n == 0) return LDDMM_OK;
    LDDMM_REQUIRE(y && out && ay && betay, "null target arrays");
    hipStream_t st = static_cast<hipStream_t>(stream);
    if (m == 0) return launch_fill(out, size_t(n), 0.0, st);
    LDDMM_REQUIRE(x && ax && betax, "null source arrays");
    const bool ms = general_path(kd);
    if (kd->prec == LDDMM_PREC_F32C)
        return ms ? ktensor_t<float, true>(kd, y, n, x, m, dim, ay, ax, betay, betax, kb, out, st)
                  : ktensor_t<float, false>(kd, y, n, x, m, dim, ay, ax, betay, betax, kb, out, st);
    return ms ? ktensor_t<double, true>(kd, y, n, x, m, dim, ay, ax, betay, betax, kb, out, st)
              : ktensor_t<double, false>(kd, y, n, x, m, dim, ay, ax, betay, betax, kb, out, st);
}

extern "C" int lddmm_applyDiffK(const lddmm_kernel* kd, const double* x, int64_t n, int32_t dim,
                                const double* a1, const double* a2, int32_t k, int32_t option,
                                double* out, void* stream) {
    LDDMM_TRY(check_common(kd, n, n, dim));
    LDDMM_REQUIRE(k >= 1, "k must be >= 1");
    LDDMM_REQUIRE(option >= LDDMM_DIFFK_1 && option <= LDDMM_DIFFK_1AND2, "bad option %d", option);
    if (n == 0) return LDDMM_OK;
    LDDMM_REQUIRE(x && a1 && a2 && out, "null pointer");
    hipStream_t st = static_cast<hipStream_t>(stream);
    const bool ms = general_path(kd);
    const bool f32 = kd->prec == LDDMM_PREC_F32C;
#define LDDMM_DK(T, O, M) diffK_t<T, O, M>(kd, x, n, dim, a1, a2, k, out, st)
#define LDDMM_DK_OPT(T, M)                                                     \
    (option == LDDMM_DIFFK_1 ? LDDMM_DK(T, LDDMM_DIFFK_1, M)                   \
     : option == LDDMM_DIFFK_2 ? LDDMM_DK(T, LDDMM_DIFFK_2, M)                 \
                               : LDDMM_DK(T, LDDMM_DIFFK_1AND2, M))
    if (f32) return ms ? LDDMM_DK_OPT(float, true) : LDDMM_DK_OPT(float, false);
    return ms ? LDDMM_DK_OPT(double, true) : LDDMM_DK_OPT(double, false);
#undef LDDMM_DK_OPT
#undef LDDMM_DK
}

extern "C" int lddmm_applyDDiffK(const lddmm_kernel* kd, const double* x, int64_t n, int32_t dim,
                                 const double* a1, const double* a2, int32_t k, const double* p,
                                 int32_t option, double* out, void* stream) {
    LDDMM_TRY(check_common(kd, n, n, dim));
    LDDMM_REQUIRE(k >= 1, "k must be >= 1");
    LDDMM_REQUIRE(option >= LDDMM_DDIFFK_11 && option <= LDDMM_DDIFFK_11AND12, "bad option %d", option);
    if (n == 0) return LDDMM_OK;
    LDDMM_REQUIRE(x && a1 && a2 && p && out, "null pointer");
    hipStream_t st = static_cast<hipStream_t>(stream);
    const bool ms = general_path(kd);
    const bool f32 = kd->prec == LDDMM_PREC_F32C;
#define LDDMM_DDK(T, O, M) ddiffK_t<T, O, M>(kd, x, n, dim, a1, a2, k, p, out, st)
#define LDDMM_DDK_OPT(T, M)                                                    \
    (option == LDDMM_DDIFFK_11 ? LDDMM_DDK(T, LDDMM_DDIFFK_11, M)              \
     : option == LDDMM_DDIFFK_12 ? LDDMM_DDK(T, LDDMM_DDIFFK_12, M)            \
                                 : LDDMM_DDK(T, LDDMM_DDIFFK_11AND12, M))
    if (f32) return ms ? LDDMM_DDK_OPT(float, true) : LDDMM_DDK_OPT(float, false);
    return ms ? LDDMM_DDK_OPT(double, true) : LDDMM_DDK_OPT(double, false);
#undef LDDMM_DDK_OPT
#undef LDDMM_DDK
}

extern "C" int lddmm_applykmat(const lddmm_kernel* kd, const double* y, int64_t n, const double* x,
                               int64_t m, int32_t dim, const double* beta, int32_t kb, double* out,
                               void* stream) {
    LDDMM_TRY(check_common(kd, n, m, dim));
    LDDMM_REQUIRE(kb >= 1, "kb must be >= 1");
    if (n == 0) return LDDMM_OK;
    LDDMM_REQUIRE(y && out, "null y/out");
    hipStream_t st = static_cast<hipStream_t>(stream);
    if (m == 0) return launch_fill(out, size_t(n) * kb, 0.0, st);
    LDDMM_REQUIRE(x && beta, "null x/beta");
    ScaleSet<double> sc = make_scales<double>(kd, nullptr);
    hipLaunchKernelGGL((k_dense<false>), dim3(unsigned((n + 3) / 4)), dim3(256), 0, st, y, int(n), x,
                       int(m), dim, beta, kb, sc, out);
    LDDMM_CHECK_LAUNCH();
    return LDDMM_OK;
}

extern "C" int lddmm_applykdiffmat(const lddmm_kernel* kd, const double* y, int64_t n, const double* x,
                                   int64_t m, int32_t dim, const double* beta, double* out,
                                   void* stream) {
    LDDMM_TRY(check_common(kd, n, m, dim));
    if (n == 0) return LDDMM_OK;
    LDDMM_REQUIRE(y && out, "null y/out");
    hipStream_t st = static_cast<hipStream_t>(stream);
    if (m == 0) return launch_fill(out, size_t(n) * dim, 0.0, st);
    LDDMM_REQUIRE(x && beta, "null x/beta");
    ScaleSet<double> sc = make_scales<double>(kd, nullptr);
    hipLaunchKernelGGL((k_dense<true>), dim3(unsigned((n + 3) / 4)), dim3(256), 0, st, y, int(n), x,
                       int(m), dim, beta, 1, sc, out);
    LDDMM_CHECK_LAUNCH();
    return LDDMM_OK;
}
