// Gaussian kernel sums on point sets: applyK, applyDiffKT, applyktensor,
// applyDiffK, applyDDiffK (pair machinery in pair.hpp) and the dense-weight
// sums applykmat / applykdiffmat.
#include "pair.hpp"

#include <algorithm>
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

__global__ void __launch_bounds__(1024) k_center(const double* y, int n, const double* x, int m, int dim,
                                                 double* c0) {
    __shared__ double lo[3][1024], hi[3][1024];
    int t = threadIdx.x;
    double l[3] = {1e300, 1e300, 1e300}, h[3] = {-1e300, -1e300, -1e300};
    for (int i = t; i < n; i += 1024)
        for (int d = 0; d < dim; ++d) {
            double v = y[size_t(i) * dim + d];
            l[d] = fmin(l[d], v);
            h[d] = fmax(h[d], v);
        }
    for (int i = t; i < m; i += 1024)
        for (int d = 0; d < dim; ++d) {
            double v = x[size_t(i) * dim + d];
            l[d] = fmin(l[d], v);
            h[d] = fmax(h[d], v);
        }
    for (int d = 0; d < 3; ++d) { lo[d][t] = l[d]; hi[d][t] = h[d]; }
    __syncthreads();
    for (int s = 512; s > 0; s >>= 1) {
        if (t < s)
            for (int d = 0; d < 3; ++d) {
                lo[d][t] = fmin(lo[d][t], lo[d][t + s]);
                hi[d][t] = fmax(hi[d][t], hi[d][t + s]);
            }
        __syncthreads();
    }
    if (t < 4) {
        double c = 0.0;
        // an empty or non-finite box centres at 0
        if (t < dim && lo[t][0] <= hi[t][0]) c = 0.5 * (lo[t][0] + hi[t][0]);
        if (!isfinite(c)) c = 0.0;
        c0[t] = c;
    }
}

int launch_center(const double* y, int n, const double* x, int m, int dim, double* c0, hipStream_t st) {
    hipLaunchKernelGGL(k_center, dim3(1), dim3(1024), 0, st, y, n, x, m, dim, c0);
    LDDMM_CHECK_LAUNCH();
    return LDDMM_OK;
}

PairPlan plan_pair(int n, int m, int R, int num_cu, int occ) {
    PairPlan p;
    p.bx = (n + kBlock * R - 1) / (kBlock * R);
    if (m <= 0 || n <= 0) { p.nchunk = 0; p.chunk = kTile; return p; }
    const long cap = std::max(1L, long(num_cu) * std::max(1, occ));
    const int maxS = std::max(1, (m + kTile - 1) / kTile);
    double best = 1e300;
    for (int S = 1; S <= maxS; ++S) {
        int chunk = ((m + S - 1) / S + kTile - 1) / kTile * kTile;
        int se = (m + chunk - 1) / chunk;
        if (se != S) continue;
        long blocks = long(p.bx) * se;
        long rounds = (blocks + cap - 1) / cap;
        // per-block fixed cost (target loads, partial stores) ~ half a tile
        double cost = double(rounds) * double(chunk + kTile / 2) + 0.25 * se;
        if (cost < best) { best = cost; p.nchunk = se; p.chunk = chunk; }
    }
    return p;
}

template <typename T>
__device__ __forceinline__ T fmaT(T a, T b, T c) { return fma(a, b, c); }

// ---------------------------------------------------------------------------
// applyK:  out_i = sum_j g(r2) a_j          (base/kernelFunctions_util.py:451-466)
// ---------------------------------------------------------------------------
template <int KW, bool MS>
struct OpApplyK {
    static constexpr int NCH = KW;
    static constexpr int REC = 8;
    using Chunks = ContiguousChunks;
    template <typename T>
    struct Args {
        const double *y, *x, *a, *c0;
        int dim, lda, col0;
        ScaleSet<T> sc;
    };
    template <typename T>
    struct Tgt { T u0, u1, u2; };
    struct Raw { double x0, x1, x2, a[KW]; };

    template <typename T>
    __device__ static void load_target(const Args<T>& A, int i, Tgt<T>& t) {
        t.u0 = coord<T>(A.y, i, A.dim, 0, A.c0, A.sc.cs);
        t.u1 = coord<T>(A.y, i, A.dim, 1, A.c0, A.sc.cs);
        t.u2 = coord<T>(A.y, i, A.dim, 2, A.c0, A.sc.cs);
    }
    template <typename T>
    __device__ static void load_raw(const Args<T>& A, int j, Raw& r) {
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
    template <typename T>
    __device__ static void store_rec(const Args<T>& A, const Raw& r, T* rec) {
        rec[0] = T((r.x0 - A.c0[0]) * A.sc.cs);
        rec[1] = T((r.x1 - A.c0[1]) * A.sc.cs);
        rec[2] = T((r.x2 - A.c0[2]) * A.sc.cs);
        rec[3] = T(0);
#pragma unroll
        for (int c = 0; c < 4; ++c) rec[4 + c] = c < KW ? T(r.a[c < KW ? c : 0]) : T(0);
    }
    template <typename T>
    __device__ static void pair(const Args<T>& A, const Tgt<T>& t, const T* rec, T* acc) {
        Vec4<T> p = lds4(rec);
        T dx = t.u0 - p.x, dy = t.u1 - p.y, dz = t.u2 - p.z;
        T r2 = fmaT(dz, dz, fmaT(dy, dy, dx * dx));
        T K = gK<T, MS>(A.sc, r2);
        Vec4<T> w = lds4(rec + 4);
        acc[0] = fmaT(K, w.x, acc[0]);
        if constexpr (KW > 1) acc[1] = fmaT(K, w.y, acc[1]);
        if constexpr (KW > 2) acc[2] = fmaT(K, w.z, acc[2]);
        if constexpr (KW > 3) acc[3] = fmaT(K, w.w, acc[3]);
    }
};

// ---------------------------------------------------------------------------
// applyDiffKT:  out_i = sum_j g'(r2) c_ij (y_i - x_j)
//   plain: c_ij = p_i.a_j;  sym/lddmm: c_ij = p_i.a_j + a_i.q_j with
//   q_j = p_j - 2 r a_j (r = 0 for sym)      (base/kernelFunctions_util.py:773-793)
// ---------------------------------------------------------------------------
template <int KW, bool SYM, bool MS>
struct OpDiffKT {
    static constexpr int NCH = 3;
    static constexpr int REC = SYM ? 12 : 8;
    using Chunks = ContiguousChunks;
    template <typename T>
    struct Args {
        const double *y, *x, *p, *a, *c0;
        int dim, lda, col0;
        double r2w;       // 2 * regweight (lddmm) or 0 (sym)
        ScaleSet<T> sc;
    };
    template <typename T>
    struct Tgt { T u0, u1, u2, p[KW], a[SYM ? KW : 1]; };
    struct Raw { double x0, x1, x2, a[KW], p[SYM ? KW : 1]; };

    template <typename T>
    __device__ static void load_target(const Args<T>& A, int i, Tgt<T>& t) {
        t.u0 = coord<T>(A.y, i, A.dim, 0, A.c0, A.sc.cs);
        t.u1 = coord<T>(A.y, i, A.dim, 1, A.c0, A.sc.cs);
        t.u2 = coord<T>(A.y, i, A.dim, 2, A.c0, A.sc.cs);
        const double* pp = A.p + size_t(i) * A.lda + A.col0;
#pragma unroll
        for (int c = 0; c < KW; ++c) t.p[c] = T(pp[c]);
        if constexpr (SYM) {
            const double* ap = A.a + size_t(i) * A.lda + A.col0;
#pragma unroll
            for (int c = 0; c < KW; ++c) t.a[c] = T(ap[c]);
        }
    }
    template <typename T>
    __device__ static void load_raw(const Args<T>& A, int j, Raw& r) {
        if (j < 0) {
            r.x0 = A.c0[0]; r.x1 = A.c0[1]; r.x2 = A.c0[2];
#pragma unroll
            for (int c = 0; c < KW; ++c) r.a[c] = 0.0;
            if constexpr (SYM) {
#pragma unroll
                for (int c = 0; c < KW; ++c) r.p[c] = 0.0;
            }
            return;
        }
        const double* xp = A.x + size_t(j) * A.dim;
        r.x0 = xp[0];
        r.x1 = A.dim > 1 ? xp[1] : 0.0;
        r.x2 = A.dim > 2 ? xp[2] : 0.0;
        const double* ap = A.a + size_t(j) * A.lda + A.col0;
#pragma unroll
        for (int c = 0; c < KW; ++c) r.a[c] = ap[c];
        if constexpr (SYM) {
            const double* pp = A.p + size_t(j) * A.lda + A.col0;
#pragma unroll
            for (int c = 0; c < KW; ++c) r.p[c] = pp[c];
        }
    }
    template <typename T>
    __device__ static void store_rec(const Args<T>& A, const Raw& r, T* rec) {
        rec[0] = T((r.x0 - A.c0[0]) * A.sc.cs);
        rec[1] = T((r.x1 - A.c0[1]) * A.sc.cs);
        rec[2] = T((r.x2 - A.c0[2]) * A.sc.cs);
        rec[3] = T(0);
#pragma unroll
        for (int c = 0; c < 4; ++c) rec[4 + c] = c < KW ? T(r.a[c < KW ? c : 0]) : T(0);
        if constexpr (SYM) {
#pragma unroll
            for (int c = 0; c < 4; ++c) {
                int cc = c < KW ? c : 0;
                rec[8 + c] = c < KW ? T(r.p[cc] - A.r2w * r.a[cc]) : T(0);
            }
        }
    }
    template <typename T>
    __device__ static void pair(const Args<T>& A, const Tgt<T>& t, const T* rec, T* acc) {
        Vec4<T> p = lds4(rec);
        T dx = t.u0 - p.x, dy = t.u1 - p.y, dz = t.u2 - p.z;
        T r2 = fmaT(dz, dz, fmaT(dy, dy, dx * dx));
        T g = gD<T, MS>(A.sc, r2);
        Vec4<T> a = lds4(rec + 4);
        T c = t.p[0] * a.x;
        if constexpr (KW > 1) c = fmaT(t.p[1], a.y, c);
        if constexpr (KW > 2) c = fmaT(t.p[2], a.z, c);
        if constexpr (KW > 3) c = fmaT(t.p[3], a.w, c);
        if constexpr (SYM) {
            Vec4<T> q = lds4(rec + 8);
            c = fmaT(t.a[0], q.x, c);
            if constexpr (KW > 1) c = fmaT(t.a[1], q.y, c);
            if constexpr (KW > 2) c = fmaT(t.a[2], q.z, c);
            if constexpr (KW > 3) c = fmaT(t.a[3], q.w, c);
        }
        T w = g * c;
        acc[0] = fmaT(w, dx, acc[0]);
        acc[1] = fmaT(w, dy, acc[1]);
        acc[2] = fmaT(w, dz, acc[2]);
    }
};

// ---------------------------------------------------------------------------
// applyktensor: out_i = sum_j g(r2)(ay_i ax_j + (by_i.bx_j)^2)
//                                          (base/kernelFunctions_util.py:1475-1506)
// ---------------------------------------------------------------------------
template <int KB, bool MS>
struct OpKTensor {
    static constexpr int NCH = 1;
    static constexpr int REC = 8;
    using Chunks = ContiguousChunks;
    template <typename T>
    struct Args {
        const double *y, *x, *ay, *ax, *by, *bx, *c0;
        int dim, kb;
        ScaleSet<T> sc;
    };
    template <typename T>
    struct Tgt { T u0, u1, u2, ay, b[KB]; };
    struct Raw { double x0, x1, x2, ax, b[KB]; };

    template <typename T>
    __device__ static void load_target(const Args<T>& A, int i, Tgt<T>& t) {
        t.u0 = coord<T>(A.y, i, A.dim, 0, A.c0, A.sc.cs);
        t.u1 = coord<T>(A.y, i, A.dim, 1, A.c0, A.sc.cs);
        t.u2 = coord<T>(A.y, i, A.dim, 2, A.c0, A.sc.cs);
        t.ay = T(A.ay[i]);
#pragma unroll
        for (int c = 0; c < KB; ++c) t.b[c] = T(A.by[size_t(i) * A.kb + c]);
    }
    template <typename T>
    __device__ static void load_raw(const Args<T>& A, int j, Raw& r) {
        if (j < 0) {
            r.x0 = A.c0[0]; r.x1 = A.c0[1]; r.x2 = A.c0[2]; r.ax = 0.0;
#pragma unroll
            for (int c = 0; c < KB; ++c) r.b[c] = 0.0;
            return;
        }
        const double* xp = A.x + size_t(j) * A.dim;
        r.x0 = xp[0];
        r.x1 = A.dim > 1 ? xp[1] : 0.0;
        r.x2 = A.dim > 2 ? xp[2] : 0.0;
        r.ax = A.ax[j];
#pragma unroll
        for (int c = 0; c < KB; ++c) r.b[c] = A.bx[size_t(j) * A.kb + c];
    }
    template <typename T>
    __device__ static void store_rec(const Args<T>& A, const Raw& r, T* rec) {
        rec[0] = T((r.x0 - A.c0[0]) * A.sc.cs);
        rec[1] = T((r.x1 - A.c0[1]) * A.sc.cs);
        rec[2] = T((r.x2 - A.c0[2]) * A.sc.cs);
        rec[3] = T(r.ax);
#pragma unroll
        for (int c = 0; c < 4; ++c) rec[4 + c] = c < KB ? T(r.b[c < KB ? c : 0]) : T(0);
    }
    template <typename T>
    __device__ static void pair(const Args<T>& A, const Tgt<T>& t, const T* rec, T* acc) {
        Vec4<T> p = lds4(rec);
        T dx = t.u0 - p.x, dy = t.u1 - p.y, dz = t.u2 - p.z;
        T r2 = fmaT(dz, dz, fmaT(dy, dy, dx * dx));
        T K = gK<T, MS>(A.sc, r2);
        Vec4<T> b = lds4(rec + 4);
        T g = t.b[0] * b.x;
        if constexpr (KB > 1) g = fmaT(t.b[1], b.y, g);
        if constexpr (KB > 2) g = fmaT(t.b[2], b.z, g);
        if constexpr (KB > 3) g = fmaT(t.b[3], b.w, g);
        acc[0] = fmaT(K, fmaT(g, g, t.ay * p.w), acc[0]);
    }
};

// ---------------------------------------------------------------------------
// applyDiffK: out_i = sum_l -g'(r2) (d_il . A_il) a2_l with A = a1_i (opt 1),
// a1_l (opt 2, overall sign flipped), a1_i - a1_l (opt 1and2)
//                                          (base/kernelFunctions_util.py:655-680)
// ---------------------------------------------------------------------------
template <int KW, int OPT, bool MS>
struct OpDiffK {
    static constexpr int NCH = KW;
    static constexpr int REC = 12;
    using Chunks = ContiguousChunks;
    template <typename T>
    struct Args {
        const double *x, *a1, *a2, *c0;
        int dim, lda, col0;
        ScaleSet<T> sc;
    };
    template <typename T>
    struct Tgt { T u0, u1, u2, b0, b1, b2; };
    struct Raw { double x0, x1, x2, b0, b1, b2, a[KW]; };

    template <typename T>
    __device__ static void load_target(const Args<T>& A, int i, Tgt<T>& t) {
        t.u0 = coord<T>(A.x, i, A.dim, 0, A.c0, A.sc.cs);
        t.u1 = coord<T>(A.x, i, A.dim, 1, A.c0, A.sc.cs);
        t.u2 = coord<T>(A.x, i, A.dim, 2, A.c0, A.sc.cs);
        const double* bp = A.a1 + size_t(i) * A.dim;
        t.b0 = T(bp[0]);
        t.b1 = A.dim > 1 ? T(bp[1]) : T(0);
        t.b2 = A.dim > 2 ? T(bp[2]) : T(0);
    }
    template <typename T>
    __device__ static void load_raw(const Args<T>& A, int j, Raw& r) {
        if (j < 0) {
            r.x0 = A.c0[0]; r.x1 = A.c0[1]; r.x2 = A.c0[2];
            r.b0 = r.b1 = r.b2 = 0.0;
#pragma unroll
            for (int c = 0; c < KW; ++c) r.a[c] = 0.0;
            return;
        }
        const double* xp = A.x + size_t(j) * A.dim;
        const double* bp = A.a1 + size_t(j) * A.dim;
        r.x0 = xp[0];
        r.x1 = A.dim > 1 ? xp[1] : 0.0;
        r.x2 = A.dim > 2 ? xp[2] : 0.0;
        r.b0 = bp[0];
        r.b1 = A.dim > 1 ? bp[1] : 0.0;
        r.b2 = A.dim > 2 ? bp[2] : 0.0;
        const double* ap = A.a2 + size_t(j) * A.lda + A.col0;
#pragma unroll
        for (int c = 0; c < KW; ++c) r.a[c] = ap[c];
    }
    template <typename T>
    __device__ static void store_rec(const Args<T>& A, const Raw& r, T* rec) {
        rec[0] = T((r.x0 - A.c0[0]) * A.sc.cs);
        rec[1] = T((r.x1 - A.c0[1]) * A.sc.cs);
        rec[2] = T((r.x2 - A.c0[2]) * A.sc.cs);
        rec[3] = T(0);
        rec[4] = T(r.b0); rec[5] = T(r.b1); rec[6] = T(r.b2); rec[7] = T(0);
#pragma unroll
        for (int c = 0; c < 4; ++c) rec[8 + c] = c < KW ? T(r.a[c < KW ? c : 0]) : T(0);
    }
    template <typename T>
    __device__ static void pair(const Args<T>& A, const Tgt<T>& t, const T* rec, T* acc) {
        Vec4<T> p = lds4(rec);
        T dx = t.u0 - p.x, dy = t.u1 - p.y, dz = t.u2 - p.z;
        T r2 = fmaT(dz, dz, fmaT(dy, dy, dx * dx));
        T g = gD<T, MS>(A.sc, r2);
        Vec4<T> b = lds4(rec + 4);
        T s;
        if constexpr (OPT == LDDMM_DIFFK_1) s = fmaT(dz, t.b2, fmaT(dy, t.b1, dx * t.b0));
        else if constexpr (OPT == LDDMM_DIFFK_2) s = fmaT(dz, b.z, fmaT(dy, b.y, dx * b.x));
        else s = fmaT(dz, t.b2 - b.z, fmaT(dy, t.b1 - b.y, dx * (t.b0 - b.x)));
        T w = g * s;
        Vec4<T> a = lds4(rec + 8);
        acc[0] = fmaT(w, a.x, acc[0]);
        if constexpr (KW > 1) acc[1] = fmaT(w, a.y, acc[1]);
        if constexpr (KW > 2) acc[2] = fmaT(w, a.z, acc[2]);
        if constexpr (KW > 3) acc[3] = fmaT(w, a.w, acc[3]);
    }
};

// ---------------------------------------------------------------------------
// applyDDiffK: out_i = sum_l ((d.q) d gamma - q) g(r2) (a1_i . a2_l)
//   q = p_i (11), -p_l (12), p_i - p_l (11and12)
//                                          (base/kernelFunctions_util.py:1092-1129)
// ---------------------------------------------------------------------------
template <int KW, int OPT, bool MS>
struct OpDDiffK {
    static constexpr int NCH = 3;
    static constexpr int REC = 12;
    using Chunks = ContiguousChunks;
    template <typename T>
    struct Args {
        const double *x, *a1, *a2, *p, *c0;
        int dim, lda, col0;
        T gamma;
        ScaleSet<T> sc;
    };
    template <typename T>
    struct Tgt { T u0, u1, u2, a[KW], p0, p1, p2; };
    struct Raw { double x0, x1, x2, a[KW], p0, p1, p2; };

    template <typename T>
    __device__ static void load_target(const Args<T>& A, int i, Tgt<T>& t) {
        t.u0 = coord<T>(A.x, i, A.dim, 0, A.c0, A.sc.cs);
        t.u1 = coord<T>(A.x, i, A.dim, 1, A.c0, A.sc.cs);
        t.u2 = coord<T>(A.x, i, A.dim, 2, A.c0, A.sc.cs);
        const double* ap = A.a1 + size_t(i) * A.lda + A.col0;
#pragma unroll
        for (int c = 0; c < KW; ++c) t.a[c] = T(ap[c]);
        const double* pp = A.p + size_t(i) * A.dim;
        t.p0 = T(pp[0]);
        t.p1 = A.dim > 1 ? T(pp[1]) : T(0);
        t.p2 = A.dim > 2 ? T(pp[2]) : T(0);
    }
    template <typename T>
    __device__ static void load_raw(const Args<T>& A, int j, Raw& r) {
        if (j < 0) {
            r.x0 = A.c0[0]; r.x1 = A.c0[1]; r.x2 = A.c0[2];
            r.p0 = r.p1 = r.p2 = 0.0;
#pragma unroll
            for (int c = 0; c < KW; ++c) r.a[c] = 0.0;
            return;
        }
        const double* xp = A.x + size_t(j) * A.dim;
        const double* pp = A.p + size_t(j) * A.dim;
        r.x0 = xp[0];
        r.x1 = A.dim > 1 ? xp[1] : 0.0;
        r.x2 = A.dim > 2 ? xp[2] : 0.0;
        r.p0 = pp[0];
        r.p1 = A.dim > 1 ? pp[1] : 0.0;
        r.p2 = A.dim > 2 ? pp[2] : 0.0;
        const double* ap = A.a2 + size_t(j) * A.lda + A.col0;
#pragma unroll
        for (int c = 0; c < KW; ++c) r.a[c] = ap[c];
    }
    template <typename T>
    __device__ static void store_rec(const Args<T>& A, const Raw& r, T* rec) {
        rec[0] = T((r.x0 - A.c0[0]) * A.sc.cs);
        rec[1] = T((r.x1 - A.c0[1]) * A.sc.cs);
        rec[2] = T((r.x2 - A.c0[2]) * A.sc.cs);
        rec[3] = T(0);
#pragma unroll
        for (int c = 0; c < 4; ++c) rec[4 + c] = c < KW ? T(r.a[c < KW ? c : 0]) : T(0);
        rec[8] = T(r.p0); rec[9] = T(r.p1); rec[10] = T(r.p2); rec[11] = T(0);
    }
    template <typename T>
    __device__ static void pair(const Args<T>& A, const Tgt<T>& t, const T* rec, T* acc) {
        Vec4<T> p = lds4(rec);
        T dx = t.u0 - p.x, dy = t.u1 - p.y, dz = t.u2 - p.z;
        T r2 = fmaT(dz, dz, fmaT(dy, dy, dx * dx));
        Vec4<T> a = lds4(rec + 4);
        T aa = t.a[0] * a.x;
        if constexpr (KW > 1) aa = fmaT(t.a[1], a.y, aa);
        if constexpr (KW > 2) aa = fmaT(t.a[2], a.z, aa);
        if constexpr (KW > 3) aa = fmaT(t.a[3], a.w, aa);
        Vec4<T> pl = lds4(rec + 8);
        T q0, q1, q2;
        if constexpr (OPT == LDDMM_DDIFFK_11) { q0 = t.p0; q1 = t.p1; q2 = t.p2; }
        else if constexpr (OPT == LDDMM_DDIFFK_12) { q0 = -pl.x; q1 = -pl.y; q2 = -pl.z; }
        else { q0 = t.p0 - pl.x; q1 = t.p1 - pl.y; q2 = t.p2 - pl.z; }
        T dq = fmaT(dz, q2, fmaT(dy, q1, dx * q0));
        if constexpr (!MS) {
            T K = fast_exp2(-r2) * aa;
            T tq = A.gamma * dq;
            acc[0] = fmaT(K, fmaT(tq, dx, -q0), acc[0]);
            acc[1] = fmaT(K, fmaT(tq, dy, -q1), acc[1]);
            acc[2] = fmaT(K, fmaT(tq, dz, -q2), acc[2]);
        } else {
            T gdd, gd;
            gDDD<T, MS>(A.sc, r2, gdd, gd);
            T tq = dq * gdd * aa, tp = gd * aa;
            acc[0] = fmaT(tq, dx, fmaT(-tp, q0, acc[0]));
            acc[1] = fmaT(tq, dy, fmaT(-tp, q1, acc[1]));
            acc[2] = fmaT(tq, dz, fmaT(-tp, q2, acc[2]));
        }
    }
};

// ---------------------------------------------------------------------------
// Host drivers
// ---------------------------------------------------------------------------
namespace {

int check_common(const lddmm_kernel* kd, int64_t n, int64_t m, int dim) {
    LDDMM_TRY(validate_kernel(kd));
    LDDMM_REQUIRE(n >= 0 && m >= 0, "negative sizes");
    LDDMM_REQUIRE(n < (int64_t(1) << 31) && m < (int64_t(1) << 31), "sizes beyond int32");
    LDDMM_REQUIRE(dim >= 1 && dim <= 3, "dim %d outside [1, 3]", dim);
    return LDDMM_OK;
}

template <class F>
int with_kw(int kw, F&& f) {
    switch (kw) {
        case 1: return f(std::integral_constant<int, 1>{});
        case 2: return f(std::integral_constant<int, 2>{});
        case 3: return f(std::integral_constant<int, 3>{});
        default: return f(std::integral_constant<int, 4>{});
    }
}

// targets per lane: more targets amortise each LDS record over more pairs
inline int pick_R(int n) { return n >= 8192 ? 2 : 1; }

// One pair launch + reduce of an op whose Args are already filled (c0 set).
template <typename T, class OP>
struct Job {
    typename OP::template Args<T> args;
    PairPlan plan;
    int R;
};

template <typename T, class OP>
int plan_job(Job<T, OP>& j, int n, int m, bool allow_r2) {
    j.R = allow_r2 ? pick_R(n) : 1;
    if (j.R == 2) return plan_for<T, OP, 2>(n, m, &j.plan);
    return plan_for<T, OP, 1>(n, m, &j.plan);
}

template <typename T, class OP>
int run_job(const Job<T, OP>& j, int n, int m, Partials<T> part, hipStream_t st) {
    if (j.R == 2) return run_pair<T, OP, 2>(j.args, n, m, j.plan, part, st);
    return run_pair<T, OP, 1>(j.args, n, m, j.plan, part, st);
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
            return run_reduce(part, j.plan.nchunk, KW, KW, n, cs, out, k, 4 * g, 0, nullptr, 0.0, st);
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
            return run_reduce(part, j.plan.nchunk, 3, dim, n, cs, out, dim, 0, g > 0 ? 1 : 0, nullptr, 0.0, st);
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
        return run_reduce(part, j.plan.nchunk, 1, 1, n, cs, out, 1, 0, 0, nullptr, 0.0, st);
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
            return run_reduce(part, j.plan.nchunk, KW, KW, n, cs, out, k, 4 * g, 0, nullptr, 0.0, st);
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
            return run_reduce(part, j.plan.nchunk, 3, dim, n, cs, out, dim, 0, g > 0 ? 1 : 0, nullptr, 0.0, st);
        });
        if (rc) return rc;
    }
    return LDDMM_OK;
}

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
                double g = 0.0;
                for (int q = 0; q < sc.ns; ++q) g += sc.wD[q] * exp2(-sc.beta[q] * r2);
                double b = beta[size_t(i) * m + j] * g;
                for (int c = 0; c < 4 && c0 + c < dim; ++c) acc[c] -= b * dd[c0 + c];
            } else {
                double g = 0.0;
                for (int q = 0; q < sc.ns; ++q) g += sc.wK[q] * exp2(-sc.beta[q] * r2);
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
    const bool ms = kd->nscale > 1;
    if (kd->prec == LDDMM_PREC_F32C)
        return ms ? applyK_t<float, true>(kd, y, n, x, m, dim, a, k, out, st)
                  : applyK_t<float, false>(kd, y, n, x, m, dim, a, k, out, st);
    return ms ? applyK_t<double, true>(kd, y, n, x, m, dim, a, k, out, st)
              : applyK_t<double, false>(kd, y, n, x, m, dim, a, k, out, st);
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
    const bool ms = kd->nscale > 1;
    const bool sym = mode != LDDMM_DIFFKT_PLAIN;
    const double r2w = mode == LDDMM_DIFFKT_LDDMM ? 2.0 * regweight : 0.0;
#define LDDMM_DKT(T, S, M) diffKT_t<T, S, M>(kd, y, n, x, m, dim, p, a, k, r2w, out, st)
    if (kd->prec == LDDMM_PREC_F32C) {
        if (sym) return ms ? LDDMM_DKT(float, true, true) : LDDMM_DKT(float, true, false);
        return ms ? LDDMM_DKT(float, false, true) : LDDMM_DKT(float, false, false);
    }
    if (sym) return ms ? LDDMM_DKT(double, true, true) : LDDMM_DKT(double, true, false);
    return ms ? LDDMM_DKT(double, false, true) : LDDMM_DKT(double, false, false);
#undef LDDMM_DKT
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
    const bool ms = kd->nscale > 1;
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
    const bool ms = kd->nscale > 1;
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
    const bool ms = kd->nscale > 1;
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
