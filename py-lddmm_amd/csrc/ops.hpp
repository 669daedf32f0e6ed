// Pair-sum operators (pair functions, LDS records, target state) of the
// Gaussian kernel sums, shared by the single-call entry points and by the
// device-resident shooting loops.
#pragma once

#include "pair.hpp"

#include <type_traits>

namespace lddmm {

template <typename T>
__device__ __forceinline__ T fmaT(T a, T b, T c) { return fma(a, b, c); }

// ---------------------------------------------------------------------------
// applyK:  out_i = sum_j g(r2) a_j          (base/kernelFunctions_util.py:451-466)
// ---------------------------------------------------------------------------
template <int KW, bool MS>
struct OpApplyK {
    static constexpr int NCH = KW;
    static constexpr int REC = 8;
    static constexpr bool kMultiScale = MS;
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
#pragma unroll
        for (int c = 0; c < KW; ++c) rec[3 + c] = T(r.a[c]);
    }
    // record: u (3) then a_j (KW), packed; the pair reads only those values
    template <typename T>
    __device__ static void pair(const Args<T>& A, const Tgt<T>& t, const T* rec, T* acc) {
        T v[3 + KW];
        lds_n<3 + KW>(rec, v);
        T dx = t.u0 - v[0], dy = t.u1 - v[1], dz = t.u2 - v[2];
        T r2 = fmaT(dz, dz, fmaT(dy, dy, dx * dx));
        T K = gK<T, MS>(A.sc, r2);
#pragma unroll
        for (int c = 0; c < KW; ++c) acc[c] = fmaT(K, v[3 + c], acc[c]);
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
    static constexpr bool kMultiScale = MS;
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
#pragma unroll
        for (int c = 0; c < KW; ++c) rec[3 + c] = T(r.a[c]);
        if constexpr (SYM) {
#pragma unroll
            for (int c = 0; c < KW; ++c) rec[3 + KW + c] = T(r.p[c] - A.r2w * r.a[c]);
        }
    }
    // record: u (3), a_j (KW) and for sym / lddmm q_j (KW), packed
    static constexpr int NV = 3 + (SYM ? 2 : 1) * KW;
    template <typename T>
    __device__ static void pair(const Args<T>& A, const Tgt<T>& t, const T* rec, T* acc) {
        T v[NV];
        lds_n<NV>(rec, v);
        T dx = t.u0 - v[0], dy = t.u1 - v[1], dz = t.u2 - v[2];
        T r2 = fmaT(dz, dz, fmaT(dy, dy, dx * dx));
        T g = gD<T, MS>(A.sc, r2);
        T c = t.p[0] * v[3];
#pragma unroll
        for (int k = 1; k < KW; ++k) c = fmaT(t.p[k], v[3 + k], c);
        if constexpr (SYM) {
#pragma unroll
            for (int k = 0; k < KW; ++k) c = fmaT(t.a[k], v[3 + KW + k], c);
        }
        T w = g * c;
        acc[0] = fmaT(w, dx, acc[0]);
        acc[1] = fmaT(w, dy, acc[1]);
        acc[2] = fmaT(w, dz, acc[2]);
    }
};

// ---------------------------------------------------------------------------
// One step of the first-order adjoint with the Euclidean gradient, fused:
// channels 0..2 the lddmm covector sum of OpDiffKT<KW, true>
// (sum_j g'(r2) (p_i.a_j + a_i.q_j) (y_i - x_j), q_j = p_j - 2 r a_j), channels
// 3..3+KW the kernel sum of q: sum_j g(r2) q_j = -K(x) (2 r a - p), whose
// negative is the Euclidean gradient direction K(x_t)(2 r a_t - p_{t+1})
// (base/surfaceMatching.py:860-891, base/pointEvolution.py:167-236).  The pair's
// distance and, for one Gaussian scale, its exponential serve both sums.
// ---------------------------------------------------------------------------
template <int KW, bool MS>
struct OpAdjoint {
    using Base = OpDiffKT<KW, true, MS>;
    static constexpr int NCH = 3 + KW;
    static constexpr bool kMultiScale = MS;     // general (multi-scale / Laplacian) radial factors
    static constexpr int REC = Base::REC;
    using Chunks = ContiguousChunks;
    template <typename T>
    using Args = typename Base::template Args<T>;
    template <typename T>
    using Tgt = typename Base::template Tgt<T>;
    using Raw = typename Base::Raw;

    template <typename T>
    __device__ static void load_target(const Args<T>& A, int i, Tgt<T>& t) { Base::load_target(A, i, t); }
    template <typename T>
    __device__ static void load_raw(const Args<T>& A, int j, Raw& r) { Base::load_raw(A, j, r); }
    template <typename T>
    __device__ static void store_rec(const Args<T>& A, const Raw& r, T* rec) { Base::store_rec(A, r, rec); }
    template <typename T>
    __device__ static void pair(const Args<T>& A, const Tgt<T>& t, const T* rec, T* acc) {
        T v[Base::NV];                          // u, a_j, q_j (the base op's packed record)
        lds_n<Base::NV>(rec, v);
        T dx = t.u0 - v[0], dy = t.u1 - v[1], dz = t.u2 - v[2];
        T r2 = fmaT(dz, dz, fmaT(dy, dy, dx * dx));
        T g = gD<T, MS>(A.sc, r2);
        T K;
        if constexpr (MS) K = gK<T, MS>(A.sc, r2);
        else K = g;
        T c = t.p[0] * v[3];
#pragma unroll
        for (int k = 1; k < KW; ++k) c = fmaT(t.p[k], v[3 + k], c);
#pragma unroll
        for (int k = 0; k < KW; ++k) c = fmaT(t.a[k], v[3 + KW + k], c);
        T w = g * c;
        acc[0] = fmaT(w, dx, acc[0]);
        acc[1] = fmaT(w, dy, acc[1]);
        acc[2] = fmaT(w, dz, acc[2]);
#pragma unroll
        for (int k = 0; k < KW; ++k) acc[3 + k] = fmaT(K, v[3 + KW + k], acc[3 + k]);
    }
};

// ---------------------------------------------------------------------------
// Divergence of the velocity field K(., x) a at the targets:
//   out_i = sum_j g'(r2) (y_i - x_j).a_j         (base/kernelFunctions_util.py:596-620)
// the radial factor and its scaling are those of OpDiffKT (same derivative).
// ---------------------------------------------------------------------------
template <bool MS>
struct OpDiv {
    static constexpr int NCH = 1;
    static constexpr int REC = 8;
    using Chunks = ContiguousChunks;
    template <typename T>
    struct Args {
        const double *y, *x, *a, *c0;
        int dim;
        ScaleSet<T> sc;
    };
    template <typename T>
    struct Tgt { T u0, u1, u2; };
    struct Raw { double x0, x1, x2, a0, a1, a2; };

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
            r.a0 = r.a1 = r.a2 = 0.0;
            return;
        }
        const double* xp = A.x + size_t(j) * A.dim;
        const double* ap = A.a + size_t(j) * A.dim;
        r.x0 = xp[0];
        r.x1 = A.dim > 1 ? xp[1] : 0.0;
        r.x2 = A.dim > 2 ? xp[2] : 0.0;
        r.a0 = ap[0];
        r.a1 = A.dim > 1 ? ap[1] : 0.0;
        r.a2 = A.dim > 2 ? ap[2] : 0.0;
    }
    template <typename T>
    __device__ static void store_rec(const Args<T>& A, const Raw& r, T* rec) {
        rec[0] = T((r.x0 - A.c0[0]) * A.sc.cs);
        rec[1] = T((r.x1 - A.c0[1]) * A.sc.cs);
        rec[2] = T((r.x2 - A.c0[2]) * A.sc.cs);
        rec[3] = T(0);
        rec[4] = T(r.a0); rec[5] = T(r.a1); rec[6] = T(r.a2); rec[7] = T(0);
    }
    template <typename T>
    __device__ static void pair(const Args<T>& A, const Tgt<T>& t, const T* rec, T* acc) {
        Vec4<T> p = lds4(rec);
        T dx = t.u0 - p.x, dy = t.u1 - p.y, dz = t.u2 - p.z;
        T r2 = fmaT(dz, dz, fmaT(dy, dy, dx * dx));
        T g = gD<T, MS>(A.sc, r2);
        Vec4<T> a = lds4(rec + 4);
        T s = fmaT(dz, a.z, fmaT(dy, a.y, dx * a.x));
        acc[0] = fmaT(g, s, acc[0]);
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
// Geodesic (EPDiff) step, one Gaussian scale, fused: the sums of one forward
// step (base/pointEvolution.py:320-379, :385-472) share the pair's distance and
// exponential g = exp(-|x_i - x_j|^2 / 2s^2):
//   channels 0..KW:     sum_j g a_j                          (K(x) a)
//   channels KW..KW+3:  sum_j g (a_i.a_j) (x_i - x_j)        (DiffKT(x, a, a), plain)
// ---------------------------------------------------------------------------
template <int KW>
struct OpEPDiffFwd {
    static constexpr int NCH = KW + 3;
    static constexpr int REC = 8;
    using Chunks = ContiguousChunks;
    template <typename T>
    struct Args {
        const double *x, *a, *c0;
        ScaleSet<T> sc;
    };
    template <typename T>
    struct Tgt { T u0, u1, u2, a[KW]; };
    struct Raw { double x0, x1, x2, a[KW]; };

    template <typename T>
    __device__ static void load_target(const Args<T>& A, int i, Tgt<T>& t) {
        t.u0 = coord<T>(A.x, i, KW, 0, A.c0, A.sc.cs);
        t.u1 = coord<T>(A.x, i, KW, 1, A.c0, A.sc.cs);
        t.u2 = coord<T>(A.x, i, KW, 2, A.c0, A.sc.cs);
#pragma unroll
        for (int c = 0; c < KW; ++c) t.a[c] = T(A.a[size_t(i) * KW + c]);
    }
    template <typename T>
    __device__ static void load_raw(const Args<T>& A, int j, Raw& r) {
        if (j < 0) {
            r.x0 = A.c0[0]; r.x1 = A.c0[1]; r.x2 = A.c0[2];
#pragma unroll
            for (int c = 0; c < KW; ++c) r.a[c] = 0.0;
            return;
        }
        const double* xp = A.x + size_t(j) * KW;
        r.x0 = xp[0];
        r.x1 = KW > 1 ? xp[1] : 0.0;
        r.x2 = KW > 2 ? xp[2] : 0.0;
#pragma unroll
        for (int c = 0; c < KW; ++c) r.a[c] = A.a[size_t(j) * KW + c];
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
        T g = fast_exp2(-r2);
        Vec4<T> a = lds4(rec + 4);
        acc[0] = fmaT(g, a.x, acc[0]);
        if constexpr (KW > 1) acc[1] = fmaT(g, a.y, acc[1]);
        if constexpr (KW > 2) acc[2] = fmaT(g, a.z, acc[2]);
        T c = t.a[0] * a.x;
        if constexpr (KW > 1) c = fmaT(t.a[1], a.y, c);
        if constexpr (KW > 2) c = fmaT(t.a[2], a.z, c);
        T w = g * c;
        acc[KW] = fmaT(w, dx, acc[KW]);
        acc[KW + 1] = fmaT(w, dy, acc[KW + 1]);
        acc[KW + 2] = fmaT(w, dz, acc[KW + 2]);
    }
};

// ---------------------------------------------------------------------------
// Geodesic adjoint step, one Gaussian scale, fused: the four sums of one
// backward step of secondOrderCovector (base/pointEvolution.py:489-552)
//   zpx = DiffKT_sym(x, px, a) - DDiffK11and12(x, a, a, pa)
//   zpa = K(x) px - DiffK1and2(x, pa, a)
// from one distance and exponential per pair; each sum's per-pair arithmetic
// is that of OpDiffKT<KW, true>, OpDDiffK<KW, 11and12>, OpApplyK and
// OpDiffK<KW, 1and2> (the directional term d.(pa_i - pa_j) is shared by the
// last two).  Channels: [0, 3) DiffKT_sym, [3, 6) DDiffK, [6, 6+KW) K px,
// [6+KW, 6+2KW) DiffK1and2.
// ---------------------------------------------------------------------------
template <int KW>
struct OpEPDiffAdj {
    static_assert(KW >= 1 && KW <= 3, "records pack three 3-vectors after the point");
    static constexpr int NCH = 6 + 2 * KW;
    // record: u (3), a_j (3), px_j (3), pa_j (3), read as three 4-vectors
    static constexpr int REC = 12;
    using Chunks = ContiguousChunks;
    template <typename T>
    struct Args {
        const double *x, *a, *px, *pa, *c0;
        T gamma;
        ScaleSet<T> sc;
    };
    template <typename T>
    struct Tgt { T u0, u1, u2, px[KW], a[KW], pa0, pa1, pa2; };
    struct Raw { double x0, x1, x2, a[KW], px[KW], pa[KW]; };

    template <typename T>
    __device__ static void load_target(const Args<T>& A, int i, Tgt<T>& t) {
        t.u0 = coord<T>(A.x, i, KW, 0, A.c0, A.sc.cs);
        t.u1 = coord<T>(A.x, i, KW, 1, A.c0, A.sc.cs);
        t.u2 = coord<T>(A.x, i, KW, 2, A.c0, A.sc.cs);
#pragma unroll
        for (int c = 0; c < KW; ++c) {
            t.px[c] = T(A.px[size_t(i) * KW + c]);
            t.a[c] = T(A.a[size_t(i) * KW + c]);
        }
        const double* pp = A.pa + size_t(i) * KW;
        t.pa0 = T(pp[0]);
        t.pa1 = KW > 1 ? T(pp[1]) : T(0);
        t.pa2 = KW > 2 ? T(pp[2]) : T(0);
    }
    template <typename T>
    __device__ static void load_raw(const Args<T>& A, int j, Raw& r) {
        if (j < 0) {
            r.x0 = A.c0[0]; r.x1 = A.c0[1]; r.x2 = A.c0[2];
#pragma unroll
            for (int c = 0; c < KW; ++c) { r.a[c] = 0.0; r.px[c] = 0.0; r.pa[c] = 0.0; }
            return;
        }
        const double* xp = A.x + size_t(j) * KW;
        r.x0 = xp[0];
        r.x1 = KW > 1 ? xp[1] : 0.0;
        r.x2 = KW > 2 ? xp[2] : 0.0;
#pragma unroll
        for (int c = 0; c < KW; ++c) {
            r.a[c] = A.a[size_t(j) * KW + c];
            r.px[c] = A.px[size_t(j) * KW + c];
            r.pa[c] = A.pa[size_t(j) * KW + c];
        }
    }
    template <typename T>
    __device__ static void store_rec(const Args<T>& A, const Raw& r, T* rec) {
        rec[0] = T((r.x0 - A.c0[0]) * A.sc.cs);
        rec[1] = T((r.x1 - A.c0[1]) * A.sc.cs);
        rec[2] = T((r.x2 - A.c0[2]) * A.sc.cs);
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const int cc = c < KW ? c : 0;
            rec[3 + c] = c < KW ? T(r.a[cc]) : T(0);
            rec[6 + c] = c < KW ? T(r.px[cc]) : T(0);
            rec[9 + c] = c < KW ? T(r.pa[cc]) : T(0);
        }
    }
    template <typename T>
    __device__ static void pair(const Args<T>& A, const Tgt<T>& t, const T* rec, T* acc) {
        const Vec4<T> r0 = lds4(rec), r1 = lds4(rec + 4), r2v = lds4(rec + 8);
        T dx = t.u0 - r0.x, dy = t.u1 - r0.y, dz = t.u2 - r0.z;
        T r2 = fmaT(dz, dz, fmaT(dy, dy, dx * dx));
        T g = fast_exp2(-r2);
        const Vec4<T> a{r0.w, r1.x, r1.y, T(0)};      // a_j
        const Vec4<T> qx{r1.z, r1.w, r2v.x, T(0)};    // px_j
        const Vec4<T> pl{r2v.y, r2v.z, r2v.w, T(0)};  // pa_j
        // DiffKT sym: c = px_i.a_j + a_i.px_j
        T c = t.px[0] * a.x;
        if constexpr (KW > 1) c = fmaT(t.px[1], a.y, c);
        if constexpr (KW > 2) c = fmaT(t.px[2], a.z, c);
        c = fmaT(t.a[0], qx.x, c);
        if constexpr (KW > 1) c = fmaT(t.a[1], qx.y, c);
        if constexpr (KW > 2) c = fmaT(t.a[2], qx.z, c);
        T w = g * c;
        acc[0] = fmaT(w, dx, acc[0]);
        acc[1] = fmaT(w, dy, acc[1]);
        acc[2] = fmaT(w, dz, acc[2]);
        // DDiffK 11and12: K aa (gamma (d.q) d - q), q = pa_i - pa_j
        T aa = t.a[0] * a.x;
        if constexpr (KW > 1) aa = fmaT(t.a[1], a.y, aa);
        if constexpr (KW > 2) aa = fmaT(t.a[2], a.z, aa);
        const T q0 = t.pa0 - pl.x, q1 = t.pa1 - pl.y, q2 = t.pa2 - pl.z;
        const T dq = fmaT(dz, q2, fmaT(dy, q1, dx * q0));
        const T K = g * aa;
        const T tq = A.gamma * dq;
        acc[3] = fmaT(K, fmaT(tq, dx, -q0), acc[3]);
        acc[4] = fmaT(K, fmaT(tq, dy, -q1), acc[4]);
        acc[5] = fmaT(K, fmaT(tq, dz, -q2), acc[5]);
        // K(x) px
        acc[6] = fmaT(g, qx.x, acc[6]);
        if constexpr (KW > 1) acc[7] = fmaT(g, qx.y, acc[7]);
        if constexpr (KW > 2) acc[8] = fmaT(g, qx.z, acc[8]);
        // DiffK1and2(x, pa, a): g (d.(pa_i - pa_j)) a_j
        const T w4 = g * dq;
        acc[6 + KW] = fmaT(w4, a.x, acc[6 + KW]);
        if constexpr (KW > 1) acc[7 + KW] = fmaT(w4, a.y, acc[7 + KW]);
        if constexpr (KW > 2) acc[8 + KW] = fmaT(w4, a.z, acc[8 + KW]);
    }
};

// ---------------------------------------------------------------------------
// Host helpers
// ---------------------------------------------------------------------------
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

// One planned pair launch of an op.
template <typename T, class OP>
struct Job {
    typename OP::template Args<T> args;
    PairPlan plan;
    int R = 1;
};

template <typename T, class OP>
int plan_job(Job<T, OP>& j, int n, int m, bool allow_r2, int max_nchunk = 1 << 30) {
    j.R = allow_r2 ? pick_R(n) : 1;
    if (j.R == 2) return plan_for<T, OP, 2>(n, m, &j.plan, max_nchunk);
    return plan_for<T, OP, 1>(n, m, &j.plan, max_nchunk);
}

template <typename T, class OP>
int run_job(const Job<T, OP>& j, int n, int m, Partials<T> part, hipStream_t st) {
    if (j.R == 2) return run_pair<T, OP, 2>(j.args, n, m, j.plan, part, st);
    return run_pair<T, OP, 1>(j.args, n, m, j.plan, part, st);
}

}  // namespace lddmm
