// Varifold operators shared by the single-shape data terms and the
// population kernels: face / segment records, the two-shape source layout and
// the energy / gradient pair functions (py-lddmm base/surfaceDistances.py:286-468,
// base/curveDistances.py:186-336).
#pragma once

#include "pair.hpp"

namespace lddmm {

// Per-face record in float64: c (3), a*w (area times face weight), n (3), w.
struct FaceRec { double c0, c1, c2, aw, n0, n1, n2, fw; };

// Sources are the faces of surface 1 followed by those of surface 2; the
// weight of a surface-2 face is multiplied by `w2`.  In gradient mode chunks
// never straddle the two surfaces (chunks [0, c1) cover surface 1).
struct VarifoldArgsBase {
    const FaceRec* f1;
    const FaceRec* f2;
    int n1, n2;
    double w2;
    int c1;            // number of surface-1 chunks (gradient mode)
    const double* c0;
    double sd;         // sqrt of the orientation weight d in f(u) = 1 + d u^2 (1 for surfaces)
};

struct TwoSurfaceChunks {
    template <class A>
    __device__ static void range(const A& a, int chunk_id, int chunk, int m, int& j0, int& j1) {
        if (chunk_id < a.c1) {
            j0 = chunk_id * chunk;
            j1 = min(a.n1, j0 + chunk);
        } else {
            j0 = a.n1 + (chunk_id - a.c1) * chunk;
            j1 = min(a.n1 + a.n2, j0 + chunk);
        }
    }
};

struct VarifoldRaw { double c0, c1, c2, a, n0, n1, n2; };

template <typename T>
__device__ __forceinline__ void vf_load_raw(const VarifoldArgsBase& A, int j, VarifoldRaw& r) {
    if (j < 0) {
        r = VarifoldRaw{A.c0[0], A.c0[1], A.c0[2], 0.0, 0.0, 0.0, 0.0};
        return;
    }
    const FaceRec& f = j < A.n1 ? A.f1[j] : A.f2[j - A.n1];
    double w = j < A.n1 ? 1.0 : A.w2;
    r = VarifoldRaw{f.c0, f.c1, f.c2, f.aw * w, f.n0, f.n1, f.n2};
}

template <typename T>
__device__ __forceinline__ void vf_store(const VarifoldArgsBase& A, double cs, const VarifoldRaw& r, T* rec) {
    rec[0] = T((r.c0 - A.c0[0]) * cs);
    rec[1] = T((r.c1 - A.c0[1]) * cs);
    rec[2] = T((r.c2 - A.c0[2]) * cs);
    rec[3] = T(r.a);
    rec[4] = T(r.n0);
    rec[5] = T(r.n1);
    rec[6] = T(r.n2);
    rec[7] = T(0);
}

template <typename T>
struct VfTgt { T u0, u1, u2, n0, n1, n2; };

template <typename T>
__device__ __forceinline__ void vf_target(const VarifoldArgsBase& A, double cs, int i, VfTgt<T>& t) {
    const FaceRec& f = A.f1[i];
    t.u0 = T((f.c0 - A.c0[0]) * cs);
    t.u1 = T((f.c1 - A.c0[1]) * cs);
    t.u2 = T((f.c2 - A.c0[2]) * cs);
    // the target normal carries sqrt(d): c = sqrt(d) n_i.n_j, so c^2 = d (n_i.n_j)^2
    t.n0 = T(f.n0 * A.sd);
    t.n1 = T(f.n1 * A.sd);
    t.n2 = T(f.n2 * A.sd);
}

template <typename T>
__device__ __forceinline__ T fma_(T a, T b, T c) { return fma(a, b, c); }

// Energy rows: row_i = sum_j g(r2) a_j (1 + d (n_i.n_j)^2)
template <bool MS>
struct OpVarifoldE {
    static constexpr int NCH = 1;
    static constexpr int REC = 8;
    static constexpr bool kMultiScale = MS;
    using Chunks = ContiguousChunks;
    template <typename T>
    struct Args : VarifoldArgsBase { ScaleSet<T> sc; };
    template <typename T> using Tgt = VfTgt<T>;
    using Raw = VarifoldRaw;
    template <typename T>
    __device__ static void load_target(const Args<T>& A, int i, Tgt<T>& t) { vf_target(A, A.sc.cs, i, t); }
    template <typename T>
    __device__ static void load_raw(const Args<T>& A, int j, Raw& r) { vf_load_raw<T>(A, j, r); }
    template <typename T>
    __device__ static void store_rec(const Args<T>& A, const Raw& r, T* rec) { vf_store(A, A.sc.cs, r, rec); }
    template <typename T>
    __device__ static void pair(const Args<T>& A, const Tgt<T>& t, const T* rec, T* acc) {
        Vec4<T> p = lds4(rec);
        T dx = t.u0 - p.x, dy = t.u1 - p.y, dz = t.u2 - p.z;
        T r2 = fma_(dz, dz, fma_(dy, dy, dx * dx));
        T K = gK<T, MS>(A.sc, r2);
        Vec4<T> nj = lds4(rec + 4);
        T c = fma_(t.n2, nj.z, fma_(t.n1, nj.y, t.n0 * nj.x));
        acc[0] = fma_(K * p.w, fma_(c, c, T(1)), acc[0]);
    }
};

// Gradient sums with c = sqrt(d) n_i.n_j: S1 = sum g a_j c n_j, A = sum g a_j,
// W = sum g a_j (1 + c^2) (= A + B with B = sum g a_j c^2, which the
// epilogues recover as W - A), D = sum g' a_j (1 + c^2) (u_i - u_j).
// Channels: [0, 3) S1, 3 A, 4 W, [5, 8) D.
template <bool MS>
struct OpVarifoldG {
    static constexpr int NCH = 8;
    static constexpr int REC = 8;
    static constexpr bool kMultiScale = MS;
    using Chunks = TwoSurfaceChunks;
    template <typename T>
    struct Args : VarifoldArgsBase { ScaleSet<T> sc; };
    template <typename T> using Tgt = VfTgt<T>;
    using Raw = VarifoldRaw;
    template <typename T>
    __device__ static void load_target(const Args<T>& A, int i, Tgt<T>& t) { vf_target(A, A.sc.cs, i, t); }
    template <typename T>
    __device__ static void load_raw(const Args<T>& A, int j, Raw& r) { vf_load_raw<T>(A, j, r); }
    template <typename T>
    __device__ static void store_rec(const Args<T>& A, const Raw& r, T* rec) { vf_store(A, A.sc.cs, r, rec); }
    template <typename T>
    __device__ static void pair(const Args<T>& A, const Tgt<T>& t, const T* rec, T* acc) {
        Vec4<T> p = lds4(rec);
        T dx = t.u0 - p.x, dy = t.u1 - p.y, dz = t.u2 - p.z;
        T r2 = fma_(dz, dz, fma_(dy, dy, dx * dx));
        T gk, gd;
        gKD<T, MS>(A.sc, r2, gk, gd);
        Vec4<T> nj = lds4(rec + 4);
        T c = fma_(t.n2, nj.z, fma_(t.n1, nj.y, t.n0 * nj.x));
        T Ka = gk * p.w;
        T Kac = Ka * c;
        acc[0] = fma_(Kac, nj.x, acc[0]);
        acc[1] = fma_(Kac, nj.y, acc[1]);
        acc[2] = fma_(Kac, nj.z, acc[2]);
        const T wk = fma_(Kac, c, Ka);         // Ka (1 + c^2)
        acc[3] += Ka;
        acc[4] += wk;
        T w;
        if constexpr (!MS) w = wk;
        else { T Kd = gd * p.w; w = fma_(Kd * c, c, Kd); }
        acc[5] = fma_(w, dx, acc[5]);
        acc[6] = fma_(w, dy, acc[6]);
        acc[7] = fma_(w, dz, acc[7]);
    }
};

template <typename T>
__device__ __forceinline__ double sum_range(const Partials<T>& p, int k0, int k1, int nch, int i, int c) {
    if constexpr (kCompensated<T>) {
        float h = 0.f, l = 0.f;
        for (int k = k0; k < k1; ++k) {
            size_t o = (size_t(k) * nch + c) * size_t(p.n) + i;
            two_sum_acc(h, l, p.hi[o]);
            l += p.lo[o];
        }
        return double(h) + double(l);
    } else {
        double s = 0.0;
        for (int k = k0; k < k1; ++k) s += p.hi[(size_t(k) * nch + c) * size_t(p.n) + i];
        return s;
    }
}

// Face features of faces [0, nf) -> records / centres / weighted surfels / weights.
__global__ void k_face_features(const double* __restrict__ v, const int32_t* __restrict__ f, int nf,
                                const double* __restrict__ vw, FaceRec* __restrict__ rec,
                                double* __restrict__ centers, double* __restrict__ wsurfel,
                                double* __restrict__ fweights);

// px[v] = sum over incident (face, corner) of dz1_f - (x_{corner+2} - x_{corner+1}) x z1_f.
// With vpair/active set, vertices of inactive pairs are left untouched.
__global__ void k_face_gather(const double* __restrict__ x, const int32_t* __restrict__ f,
                              const int32_t* __restrict__ ptr, const int32_t* __restrict__ idx, int nv,
                              const double* __restrict__ dz1, const double* __restrict__ z1,
                              double* __restrict__ px, const int32_t* __restrict__ vpair,
                              const uint8_t* __restrict__ active);

}  // namespace lddmm
