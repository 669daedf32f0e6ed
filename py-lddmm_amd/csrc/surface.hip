// Surface data terms on the GPU: face features, the orientation-free
// varifold energy and its fused gradient, and the deterministic face ->
// vertex gather (py-lddmm base/surfaceDistances.py:286-468, :87-139).
#include "pair.hpp"

#include <algorithm>

namespace lddmm {

// Per-face record in float64: c (3), a*w (area times face weight), n (3), w.
struct FaceRec { double c0, c1, c2, aw, n0, n1, n2, fw; };

// Face features (base/surfaces.py:326-335, base/surfaceDistances.py:340-343):
// c = (x1+x2+x3)/3, S = (x2-x1)x(x3-x1)/2, a = sqrt(|S|^2 + 1e-10), n = S/a,
// face weight w = mean of the three vertex weights.
__global__ void k_face_features(const double* __restrict__ v, const int32_t* __restrict__ f, int nf,
                                const double* __restrict__ vw, FaceRec* __restrict__ rec,
                                double* __restrict__ centers, double* __restrict__ wsurfel,
                                double* __restrict__ fweights) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nf) return;
    int i0 = f[3 * i], i1 = f[3 * i + 1], i2 = f[3 * i + 2];
    double x1[3], x2[3], x3[3];
    for (int d = 0; d < 3; ++d) {
        x1[d] = v[3 * size_t(i0) + d];
        x2[d] = v[3 * size_t(i1) + d];
        x3[d] = v[3 * size_t(i2) + d];
    }
    double e1[3] = {x2[0] - x1[0], x2[1] - x1[1], x2[2] - x1[2]};
    double e2[3] = {x3[0] - x1[0], x3[1] - x1[1], x3[2] - x1[2]};
    double S[3] = {0.5 * (e1[1] * e2[2] - e1[2] * e2[1]), 0.5 * (e1[2] * e2[0] - e1[0] * e2[2]),
                   0.5 * (e1[0] * e2[1] - e1[1] * e2[0])};
    double w = vw ? (vw[i0] + vw[i1] + vw[i2]) / 3.0 : 1.0;
    double c[3] = {(x1[0] + x2[0] + x3[0]) / 3.0, (x1[1] + x2[1] + x3[1]) / 3.0,
                   (x1[2] + x2[2] + x3[2]) / 3.0};
    if (rec) {
        double a = sqrt(S[0] * S[0] + S[1] * S[1] + S[2] * S[2] + 1e-10);
        rec[i] = FaceRec{c[0], c[1], c[2], a * w, S[0] / a, S[1] / a, S[2] / a, w};
    }
    if (centers)
        for (int d = 0; d < 3; ++d) centers[3 * size_t(i) + d] = c[d];
    if (wsurfel)
        for (int d = 0; d < 3; ++d) wsurfel[3 * size_t(i) + d] = S[d] * w;
    if (fweights) fweights[i] = w;
}

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
    t.n0 = T(f.n0);
    t.n1 = T(f.n1);
    t.n2 = T(f.n2);
}

template <typename T>
__device__ __forceinline__ T fma_(T a, T b, T c) { return fma(a, b, c); }

// Energy rows: row_i = sum_j g(r2) a_j (1 + (n_i.n_j)^2)
template <bool MS>
struct OpVarifoldE {
    static constexpr int NCH = 1;
    static constexpr int REC = 8;
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

// Gradient sums: S1 = sum g a_j c n_j, A = sum g a_j, B = sum g a_j c^2,
// D = sum g' a_j (1 + c^2) (u_i - u_j)
template <bool MS>
struct OpVarifoldG {
    static constexpr int NCH = 8;
    static constexpr int REC = 8;
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
        T Kac2 = Kac * c;
        acc[3] += Ka;
        acc[4] += Kac2;
        T w;
        if constexpr (!MS) w = Ka + Kac2;
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

// Energy: E = fK * sum_i aw_i row_i, one block, fixed order (deterministic).
template <typename T>
__global__ void __launch_bounds__(1024) k_vf_energy(Partials<T> part, int nchunk, const FaceRec* f1, int n1,
                                                    double fK, double* energy) {
    __shared__ double red[1024];
    double s = 0.0;
    for (int i = threadIdx.x; i < n1; i += 1024) s += f1[i].aw * sum_range(part, 0, nchunk, 1, i, 0);
    red[threadIdx.x] = s;
    __syncthreads();
    for (int o = 512; o > 0; o >>= 1) {
        if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) *energy = fK * red[0];
}

// Gradient epilogue per face: z1, dz1 and the energy row.
template <typename T>
__global__ void k_vf_grad_epi(Partials<T> part, int c1, int nchunk, const FaceRec* f1, int n1, double fK,
                              double fD, double* z1, double* dz1, double* erow) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n1) return;
    double S[3], D[3];
    for (int c = 0; c < 3; ++c) {
        S[c] = sum_range(part, 0, nchunk, 8, i, c);
        D[c] = sum_range(part, 0, nchunk, 8, i, 5 + c);
    }
    double A1 = sum_range(part, 0, c1, 8, i, 3), A2 = sum_range(part, c1, nchunk, 8, i, 3);
    double B1 = sum_range(part, 0, c1, 8, i, 4), B2 = sum_range(part, c1, nchunk, 8, i, 4);
    const FaceRec f = f1[i];
    double AmB = (A1 + A2) - (B1 + B2);
    double n[3] = {f.n0, f.n1, f.n2};
    for (int c = 0; c < 3; ++c) {
        z1[3 * size_t(i) + c] = f.fw * fK * (2.0 * S[c] + AmB * n[c]);
        dz1[3 * size_t(i) + c] = (2.0 / 3.0) * fD * f.aw * D[c];
    }
    // surface-2 sources carry weight -a_j here; the cross term of the energy is -2<S1,S2>
    erow[i] = f.aw * fK * ((A1 + B1) + 2.0 * (A2 + B2));
}

__global__ void __launch_bounds__(1024) k_sum_rows(const double* rows, int n, double* out) {
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

// px[v] = sum over incident (face, corner): dz1_f - (x_{corner+2} - x_{corner+1}) x z1_f
__global__ void k_face_gather(const double* __restrict__ x, const int32_t* __restrict__ f,
                              const int32_t* __restrict__ ptr, const int32_t* __restrict__ idx, int nv,
                              const double* __restrict__ dz1, const double* __restrict__ z1,
                              double* __restrict__ px) {
    int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= nv) return;
    double acc[3] = {0.0, 0.0, 0.0};
    for (int e = ptr[v]; e < ptr[v + 1]; ++e) {
        int code = idx[e];
        int fi = code / 3, c = code - 3 * fi;
        int va = f[3 * fi + (c + 1) % 3], vb = f[3 * fi + (c + 2) % 3];
        double d[3], z[3];
        for (int k = 0; k < 3; ++k) {
            d[k] = x[3 * size_t(vb) + k] - x[3 * size_t(va) + k];
            z[k] = z1[3 * size_t(fi) + k];
        }
        acc[0] += dz1[3 * size_t(fi) + 0] - (d[1] * z[2] - d[2] * z[1]);
        acc[1] += dz1[3 * size_t(fi) + 1] - (d[2] * z[0] - d[0] * z[2]);
        acc[2] += dz1[3 * size_t(fi) + 2] - (d[0] * z[1] - d[1] * z[0]);
    }
    for (int k = 0; k < 3; ++k) px[3 * size_t(v) + k] = acc[k];
}

namespace {

int check_surface(const lddmm_surface* s, bool need_adj, const char* what) {
    LDDMM_REQUIRE(s != nullptr, "%s: null surface", what);
    LDDMM_REQUIRE(s->nv >= 0 && s->nf >= 0, "%s: negative sizes", what);
    LDDMM_REQUIRE(s->nf < (int64_t(1) << 30) && s->nv < (int64_t(1) << 31), "%s: sizes beyond int32", what);
    LDDMM_REQUIRE(s->nf == 0 || (s->vertices && s->faces), "%s: null vertices / faces", what);
    if (need_adj) LDDMM_REQUIRE(s->nv == 0 || (s->adj_ptr && s->adj_idx), "%s: adjacency required", what);
    return LDDMM_OK;
}

int launch_features(const lddmm_surface* s, FaceRec* rec, double* c, double* ws, double* fw, hipStream_t st) {
    if (s->nf == 0) return LDDMM_OK;
    hipLaunchKernelGGL(k_face_features, dim3(unsigned((s->nf + 255) / 256)), dim3(256), 0, st, s->vertices,
                       s->faces, int(s->nf), s->vweights, rec, c, ws, fw);
    LDDMM_CHECK_LAUNCH();
    return LDDMM_OK;
}

// Centre of the face centres' bounding box: the FaceRec layout starts with c (3) of 8 doubles,
// so the union box is taken over a strided view of both record arrays.
__global__ void __launch_bounds__(1024) k_center_faces(const FaceRec* a, int na, const FaceRec* b, int nb,
                                                       double* c0) {
    __shared__ double lo[3][1024], hi[3][1024];
    int t = threadIdx.x;
    double l[3] = {1e300, 1e300, 1e300}, h[3] = {-1e300, -1e300, -1e300};
    // strided sample of at most 2048 records (any central point serves)
    const int tot = na + nb;
    const int stp = tot > 2048 ? (tot + 2047) / 2048 : 1;
    for (int i = t * stp; i < tot; i += 1024 * stp) {
        const FaceRec& f = i < na ? a[i] : b[i - na];
        double c[3] = {f.c0, f.c1, f.c2};
        for (int d = 0; d < 3; ++d) { l[d] = fmin(l[d], c[d]); h[d] = fmax(h[d], c[d]); }
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
        if (t < 3 && lo[t][0] <= hi[t][0]) c = 0.5 * (lo[t][0] + hi[t][0]);
        if (!isfinite(c)) c = 0.0;
        c0[t] = c;
    }
}

template <typename T, bool MS>
int varifold_t(const lddmm_kernel* kd, const lddmm_surface* s1, const lddmm_surface* s2, double* energy,
               double* grad, hipStream_t st) {
    SingleScaleConsts ss;
    ScaleSet<T> sc = make_scales<T>(kd, &ss);
    const int n1 = int(s1->nf), n2 = s2 ? int(s2->nf) : 0;
    const double fK = 1.0;
    const double fD = MS ? -1.0 : ss.diff;
    const int R1 = n1 >= 8192 && !MS ? 2 : 1;

    PairPlan plan;
    size_t need = 256 + size_t(n1 + n2) * sizeof(FaceRec) + 256;
    if (grad) {
        DeviceInfo di;
        LDDMM_TRY(device_info(&di));
        int occ = blocks_per_cu(R1 == 2 ? reinterpret_cast<const void*>(&k_pair<T, OpVarifoldG<MS>, 2>)
                                        : reinterpret_cast<const void*>(&k_pair<T, OpVarifoldG<MS>, 1>),
                                kBlock, 0);
        plan = plan_pair(n1, n1 + n2, R1, di.num_cu, occ);
        // chunks aligned to the surface boundary
        int c1 = (n1 + plan.chunk - 1) / plan.chunk, c2 = (n2 + plan.chunk - 1) / plan.chunk;
        plan.nchunk = c1 + c2;
        need += partial_bytes<T>(plan, 8, n1) + size_t(n1) * (7 * sizeof(double)) + 1024;
    } else {
        if (R1 == 2) LDDMM_TRY((plan_for<T, OpVarifoldE<MS>, 2>(n1, n1 + n2, &plan)));
        else LDDMM_TRY((plan_for<T, OpVarifoldE<MS>, 1>(n1, n1 + n2, &plan)));
        need += partial_bytes<T>(plan, 1, n1);
    }
    void* ws;
    LDDMM_TRY(workspace(need, st, &ws));
    char* base = static_cast<char*>(ws);
    Carve cv;
    double* c0 = reinterpret_cast<double*>(base + cv.take(32));
    FaceRec* r1 = reinterpret_cast<FaceRec*>(base + cv.take(size_t(n1) * sizeof(FaceRec)));
    FaceRec* r2 = reinterpret_cast<FaceRec*>(base + cv.take(size_t(std::max(n2, 1)) * sizeof(FaceRec)));
    LDDMM_TRY(launch_features(s1, r1, nullptr, nullptr, nullptr, st));
    if (s2) LDDMM_TRY(launch_features(s2, r2, nullptr, nullptr, nullptr, st));
    hipLaunchKernelGGL(k_center_faces, dim3(1), dim3(1024), 0, st, r1, n1, r2, n2, c0);
    LDDMM_CHECK_LAUNCH();

    if (!grad) {
        typename OpVarifoldE<MS>::template Args<T> A;
        A.f1 = r1; A.f2 = r2; A.n1 = n1; A.n2 = n2; A.w2 = -2.0; A.c1 = 0; A.c0 = c0; A.sc = sc;
        Partials<T> part = carve_partials<T>(base, cv, plan, 1, n1);
        if (R1 == 2) LDDMM_TRY((run_pair<T, OpVarifoldE<MS>, 2>(A, n1, n1 + n2, plan, part, st)));
        else LDDMM_TRY((run_pair<T, OpVarifoldE<MS>, 1>(A, n1, n1 + n2, plan, part, st)));
        hipLaunchKernelGGL((k_vf_energy<T>), dim3(1), dim3(1024), 0, st, part, plan.nchunk, r1, n1, fK, energy);
        LDDMM_CHECK_LAUNCH();
        return LDDMM_OK;
    }

    typename OpVarifoldG<MS>::template Args<T> A;
    const int c1 = (n1 + plan.chunk - 1) / plan.chunk;
    A.f1 = r1; A.f2 = r2; A.n1 = n1; A.n2 = n2; A.w2 = -1.0; A.c1 = c1; A.c0 = c0; A.sc = sc;
    Partials<T> part = carve_partials<T>(base, cv, plan, 8, n1);
    double* z1 = reinterpret_cast<double*>(base + cv.take(size_t(n1) * 3 * sizeof(double)));
    double* dz1 = reinterpret_cast<double*>(base + cv.take(size_t(n1) * 3 * sizeof(double)));
    double* erow = reinterpret_cast<double*>(base + cv.take(size_t(n1) * sizeof(double)));
    if (R1 == 2) LDDMM_TRY((run_pair<T, OpVarifoldG<MS>, 2>(A, n1, n1 + n2, plan, part, st)));
    else LDDMM_TRY((run_pair<T, OpVarifoldG<MS>, 1>(A, n1, n1 + n2, plan, part, st)));
    hipLaunchKernelGGL((k_vf_grad_epi<T>), dim3(unsigned((n1 + 255) / 256)), dim3(256), 0, st, part, c1,
                       plan.nchunk, r1, n1, fK, fD, z1, dz1, erow);
    LDDMM_CHECK_LAUNCH();
    if (energy) {
        hipLaunchKernelGGL(k_sum_rows, dim3(1), dim3(1024), 0, st, erow, n1, energy);
        LDDMM_CHECK_LAUNCH();
    }
    if (s1->nv > 0) {
        hipLaunchKernelGGL(k_face_gather, dim3(unsigned((s1->nv + 255) / 256)), dim3(256), 0, st, s1->vertices,
                           s1->faces, s1->adj_ptr, s1->adj_idx, int(s1->nv), dz1, z1, grad);
        LDDMM_CHECK_LAUNCH();
    }
    return LDDMM_OK;
}

}  // namespace
}  // namespace lddmm

using namespace lddmm;

extern "C" int lddmm_surface_features(const lddmm_surface* s, double* centers, double* wsurfel,
                                      double* face_weights, void* stream) {
    LDDMM_TRY(check_surface(s, false, "surface_features"));
    return launch_features(s, nullptr, centers, wsurfel, face_weights, static_cast<hipStream_t>(stream));
}

extern "C" int lddmm_surface_face_gather(const lddmm_surface* s, const double* dz1, const double* z1,
                                         double* px, void* stream) {
    LDDMM_TRY(check_surface(s, true, "surface_face_gather"));
    if (s->nv == 0) return LDDMM_OK;
    LDDMM_REQUIRE(px && (s->nf == 0 || (dz1 && z1)), "surface_face_gather: null arrays");
    hipLaunchKernelGGL(k_face_gather, dim3(unsigned((s->nv + 255) / 256)), dim3(256), 0,
                       static_cast<hipStream_t>(stream), s->vertices, s->faces, s->adj_ptr, s->adj_idx,
                       int(s->nv), dz1, z1, px);
    LDDMM_CHECK_LAUNCH();
    return LDDMM_OK;
}

extern "C" int lddmm_surface_varifold(const lddmm_kernel* kd, const lddmm_surface* s1,
                                      const lddmm_surface* tgt, double* energy, double* grad, void* stream) {
    LDDMM_TRY(validate_kernel(kd));
    LDDMM_TRY(check_surface(s1, grad != nullptr, "surface_varifold"));
    if (tgt) LDDMM_TRY(check_surface(tgt, false, "surface_varifold (target)"));
    LDDMM_REQUIRE(energy || grad, "surface_varifold: nothing requested");
    hipStream_t st = static_cast<hipStream_t>(stream);
    if (s1->nf == 0) {
        if (energy) LDDMM_TRY(launch_fill(energy, 1, 0.0, st));
        if (grad) LDDMM_TRY(launch_fill(grad, size_t(s1->nv) * 3, 0.0, st));
        return LDDMM_OK;
    }
    const bool ms = kd->nscale > 1;
    if (kd->prec == LDDMM_PREC_F32C)
        return ms ? varifold_t<float, true>(kd, s1, tgt, energy, grad, st)
                  : varifold_t<float, false>(kd, s1, tgt, energy, grad, st);
    return ms ? varifold_t<double, true>(kd, s1, tgt, energy, grad, st)
              : varifold_t<double, false>(kd, s1, tgt, energy, grad, st);
}
