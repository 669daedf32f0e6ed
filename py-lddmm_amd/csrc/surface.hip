// Surface data terms on the GPU: face features, the orientation-free
// varifold energy and its fused gradient, and the deterministic face ->
// vertex gather (py-lddmm base/surfaceDistances.py:286-468, :87-139).
#include "varifold.hpp"
#include "varifold_mfma.hpp"

#include <algorithm>
#include <atomic>
#include <type_traits>
#include <cmath>

namespace lddmm {

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

// Energy: E = fK * sum_i aw_i row_i, one block, fixed order (deterministic).
template <typename T>
__global__ void __launch_bounds__(256) k_vf_energy_rows(Partials<T> part, int nchunk, const FaceRec* f1, int n1,
                                                        double fK, double* rows) {
    __shared__ double red[kEpiLanes][kEpiRows];
    const int fl = threadIdx.x % kEpiRows, lane = threadIdx.x / kEpiRows;
    const int i = blockIdx.x * kEpiRows + fl;
    red[lane][fl] = i < n1 ? sum_lane(part, 0, nchunk, 1, i, 0, lane) : 0.0;
    __syncthreads();
    if (lane == 0 && i < n1) rows[i] = fK * f1[i].aw * (((red[0][fl] + red[1][fl]) + red[2][fl]) + red[3][fl]);
}

// Gradient epilogue per face: z1, dz1 and the energy row.
// Surfaces (base/surfaceDistances.py:394-468): zf = 1, sd = 1, dzf = 2/3 (centre
// = mean of 3 vertices).  Curves (base/curveDistances.py:280-336): zf = 2,
// sd = sqrt(d), dzf = 1 (no 1/3 factor; the midpoint's 1/2 cancels the 2 of
// the symmetric term).  z1 = fw zf (2 sd S + (A - B) n) is the weighted z1.
// 64 faces per workgroup, 4 chunk lanes per face.
template <typename T>
__global__ void __launch_bounds__(256) k_vf_grad_epi(Partials<T> part, int c1, int nchunk, const FaceRec* f1, int n1,
                                                     double fK, double fD, double zf, double sd, double dzf,
                                                     double* z1, double* dz1, double* erow) {
    constexpr int NV = 10;       // S[3], D[3], A1, A2, W1, W2 (W = A + B)
    __shared__ double red[NV][kEpiLanes][kEpiRows];
    const int fl = threadIdx.x % kEpiRows, lane = threadIdx.x / kEpiRows;
    const int i = blockIdx.x * kEpiRows + fl;
    double v[NV] = {};
    if (i < n1) {
        for (int c = 0; c < 3; ++c) {
            v[c] = sum_lane(part, 0, nchunk, 8, i, c, lane);
            v[3 + c] = sum_lane(part, 0, nchunk, 8, i, 5 + c, lane);
        }
        v[6] = sum_lane(part, 0, c1, 8, i, 3, lane);
        v[7] = sum_lane(part, c1, nchunk, 8, i, 3, lane);
        v[8] = sum_lane(part, 0, c1, 8, i, 4, lane);
        v[9] = sum_lane(part, c1, nchunk, 8, i, 4, lane);
    }
#pragma unroll
    for (int q = 0; q < NV; ++q) red[q][lane][fl] = v[q];
    __syncthreads();
    if (lane != 0 || i >= n1) return;
    double t[NV];
#pragma unroll
    for (int q = 0; q < NV; ++q) t[q] = ((red[q][0][fl] + red[q][1][fl]) + red[q][2][fl]) + red[q][3][fl];
    const double A1 = t[6], A2 = t[7], W1 = t[8], W2 = t[9];
    const FaceRec f = f1[i];
    const double AmB = 2.0 * (A1 + A2) - (W1 + W2);
    const double n[3] = {f.n0, f.n1, f.n2};
    for (int c = 0; c < 3; ++c) {
        z1[3 * size_t(i) + c] = f.fw * fK * zf * (2.0 * sd * t[c] + AmB * n[c]);
        dz1[3 * size_t(i) + c] = dzf * fD * f.aw * t[3 + c];
    }
    // surface-2 sources carry weight -a_j here; the cross term of the energy is -2<S1,S2>
    erow[i] = f.aw * fK * (W1 + 2.0 * W2);
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
                              double* __restrict__ px, const int32_t* __restrict__ vpair,
                              const uint8_t* __restrict__ active) {
    int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= nv) return;
    if (vpair && !active[vpair[v]]) return;
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

std::atomic<int> g_vf_engine{LDDMM_VARIFOLD_VALU};

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
constexpr int kCenterFaceThreads = 256;
constexpr int kCenterFacePer = 2048 / kCenterFaceThreads;   // sampled records per lane

__global__ void __launch_bounds__(kCenterFaceThreads) k_center_faces(const FaceRec* a, int na, const FaceRec* b,
                                                                    int nb, double* c0) {
    constexpr int W = kCenterFaceThreads / 64;
    __shared__ double lo[3][W], hi[3][W];
    const int t = threadIdx.x;
    double l[3] = {1e300, 1e300, 1e300}, h[3] = {-1e300, -1e300, -1e300};
    // strided sample of at most 2048 records (any central point serves); a
    // lane's loads are all issued before any is used
    const int tot = na + nb;
    const int stp = tot > 2048 ? (tot + 2047) / 2048 : 1;
    double c[kCenterFacePer][3];
#pragma unroll
    for (int k = 0; k < kCenterFacePer; ++k) {
        const int64_t i = (int64_t(t) + int64_t(k) * kCenterFaceThreads) * stp;
        const FaceRec* f = i < na ? a + i : (i < tot ? b + (i - na) : nullptr);
        c[k][0] = f ? f->c0 : 0.0;             // unused past the sample
        c[k][1] = f ? f->c1 : 0.0;
        c[k][2] = f ? f->c2 : 0.0;
    }
#pragma unroll
    for (int k = 0; k < kCenterFacePer; ++k) {
        const int64_t i = (int64_t(t) + int64_t(k) * kCenterFaceThreads) * stp;
        if (i < tot) {
#pragma unroll
            for (int d = 0; d < 3; ++d) { l[d] = fmin(l[d], c[k][d]); h[d] = fmax(h[d], c[k][d]); }
        }
    }
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
    if (t < 3) {
#pragma unroll
        for (int w = 1; w < W; ++w) { lo[t][0] = fmin(lo[t][0], lo[t][w]); hi[t][0] = fmax(hi[t][0], hi[t][w]); }
    }
    if (t < 4) {
        double c = 0.0;
        if (t < 3 && lo[t][0] <= hi[t][0]) c = 0.5 * (lo[t][0] + hi[t][0]);
        if (!isfinite(c)) c = 0.0;
        c0[t] = c;
    }
}

// Shape of the varifold for one geometry (surface faces or curve segments).
struct VfShape {
    double sd;    // sqrt(d) of f(u) = 1 + d u^2
    double zf;    // factor of z1
    double dzf;   // factor of dz1 (centre derivative)
};

// Shared driver: `feat(r1, r2)` fills the records of both shapes, `gather(dz1, z1)`
// scatters the per-element gradient to the vertices.
template <typename T, bool MS, class Feat, class Gather>
int varifold_core(const lddmm_kernel* kd, int n1, int n2, bool has_tgt, VfShape shp, Feat&& feat,
                  Gather&& gather, double* energy, bool want_grad, hipStream_t st) {
    SingleScaleConsts ss;
    ScaleSet<T> sc = make_scales<T>(kd, &ss);
    const double fK = 1.0;
    const double fD = MS ? -1.0 : ss.diff;
    const int R1 = n1 >= 8192 && !MS ? 2 : 1;
    // matrix-core energy engine: float32, single-scale Gaussian only
    int vf_engine = LDDMM_VARIFOLD_VALU;
    if constexpr (std::is_same<T, float>::value && !MS)
        if (kd->kind == LDDMM_KERNEL_GAUSS) vf_engine = g_vf_engine.load();

    PairPlan plan;
    size_t need = 256 + size_t(n1 + n2) * sizeof(FaceRec) + 256;
    if (want_grad) {
        DeviceInfo di;
        LDDMM_TRY(device_info(&di));
        int occ = blocks_per_cu(R1 == 2 ? reinterpret_cast<const void*>(&k_pair<T, OpVarifoldG<MS>, 2>)
                                        : reinterpret_cast<const void*>(&k_pair<T, OpVarifoldG<MS>, 1>),
                                kBlock, 0);
        plan = plan_pair(n1, n1 + n2, R1, di.num_cu, occ);
        // chunks aligned to the boundary between the two shapes
        int c1 = (n1 + plan.chunk - 1) / plan.chunk, c2 = (n2 + plan.chunk - 1) / plan.chunk;
        plan.nchunk = c1 + c2;
        need += partial_bytes<T>(plan, 8, n1) + size_t(n1) * (7 * sizeof(double)) + 1024;
    } else {
        if (vf_engine) {
            if (vf_engine == LDDMM_VARIFOLD_MFMA_COS_R2) LDDMM_TRY((plan_vf_energy_mfma<2, true>(n1, n1 + n2, &plan)));
            else LDDMM_TRY((plan_vf_energy_mfma<2, false>(n1, n1 + n2, &plan)));
        } else if (R1 == 2) {
            LDDMM_TRY((plan_for<T, OpVarifoldE<MS>, 2>(n1, n1 + n2, &plan)));
        } else {
            LDDMM_TRY((plan_for<T, OpVarifoldE<MS>, 1>(n1, n1 + n2, &plan)));
        }
        need += partial_bytes<T>(plan, 1, n1) + size_t(n1) * sizeof(double) + 256;
    }
    void* ws;
    LDDMM_TRY(workspace(need, st, &ws));
    char* base = static_cast<char*>(ws);
    Carve cv;
    double* c0 = reinterpret_cast<double*>(base + cv.take(32));
    FaceRec* r1 = reinterpret_cast<FaceRec*>(base + cv.take(size_t(n1) * sizeof(FaceRec)));
    FaceRec* r2 = reinterpret_cast<FaceRec*>(base + cv.take(size_t(std::max(n2, 1)) * sizeof(FaceRec)));
    LDDMM_TRY(feat(r1, has_tgt ? r2 : nullptr));
    hipLaunchKernelGGL(k_center_faces, dim3(1), dim3(kCenterFaceThreads), 0, st, r1, n1, r2, n2, c0);
    LDDMM_CHECK_LAUNCH();

    if (!want_grad) {
        typename OpVarifoldE<MS>::template Args<T> A;
        A.f1 = r1; A.f2 = r2; A.n1 = n1; A.n2 = n2; A.w2 = -2.0; A.c1 = 0; A.c0 = c0; A.sd = shp.sd; A.sc = sc;
        Partials<T> part = carve_partials<T>(base, cv, plan, 1, n1);
        if constexpr (std::is_same<T, float>::value && !MS) {
            if (vf_engine == LDDMM_VARIFOLD_MFMA_COS_R2)
                LDDMM_TRY((run_vf_energy_mfma<2, true>(A, n1, n1 + n2, plan, part, st)));
            else if (vf_engine == LDDMM_VARIFOLD_MFMA_COS)
                LDDMM_TRY((run_vf_energy_mfma<2, false>(A, n1, n1 + n2, plan, part, st)));
        }
        if (!vf_engine) {
            if (R1 == 2) LDDMM_TRY((run_pair<T, OpVarifoldE<MS>, 2>(A, n1, n1 + n2, plan, part, st)));
            else LDDMM_TRY((run_pair<T, OpVarifoldE<MS>, 1>(A, n1, n1 + n2, plan, part, st)));
        }
        double* rows = reinterpret_cast<double*>(base + cv.take(size_t(n1) * sizeof(double)));
        hipLaunchKernelGGL((k_vf_energy_rows<T>), dim3(unsigned((n1 + kEpiRows - 1) / kEpiRows)), dim3(256), 0, st,
                           part, plan.nchunk, r1, n1, fK, rows);
        LDDMM_CHECK_LAUNCH();
        hipLaunchKernelGGL(k_sum_rows, dim3(1), dim3(1024), 0, st, rows, n1, energy);
        LDDMM_CHECK_LAUNCH();
        return LDDMM_OK;
    }

    typename OpVarifoldG<MS>::template Args<T> A;
    const int c1 = (n1 + plan.chunk - 1) / plan.chunk;
    A.f1 = r1; A.f2 = r2; A.n1 = n1; A.n2 = n2; A.w2 = -1.0; A.c1 = c1; A.c0 = c0; A.sd = shp.sd; A.sc = sc;
    Partials<T> part = carve_partials<T>(base, cv, plan, 8, n1);
    double* z1 = reinterpret_cast<double*>(base + cv.take(size_t(n1) * 3 * sizeof(double)));
    double* dz1 = reinterpret_cast<double*>(base + cv.take(size_t(n1) * 3 * sizeof(double)));
    double* erow = reinterpret_cast<double*>(base + cv.take(size_t(n1) * sizeof(double)));
    if (R1 == 2) LDDMM_TRY((run_pair<T, OpVarifoldG<MS>, 2>(A, n1, n1 + n2, plan, part, st)));
    else LDDMM_TRY((run_pair<T, OpVarifoldG<MS>, 1>(A, n1, n1 + n2, plan, part, st)));
    hipLaunchKernelGGL((k_vf_grad_epi<T>), dim3(unsigned((n1 + kEpiRows - 1) / kEpiRows)), dim3(256), 0, st, part, c1,
                       plan.nchunk, r1, n1, fK, fD, shp.zf, shp.sd, shp.dzf, z1, dz1, erow);
    LDDMM_CHECK_LAUNCH();
    if (energy) {
        hipLaunchKernelGGL(k_sum_rows, dim3(1), dim3(1024), 0, st, erow, n1, energy);
        LDDMM_CHECK_LAUNCH();
    }
    return gather(dz1, z1);
}

template <typename T, bool MS>
int varifold_t(const lddmm_kernel* kd, const lddmm_surface* s1, const lddmm_surface* s2, double* energy,
               double* grad, hipStream_t st) {
    auto feat = [&](FaceRec* r1, FaceRec* r2) -> int {
        LDDMM_TRY(launch_features(s1, r1, nullptr, nullptr, nullptr, st));
        if (r2) LDDMM_TRY(launch_features(s2, r2, nullptr, nullptr, nullptr, st));
        return LDDMM_OK;
    };
    auto gather = [&](const double* dz1, const double* z1) -> int {
        if (s1->nv > 0) {
            hipLaunchKernelGGL(k_face_gather, dim3(unsigned((s1->nv + 255) / 256)), dim3(256), 0, st,
                               s1->vertices, s1->faces, s1->adj_ptr, s1->adj_idx, int(s1->nv), dz1, z1, grad,
                               nullptr, nullptr);
            LDDMM_CHECK_LAUNCH();
        }
        return LDDMM_OK;
    };
    return varifold_core<T, MS>(kd, int(s1->nf), s2 ? int(s2->nf) : 0, s2 != nullptr,
                                VfShape{1.0, 1.0, 2.0 / 3.0}, feat, gather, energy, grad != nullptr, st);
}

// ---------------------------------------------------------------- curves

// Segment features (base/curves.py:117-125, base/curveDistances.py:186-336):
// c = (x1+x2)/2, l = x2-x1, w = mean of the two vertex weights.  Energy form
// (varifoldNorm0/Def) folds w into l before the norm: a = sqrt(|w l|^2 + 1e-10);
// gradient form: a = sqrt(|l|^2 + 1e-10), record weight a*w.  2-D curves are
// padded with a zero third coordinate.
__global__ void k_seg_features(const double* __restrict__ v, int dim, const int32_t* __restrict__ sg, int ns,
                               const double* __restrict__ vw, int energy_form, FaceRec* __restrict__ rec,
                               double* __restrict__ centers, double* __restrict__ linel,
                               double* __restrict__ sweights) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ns) return;
    int i0 = sg[2 * i], i1 = sg[2 * i + 1];
    double x1[3] = {0.0, 0.0, 0.0}, x2[3] = {0.0, 0.0, 0.0};
    for (int d = 0; d < dim; ++d) {
        x1[d] = v[size_t(dim) * i0 + d];
        x2[d] = v[size_t(dim) * i1 + d];
    }
    double w = vw ? (vw[i0] + vw[i1]) / 2.0 : 1.0;
    double l[3] = {x2[0] - x1[0], x2[1] - x1[1], x2[2] - x1[2]};
    double c[3] = {(x1[0] + x2[0]) / 2.0, (x1[1] + x2[1]) / 2.0, (x1[2] + x2[2]) / 2.0};
    if (rec) {
        double ll = l[0] * l[0] + l[1] * l[1] + l[2] * l[2];
        if (energy_form) {
            double a = sqrt(w * w * ll + 1e-10);
            rec[i] = FaceRec{c[0], c[1], c[2], a, w * l[0] / a, w * l[1] / a, w * l[2] / a, w};
        } else {
            double a = sqrt(ll + 1e-10);
            rec[i] = FaceRec{c[0], c[1], c[2], a * w, l[0] / a, l[1] / a, l[2] / a, w};
        }
    }
    for (int d = 0; d < dim; ++d) {
        if (centers) centers[size_t(dim) * i + d] = c[d];
        if (linel) linel[size_t(dim) * i + d] = l[d];
    }
    if (sweights) sweights[i] = w;
}

// px[v] = sum over incident (segment, end): dz1_s - wz1_s (end 0) or dz1_s + wz1_s (end 1)
// (base/curveDistances.py:97-103, :324-326).  dz1 / wz1 rows have stride ld.
__global__ void k_seg_gather(const int32_t* __restrict__ ptr, const int32_t* __restrict__ idx, int nv, int dim,
                             const double* __restrict__ dz1, const double* __restrict__ wz1, int ld,
                             double* __restrict__ px) {
    int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= nv) return;
    double acc[3] = {0.0, 0.0, 0.0};
    for (int e = ptr[v]; e < ptr[v + 1]; ++e) {
        int code = idx[e];
        int si = code >> 1;
        double sg = (code & 1) ? 1.0 : -1.0;
        for (int k = 0; k < dim; ++k) acc[k] += dz1[size_t(ld) * si + k] + sg * wz1[size_t(ld) * si + k];
    }
    for (int k = 0; k < dim; ++k) px[size_t(dim) * v + k] = acc[k];
}

int check_curve(const lddmm_curve* c, bool need_adj, const char* what) {
    LDDMM_REQUIRE(c != nullptr, "%s: null curve", what);
    LDDMM_REQUIRE(c->nv >= 0 && c->ns >= 0, "%s: negative sizes", what);
    LDDMM_REQUIRE(c->dim == 2 || c->dim == 3, "%s: dim must be 2 or 3 (got %d)", what, int(c->dim));
    LDDMM_REQUIRE(c->ns < (int64_t(1) << 30) && c->nv < (int64_t(1) << 31), "%s: sizes beyond int32", what);
    LDDMM_REQUIRE(c->ns == 0 || (c->vertices && c->segments), "%s: null vertices / segments", what);
    if (need_adj) LDDMM_REQUIRE(c->nv == 0 || (c->adj_ptr && c->adj_idx), "%s: adjacency required", what);
    return LDDMM_OK;
}

int launch_seg_features(const lddmm_curve* c, int energy_form, FaceRec* rec, double* ctr, double* ll, double* sw,
                        hipStream_t st) {
    if (c->ns == 0) return LDDMM_OK;
    hipLaunchKernelGGL(k_seg_features, dim3(unsigned((c->ns + 255) / 256)), dim3(256), 0, st, c->vertices,
                       int(c->dim), c->segments, int(c->ns), c->vweights, energy_form, rec, ctr, ll, sw);
    LDDMM_CHECK_LAUNCH();
    return LDDMM_OK;
}

template <typename T, bool MS>
int curve_varifold_t(const lddmm_kernel* kd, const lddmm_curve* c1, const lddmm_curve* c2, double d,
                     double* energy, double* grad, hipStream_t st) {
    const int eform = grad ? 0 : 1;
    auto feat = [&](FaceRec* r1, FaceRec* r2) -> int {
        LDDMM_TRY(launch_seg_features(c1, eform, r1, nullptr, nullptr, nullptr, st));
        if (r2) LDDMM_TRY(launch_seg_features(c2, eform, r2, nullptr, nullptr, nullptr, st));
        return LDDMM_OK;
    };
    auto gather = [&](const double* dz1, const double* z1) -> int {
        if (c1->nv > 0) {
            hipLaunchKernelGGL(k_seg_gather, dim3(unsigned((c1->nv + 255) / 256)), dim3(256), 0, st, c1->adj_ptr,
                               c1->adj_idx, int(c1->nv), int(c1->dim), dz1, z1, 3, grad);
            LDDMM_CHECK_LAUNCH();
        }
        return LDDMM_OK;
    };
    return varifold_core<T, MS>(kd, int(c1->ns), c2 ? int(c2->ns) : 0, c2 != nullptr,
                                VfShape{std::sqrt(d), 2.0, 1.0}, feat, gather, energy, grad != nullptr, st);
}

}  // namespace
}  // namespace lddmm

using namespace lddmm;

extern "C" int lddmm_set_varifold_engine(int32_t engine) {
    if (engine != LDDMM_VARIFOLD_VALU && engine != LDDMM_VARIFOLD_MFMA_COS && engine != LDDMM_VARIFOLD_MFMA_COS_R2) {
        lddmm::set_error("unknown varifold engine %d", engine);
        return LDDMM_EINVAL;
    }
    lddmm::g_vf_engine.store(engine);
    return LDDMM_OK;
}

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
                       int(s->nv), dz1, z1, px, nullptr, nullptr);
    LDDMM_CHECK_LAUNCH();
    return LDDMM_OK;
}

extern "C" int lddmm_surface_varifold(const lddmm_kernel* kd, const lddmm_surface* s1,
                                      const lddmm_surface* tgt, double* energy, double* grad, void* stream) {
    LDDMM_TRY(validate_kernel(kd, kValDataTerm));
    LDDMM_TRY(check_surface(s1, grad != nullptr, "surface_varifold"));
    if (tgt) LDDMM_TRY(check_surface(tgt, false, "surface_varifold (target)"));
    LDDMM_REQUIRE(energy || grad, "surface_varifold: nothing requested");
    hipStream_t st = static_cast<hipStream_t>(stream);
    if (s1->nf == 0) {
        if (energy) LDDMM_TRY(launch_fill(energy, 1, 0.0, st));
        if (grad) LDDMM_TRY(launch_fill(grad, size_t(s1->nv) * 3, 0.0, st));
        return LDDMM_OK;
    }
    const bool ms = general_path(kd);
    if (kd->prec == LDDMM_PREC_F32C)
        return ms ? varifold_t<float, true>(kd, s1, tgt, energy, grad, st)
                  : varifold_t<float, false>(kd, s1, tgt, energy, grad, st);
    return ms ? varifold_t<double, true>(kd, s1, tgt, energy, grad, st)
              : varifold_t<double, false>(kd, s1, tgt, energy, grad, st);
}

extern "C" int lddmm_curve_features(const lddmm_curve* c, double* centers, double* linel, double* seg_weights,
                                    void* stream) {
    LDDMM_TRY(check_curve(c, false, "curve_features"));
    return launch_seg_features(c, 0, nullptr, centers, linel, seg_weights, static_cast<hipStream_t>(stream));
}

extern "C" int lddmm_curve_seg_gather(const lddmm_curve* c, const double* dz1, const double* wz1, double* px,
                                      void* stream) {
    LDDMM_TRY(check_curve(c, true, "curve_seg_gather"));
    if (c->nv == 0) return LDDMM_OK;
    LDDMM_REQUIRE(px && (c->ns == 0 || (dz1 && wz1)), "curve_seg_gather: null arrays");
    hipLaunchKernelGGL(k_seg_gather, dim3(unsigned((c->nv + 255) / 256)), dim3(256), 0,
                       static_cast<hipStream_t>(stream), c->adj_ptr, c->adj_idx, int(c->nv), int(c->dim), dz1, wz1,
                       int(c->dim), px);
    LDDMM_CHECK_LAUNCH();
    return LDDMM_OK;
}

extern "C" int lddmm_curve_varifold(const lddmm_kernel* kd, const lddmm_curve* c1, const lddmm_curve* tgt,
                                    double weight, double* energy, double* grad, void* stream) {
    LDDMM_TRY(validate_kernel(kd, kValDataTerm));
    LDDMM_TRY(check_curve(c1, grad != nullptr, "curve_varifold"));
    if (tgt) {
        LDDMM_TRY(check_curve(tgt, false, "curve_varifold (target)"));
        LDDMM_REQUIRE(tgt->dim == c1->dim, "curve_varifold: dimension mismatch (%d vs %d)", int(c1->dim),
                      int(tgt->dim));
    }
    LDDMM_REQUIRE(energy || grad, "curve_varifold: nothing requested");
    LDDMM_REQUIRE(weight >= 0.0 && std::isfinite(weight), "curve_varifold: weight must be finite and >= 0");
    hipStream_t st = static_cast<hipStream_t>(stream);
    if (c1->ns == 0) {
        if (energy) LDDMM_TRY(launch_fill(energy, 1, 0.0, st));
        if (grad) LDDMM_TRY(launch_fill(grad, size_t(c1->nv) * size_t(c1->dim), 0.0, st));
        return LDDMM_OK;
    }
    const bool ms = general_path(kd);
    if (kd->prec == LDDMM_PREC_F32C)
        return ms ? curve_varifold_t<float, true>(kd, c1, tgt, weight, energy, grad, st)
                  : curve_varifold_t<float, false>(kd, c1, tgt, weight, energy, grad, st);
    return ms ? curve_varifold_t<double, true>(kd, c1, tgt, weight, energy, grad, st)
              : curve_varifold_t<double, false>(kd, c1, tgt, weight, energy, grad, st);
}
