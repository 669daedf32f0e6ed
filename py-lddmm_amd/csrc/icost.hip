// Internal (regularity) costs of a velocity field phi on a triangulated
// surface, restating py-lddmm base/surfaceDistances.py:
//   h1           normGrad / diffNormGrad                  (:631-645, :680-720)
//   elastic      elasticNorm_nb / diffElasticNorm_nb      (:779-856)
//   displacement normDisplacement / diffNormDisplacement  (:731-777)
// One thread per face computes every listed cost, its energy and the
// per-corner derivatives in phi and in the vertices (float64); a second
// kernel gathers the corner records per vertex through the CSR
// vertex -> (face, corner) adjacency (deterministic, no atomics), and the
// energy is reduced per workgroup, then in workgroup order.
#include "icost.hpp"

#include <algorithm>

namespace lddmm {
namespace {

struct V3 { double x, y, z; };
__device__ __forceinline__ V3 operator+(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ V3 operator-(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ V3 operator*(double s, V3 a) { return {s * a.x, s * a.y, s * a.z}; }
__device__ __forceinline__ double dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ V3 cross(V3 a, V3 b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
__device__ __forceinline__ V3 ld3(const double* p, int i) { return {p[3 * i], p[3 * i + 1], p[3 * i + 2]}; }
__device__ __forceinline__ void st3(double* p, size_t i, V3 v) {
    p[3 * i] = v.x;
    p[3 * i + 1] = v.y;
    p[3 * i + 2] = v.z;
}

__global__ void __launch_bounds__(256)
k_ic_face(int nf, const int32_t* __restrict__ faces, const double* __restrict__ x, const double* __restrict__ phi,
          IcList L, double escale, double* __restrict__ part, double* __restrict__ rec_phi,
          double* __restrict__ rec_x) {
    __shared__ double red[256];
    const int f = blockIdx.x * 256 + threadIdx.x;
    double e = 0.0;
    if (f < nf) {
        const int i1 = faces[3 * f], i2 = faces[3 * f + 1], i3 = faces[3 * f + 2];
        const V3 v1 = ld3(x, i1), v2 = ld3(x, i2), v3 = ld3(x, i3);
        const V3 p1 = ld3(phi, i1), p2 = ld3(phi, i2), p3 = ld3(phi, i3);
        const V3 S = 0.5 * cross(v2 - v1, v3 - v1);
        const double A = sqrt(dot(S, S));
        const double l1 = dot(v2 - v3, v2 - v3), l2 = dot(v1 - v3, v1 - v3), l3 = dot(v1 - v2, v1 - v2);
        const double q1 = dot(p2 - p1, p3 - p1), q2 = dot(p3 - p2, p1 - p2), q3 = dot(p1 - p3, p2 - p3);
        const double u = l1 * q1 + l2 * q2 + l3 * q3;
        const V3 c1 = cross(v2 - v3, S), c2 = cross(v3 - v1, S), c3 = cross(v1 - v2, S);
        V3 gp[3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
        V3 gx[3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
        for (int k = 0; k < L.n; ++k) {
            const double w = L.w[k];
            if (L.kinds[k] == LDDMM_IC_DISPLACEMENT) {
                const V3 pc = (1.0 / 3.0) * (p1 + p2 + p3);
                const double pp = dot(pc, pc);
                e += w * pp * A;
                const V3 g = (w * 2.0 / 3.0 * A) * pc;
                gp[0] = gp[0] + g;
                gp[1] = gp[1] + g;
                gp[2] = gp[2] + g;
                const double s = 0.5 * w * pp / A;
                gx[0] = gx[0] + s * c1;
                gx[1] = gx[1] + s * c2;
                gx[2] = gx[2] + s * c3;
            } else {
                const bool el = L.kinds[k] == LDDMM_IC_ELASTIC;
                const double corr = el ? dot(v1 - v3, p1 - p2) - dot(v1 - v2, p1 - p3) : 0.0;
                const double ue = u - 0.5 * corr * corr;
                e += w * ue / (4.0 * A);
                const double a = 2.0 * A, a2 = 4.0 * A;
                const double ca = 0.5 * corr / a;
                V3 r1 = (1.0 / a2) * (l1 * (p2 + p3 - 2.0 * p1) + (l2 - l3) * (p2 - p3));
                V3 r2 = (1.0 / a2) * (l2 * (p1 + p3 - 2.0 * p2) + (l1 - l3) * (p1 - p3));
                V3 r3 = (1.0 / a2) * (l3 * (p1 + p2 - 2.0 * p3) + (l2 - l1) * (p2 - p1));
                if (el) {
                    r1 = r1 + ca * (v2 - v3);
                    r2 = r2 + ca * (v3 - v1);
                    r3 = r3 + ca * (v1 - v2);
                }
                gp[0] = gp[0] - w * r1;
                gp[1] = gp[1] - w * r2;
                gp[2] = gp[2] - w * r3;
                const double uu = ue / (a * a * a);
                V3 x1 = (2.0 / a2) * (q2 * (v1 - v3) + q3 * (v1 - v2)) - uu * c1;
                V3 x2 = (2.0 / a2) * (q3 * (v2 - v1) + q1 * (v2 - v3)) - uu * c2;
                V3 x3 = (2.0 / a2) * (q1 * (v3 - v2) + q2 * (v3 - v1)) - uu * c3;
                if (el) {
                    x1 = x1 - ca * (p3 - p2);
                    x2 = x2 - ca * (p1 - p3);
                    x3 = x3 - ca * (p2 - p1);
                }
                gx[0] = gx[0] + w * x1;
                gx[1] = gx[1] + w * x2;
                gx[2] = gx[2] + w * x3;
            }
        }
        if (rec_phi)
            for (int c = 0; c < 3; ++c) st3(rec_phi, size_t(f) * 3 + c, gp[c]);
        if (rec_x)
            for (int c = 0; c < 3; ++c) st3(rec_x, size_t(f) * 3 + c, gx[c]);
    }
    if (!part) return;
    red[threadIdx.x] = escale * e;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

__global__ void __launch_bounds__(1024) k_ic_sum(const double* __restrict__ part, int n, double* __restrict__ out) {
    __shared__ double red[1024];
    double s = 0.0;
    for (int i = threadIdx.x; i < n; i += 1024) s += part[i];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int o = 512; o > 0; o >>= 1) {
        if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = red[0];
}

// out[v] = scale * sum of the corner records of v (adjacency order).
__global__ void __launch_bounds__(256)
k_ic_gather(int nv, const int32_t* __restrict__ adj_ptr, const int32_t* __restrict__ adj_idx,
            const double* __restrict__ rec, double scale, double* __restrict__ out) {
    const int v = blockIdx.x * 256 + threadIdx.x;
    if (v >= nv) return;
    double s0 = 0.0, s1 = 0.0, s2 = 0.0;
    for (int k = adj_ptr[v]; k < adj_ptr[v + 1]; ++k) {
        const size_t e = size_t(adj_idx[k]);
        s0 += rec[3 * e];
        s1 += rec[3 * e + 1];
        s2 += rec[3 * e + 2];
    }
    out[3 * size_t(v)] = scale * s0;
    out[3 * size_t(v) + 1] = scale * s1;
    out[3 * size_t(v) + 2] = scale * s2;
}

}  // namespace

int ic_check(const lddmm_internal_cost* ic, IcList* out) {
    LDDMM_REQUIRE(ic != nullptr, "internal cost: null description");
    LDDMM_REQUIRE(ic->ncost >= 1 && ic->ncost <= LDDMM_MAX_ICOSTS, "internal cost: ncost %d outside [1, %d]",
                  ic->ncost, LDDMM_MAX_ICOSTS);
    IcList L{};
    L.n = ic->ncost;
    for (int k = 0; k < ic->ncost; ++k) {
        LDDMM_REQUIRE(ic->kinds[k] == LDDMM_IC_H1 || ic->kinds[k] == LDDMM_IC_ELASTIC ||
                          ic->kinds[k] == LDDMM_IC_DISPLACEMENT,
                      "internal cost: bad kind %d", ic->kinds[k]);
        LDDMM_REQUIRE(std::isfinite(ic->weights[k]), "internal cost: weight %d not finite", k);
        L.kinds[k] = ic->kinds[k];
        L.w[k] = ic->weights[k];
    }
    *out = L;
    return LDDMM_OK;
}

size_t ic_scratch_doubles(int64_t nf, bool grads) {
    size_t nb = size_t((nf + 255) / 256);
    return ((nb + 31) & ~size_t(31)) + (grads ? size_t(18) * size_t(nf) : 0);
}

int ic_eval(const IcList& L, int nv, int nf, const int32_t* faces, const int32_t* adj_ptr, const int32_t* adj_idx,
            const double* x, const double* phi, double escale, double* energy, double gscale, double* grad_phi,
            double* grad_x, double* scratch, hipStream_t st) {
    if (nf == 0) {
        if (energy) LDDMM_TRY(launch_fill(energy, 1, 0.0, st));
        if (grad_phi) LDDMM_TRY(launch_fill(grad_phi, size_t(nv) * 3, 0.0, st));
        if (grad_x) LDDMM_TRY(launch_fill(grad_x, size_t(nv) * 3, 0.0, st));
        return LDDMM_OK;
    }
    const int nb = (nf + 255) / 256;
    double* part = scratch;
    double* rec_phi = scratch + ((size_t(nb) + 31) & ~size_t(31));
    double* rec_x = rec_phi + size_t(9) * nf;
    hipLaunchKernelGGL(k_ic_face, dim3(unsigned(nb)), dim3(256), 0, st, nf, faces, x, phi, L, escale,
                       energy ? part : nullptr, grad_phi ? rec_phi : nullptr, grad_x ? rec_x : nullptr);
    LDDMM_CHECK_LAUNCH();
    if (energy) {
        hipLaunchKernelGGL(k_ic_sum, dim3(1), dim3(1024), 0, st, part, nb, energy);
        LDDMM_CHECK_LAUNCH();
    }
    const unsigned gv = unsigned((nv + 255) / 256);
    if (grad_phi && nv > 0) {
        hipLaunchKernelGGL(k_ic_gather, dim3(gv), dim3(256), 0, st, nv, adj_ptr, adj_idx, rec_phi, gscale, grad_phi);
        LDDMM_CHECK_LAUNCH();
    }
    if (grad_x && nv > 0) {
        hipLaunchKernelGGL(k_ic_gather, dim3(gv), dim3(256), 0, st, nv, adj_ptr, adj_idx, rec_x, gscale, grad_x);
        LDDMM_CHECK_LAUNCH();
    }
    return LDDMM_OK;
}

}  // namespace lddmm

using namespace lddmm;

extern "C" int lddmm_surface_internal_cost(const lddmm_surface* s, const double* phi, const lddmm_internal_cost* ic,
                                           double* energy, double* grad_phi, double* grad_x, void* stream) {
    IcList L;
    LDDMM_TRY(ic_check(ic, &L));
    LDDMM_REQUIRE(s != nullptr, "surface_internal_cost: null surface");
    LDDMM_REQUIRE(s->nv >= 0 && s->nf >= 0 && s->nf < (int64_t(1) << 30) && s->nv < (int64_t(1) << 31),
                  "surface_internal_cost: bad sizes");
    LDDMM_REQUIRE(energy || grad_phi || grad_x, "surface_internal_cost: nothing requested");
    LDDMM_REQUIRE(s->nf == 0 || (s->vertices && s->faces && phi), "surface_internal_cost: null arrays");
    const bool grads = grad_phi || grad_x;
    LDDMM_REQUIRE(!grads || s->nv == 0 || (s->adj_ptr && s->adj_idx), "surface_internal_cost: adjacency required");
    hipStream_t st = static_cast<hipStream_t>(stream);
    void* ws = nullptr;
    LDDMM_TRY(workspace(ic_scratch_doubles(s->nf, grads) * sizeof(double) + 256, st, &ws));
    return ic_eval(L, int(s->nv), int(s->nf), s->faces, s->adj_ptr, s->adj_idx, s->vertices, phi, 1.0, energy, 1.0,
                   grad_phi, grad_x, static_cast<double*>(ws), st);
}
