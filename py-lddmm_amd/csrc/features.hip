// Input features of the learned warm start that need no vertex correspondence:
// the signed distance from each template vertex to the target surface along
// the vertex normal (ray-triangle intersection, Moller-Trumbore), float64.
//
// One thread per ray; the target's triangles are staged in LDS as
// (v0, e1 = v1 - v0, e2 = v2 - v0) records, 256 per tile, and every thread of
// the workgroup tests its ray against the whole tile (wave-uniform LDS reads).
// Over all triangles the hit with the smallest |t| is kept, the lowest face
// index on ties (triangles are visited in index order and only a strictly
// smaller |t| replaces the current hit), so the result is deterministic.
#include "common.hpp"

namespace lddmm {

namespace {

constexpr int RC_BLOCK = 256;
constexpr int RC_TILE = 256;
// barycentric slack: a ray through a shared edge or vertex is counted by
// every adjacent triangle (they agree on t to rounding)
constexpr double RC_BARY_TOL = 1e-9;

__global__ void __launch_bounds__(RC_BLOCK) k_normal_raycast(const double* __restrict__ orig,
                                                             const double* __restrict__ dir, int64_t n,
                                                             const double* __restrict__ tv, int64_t nv,
                                                             const int32_t* __restrict__ tf, int64_t nf,
                                                             double* __restrict__ t_out,
                                                             int32_t* __restrict__ hit) {
    __shared__ double tri[RC_TILE][9];
    const int64_t i = int64_t(blockIdx.x) * RC_BLOCK + threadIdx.x;
    const bool live = i < n;
    double ox = 0, oy = 0, oz = 0, dx = 0, dy = 0, dz = 0;
    if (live) {
        ox = orig[3 * i]; oy = orig[3 * i + 1]; oz = orig[3 * i + 2];
        dx = dir[3 * i]; dy = dir[3 * i + 1]; dz = dir[3 * i + 2];
    }
    double best = 0.0, best_abs = INFINITY;
    int found = 0;
    for (int64_t f0 = 0; f0 < nf; f0 += RC_TILE) {
        const int cnt = int(nf - f0 < RC_TILE ? nf - f0 : RC_TILE);
        __syncthreads();
        if (threadIdx.x < cnt) {
            const int64_t f = f0 + threadIdx.x;
            const int32_t a = tf[3 * f], b = tf[3 * f + 1], c = tf[3 * f + 2];
            double* r = tri[threadIdx.x];
            if (a < 0 || b < 0 || c < 0 || a >= nv || b >= nv || c >= nv) {
                // an index outside [0, nv): a degenerate triangle (det = 0, never hit)
                for (int k = 0; k < 9; ++k) r[k] = 0.0;
            } else {
                const double ax = tv[3 * a], ay = tv[3 * a + 1], az = tv[3 * a + 2];
                r[0] = ax; r[1] = ay; r[2] = az;
                r[3] = tv[3 * b] - ax; r[4] = tv[3 * b + 1] - ay; r[5] = tv[3 * b + 2] - az;
                r[6] = tv[3 * c] - ax; r[7] = tv[3 * c + 1] - ay; r[8] = tv[3 * c + 2] - az;
            }
        }
        __syncthreads();
        if (!live) continue;
        for (int k = 0; k < cnt; ++k) {
            const double* r = tri[k];
            const double e1x = r[3], e1y = r[4], e1z = r[5];
            const double e2x = r[6], e2y = r[7], e2z = r[8];
            // p = d x e2, det = e1 . p
            const double px = dy * e2z - dz * e2y, py = dz * e2x - dx * e2z, pz = dx * e2y - dy * e2x;
            const double det = e1x * px + e1y * py + e1z * pz;
            if (det == 0.0) continue;                       // ray parallel to the triangle's plane
            const double inv = 1.0 / det;
            const double sx = ox - r[0], sy = oy - r[1], sz = oz - r[2];
            const double u = (sx * px + sy * py + sz * pz) * inv;
            if (u < -RC_BARY_TOL || u > 1.0 + RC_BARY_TOL) continue;
            // q = s x e1
            const double qx = sy * e1z - sz * e1y, qy = sz * e1x - sx * e1z, qz = sx * e1y - sy * e1x;
            const double v = (dx * qx + dy * qy + dz * qz) * inv;
            if (v < -RC_BARY_TOL || u + v > 1.0 + RC_BARY_TOL) continue;
            const double t = (e2x * qx + e2y * qy + e2z * qz) * inv;
            const double at = fabs(t);
            if (at < best_abs) {
                best_abs = at;
                best = t;
                found = 1;
            }
        }
    }
    if (live) {
        t_out[i] = found ? best : 0.0;
        hit[i] = found;
    }
}

}  // namespace
}  // namespace lddmm

extern "C" {

int lddmm_normal_raycast(const double* orig, const double* dir, int64_t n, const double* tv, int64_t nv,
                         const int32_t* tf, int64_t nf, double* t_out, int32_t* hit, void* stream) {
    using namespace lddmm;
    LDDMM_REQUIRE(n >= 0 && nv >= 0 && nf >= 0, "normal_raycast: negative size");
    LDDMM_REQUIRE(n < (int64_t(1) << 31) * RC_BLOCK, "normal_raycast: too many rays");
    LDDMM_REQUIRE(nf == 0 || nv > 0, "normal_raycast: faces without vertices");
    if (n == 0) return LDDMM_OK;
    hipLaunchKernelGGL(k_normal_raycast, dim3(unsigned((n + RC_BLOCK - 1) / RC_BLOCK)), dim3(RC_BLOCK), 0,
                       static_cast<hipStream_t>(stream), orig, dir, n, tv, nv, tf, nf, t_out, hit);
    LDDMM_CHECK_LAUNCH();
    return LDDMM_OK;
}

}  // extern "C"
