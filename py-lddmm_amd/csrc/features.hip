// Input features of the learned warm start that need no vertex correspondence:
// the signed distance from each template vertex to the target surface along
// the vertex normal (ray-triangle intersection, Moller-Trumbore), float64.
//
// One thread per ray and face chunk (grid: ray blocks x face chunks, so that a
// few thousand rays still spread over the CUs); the chunk's triangles are
// staged in LDS as (v0, e1 = v1 - v0, e2 = v2 - v0) records, 256 per tile,
// and every thread of the workgroup tests its ray against the whole tile
// (wave-uniform LDS reads).  Within a chunk the hit with the smallest |t| is
// kept, the lowest face index on ties (triangles visited in index order, only
// a strictly smaller |t| replaces the current hit); the chunks' hits are then
// combined in chunk order with the same strict rule, which gives the same
// answer as one pass over all faces.
#include "common.hpp"

namespace lddmm {

namespace {

constexpr int RC_BLOCK = 256;
constexpr int RC_TILE = 256;
// barycentric slack: a ray through a shared edge or vertex is counted by
// every adjacent triangle (they agree on t to rounding)
constexpr double RC_BARY_TOL = 1e-9;

// faces [blockIdx.y * fchunk, min(nf, (blockIdx.y + 1) * fchunk)); results of
// chunk c go to t_out[c * n + i], hit[c * n + i]
__global__ void __launch_bounds__(RC_BLOCK) k_normal_raycast(const double* __restrict__ orig,
                                                             const double* __restrict__ dir, int64_t n,
                                                             const double* __restrict__ tv, int64_t nv,
                                                             const int32_t* __restrict__ tf, int64_t nf,
                                                             int64_t fchunk, double* __restrict__ t_out,
                                                             int32_t* __restrict__ hit) {
    __shared__ double tri[RC_TILE][9];
    const int64_t i = int64_t(blockIdx.x) * RC_BLOCK + threadIdx.x;
    const bool live = i < n;
    const int64_t fb = int64_t(blockIdx.y) * fchunk;
    const int64_t fe = fb + fchunk < nf ? fb + fchunk : nf;
    double ox = 0, oy = 0, oz = 0, dx = 0, dy = 0, dz = 0;
    if (live) {
        ox = orig[3 * i]; oy = orig[3 * i + 1]; oz = orig[3 * i + 2];
        dx = dir[3 * i]; dy = dir[3 * i + 1]; dz = dir[3 * i + 2];
    }
    double best = 0.0, best_abs = INFINITY;
    int found = 0;
    for (int64_t f0 = fb; f0 < fe; f0 += RC_TILE) {
        const int cnt = int(fe - f0 < RC_TILE ? fe - f0 : RC_TILE);
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
        const size_t o = size_t(blockIdx.y) * size_t(n) + size_t(i);
        t_out[o] = found ? best : 0.0;
        hit[o] = found;
    }
}

// Per ray: the chunks' hits in chunk order, a strictly smaller |t| replacing.
__global__ void __launch_bounds__(RC_BLOCK) k_raycast_pick(int64_t n, int nchunk, const double* __restrict__ pt,
                                                           const int32_t* __restrict__ ph,
                                                           double* __restrict__ t_out, int32_t* __restrict__ hit) {
    const int64_t i = int64_t(blockIdx.x) * RC_BLOCK + threadIdx.x;
    if (i >= n) return;
    double best = 0.0, best_abs = INFINITY;
    int found = 0;
    for (int c = 0; c < nchunk; ++c) {
        const size_t o = size_t(c) * size_t(n) + size_t(i);
        if (ph[o] && fabs(pt[o]) < best_abs) {
            best_abs = fabs(pt[o]);
            best = pt[o];
            found = 1;
        }
    }
    t_out[i] = found ? best : 0.0;
    hit[i] = found;
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
    hipStream_t st = static_cast<hipStream_t>(stream);
    const int64_t bx = (n + RC_BLOCK - 1) / RC_BLOCK;
    // face chunks of whole tiles, enough for about four workgroups per CU
    int64_t nchunk = 1;
    if (nf > RC_TILE) {
        DeviceInfo di;
        LDDMM_TRY(device_info(&di));
        const int64_t want = (4 * int64_t(di.num_cu) + bx - 1) / bx;
        const int64_t tiles = (nf + RC_TILE - 1) / RC_TILE;
        nchunk = want < tiles ? want : tiles;
        if (nchunk > 65535) nchunk = 65535;
        if (nchunk < 1) nchunk = 1;
    }
    const int64_t tiles = (nf + RC_TILE - 1) / RC_TILE;
    const int64_t fchunk = nchunk > 1 ? (tiles + nchunk - 1) / nchunk * RC_TILE : (nf > 0 ? nf : 1);
    nchunk = nf > 0 ? (nf + fchunk - 1) / fchunk : 1;
    if (nchunk == 1) {
        hipLaunchKernelGGL(k_normal_raycast, dim3(unsigned(bx), 1), dim3(RC_BLOCK), 0, st, orig, dir, n, tv, nv, tf,
                           nf, fchunk, t_out, hit);
        LDDMM_CHECK_LAUNCH();
        return LDDMM_OK;
    }
    void* ws;
    const size_t cells = size_t(nchunk) * size_t(n);
    LDDMM_TRY(workspace(cells * (sizeof(double) + sizeof(int32_t)) + 512, st, &ws));
    double* pt = static_cast<double*>(ws);
    int32_t* ph = reinterpret_cast<int32_t*>(static_cast<char*>(ws) + (cells * sizeof(double) + 255) / 256 * 256);
    hipLaunchKernelGGL(k_normal_raycast, dim3(unsigned(bx), unsigned(nchunk)), dim3(RC_BLOCK), 0, st, orig, dir, n,
                       tv, nv, tf, nf, fchunk, pt, ph);
    LDDMM_CHECK_LAUNCH();
    hipLaunchKernelGGL(k_raycast_pick, dim3(unsigned(bx)), dim3(RC_BLOCK), 0, st, n, int(nchunk), pt, ph, t_out, hit);
    LDDMM_CHECK_LAUNCH();
    return LDDMM_OK;
}

}  // extern "C"
