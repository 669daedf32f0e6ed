// Point-cloud network kernels (forward, backward, Adam) for the learned warm
// start: dense layers with BF16 weights and FP32 accumulation, per-segment
// max-pooling, per-segment bias (the broadcast global feature), and a fused
// Adam step that keeps FP32 master weights and refreshes the BF16 copies.
// Points of one template-target pair form one segment; seg[p] gives it.
#include "common.hpp"

#include <hip/hip_bf16.h>
#include <algorithm>

namespace lddmm {

namespace {

__device__ __forceinline__ float bf2f(uint16_t v) { return __uint_as_float(uint32_t(v) << 16); }

// round-to-nearest-even float -> bf16 (finite inputs)
__device__ __forceinline__ uint16_t f2bf(float f) {
    uint32_t u = __float_as_uint(f);
    uint32_t r = u + 0x7FFFu + ((u >> 16) & 1u);
    return uint16_t(r >> 16);
}

constexpr int TM = 64;   // rows (points) per tile
constexpr int TN = 64;   // output channels per tile
constexpr int TK = 16;   // reduction slice

// Y[P, N] = act(X[P, K] W + b + segbias[seg[p]]);  W is [K, N] bf16 (or [N, K] when transW).
__global__ void __launch_bounds__(256) k_linear(const float* __restrict__ X, int P, int K,
                                                const uint16_t* __restrict__ W, int transW,
                                                const float* __restrict__ b, const float* __restrict__ segbias,
                                                const int32_t* __restrict__ seg, int N, int relu,
                                                float* __restrict__ Y) {
    __shared__ float xs[TK][TM + 1];
    __shared__ float ws[TK][TN + 1];
    const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;   // 16 x 16 threads, 4 x 4 outputs each
    const int row0 = blockIdx.y * TM, col0 = blockIdx.x * TN;
    float acc[4][4] = {};
    for (int k0 = 0; k0 < K; k0 += TK) {
        for (int e = threadIdx.x; e < TK * TM; e += 256) {
            int kk = e % TK, rr = e / TK;
            int r = row0 + rr, k = k0 + kk;
            xs[kk][rr] = (r < P && k < K) ? X[size_t(r) * K + k] : 0.f;
        }
        for (int e = threadIdx.x; e < TK * TN; e += 256) {
            int cc = e % TN, kk = e / TN;
            int c = col0 + cc, k = k0 + kk;
            float w = 0.f;
            if (c < N && k < K) w = bf2f(transW ? W[size_t(c) * K + k] : W[size_t(k) * N + c]);
            ws[kk][cc] = w;
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < TK; ++kk) {
            float xv[4], wv[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) xv[i] = xs[kk][ty * 4 + i];
#pragma unroll
            for (int j = 0; j < 4; ++j) wv[j] = ws[kk][tx * 4 + j];
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int j = 0; j < 4; ++j) acc[i][j] = fmaf(xv[i], wv[j], acc[i][j]);
        }
        __syncthreads();
    }
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        int r = row0 + ty * 4 + i;
        if (r >= P) continue;
        int s = seg ? seg[r] : 0;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            int c = col0 + tx * 4 + j;
            if (c >= N) continue;
            float v = acc[i][j];
            if (b) v += b[c];
            if (segbias) v += segbias[size_t(s) * N + c];
            if (relu) v = fmaxf(v, 0.f);
            Y[size_t(r) * N + c] = v;
        }
    }
}

// dZ = dY * (Y > 0)   (ReLU backward on post-activation values; in place allowed)
__global__ void k_relu_bwd(const float* __restrict__ Y, const float* dY, float* dZ, size_t count) {
    size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < count) dZ[i] = Y[i] > 0.f ? dY[i] : 0.f;
}

// Partial dW[K, N] = X^T dZ over point chunks: part[chunk][K][N]
__global__ void __launch_bounds__(256) k_wgrad_partial(const float* __restrict__ X, const float* __restrict__ dZ,
                                                       int P, int K, int N, int rows_per_chunk,
                                                       float* __restrict__ part) {
    __shared__ float xs[TK][TM + 1];   // [point][k] slice, reused as points x K-tile
    __shared__ float zs[TK][TN + 1];
    const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
    const int k0 = blockIdx.y * TM, n0 = blockIdx.x * TN;
    const int p0 = blockIdx.z * rows_per_chunk, p1 = min(P, p0 + rows_per_chunk);
    float acc[4][4] = {};
    for (int pb = p0; pb < p1; pb += TK) {
        for (int e = threadIdx.x; e < TK * TM; e += 256) {
            int kk = e % TM, pp = e / TM;
            int p = pb + pp, k = k0 + kk;
            xs[pp][kk] = (p < p1 && k < K) ? X[size_t(p) * K + k] : 0.f;
        }
        for (int e = threadIdx.x; e < TK * TN; e += 256) {
            int nn = e % TN, pp = e / TN;
            int p = pb + pp, n = n0 + nn;
            zs[pp][nn] = (p < p1 && n < N) ? dZ[size_t(p) * N + n] : 0.f;
        }
        __syncthreads();
#pragma unroll
        for (int pp = 0; pp < TK; ++pp) {
            float xv[4], zv[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) xv[i] = xs[pp][ty * 4 + i];
#pragma unroll
            for (int j = 0; j < 4; ++j) zv[j] = zs[pp][tx * 4 + j];
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int j = 0; j < 4; ++j) acc[i][j] = fmaf(xv[i], zv[j], acc[i][j]);
        }
        __syncthreads();
    }
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        int k = k0 + ty * 4 + i;
        if (k >= K) continue;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            int n = n0 + tx * 4 + j;
            if (n < N) part[(size_t(blockIdx.z) * K + k) * N + n] = acc[i][j];
        }
    }
}

// out[e] = sum_c part[c][e]  (fixed order)
__global__ void k_sum_parts(const float* __restrict__ part, int nparts, size_t count, float* __restrict__ out) {
    size_t e = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (e >= count) return;
    float s = 0.f;
    for (int c = 0; c < nparts; ++c) s += part[size_t(c) * count + e];
    out[e] = s;
}

// Per-segment column reductions over contiguous point ranges [off[s], off[s+1]),
// in two deterministic passes.  Pass 1: workgroup (c, s) covers chunk c of
// the SEG_CHUNKS equal row chunks of segment s; its 256 lanes form
// 256/NC row lanes x NC columns (NC = min(N, 256), column blocks looped), each
// lane strides the chunk's rows, and the row lanes are combined in LDS in
// lane order.  Pass 2 folds the chunks of a segment in chunk order.
constexpr int SEG_CHUNKS = 32;
constexpr int SEG_THREADS = 256;

__device__ __forceinline__ void seg_chunk_rows(const int32_t* off, int s, int c, int nch, int* r0, int* r1) {
    const int b0 = off[s], b1 = off[s + 1];
    const int per = (b1 - b0 + nch - 1) / nch;
    *r0 = b0 + c * per;
    *r1 = min(b1, *r0 + per);
}

// colsum pass 1: part[(s * nch + c) * N + n]
__global__ void __launch_bounds__(SEG_THREADS) k_seg_colsum_part(const float* __restrict__ dZ,
                                                                 const int32_t* __restrict__ off, int N, int nch,
                                                                 float* __restrict__ part) {
    __shared__ float red[SEG_THREADS];
    const int s = blockIdx.y, c = blockIdx.x;
    const int NC = N < SEG_THREADS ? N : SEG_THREADS;
    const int R = SEG_THREADS / NC;
    const int lane = threadIdx.x / NC, col = threadIdx.x % NC;
    int r0, r1;
    seg_chunk_rows(off, s, c, nch, &r0, &r1);
    for (int n0 = 0; n0 < N; n0 += NC) {
        const int n = n0 + col;
        float acc = 0.f;
        if (lane < R && n < N)
            for (int p = r0 + lane; p < r1; p += R) acc += dZ[size_t(p) * N + n];
        red[threadIdx.x] = acc;
        __syncthreads();
        if (lane == 0 && n < N) {
            float t = 0.f;
            for (int l = 0; l < R; ++l) t += red[l * NC + col];
            part[(size_t(s) * nch + c) * N + n] = t;
        }
        __syncthreads();
    }
}

// colsum pass 2: out[s][n] = sum_c part[s][c][n]
__global__ void k_seg_colsum_fold(const float* __restrict__ part, int nseg, int nch, int N, float* __restrict__ out) {
    const size_t e = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (e >= size_t(nseg) * N) return;
    const size_t s = e / N, n = e % N;
    float t = 0.f;
    for (int c = 0; c < nch; ++c) t += part[(s * nch + c) * N + n];
    out[e] = t;
}

// (v, p) beats (best, bi): larger value, or equal value at a smaller point index
__device__ __forceinline__ bool max_first(float v, int p, float best, int bi) {
    return v > best || (v == best && p < bi);
}

// segmax pass 1: (pv, pi)[(s * nch + c) * N + n] = max and first argmax over the chunk
__global__ void __launch_bounds__(SEG_THREADS) k_segmax_part(const float* __restrict__ H,
                                                             const int32_t* __restrict__ off, int N, int nch,
                                                             float* __restrict__ pv, int32_t* __restrict__ pi) {
    __shared__ float rv[SEG_THREADS];
    __shared__ int ri[SEG_THREADS];
    const int s = blockIdx.y, c = blockIdx.x;
    const int NC = N < SEG_THREADS ? N : SEG_THREADS;
    const int R = SEG_THREADS / NC;
    const int lane = threadIdx.x / NC, col = threadIdx.x % NC;
    int r0, r1;
    seg_chunk_rows(off, s, c, nch, &r0, &r1);
    for (int n0 = 0; n0 < N; n0 += NC) {
        const int n = n0 + col;
        float best = -INFINITY;
        int bi = INT32_MAX;
        if (lane < R && n < N)
            for (int p = r0 + lane; p < r1; p += R) {
                const float v = H[size_t(p) * N + n];
                if (max_first(v, p, best, bi)) { best = v; bi = p; }
            }
        rv[threadIdx.x] = best;
        ri[threadIdx.x] = bi;
        __syncthreads();
        if (lane == 0 && n < N) {
            float b = -INFINITY;
            int i = INT32_MAX;
            for (int l = 0; l < R; ++l)
                if (max_first(rv[l * NC + col], ri[l * NC + col], b, i)) { b = rv[l * NC + col]; i = ri[l * NC + col]; }
            pv[(size_t(s) * nch + c) * N + n] = b;
            pi[(size_t(s) * nch + c) * N + n] = i;
        }
        __syncthreads();
    }
}

// segmax pass 2: g[s][n], arg[s][n] over the chunks (0 and off[s] for an empty segment)
__global__ void k_segmax_fold(const float* __restrict__ pv, const int32_t* __restrict__ pi,
                              const int32_t* __restrict__ off, int nseg, int nch, int N, float* __restrict__ g,
                              int32_t* __restrict__ arg) {
    const size_t e = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (e >= size_t(nseg) * N) return;
    const size_t s = e / N, n = e % N;
    float b = -INFINITY;
    int i = INT32_MAX;
    for (int c = 0; c < nch; ++c) {
        const float v = pv[(s * nch + c) * N + n];
        const int p = pi[(s * nch + c) * N + n];
        if (max_first(v, p, b, i)) { b = v; i = p; }
    }
    // an empty segment, or one whose values never compare (all -inf / NaN): the first row
    const bool empty = off[s + 1] <= off[s];
    g[e] = empty ? 0.f : b;
    arg[e] = (empty || i == INT32_MAX) ? off[s] : i;
}

// dH[arg[s][n]][n] += dg[s][n]  (one writer per (s, n): no races)
__global__ void k_segmax_bwd(const float* __restrict__ dg, const int32_t* __restrict__ off,
                             const int32_t* __restrict__ arg, int N, float* __restrict__ dH) {
    int s = blockIdx.x;
    if (off[s + 1] <= off[s]) return;
    for (int n = threadIdx.x; n < N; n += blockDim.x) dH[size_t(arg[size_t(s) * N + n]) * N + n] += dg[size_t(s) * N + n];
}

__global__ void k_adam(float* __restrict__ w, uint16_t* __restrict__ wbf, const float* __restrict__ g,
                       float* __restrict__ m, float* __restrict__ v, size_t count, float lr, float b1, float b2,
                       float eps, float c1, float c2, float wd) {
    size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= count) return;
    float gi = g[i] + wd * w[i];
    float mi = b1 * m[i] + (1.f - b1) * gi;
    float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    float wi = w[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
    w[i] = wi;
    if (wbf) wbf[i] = f2bf(wi);
}

__global__ void k_to_bf16(const float* __restrict__ w, uint16_t* __restrict__ wbf, size_t count) {
    size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < count) wbf[i] = f2bf(w[i]);
}

inline unsigned nblk(size_t count) { return unsigned((count + 255) / 256); }

}  // namespace
}  // namespace lddmm

using namespace lddmm;

extern "C" {

int lddmm_nn_linear(const float* X, int64_t P, int32_t K, const uint16_t* W, int32_t transW, const float* b,
                    const float* segbias, const int32_t* seg, int32_t N, int32_t relu, float* Y, void* stream) {
    LDDMM_REQUIRE(P >= 0 && K > 0 && N > 0, "nn_linear: bad sizes");
    LDDMM_REQUIRE(P < (int64_t(1) << 31), "nn_linear: P beyond int32");
    LDDMM_REQUIRE(!segbias || seg, "nn_linear: segbias needs seg");
    if (P == 0) return LDDMM_OK;
    LDDMM_REQUIRE(X && W && Y, "nn_linear: null pointer");
    dim3 grid(unsigned((N + TN - 1) / TN), unsigned((P + TM - 1) / TM));
    hipLaunchKernelGGL(k_linear, grid, dim3(256), 0, static_cast<hipStream_t>(stream), X, int(P), K, W, transW, b,
                       segbias, seg, N, relu, Y);
    LDDMM_CHECK_LAUNCH();
    return LDDMM_OK;
}

int lddmm_nn_relu_bwd(const float* Y, const float* dY, float* dZ, int64_t count, void* stream) {
    if (count <= 0) return LDDMM_OK;
    hipLaunchKernelGGL(k_relu_bwd, dim3(nblk(size_t(count))), dim3(256), 0, static_cast<hipStream_t>(stream), Y, dY,
                       dZ, size_t(count));
    LDDMM_CHECK_LAUNCH();
    return LDDMM_OK;
}

// dW[K, N] = X^T dZ (float32), reduced over point chunks in a fixed order.
int lddmm_nn_wgrad(const float* X, const float* dZ, int64_t P, int32_t K, int32_t N, float* dW, void* stream) {
    LDDMM_REQUIRE(P >= 0 && K > 0 && N > 0 && P < (int64_t(1) << 31), "nn_wgrad: bad sizes");
    hipStream_t st = static_cast<hipStream_t>(stream);
    const int chunks = int(std::min<int64_t>(64, std::max<int64_t>(1, (P + 1023) / 1024)));
    const int rows = int((P + chunks - 1) / chunks);
    void* ws;
    LDDMM_TRY(workspace(size_t(chunks) * K * N * sizeof(float) + 256, st, &ws));
    float* part = static_cast<float*>(ws);
    dim3 grid(unsigned((N + TN - 1) / TN), unsigned((K + TM - 1) / TM), unsigned(chunks));
    hipLaunchKernelGGL(k_wgrad_partial, grid, dim3(256), 0, st, X, dZ, int(P), K, N, rows, part);
    LDDMM_CHECK_LAUNCH();
    hipLaunchKernelGGL(k_sum_parts, dim3(nblk(size_t(K) * N)), dim3(256), 0, st, part, chunks, size_t(K) * N, dW);
    LDDMM_CHECK_LAUNCH();
    return LDDMM_OK;
}

int lddmm_nn_seg_colsum(const float* dZ, const int32_t* off, int32_t nseg, int32_t N, float* out, void* stream) {
    LDDMM_REQUIRE(nseg >= 0 && N > 0 && nseg < 65536, "nn_seg_colsum: bad sizes");
    if (nseg == 0) return LDDMM_OK;
    hipStream_t st = static_cast<hipStream_t>(stream);
    const int nch = SEG_CHUNKS;
    void* ws;
    LDDMM_TRY(workspace(size_t(nseg) * nch * N * sizeof(float) + 256, st, &ws));
    float* part = static_cast<float*>(ws);
    hipLaunchKernelGGL(k_seg_colsum_part, dim3(unsigned(nch), unsigned(nseg)), dim3(SEG_THREADS), 0, st, dZ, off, N,
                       nch, part);
    LDDMM_CHECK_LAUNCH();
    hipLaunchKernelGGL(k_seg_colsum_fold, dim3(nblk(size_t(nseg) * N)), dim3(256), 0, st, part, nseg, nch, N, out);
    LDDMM_CHECK_LAUNCH();
    return LDDMM_OK;
}

int lddmm_nn_segmax(const float* H, const int32_t* off, int32_t nseg, int32_t N, float* g, int32_t* arg,
                    void* stream) {
    LDDMM_REQUIRE(nseg >= 0 && N > 0 && nseg < 65536, "nn_segmax: bad sizes");
    if (nseg == 0) return LDDMM_OK;
    hipStream_t st = static_cast<hipStream_t>(stream);
    const int nch = SEG_CHUNKS;
    void* ws;
    const size_t cnt = size_t(nseg) * nch * N;
    LDDMM_TRY(workspace(cnt * (sizeof(float) + sizeof(int32_t)) + 512, st, &ws));
    float* pv = static_cast<float*>(ws);
    int32_t* pi = reinterpret_cast<int32_t*>(static_cast<char*>(ws) + ((cnt * sizeof(float) + 255) & ~size_t(255)));
    hipLaunchKernelGGL(k_segmax_part, dim3(unsigned(nch), unsigned(nseg)), dim3(SEG_THREADS), 0, st, H, off, N, nch,
                       pv, pi);
    LDDMM_CHECK_LAUNCH();
    hipLaunchKernelGGL(k_segmax_fold, dim3(nblk(size_t(nseg) * N)), dim3(256), 0, st, pv, pi, off, nseg, nch, N, g,
                       arg);
    LDDMM_CHECK_LAUNCH();
    return LDDMM_OK;
}

int lddmm_nn_segmax_bwd(const float* dg, const int32_t* off, const int32_t* arg, int32_t nseg, int32_t N, float* dH,
                        void* stream) {
    LDDMM_REQUIRE(nseg >= 0 && N > 0, "nn_segmax_bwd: bad sizes");
    if (nseg == 0) return LDDMM_OK;
    hipLaunchKernelGGL(k_segmax_bwd, dim3(unsigned(nseg)), dim3(128), 0, static_cast<hipStream_t>(stream), dg, off, arg,
                       N, dH);
    LDDMM_CHECK_LAUNCH();
    return LDDMM_OK;
}

int lddmm_nn_adam(float* w, uint16_t* wbf, const float* g, float* m, float* v, int64_t count, float lr, float b1,
                  float b2, float eps, int32_t step, float wd, void* stream) {
    LDDMM_REQUIRE(count >= 0 && step >= 1, "nn_adam: bad arguments");
    if (count == 0) return LDDMM_OK;
    float c1 = 1.f - powf(b1, float(step)), c2 = 1.f - powf(b2, float(step));
    hipLaunchKernelGGL(k_adam, dim3(nblk(size_t(count))), dim3(256), 0, static_cast<hipStream_t>(stream), w, wbf, g, m,
                       v, size_t(count), lr, b1, b2, eps, c1, c2, wd);
    LDDMM_CHECK_LAUNCH();
    return LDDMM_OK;
}

int lddmm_nn_to_bf16(const float* w, uint16_t* wbf, int64_t count, void* stream) {
    if (count <= 0) return LDDMM_OK;
    hipLaunchKernelGGL(k_to_bf16, dim3(nblk(size_t(count))), dim3(256), 0, static_cast<hipStream_t>(stream), w, wbf,
                       size_t(count));
    LDDMM_CHECK_LAUNCH();
    return LDDMM_OK;
}

}  // extern "C"
