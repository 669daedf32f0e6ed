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

// Per-segment column sums: out[s][n] = sum_{p in segment s} dZ[p][n]; segments are
// contiguous point ranges [off[s], off[s+1]).
__global__ void k_seg_colsum(const float* __restrict__ dZ, const int32_t* __restrict__ off, int N,
                             float* __restrict__ out) {
    int s = blockIdx.x;
    for (int n = threadIdx.x; n < N; n += blockDim.x) {
        float acc = 0.f;
        for (int p = off[s]; p < off[s + 1]; ++p) acc += dZ[size_t(p) * N + n];
        out[size_t(s) * N + n] = acc;
    }
}

// g[s][n] = max_{p in s} H[p][n], arg[s][n] = first p attaining it
__global__ void k_segmax(const float* __restrict__ H, const int32_t* __restrict__ off, int N, float* __restrict__ g,
                         int32_t* __restrict__ arg) {
    int s = blockIdx.x;
    for (int n = threadIdx.x; n < N; n += blockDim.x) {
        float best = -INFINITY;
        int bi = off[s];
        for (int p = off[s]; p < off[s + 1]; ++p) {
            float v = H[size_t(p) * N + n];
            if (v > best) { best = v; bi = p; }
        }
        g[size_t(s) * N + n] = off[s + 1] > off[s] ? best : 0.f;
        arg[size_t(s) * N + n] = bi;
    }
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
    LDDMM_REQUIRE(nseg >= 0 && N > 0, "nn_seg_colsum: bad sizes");
    if (nseg == 0) return LDDMM_OK;
    hipLaunchKernelGGL(k_seg_colsum, dim3(unsigned(nseg)), dim3(128), 0, static_cast<hipStream_t>(stream), dZ, off, N,
                       out);
    LDDMM_CHECK_LAUNCH();
    return LDDMM_OK;
}

int lddmm_nn_segmax(const float* H, const int32_t* off, int32_t nseg, int32_t N, float* g, int32_t* arg,
                    void* stream) {
    LDDMM_REQUIRE(nseg >= 0 && N > 0, "nn_segmax: bad sizes");
    if (nseg == 0) return LDDMM_OK;
    hipLaunchKernelGGL(k_segmax, dim3(unsigned(nseg)), dim3(128), 0, static_cast<hipStream_t>(stream), H, off, N, g,
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
