// Shared host/device helpers of the lddmm_amd HIP library (gfx950).
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstddef>
#include <cstdio>
#include <cmath>
#include <string>

#include "../../include/lddmm_amd.h"

namespace lddmm {

// ---------------------------------------------------------------------------
// Errors
// ---------------------------------------------------------------------------
void set_error(const char* fmt, ...);
int hip_fail(hipError_t e, const char* what, const char* file, int line);

#define LDDMM_HIP(call)                                                        \
    do {                                                                       \
        hipError_t e_ = (call);                                                \
        if (e_ != hipSuccess) return ::lddmm::hip_fail(e_, #call, __FILE__, __LINE__); \
    } while (0)

#define LDDMM_CHECK_LAUNCH() LDDMM_HIP(hipGetLastError())

#define LDDMM_REQUIRE(cond, ...)                                               \
    do {                                                                       \
        if (!(cond)) { ::lddmm::set_error(__VA_ARGS__); return LDDMM_EINVAL; } \
    } while (0)

// Runs `expr` (an int-returning call) and propagates a non-zero result.
#define LDDMM_TRY(expr)                                                        \
    do { int rc_ = (expr); if (rc_ != LDDMM_OK) return rc_; } while (0)

// Largest point count of a pair sum: the kernels index targets and sources in
// int, and a target block's first index blockIdx.x * (256 lanes * 2 targets)
// must stay below 2^31 for the last block too.
constexpr int64_t kMaxPoints = (int64_t(1) << 31) - 4096;

// ---------------------------------------------------------------------------
// Device facts and workspace
// ---------------------------------------------------------------------------
struct DeviceInfo {
    int device = -1;
    int num_cu = 0;
};
int device_info(DeviceInfo* out);

// Grow-only per-device scratch.  The returned pointer stays valid until the
// next call that needs more bytes on the same device; growth synchronises
// `stream` before the old block is released.
int workspace(size_t bytes, hipStream_t stream, void** ptr);

// Bump allocator over one workspace request (256-byte aligned slices).
struct Carve {
    size_t off = 0;
    size_t take(size_t bytes) {
        size_t o = off;
        off += (bytes + 255) & ~size_t(255);
        return o;
    }
};

// Pair-kernel timing hooks (enabled by lddmm_profile_enable).
bool profiling();
int profile_begin(hipStream_t st);
int profile_end(hipStream_t st);

// While an instance lives, profiling() is false on the calling thread (the
// hooks create HIP events, which entry points that must not allocate skip).
struct ProfileOff {
    ProfileOff();
    ~ProfileOff();
    bool prev;
};

// Occupancy (blocks per CU) of a kernel, cached by function address.
int blocks_per_cu(const void* func, int block, size_t dyn_lds);

// ---------------------------------------------------------------------------
// Kernel description -> per-precision scale constants
// ---------------------------------------------------------------------------
constexpr double kLog2e = 1.4426950408889634;

// Radial profiles of the general path.  With r = |x - y| / s:
//   gauss:     K = D = DD = exp(-r^2 / 2)
//   laplacian: K = P(r) e^-r, D = Q(r) e^-r, DD = R(r) e^-r
// where the derivative sums use D (first) and DD (second derivative), and
// Q = (P - P') / r, R = (Q - Q') / r for the order's polynomial P
// (base/kernelFunctions_util.py:20-52 tables c_, c1_, c2_).
template <typename T>
struct ScaleSet {
    int ns;                         // number of scales
    int kind;                       // LDDMM_KERNEL_GAUSS / LDDMM_KERNEL_LAPLACIAN
    T beta[LDDMM_MAX_SCALES];       // gauss: exp2 coefficient on |x - y|^2 (centred units)
    T rinv[LDDMM_MAX_SCALES];       // laplacian: 1 / s
    T wK[LDDMM_MAX_SCALES];         // s^KP / W
    T wD[LDDMM_MAX_SCALES];         // s^(KP-2) / W
    T wDD[LDDMM_MAX_SCALES];        // s^(KP-4) / W
    T pc[5], qc[4], rc[3];          // laplacian P, Q, R coefficients (ascending powers of r)
    double cs;                      // coordinate factor (single-scale gauss: K = exp2(-|u_i - u_j|^2), else 1)
};

// Checks a kernel description.  Flags: kValDataTerm for the fused varifold
// kernels (Gaussian kind only), kValAffine for the entry points that apply
// the kernel's affine part (the others reject affine != LDDMM_AFFINE_NONE).
enum : unsigned { kValDataTerm = 1u, kValAffine = 2u };
int validate_kernel(const lddmm_kernel* kd, unsigned flags = 0u);

// The single-scale Gaussian instantiation (scaled coordinates, exp2 of r^2)
// serves one Gaussian scale; every other kernel takes the general one.
inline bool general_path(const lddmm_kernel* kd) {
    return kd->nscale > 1 || kd->kind != LDDMM_KERNEL_GAUSS;
}

// Laplacian polynomial coefficients by order (base/kernelFunctions_util.py:20-37).
constexpr double kLapP[5][5] = {{1, 0, 0, 0, 0},
                                {1, 1, 0, 0, 0},
                                {1, 1, 1.0 / 3, 0, 0},
                                {1, 1, 0.4, 1.0 / 15, 0},
                                {1, 1, 3.0 / 7, 2.0 / 21, 1.0 / 105}};
constexpr double kLapQ[5][4] = {{0, 0, 0, 0},
                                {1, 0, 0, 0},
                                {1.0 / 3, 1.0 / 3, 0, 0},
                                {1.0 / 5, 1.0 / 5, 1.0 / 15, 0},
                                {1.0 / 7, 1.0 / 7, 2.0 / 35, 1.0 / 105}};
constexpr double kLapR[5][3] = {{0, 0, 0},
                                {0, 0, 0},
                                {1.0 / 3, 0, 0},
                                {1.0 / 15, 1.0 / 15, 0},
                                {1.0 / 35, 1.0 / 35, 1.0 / 105}};

// Constant factors a single-scale op applies at the end, in float64.
struct SingleScaleConsts {
    double sigma, cs;
    double diff;      // -1 / (s^2 cs): derivative sums (diffKT, diffK, kdiffmat)
    double ddiff;     //  1 / s^2: second-derivative sums
    double gamma;     //  1 / (s^2 cs^2): (d.q) d term in DDiffK
};

template <typename T>
ScaleSet<T> make_scales(const lddmm_kernel* kd, SingleScaleConsts* single) {
    ScaleSet<T> s{};
    s.ns = kd->nscale;
    s.kind = kd->kind;
    double W = 0;
    for (int i = 0; i < kd->nscale; ++i) W += std::pow(kd->sigma[i], kd->kp);
    for (int i = 0; i < kd->nscale; ++i) {
        double sg = kd->sigma[i];
        s.beta[i] = T(kLog2e / (2.0 * sg * sg));
        s.rinv[i] = T(1.0 / sg);
        s.wK[i] = T(std::pow(sg, kd->kp) / W);
        s.wD[i] = T(std::pow(sg, kd->kp - 2) / W);
        s.wDD[i] = T(std::pow(sg, kd->kp - 4) / W);
    }
    if (kd->kind == LDDMM_KERNEL_LAPLACIAN) {
        const int o = kd->order;
        for (int c = 0; c < 5; ++c) s.pc[c] = T(kLapP[o][c]);
        for (int c = 0; c < 4; ++c) s.qc[c] = T(kLapQ[o][c]);
        for (int c = 0; c < 3; ++c) s.rc[c] = T(kLapR[o][c]);
    }
    double sg = kd->sigma[0];
    double cs = std::sqrt(kLog2e / 2.0) / sg;
    s.cs = general_path(kd) ? 1.0 : cs;
    if (single) {
        single->sigma = sg;
        single->cs = cs;
        single->diff = -1.0 / (sg * sg * cs);
        single->ddiff = 1.0 / (sg * sg);
        single->gamma = 1.0 / (sg * sg * cs * cs);
    }
    return s;
}

// ---------------------------------------------------------------------------
// Device math
// ---------------------------------------------------------------------------
__device__ __forceinline__ float fast_exp2(float x) { return __builtin_amdgcn_exp2f(x); }

// 2^x in float64 for the Gaussian factors (x <= 0 there; any x gives a value):
// Cody-Waite style reduction x = n + f with n = rint(x) and |f| <= 1/2 exact,
// 2^f from a degree-10 polynomial (Chebyshev fit of 2^f on [-1/2, 1/2],
// relative error 2.2e-16 before rounding, at most 4.0e-16 when evaluated by
// Horner's rule with FMA in float64), scaled by 2^n with ldexp (gradual
// underflow below 2^-1022, zero below about 2^-1075).  Arguments below -1100
// are clamped (the result is zero either way), which keeps n inside int; a
// NaN argument gives 0.  kExp2Poly is produced and checked by
// exp2_coeffs.py beside this file.  15 VALU instructions, against about 20
// for the device library's exp2(double).
constexpr double kExp2Poly[11] = {
    7.072585949269223e-09, 1.0208690299958306e-07, 1.321544258792169e-06, 1.5252657260200837e-05,
    0.0001540353044173605, 0.0013333558230164974, 0.009618129107606888, 0.05550410866444772,
    0.24022650695910097, 0.69314718055995, 1.0};

__device__ __forceinline__ double fast_exp2(double x) {
    const double xc = __builtin_fmax(x, -1100.0);
    const double n = __builtin_rint(xc);
    const double f = xc - n;
    double p = kExp2Poly[0];
#pragma unroll
    for (int k = 1; k < 11; ++k) p = __builtin_fma(p, f, kExp2Poly[k]);
    return __builtin_ldexp(p, int(n));
}

// Knuth two-sum: (s, e) with s + e == a + b exactly.
template <typename T>
__device__ __forceinline__ void two_sum_acc(T& hi, T& lo, T b) {
    T s = hi + b;
    T bp = s - hi;
    T e = (hi - (s - bp)) + (b - bp);
    hi = s;
    lo += e;
}

template <typename T> struct Vec4 { T x, y, z, w; };

template <typename T>
__device__ __forceinline__ Vec4<T> lds4(const T* p);

template <>
__device__ __forceinline__ Vec4<float> lds4<float>(const float* p) {
    float4 v = *reinterpret_cast<const float4*>(p);
    return {v.x, v.y, v.z, v.w};
}
template <>
__device__ __forceinline__ Vec4<double> lds4<double>(const double* p) {
    double2 a = *reinterpret_cast<const double2*>(p);
    double2 b = *reinterpret_cast<const double2*>(p + 2);
    return {a.x, a.y, b.x, b.y};
}

// N consecutive values from p (16-byte aligned) in 16-byte reads, then one
// 8-byte read for a remaining pair (float) and a scalar read for a last
// value: a packed record costs only the LDS bytes it holds.
template <int N, typename T>
__device__ __forceinline__ void lds_n(const T* p, T* v) {
#pragma unroll
    for (int k = 0; k + 4 <= N; k += 4) {
        const Vec4<T> q = lds4(p + k);
        v[k] = q.x; v[k + 1] = q.y; v[k + 2] = q.z; v[k + 3] = q.w;
    }
    constexpr int k = N / 4 * 4;
    if constexpr (N - k >= 2) {
        if constexpr (sizeof(T) == 4) {
            const float2 q = *reinterpret_cast<const float2*>(p + k);
            v[k] = q.x; v[k + 1] = q.y;
        } else {
            const double2 q = *reinterpret_cast<const double2*>(p + k);
            v[k] = q.x; v[k + 1] = q.y;
        }
    }
    if constexpr ((N - k) % 2 == 1) v[N - 1] = p[N - 1];
}

}  // namespace lddmm
