// Tiled N-body pair-sum machinery shared by every Gaussian kernel sum.
//
// One workgroup (kBlock lanes = 4 wave64) owns kBlock*R consecutive targets
// and one chunk of sources (grid.y).  Sources are staged through LDS in tiles
// of kTile records: each lane loads one source from HBM (float64, coalesced
// AoS), centres / scales / converts it and stores one LDS record; the next
// tile's raw data is fetched into registers while the current tile is
// consumed.  The inner loop reads each record with a wave-uniform address
// (LDS broadcast) and applies the op's pair function to R targets held in
// registers, so each LDS read feeds R pairs.
//
// Precision: T = float accumulates each tile in plain float32 and folds tile
// sums into a (hi, lo) two-sum accumulator; chunk partials are combined by a
// second two-sum pass in k_reduce.  T = double accumulates in float64.
#pragma once

#include "common.hpp"

namespace lddmm {

constexpr int kBlock = 256;   // lanes per workgroup
constexpr int kTile = 256;    // sources per LDS tile (one per lane)

template <typename T>
struct Partials {
    T* hi;       // [nchunk][nch][n]
    T* lo;       // float: compensation terms, same layout; double: unused
    int n;
};

template <typename T>
constexpr bool kCompensated = sizeof(T) == 4;

// Default chunk -> source range mapping: contiguous chunks of `chunk` sources.
struct ContiguousChunks {
    template <class A>
    __device__ static void range(const A& a, int chunk_id, int chunk, int m, int& j0, int& j1) {
        j0 = chunk_id * chunk;
        j1 = min(m, j0 + chunk);
    }
};

// One workgroup's share: targets [i0, i0 + kBlock*R) clipped to `ilim`,
// sources [j0, j1); channel c of target i goes to ph[c*ldn + (i - ioff)]
// (and pl[...] for the compensation term).
template <typename T, class OP, int R>
__device__ __forceinline__ void pair_span(const typename OP::template Args<T>& args, int i0, int ilim, int j0,
                                          int j1, T* __restrict__ ph, T* __restrict__ pl, size_t ldn, int ioff,
                                          T* lds) {
    constexpr int NCH = OP::NCH;
    constexpr int REC = OP::REC;
    const int tid = threadIdx.x;
    const int ibase = i0 + tid;

    typename OP::template Tgt<T> tg[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
        int i = ibase + r * kBlock;
        OP::load_target(args, i < ilim ? i : ilim - 1, tg[r]);
    }

    T hi[R][NCH], lo[R][NCH];
#pragma unroll
    for (int r = 0; r < R; ++r)
#pragma unroll
        for (int c = 0; c < NCH; ++c) { hi[r][c] = T(0); lo[r][c] = T(0); }

    typename OP::Raw raw;
    OP::load_raw(args, j0 + tid < j1 ? j0 + tid : -1, raw);

    for (int jt = j0; jt < j1; jt += kTile) {
        __syncthreads();                       // previous tile fully consumed
        OP::store_rec(args, raw, &lds[tid * REC]);
        __syncthreads();
        {
            int jn = jt + kTile + tid;         // prefetch the next tile
            OP::load_raw(args, jn < j1 ? jn : -1, raw);
        }
        T acc[R][NCH];
#pragma unroll
        for (int r = 0; r < R; ++r)
#pragma unroll
            for (int c = 0; c < NCH; ++c) acc[r][c] = T(0);

        if (j1 - jt >= kTile) {
#pragma unroll 4
            for (int t = 0; t < kTile; ++t) {
                const T* rec = &lds[t * REC];
#pragma unroll
                for (int r = 0; r < R; ++r) OP::pair(args, tg[r], rec, acc[r]);
            }
        } else {
            // last, partial tile: records past j1 are zero-weight padding, so the
            // count is rounded up to a multiple of 4 and the rest skipped
            const int tn = (j1 - jt + 3) & ~3;
#pragma unroll 4
            for (int t = 0; t < tn; ++t) {
                const T* rec = &lds[t * REC];
#pragma unroll
                for (int r = 0; r < R; ++r) OP::pair(args, tg[r], rec, acc[r]);
            }
        }
#pragma unroll
        for (int r = 0; r < R; ++r)
#pragma unroll
            for (int c = 0; c < NCH; ++c) {
                if constexpr (kCompensated<T>) two_sum_acc(hi[r][c], lo[r][c], acc[r][c]);
                else hi[r][c] += acc[r][c];
            }
    }

#pragma unroll
    for (int r = 0; r < R; ++r) {
        int i = ibase + r * kBlock;
        if (i < ilim) {
#pragma unroll
            for (int c = 0; c < NCH; ++c) {
                size_t o = size_t(c) * ldn + size_t(i - ioff);
                ph[o] = hi[r][c];
                if constexpr (kCompensated<T>) pl[o] = lo[r][c];
            }
        }
    }
}

template <typename T, class OP, int R>
__global__ void __launch_bounds__(kBlock)
k_pair(const typename OP::template Args<T> args, int n, int m, int chunk, Partials<T> part) {
    __shared__ __attribute__((aligned(16))) T lds[kTile * OP::REC];
    int j0, j1;
    OP::Chunks::range(args, blockIdx.y, chunk, m, j0, j1);
    const size_t off = size_t(blockIdx.y) * OP::NCH * size_t(n);
    pair_span<T, OP, R>(args, blockIdx.x * (kBlock * R), n, j0, j1, part.hi + off,
                        kCompensated<T> ? part.lo + off : nullptr, size_t(n), 0, lds);
}

// Sum of chunk partials for (i, c) in float64 (two-sum in float32 first).
// Loads are issued in batches of kSumBatch independent reads (the sums stay in
// chunk order, so the result is the same as one chunk at a time).
constexpr int kSumBatch = 8;
template <typename T>
__device__ __forceinline__ double sum_chunks(const Partials<T>& p, int nchunk, int nch, int i, int c,
                                             const double* chunk_w) {
    const size_t stride = size_t(nch) * size_t(p.n);
    const size_t base = size_t(c) * size_t(p.n) + size_t(i);
    if constexpr (kCompensated<T>) {
        float h = 0.f, l = 0.f;
        for (int k0 = 0; k0 < nchunk; k0 += kSumBatch) {
            float vh[kSumBatch], vl[kSumBatch];
#pragma unroll
            for (int b = 0; b < kSumBatch; ++b) {
                const int k = k0 + b;
                vh[b] = k < nchunk ? p.hi[base + size_t(k) * stride] : 0.f;
                vl[b] = k < nchunk ? p.lo[base + size_t(k) * stride] : 0.f;
            }
#pragma unroll
            for (int b = 0; b < kSumBatch; ++b) {
                const int k = k0 + b;
                if (k < nchunk) {
                    float w = chunk_w ? float(chunk_w[k]) : 1.f;
                    two_sum_acc(h, l, vh[b] * w);
                    l += vl[b] * w;
                }
            }
        }
        return double(h) + double(l);
    } else {
        double s = 0.0;
        for (int k0 = 0; k0 < nchunk; k0 += kSumBatch) {
            double v[kSumBatch];
#pragma unroll
            for (int b = 0; b < kSumBatch; ++b) {
                const int k = k0 + b;
                v[b] = k < nchunk ? p.hi[base + size_t(k) * stride] : 0.0;
            }
#pragma unroll
            for (int b = 0; b < kSumBatch; ++b) {
                const int k = k0 + b;
                if (k < nchunk) s += chunk_w ? v[b] * chunk_w[k] : v[b];
            }
        }
        return s;
    }
}

// Chunk sums spread over kEpiLanes lanes per row: lane l adds chunks
// k0 + l, k0 + l + kEpiLanes, ...; epilogues combine the lanes in lane order.
constexpr int kEpiLanes = 4;
constexpr int kEpiRows = 64;     // rows per 256-lane epilogue workgroup

template <typename T>
__device__ __forceinline__ double sum_lane(const Partials<T>& p, int k0, int k1, int nch, int i, int c, int lane) {
    // this lane's chunks k0 + lane + kEpiLanes * q, q < cnt; loads issued kSumBatch at a time
    const int cnt = k1 - k0 > lane ? (k1 - k0 - lane + kEpiLanes - 1) / kEpiLanes : 0;
    const size_t stride = size_t(kEpiLanes) * size_t(nch) * size_t(p.n);
    const size_t base = (size_t(k0 + lane) * nch + c) * size_t(p.n) + size_t(i);
    if constexpr (kCompensated<T>) {
        float h = 0.f, l = 0.f;
        for (int q0 = 0; q0 < cnt; q0 += kSumBatch) {
            float vh[kSumBatch], vl[kSumBatch];
#pragma unroll
            for (int b = 0; b < kSumBatch; ++b) {
                const int q = q0 + b;
                vh[b] = q < cnt ? p.hi[base + size_t(q) * stride] : 0.f;
                vl[b] = q < cnt ? p.lo[base + size_t(q) * stride] : 0.f;
            }
#pragma unroll
            for (int b = 0; b < kSumBatch; ++b)
                if (q0 + b < cnt) {
                    two_sum_acc(h, l, vh[b]);
                    l += vl[b];
                }
        }
        return double(h) + double(l);
    } else {
        double s = 0.0;
        for (int q0 = 0; q0 < cnt; q0 += kSumBatch) {
            double v[kSumBatch];
#pragma unroll
            for (int b = 0; b < kSumBatch; ++b) {
                const int q = q0 + b;
                v[b] = q < cnt ? p.hi[base + size_t(q) * stride] : 0.0;
            }
#pragma unroll
            for (int b = 0; b < kSumBatch; ++b)
                if (q0 + b < cnt) s += v[b];
        }
        return s;
    }
}

struct ChanScale { double s[8]; };

// out[o] = (base ? base[o] : (accumulate ? out[o] : 0)) + alpha * scale_c * v,
// o = i*ldo + col0 + c, v = sum over chunks of channel c; partials hold `nch`
// channels per chunk, of which the first `nout` are written.
// 64 rows x 4 chunk lanes per workgroup; c = blockIdx.y.
template <typename T>
__global__ void __launch_bounds__(256) k_reduce(Partials<T> part, int nchunk, int nch, int nout, int n, ChanScale sc,
                                                double* out, int ldo, int col0, int accumulate, const double* base,
                                                double alpha) {
    __shared__ double red[kEpiLanes][kEpiRows];
    const int fl = threadIdx.x % kEpiRows, lane = threadIdx.x / kEpiRows;
    const int i = blockIdx.x * kEpiRows + fl;
    const int c = blockIdx.y;
    red[lane][fl] = (i < n && c < nout) ? sum_lane(part, 0, nchunk, nch, i, c, lane) : 0.0;
    __syncthreads();
    if (lane != 0 || i >= n || c >= nout) return;
    const double v = alpha * (((red[0][fl] + red[1][fl]) + red[2][fl]) + red[3][fl]) * sc.s[c];
    size_t o = size_t(i) * ldo + col0 + c;
    if (base) out[o] = base[o] + v;
    else if (accumulate) out[o] += v;
    else out[o] = v;
}

__global__ void k_fill(double* p, size_t count, double v);

// Midpoint of the bounding box of y[n,dim] and x[m,dim] -> c0[0..2] (c0[d] = 0 for d >= dim).
__global__ void k_center(const double* y, int n, const double* x, int m, int dim, double* c0);

int launch_center(const double* y, int n, const double* x, int m, int dim, double* c0, hipStream_t st);
int launch_fill(double* p, size_t count, double v, hipStream_t st);

// ---------------------------------------------------------------------------
// Launch planning: pick the number of source chunks so that the busiest CU
// carries as little work as possible (blocks spread evenly over the CUs,
// enough of them resident to hide latency).
// ---------------------------------------------------------------------------
struct PairPlan {
    int bx = 0;       // target blocks
    int nchunk = 0;   // source chunks (grid.y)
    int chunk = 0;    // sources per chunk: a multiple of the quantum, about 64 or more unless m is smaller
};

// quantum: chunk granularity in sources.  4 for the VALU kernels (a partial
// tile runs only its used count, rounded up to 4); a multiple of 64 for
// kernels that always process whole tiles.  max_nchunk bounds the number of
// source chunks (callers whose partial buffer has a fixed size).
PairPlan plan_pair(int n, int m, int R, int num_cu, int occ_blocks, int quantum = 4,
                   int max_nchunk = 1 << 30);

template <typename T, class OP, int R>
int run_pair(const typename OP::template Args<T>& args, int n, int m, const PairPlan& plan,
             Partials<T> part, hipStream_t st) {
    dim3 grid(plan.bx, plan.nchunk);
    const bool prof = profiling();
    if (prof) LDDMM_TRY(profile_begin(st));
    hipLaunchKernelGGL((k_pair<T, OP, R>), grid, dim3(kBlock), 0, st, args, n, m, plan.chunk, part);
    LDDMM_CHECK_LAUNCH();
    if (prof) LDDMM_TRY(profile_end(st));
    return LDDMM_OK;
}

template <typename T, class OP, int R>
int plan_for(int n, int m, PairPlan* plan, int max_nchunk = 1 << 30) {
    DeviceInfo di;
    LDDMM_TRY(device_info(&di));
    int occ = blocks_per_cu(reinterpret_cast<const void*>(&k_pair<T, OP, R>), kBlock, 0);
    *plan = plan_pair(n, m, R, di.num_cu, occ, 4, max_nchunk);
    return LDDMM_OK;
}

// Bytes carve_partials takes (each slice rounded up to Carve's 256 bytes).
template <typename T>
size_t partial_bytes(const PairPlan& p, int nch, int n) {
    size_t e = (size_t(p.nchunk) * nch * size_t(n) * sizeof(T) + 255) & ~size_t(255);
    return kCompensated<T> ? 2 * e : e;
}

template <typename T>
Partials<T> carve_partials(char* base, Carve& cv, const PairPlan& p, int nch, int n) {
    size_t e = size_t(p.nchunk) * nch * size_t(n) * sizeof(T);
    Partials<T> r;
    r.hi = reinterpret_cast<T*>(base + cv.take(e));
    r.lo = kCompensated<T> ? reinterpret_cast<T*>(base + cv.take(e)) : nullptr;
    r.n = n;
    return r;
}

template <typename T>
int run_reduce(const Partials<T>& part, int nchunk, int nch, int nout, int n, const ChanScale& sc,
               double* out, int ldo, int col0, int accumulate, const double* base, double alpha,
               hipStream_t st) {
    int nb = (n + kEpiRows - 1) / kEpiRows;
    hipLaunchKernelGGL((k_reduce<T>), dim3(nb, std::max(1, nout)), dim3(256), 0, st, part, nchunk, nch, nout, n,
                       sc, out,
                       ldo, col0, accumulate, base, alpha);
    LDDMM_CHECK_LAUNCH();
    return LDDMM_OK;
}

// ---------------------------------------------------------------------------
// Radial factors for one pair given r2 = |x - y|^2 (centred units; single
// scale: scaled so that K = exp2(-r2)).  The general path evaluates, per
// scale q, the kernel's profile (see ScaleSet) weighted by wK (kernel sums),
// wD (first-derivative sums) or wDD (second-derivative sums).
// ---------------------------------------------------------------------------
template <typename T>
__device__ __forceinline__ T lap_poly5(const T* c, T u) {
    return c[0] + u * (c[1] + u * (c[2] + u * (c[3] + u * c[4])));
}
template <typename T>
__device__ __forceinline__ T lap_poly4(const T* c, T u) {
    return c[0] + u * (c[1] + u * (c[2] + u * c[3]));
}
template <typename T>
__device__ __forceinline__ T lap_poly3(const T* c, T u) {
    return c[0] + u * (c[1] + u * c[2]);
}

template <typename T, bool MS>
__device__ __forceinline__ T gK(const ScaleSet<T>& s, T r2) {
    if constexpr (!MS) {
        return fast_exp2(-r2);
    } else {
        T g = T(0);
        if (s.kind == LDDMM_KERNEL_GAUSS) {
            for (int q = 0; q < s.ns; ++q) g += s.wK[q] * fast_exp2(-s.beta[q] * r2);
        } else {
            const T r = sqrt(r2);
            for (int q = 0; q < s.ns; ++q) {
                const T u = r * s.rinv[q];
                g += s.wK[q] * lap_poly5(s.pc, u) * fast_exp2(-T(kLog2e) * u);
            }
        }
        return g;
    }
}

template <typename T, bool MS>
__device__ __forceinline__ T gD(const ScaleSet<T>& s, T r2) {
    if constexpr (!MS) {
        return fast_exp2(-r2);
    } else {
        T g = T(0);
        if (s.kind == LDDMM_KERNEL_GAUSS) {
            for (int q = 0; q < s.ns; ++q) g += s.wD[q] * fast_exp2(-s.beta[q] * r2);
        } else {
            const T r = sqrt(r2);
            for (int q = 0; q < s.ns; ++q) {
                const T u = r * s.rinv[q];
                g += s.wD[q] * lap_poly4(s.qc, u) * fast_exp2(-T(kLog2e) * u);
            }
        }
        return g;
    }
}

// Kernel and first-derivative weightings from one set of exponentials.
template <typename T, bool MS>
__device__ __forceinline__ void gKD(const ScaleSet<T>& s, T r2, T& k, T& d) {
    if constexpr (!MS) {
        k = d = fast_exp2(-r2);
    } else {
        k = d = T(0);
        if (s.kind == LDDMM_KERNEL_GAUSS) {
            for (int q = 0; q < s.ns; ++q) {
                T e = fast_exp2(-s.beta[q] * r2);
                k += s.wK[q] * e;
                d += s.wD[q] * e;
            }
        } else {
            const T r = sqrt(r2);
            for (int q = 0; q < s.ns; ++q) {
                const T u = r * s.rinv[q];
                const T e = fast_exp2(-T(kLog2e) * u);
                k += s.wK[q] * lap_poly5(s.pc, u) * e;
                d += s.wD[q] * lap_poly4(s.qc, u) * e;
            }
        }
    }
}

// Second- and first-derivative weightings from one set of exponentials.
template <typename T, bool MS>
__device__ __forceinline__ void gDDD(const ScaleSet<T>& s, T r2, T& dd, T& d) {
    if constexpr (!MS) {
        dd = d = fast_exp2(-r2);
    } else {
        dd = d = T(0);
        if (s.kind == LDDMM_KERNEL_GAUSS) {
            for (int q = 0; q < s.ns; ++q) {
                T e = fast_exp2(-s.beta[q] * r2);
                dd += s.wDD[q] * e;
                d += s.wD[q] * e;
            }
        } else {
            const T r = sqrt(r2);
            for (int q = 0; q < s.ns; ++q) {
                const T u = r * s.rinv[q];
                const T e = fast_exp2(-T(kLog2e) * u);
                dd += s.wDD[q] * lap_poly3(s.rc, u) * e;
                d += s.wD[q] * lap_poly4(s.qc, u) * e;
            }
        }
    }
}

// Centred + scaled coordinate d of point i in a [*, dim] float64 array.
template <typename T>
__device__ __forceinline__ T coord(const double* p, int i, int dim, int d, const double* c0, double cs) {
    return d < dim ? T((p[size_t(i) * dim + d] - c0[d]) * cs) : T(0);
}

}  // namespace lddmm
