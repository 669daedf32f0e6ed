// Population registration: P template-target surface pairs of uneven sizes.
//
// Every Gaussian sum over the active pairs is one persistent launch of
// k_pair_pop.  Its work list is a segment table of items (pair, target tile
// of 256*R rows, source chunk) that the host schedules longest-first over a
// fixed number of workgroups (number of CUs x occupancy); each workgroup
// walks its own slice of the table.  Sources of an item always belong to the
// item's pair, so the op's pair function is the single-pair one, with the
// pair's own centre.  Chunk partials live in one compact buffer (per pair:
// [chunk][channel][row]) and are combined per row in float64 by the
// epilogues, as in the single-pair path.
//
// Launch shapes (R targets per lane, items per workgroup) are picked per op
// and precision by timing the candidates on the first call (autotuner).  The
// T-step forward and adjoint loops are captured into hipGraphs (one per
// buffer set) and replayed; the item tables are read from device memory at
// run time, so a graph stays valid when the active set changes.
#include "ops.hpp"
#include "varifold.hpp"

#include <algorithm>
#include <cstring>
#include <map>
#include <mutex>
#include <queue>
#include <vector>

namespace lddmm {

struct PopItem {
    int32_t r0, r1, s0, s1, pair, slot;
};
struct PopPair {
    int32_t roff, n, nchunk, c1;   // rows [roff, roff+n); chunks [0, c1) come from the first source part
    int64_t pbase;                 // partial offset in row-chunk units
};

// ------------------------------------------------------------------ kernels

template <typename T, class OP, int R>
__global__ void __launch_bounds__(kBlock)
k_pair_pop(const typename OP::template Args<T> args, const PopItem* __restrict__ items,
           const int32_t* __restrict__ bstart, const PopPair* __restrict__ pairs, const double* __restrict__ c0s,
           T* __restrict__ phi, T* __restrict__ plo, const uint8_t* __restrict__ active) {
    using Args = typename OP::template Args<T>;
    __shared__ __attribute__((aligned(16))) T lds[kTile * OP::REC];
    // The multi-scale radial factors index the scale arrays with a loop
    // counter; a private copy of the arguments indexed that way lives in
    // scratch memory, so those operators read a workgroup copy in LDS (the
    // single-scale copy below stays in registers).
    __shared__ __attribute__((aligned(16))) Args sargs;
    if constexpr (OP::kMultiScale) {
        static_assert(sizeof(Args) % 4 == 0, "argument block is copied by 32-bit words");
        const uint32_t* src = reinterpret_cast<const uint32_t*>(&args);
        uint32_t* dst = reinterpret_cast<uint32_t*>(&sargs);
        for (int k = threadIdx.x; k < int(sizeof(Args) / 4); k += kBlock) dst[k] = src[k];
    }
    const int e = bstart[blockIdx.x + 1];
    for (int it = bstart[blockIdx.x]; it < e; ++it) {
        const PopItem w = items[it];
        // items of pairs outside the active set are skipped (the item is the
        // same for every lane of the workgroup, so the branch is uniform)
        if (active && !active[w.pair]) continue;
        const PopPair pp = pairs[w.pair];
        const size_t off = (size_t(pp.pbase) + size_t(w.slot) * size_t(pp.n)) * OP::NCH;
        if constexpr (OP::kMultiScale) {
            __syncthreads();                   // previous item's reads of sargs done
            if (threadIdx.x == 0) sargs.c0 = c0s + 4 * w.pair;
            __syncthreads();
            pair_span<T, OP, R>(sargs, w.r0, w.r1, w.s0, w.s1, phi + off, kCompensated<T> ? plo + off : nullptr,
                                size_t(pp.n), pp.roff, lds);
        } else {
            Args a = args;
            a.c0 = c0s + 4 * w.pair;
            pair_span<T, OP, R>(a, w.r0, w.r1, w.s0, w.s1, phi + off, kCompensated<T> ? plo + off : nullptr,
                                size_t(pp.n), pp.roff, lds);
        }
    }
}

// Sum over chunks [k0, k1) of channel c of row i (float64 result).
template <typename T>
__device__ __forceinline__ double pop_sum(const T* phi, const T* plo, const PopPair& pp, int nch, int i, int c,
                                          int k0, int k1) {
    const size_t base = size_t(c) * pp.n + size_t(i - pp.roff);
    if constexpr (kCompensated<T>) {
        float h = 0.f, l = 0.f;
        for (int k = k0; k < k1; ++k) {
            size_t o = (size_t(pp.pbase) + size_t(k) * pp.n) * nch + base;
            two_sum_acc(h, l, phi[o]);
            l += plo[o];
        }
        return double(h) + double(l);
    } else {
        double s = 0.0;
        for (int k = k0; k < k1; ++k) s += phi[(size_t(pp.pbase) + size_t(k) * pp.n) * nch + base];
        return s;
    }
}

// out[i,c] = (base ? base[i,c] : 0) + alpha * scale * sum_c(i), active rows, 3 channels
template <typename T>
__global__ void k_pop_reduce3(const T* __restrict__ phi, const T* __restrict__ plo, const PopPair* __restrict__ pairs,
                              const int32_t* __restrict__ rowpair, const uint8_t* __restrict__ active, int nrows,
                              int nch, double scale, double alpha, const double* __restrict__ base,
                              double* __restrict__ out) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nrows) return;
    const int p = rowpair[i];
    if (!active[p]) return;
    const PopPair pp = pairs[p];
    for (int c = 0; c < 3; ++c) {
        double v = alpha * scale * pop_sum(phi, plo, pp, nch, i, c, 0, pp.nchunk);
        size_t o = size_t(i) * 3 + c;
        out[o] = (base ? base[o] : 0.0) + v;
    }
}

// Fused adjoint step (OpAdjoint partials, 6 channels), active rows:
// pc = pn + dt * fdiff * sum_{0..2}, g = -dt * sum_{3..5}.
template <typename T>
__global__ void k_pop_reduce_adj(const T* __restrict__ phi, const T* __restrict__ plo,
                                 const PopPair* __restrict__ pairs, const int32_t* __restrict__ rowpair,
                                 const uint8_t* __restrict__ active, int nrows, double fdiff, double dt,
                                 const double* __restrict__ pn, double* __restrict__ pc, double* __restrict__ g) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nrows) return;
    const int p = rowpair[i];
    if (!active[p]) return;
    const PopPair pp = pairs[p];
    for (int c = 0; c < 3; ++c) {
        const size_t o = size_t(i) * 3 + c;
        pc[o] = pn[o] + dt * fdiff * pop_sum(phi, plo, pp, 6, i, c, 0, pp.nchunk);
        g[o] = -(dt * pop_sum(phi, plo, pp, 6, i, 3 + c, 0, pp.nchunk));
    }
}

// x_next = x + dt v; rowacc_i += rdt * a_i.v_i  (v = summed partials), active rows
template <typename T>
__global__ void k_pop_shoot_epi(const T* __restrict__ phi, const T* __restrict__ plo,
                                const PopPair* __restrict__ pairs, const int32_t* __restrict__ rowpair,
                                const uint8_t* __restrict__ active, int nrows, const double* __restrict__ x,
                                const double* __restrict__ a, double dt, double* __restrict__ xn,
                                double* __restrict__ rowacc, double rdt) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nrows) return;
    const int p = rowpair[i];
    if (!active[p]) return;
    const PopPair pp = pairs[p];
    double d = 0.0;
    for (int c = 0; c < 3; ++c) {
        double v = pop_sum(phi, plo, pp, 3, i, c, 0, pp.nchunk);
        size_t o = size_t(i) * 3 + c;
        xn[o] = x[o] + dt * v;
        d += a[o] * v;
    }
    rowacc[i] += rdt * d;
}

// out = alpha * a + beta * b on the [nrows,3] rows of active pairs (a or b may be NULL)
__global__ void k_pop_axpby(int nrows, const int32_t* __restrict__ rowpair, const uint8_t* __restrict__ active,
                            double alpha, const double* __restrict__ a, double beta, const double* __restrict__ b,
                            double* __restrict__ out) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nrows) return;
    if (!active[rowpair[i]]) return;
    for (int c = 0; c < 3; ++c) {
        size_t o = size_t(i) * 3 + c;
        out[o] = (a ? alpha * a[o] : 0.0) + (b ? beta * b[o] : 0.0);
    }
}

__global__ void k_pop_zero(int nrows, const int32_t* __restrict__ rowpair, const uint8_t* __restrict__ active,
                           double* __restrict__ v) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < nrows && active[rowpair[i]]) v[i] = 0.0;
}

// Block reduction helper (256 lanes), fixed order.
__device__ __forceinline__ double block_sum256(double v, double* red) {
    red[threadIdx.x] = v;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    double r = red[0];
    __syncthreads();
    return r;
}

// out[p] = scale * sum of rows[off[p] .. off[p+1]) (one block per pair)
__global__ void __launch_bounds__(256) k_pop_segsum(const double* __restrict__ rows, const int64_t* __restrict__ off,
                                                    const uint8_t* __restrict__ active, double scale,
                                                    double* __restrict__ out) {
    __shared__ double red[256];
    const int p = blockIdx.x;
    if (!active[p]) return;
    double s = 0.0;
    for (int64_t i = off[p] + threadIdx.x; i < off[p + 1]; i += 256) s += rows[i];
    s = block_sum256(s, red);
    if (threadIdx.x == 0) out[p] = scale * s;
}

// out[p] = sum_t sum_{i in pair p} a[t,i,:].b[t,i,:] in two stages: workgroup
// (p, c) sums chunk c (kSegChunk consecutive elements of the pair's
// t-major element list) into part[p * nc + c], then one thread per pair adds
// its chunks in order (deterministic; every pair's chunks run in parallel).
constexpr int64_t kSegChunk = 16384;

__global__ void __launch_bounds__(256) k_pop_segdot_chunks(const double* __restrict__ a, const double* __restrict__ b,
                                                           int T, int64_t V, const int64_t* __restrict__ off,
                                                           const uint8_t* __restrict__ active, int nc,
                                                           double* __restrict__ part) {
    __shared__ double red[256];
    const int p = blockIdx.x;
    const int c = blockIdx.y;
    if (!active[p]) return;
    const int64_t e0 = off[p] * 3, len = off[p + 1] * 3 - e0;
    const int64_t k0 = int64_t(c) * kSegChunk, k1 = min(k0 + kSegChunk, len * T);
    if (k0 >= k1) return;                                  // beyond this pair's elements
    double s = 0.0;
    for (int64_t k = k0 + threadIdx.x; k < k1; k += 256) {
        const int64_t t = k / len;
        const size_t o = size_t(t) * size_t(V) * 3 + size_t(e0 + (k - t * len));
        s = fma(a[o], b[o], s);
    }
    s = block_sum256(s, red);
    if (threadIdx.x == 0) part[size_t(p) * nc + c] = s;
}

__global__ void __launch_bounds__(256) k_pop_segdot_fold(const double* __restrict__ part, int P, int T,
                                                         const int64_t* __restrict__ off,
                                                         const uint8_t* __restrict__ active, int nc,
                                                         double* __restrict__ out) {
    const int p = blockIdx.x * 256 + threadIdx.x;
    if (p >= P || !active[p]) return;
    const int64_t n = (off[p + 1] - off[p]) * 3 * T;
    const int64_t used = (n + kSegChunk - 1) / kSegChunk;
    double s = 0.0;
    for (int64_t c = 0; c < used; ++c) s += part[size_t(p) * nc + c];
    out[p] = s;
}

// y[t,i,:] += alpha[pair(i)] x[t,i,:]
__global__ void k_pop_segaxpy(double* __restrict__ y, const double* __restrict__ alpha, const double* __restrict__ x,
                              int T, int64_t V, const int32_t* __restrict__ rowpair,
                              const uint8_t* __restrict__ active) {
    size_t e = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    size_t tot = size_t(T) * size_t(V) * 3;
    if (e >= tot) return;
    int64_t i = int64_t((e / 3) % size_t(V));
    int p = rowpair[i];
    if (!active[p]) return;
    y[e] += alpha[p] * x[e];
}

// Bounding-box midpoint per pair of rows [off[p], off[p+1]) of x[*,3] (strided sample of <= 2048).
__global__ void __launch_bounds__(256) k_pop_center(const double* __restrict__ x, const int64_t* __restrict__ off,
                                                    const uint8_t* __restrict__ active, double* __restrict__ c0s) {
    __shared__ double lo[3][256], hi[3][256];
    const int p = blockIdx.x;
    if (!active[p]) return;
    const int64_t r0 = off[p], n = off[p + 1] - off[p];
    const int64_t stp = n > 2048 ? (n + 2047) / 2048 : 1;
    double l[3] = {1e300, 1e300, 1e300}, h[3] = {-1e300, -1e300, -1e300};
    for (int64_t k = int64_t(threadIdx.x) * stp; k < n; k += 256 * stp)
        for (int d = 0; d < 3; ++d) {
            double v = x[size_t(r0 + k) * 3 + d];
            l[d] = fmin(l[d], v);
            h[d] = fmax(h[d], v);
        }
    for (int d = 0; d < 3; ++d) { lo[d][threadIdx.x] = l[d]; hi[d][threadIdx.x] = h[d]; }
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s)
            for (int d = 0; d < 3; ++d) {
                lo[d][threadIdx.x] = fmin(lo[d][threadIdx.x], lo[d][threadIdx.x + s]);
                hi[d][threadIdx.x] = fmax(hi[d][threadIdx.x], hi[d][threadIdx.x + s]);
            }
        __syncthreads();
    }
    if (threadIdx.x < 4) {
        double c = 0.0;
        int d = threadIdx.x;
        if (d < 3 && lo[d][0] <= hi[d][0]) c = 0.5 * (lo[d][0] + hi[d][0]);
        if (!isfinite(c)) c = 0.0;
        c0s[4 * p + d] = c;
    }
}

// Same over the face records of both shapes of pair p.
__global__ void __launch_bounds__(256) k_pop_center_faces(const FaceRec* __restrict__ r1,
                                                          const int64_t* __restrict__ foff,
                                                          const FaceRec* __restrict__ r2,
                                                          const int64_t* __restrict__ tfoff,
                                                          const uint8_t* __restrict__ active,
                                                          double* __restrict__ c0s) {
    __shared__ double lo[3][256], hi[3][256];
    const int p = blockIdx.x;
    if (!active[p]) return;
    const int64_t n1 = foff[p + 1] - foff[p], n2 = tfoff[p + 1] - tfoff[p], n = n1 + n2;
    const int64_t stp = n > 2048 ? (n + 2047) / 2048 : 1;
    double l[3] = {1e300, 1e300, 1e300}, h[3] = {-1e300, -1e300, -1e300};
    for (int64_t k = int64_t(threadIdx.x) * stp; k < n; k += 256 * stp) {
        const FaceRec& f = k < n1 ? r1[foff[p] + k] : r2[tfoff[p] + (k - n1)];
        double c[3] = {f.c0, f.c1, f.c2};
        for (int d = 0; d < 3; ++d) { l[d] = fmin(l[d], c[d]); h[d] = fmax(h[d], c[d]); }
    }
    for (int d = 0; d < 3; ++d) { lo[d][threadIdx.x] = l[d]; hi[d][threadIdx.x] = h[d]; }
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s)
            for (int d = 0; d < 3; ++d) {
                lo[d][threadIdx.x] = fmin(lo[d][threadIdx.x], lo[d][threadIdx.x + s]);
                hi[d][threadIdx.x] = fmax(hi[d][threadIdx.x], hi[d][threadIdx.x + s]);
            }
        __syncthreads();
    }
    if (threadIdx.x < 4) {
        double c = 0.0;
        int d = threadIdx.x;
        if (d < 3 && lo[d][0] <= hi[d][0]) c = 0.5 * (lo[d][0] + hi[d][0]);
        if (!isfinite(c)) c = 0.0;
        c0s[4 * p + d] = c;
    }
}

// Varifold energy rows (energy mode, target faces weighted -2 in the op): erow_i = aw_i * sum
template <typename T>
__global__ void k_pop_vf_erows(const T* __restrict__ phi, const T* __restrict__ plo, const PopPair* __restrict__ pairs,
                               const int32_t* __restrict__ fpair, const uint8_t* __restrict__ active, int nf,
                               const FaceRec* __restrict__ r1, double* __restrict__ erow) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nf) return;
    const int p = fpair[i];
    if (!active[p]) return;
    const PopPair pp = pairs[p];
    erow[i] = r1[i].aw * pop_sum(phi, plo, pp, 1, i, 0, 0, pp.nchunk);
}

// Varifold gradient epilogue per template face (surface form: zf = 1, sd = 1, dzf = 2/3).
template <typename T>
__global__ void k_pop_vf_grad_epi(const T* __restrict__ phi, const T* __restrict__ plo,
                                  const PopPair* __restrict__ pairs, const int32_t* __restrict__ fpair,
                                  const uint8_t* __restrict__ active, int nf, const FaceRec* __restrict__ r1,
                                  double fD, double* __restrict__ z1, double* __restrict__ dz1,
                                  double* __restrict__ erow) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nf) return;
    const int p = fpair[i];
    if (!active[p]) return;
    const PopPair pp = pairs[p];
    double S[3], D[3];
    for (int c = 0; c < 3; ++c) {
        S[c] = pop_sum(phi, plo, pp, 8, i, c, 0, pp.nchunk);
        D[c] = pop_sum(phi, plo, pp, 8, i, 5 + c, 0, pp.nchunk);
    }
    double A1 = pop_sum(phi, plo, pp, 8, i, 3, 0, pp.c1), A2 = pop_sum(phi, plo, pp, 8, i, 3, pp.c1, pp.nchunk);
    // channel 4 holds W = A + B
    double W1 = pop_sum(phi, plo, pp, 8, i, 4, 0, pp.c1), W2 = pop_sum(phi, plo, pp, 8, i, 4, pp.c1, pp.nchunk);
    const FaceRec f = r1[i];
    double AmB = 2.0 * (A1 + A2) - (W1 + W2);
    double n[3] = {f.n0, f.n1, f.n2};
    for (int c = 0; c < 3; ++c) {
        z1[3 * size_t(i) + c] = f.fw * (2.0 * S[c] + AmB * n[c]);
        dz1[3 * size_t(i) + c] = (2.0 / 3.0) * fD * f.aw * D[c];
    }
    erow[i] = f.aw * (W1 + 2.0 * W2);
}

// ------------------------------------------------------------------- host

namespace {

enum PopFamily { kFamVert = 0, kFamFace = 1 };
enum PopOp { kOpK = 0, kOpDiffKT = 1, kOpVfE = 2, kOpVfG = 3, kOpAdj = 4 };

struct PopPlan {
    std::vector<PopItem> items;
    std::vector<int32_t> bstart;
    std::vector<PopPair> pairs;
    int64_t rowchunks = 0;
    uint64_t built_for = ~0ull;          // active-set generation the tables hold
    PopItem* d_items = nullptr;
    int32_t* d_bstart = nullptr;
    PopPair* d_pairs = nullptr;
    size_t cap_items = 0;
};

struct TuneChoice { int R = 0, f10 = 0; };

struct GraphRec { hipGraphExec_t exec = nullptr; };

}  // namespace
}  // namespace lddmm

struct lddmm_population {
    int device = -1;
    int P = 0;
    int64_t V = 0, F = 0, TF = 0;
    std::vector<int64_t> voff, foff, tfoff;
    const int32_t* faces = nullptr;
    const double* vweights = nullptr;
    const int32_t* adj_ptr = nullptr;
    const int32_t* adj_idx = nullptr;
    const double* tvert = nullptr;
    const int32_t* tfaces = nullptr;
    const double* tvw = nullptr;
    // device tables owned by the handle
    int64_t *d_voff = nullptr, *d_foff = nullptr, *d_tfoff = nullptr;
    int32_t *d_vpair = nullptr, *d_fpair = nullptr;
    uint8_t* d_active = nullptr;
    lddmm::FaceRec* d_trec = nullptr;     // target face records (fixed geometry)
    bool trec_ready = false;
    std::vector<uint8_t> active;
    uint64_t active_gen = 1;
    bool active_dirty = true;
    // plans keyed by (family, R, f10, G)
    std::map<std::vector<int64_t>, lddmm::PopPlan> plans;
    std::map<std::vector<int64_t>, lddmm::TuneChoice> tune;
    // own scratch: fixed region then partials
    char* scratch = nullptr;
    size_t scratch_bytes = 0;
    bool graphs_on = true;
    hipStream_t cap_stream = nullptr;
    double* d_segpart = nullptr;          // segment dot chunk sums (grow-only)
    size_t segpart_cap = 0;
    std::map<std::vector<uint64_t>, lddmm::GraphRec> graphs;
};

namespace lddmm {
namespace {

constexpr int kTuneR[2] = {1, 2};
constexpr int kTuneF[3] = {10, 20, 40};

int pop_sync(hipStream_t st) {
    LDDMM_HIP(hipStreamSynchronize(st));
    return LDDMM_OK;
}

template <typename X>
int dev_alloc(X** p, size_t count) {
    *p = nullptr;
    if (count == 0) count = 1;
    LDDMM_HIP(hipMalloc(reinterpret_cast<void**>(p), count * sizeof(X)));
    return LDDMM_OK;
}

template <typename X>
int dev_upload(X* d, const X* h, size_t count) {
    if (count == 0) return LDDMM_OK;
    LDDMM_HIP(hipMemcpy(d, h, count * sizeof(X), hipMemcpyHostToDevice));
    return LDDMM_OK;
}

void drop_graphs(lddmm_population* pop) {
    for (auto& kv : pop->graphs)
        if (kv.second.exec) (void)hipGraphExecDestroy(kv.second.exec);
    pop->graphs.clear();
}

// Sizes a plan is built from (host only).
struct PopShape {
    int P;
    int64_t F;                               // template faces (first target face index)
    const int64_t* voff;
    const int64_t* foff;
    const int64_t* tfoff;
    const uint8_t* active;
};

// Host LPT schedule of the segment table for one family / launch shape.
void schedule_plan(const PopShape& sh, PopPlan& pl, int fam, int R, int f10, int G) {
    const int P = sh.P;
    const int64_t* roffs = fam == kFamVert ? sh.voff : sh.foff;
    double W = 0.0;
    for (int p = 0; p < P; ++p) {
        if (!sh.active[p]) continue;
        double n = double(roffs[p + 1] - roffs[p]);
        double m = fam == kFamVert ? n : n + double(sh.tfoff[p + 1] - sh.tfoff[p]);
        W += n * m;
    }
    const double tile_rows = double(kBlock) * R;
    double wstar = std::max(W * 10.0 / (double(G) * f10), tile_rows * kTile);
    pl.items.clear();
    pl.pairs.assign(size_t(P), PopPair{0, 0, 0, 0, 0});
    pl.rowchunks = 0;
    for (int p = 0; p < P; ++p) {
        const int64_t r0 = roffs[p], n = roffs[p + 1] - roffs[p];
        PopPair& pp = pl.pairs[size_t(p)];
        pp.roff = int32_t(r0);
        pp.n = int32_t(n);
        if (!sh.active[p] || n == 0) continue;
        const int64_t tiles = (n + int64_t(tile_rows) - 1) / int64_t(tile_rows);
        int64_t parts[2][2];
        int nparts;
        if (fam == kFamVert) {
            parts[0][0] = r0; parts[0][1] = n; nparts = 1;
        } else {
            parts[0][0] = r0; parts[0][1] = n;
            parts[1][0] = sh.F + sh.tfoff[p]; parts[1][1] = sh.tfoff[p + 1] - sh.tfoff[p];
            nparts = 2;
        }
        int slot = 0;
        int c1 = 0;
        for (int q = 0; q < nparts; ++q) {
            const int64_t s0 = parts[q][0], len = parts[q][1];
            if (len > 0) {
                int64_t maxc = (len + kTile - 1) / kTile;
                int64_t chunks = int64_t(std::ceil(tile_rows * double(len) / wstar));
                chunks = std::min(std::max<int64_t>(chunks, 1), maxc);
                int64_t chunk = (len + chunks - 1) / chunks;
                chunk = (chunk + kTile - 1) / kTile * kTile;
                chunks = (len + chunk - 1) / chunk;
                for (int64_t k = 0; k < chunks; ++k)
                    for (int64_t t = 0; t < tiles; ++t) {
                        PopItem it;
                        it.r0 = int32_t(r0 + t * int64_t(tile_rows));
                        it.r1 = int32_t(std::min(r0 + n, r0 + (t + 1) * int64_t(tile_rows)));
                        it.s0 = int32_t(s0 + k * chunk);
                        it.s1 = int32_t(std::min(s0 + len, s0 + (k + 1) * chunk));
                        it.pair = p;
                        it.slot = slot + int32_t(k);
                        pl.items.push_back(it);
                    }
                slot += int32_t(chunks);
            }
            if (q == 0) c1 = slot;
        }
        pp.nchunk = slot;
        pp.c1 = fam == kFamVert ? slot : c1;
        pp.pbase = pl.rowchunks;
        pl.rowchunks += int64_t(slot) * n;
    }
    // longest processing time first over G workgroups
    const size_t ni = pl.items.size();
    std::vector<int64_t> cost(ni);
    for (size_t k = 0; k < ni; ++k)
        cost[k] = int64_t(R) * ((pl.items[k].s1 - pl.items[k].s0 + kTile - 1) / kTile) + 1;
    std::vector<size_t> order(ni);
    for (size_t k = 0; k < ni; ++k) order[k] = k;
    std::stable_sort(order.begin(), order.end(), [&](size_t a, size_t b) { return cost[a] > cost[b]; });
    using Slot = std::pair<int64_t, int>;
    std::priority_queue<Slot, std::vector<Slot>, std::greater<Slot>> heap;
    for (int g = 0; g < G; ++g) heap.push({0, g});
    std::vector<std::vector<size_t>> lists(static_cast<size_t>(G));
    for (size_t k : order) {
        Slot s = heap.top();
        heap.pop();
        lists[size_t(s.second)].push_back(k);
        heap.push({s.first + cost[k], s.second});
    }
    std::vector<PopItem> sorted;
    sorted.reserve(ni);
    pl.bstart.assign(size_t(G) + 1, 0);
    for (int g = 0; g < G; ++g) {
        pl.bstart[size_t(g)] = int32_t(sorted.size());
        for (size_t k : lists[size_t(g)]) sorted.push_back(pl.items[k]);
    }
    pl.bstart[size_t(G)] = int32_t(sorted.size());
    pl.items.swap(sorted);
}

// Checks every item against the buffers the kernels index (host only).
int validate_plan(const PopShape& sh, const PopPlan& pl, int fam, int R, int G) {
    const int64_t* roffs = fam == kFamVert ? sh.voff : sh.foff;
    const int64_t TF = sh.tfoff[sh.P];
    LDDMM_REQUIRE(pl.bstart.size() == size_t(G) + 1 && pl.bstart[0] == 0 &&
                      pl.bstart[size_t(G)] == int32_t(pl.items.size()),
                  "plan: bad workgroup table");
    for (int g = 0; g < G; ++g) LDDMM_REQUIRE(pl.bstart[size_t(g)] <= pl.bstart[size_t(g) + 1], "plan: bstart order");
    std::vector<std::vector<uint8_t>> seen(size_t(sh.P));
    for (int p = 0; p < sh.P; ++p) {
        const PopPair& pp = pl.pairs[size_t(p)];
        LDDMM_REQUIRE(pp.roff == roffs[p] && pp.n == roffs[p + 1] - roffs[p], "plan: pair %d rows", p);
        LDDMM_REQUIRE(pp.c1 >= 0 && pp.c1 <= pp.nchunk, "plan: pair %d c1", p);
        LDDMM_REQUIRE(pp.pbase >= 0 && pp.pbase + int64_t(pp.nchunk) * pp.n <= pl.rowchunks, "plan: pair %d pbase", p);
        seen[size_t(p)].assign(size_t(pp.nchunk) * size_t((pp.n + kBlock * R - 1) / (kBlock * R)), 0);
    }
    for (const PopItem& it : pl.items) {
        LDDMM_REQUIRE(it.pair >= 0 && it.pair < sh.P && sh.active[it.pair], "plan: item pair");
        const PopPair& pp = pl.pairs[size_t(it.pair)];
        LDDMM_REQUIRE(it.r0 >= pp.roff && it.r1 > it.r0 && it.r1 <= pp.roff + pp.n && it.r1 - it.r0 <= kBlock * R,
                      "plan: item rows");
        LDDMM_REQUIRE(it.slot >= 0 && it.slot < pp.nchunk, "plan: item slot");
        const int p = it.pair;
        const bool first = it.slot < pp.c1;
        int64_t lo, hi;
        if (fam == kFamVert || first) { lo = roffs[p]; hi = roffs[p + 1]; }
        else { lo = sh.F + sh.tfoff[p]; hi = sh.F + sh.tfoff[p + 1]; }
        LDDMM_REQUIRE(it.s0 >= lo && it.s1 > it.s0 && it.s1 <= hi, "plan: item sources");
        LDDMM_REQUIRE(fam == kFamVert || hi <= sh.F + TF, "plan: target index");
        size_t tile = size_t((it.r0 - pp.roff) / (kBlock * R));
        size_t k = size_t(it.slot) * size_t((pp.n + kBlock * R - 1) / (kBlock * R)) + tile;
        LDDMM_REQUIRE(k < seen[size_t(p)].size() && !seen[size_t(p)][k], "plan: duplicate item");
        seen[size_t(p)][k] = 1;
    }
    for (int p = 0; p < sh.P; ++p)
        for (uint8_t v : seen[size_t(p)]) LDDMM_REQUIRE(v, "plan: pair %d has an uncovered (tile, chunk)", p);
    return LDDMM_OK;
}

int build_plan(lddmm_population* pop, PopPlan& pl, int fam, int R, int f10, int G) {
    const int P = pop->P;
    PopShape sh{P, pop->F, pop->voff.data(), pop->foff.data(), pop->tfoff.data(), pop->active.data()};
    schedule_plan(sh, pl, fam, R, f10, G);
    LDDMM_TRY(validate_plan(sh, pl, fam, R, G));
    const size_t ni = pl.items.size();
    // device tables (grow-only; a reallocation changes the pointers, which
    // are part of every graph key)
    if (!pl.d_bstart) LDDMM_TRY(dev_alloc(&pl.d_bstart, size_t(G) + 1));
    if (!pl.d_pairs) LDDMM_TRY(dev_alloc(&pl.d_pairs, size_t(P)));
    if (ni > pl.cap_items) {
        if (pl.d_items) LDDMM_HIP(hipFree(pl.d_items));
        size_t cap = std::max(ni + ni / 2, size_t(1024));
        LDDMM_TRY(dev_alloc(&pl.d_items, cap));
        pl.cap_items = cap;
    }
    LDDMM_TRY(dev_upload(pl.d_items, pl.items.data(), ni));
    LDDMM_TRY(dev_upload(pl.d_bstart, pl.bstart.data(), size_t(G) + 1));
    LDDMM_TRY(dev_upload(pl.d_pairs, pl.pairs.data(), size_t(P)));
    pl.built_for = pop->active_gen;
    return LDDMM_OK;
}

int get_plan(lddmm_population* pop, int fam, int R, int f10, int G, hipStream_t st, PopPlan** out) {
    std::vector<int64_t> key = {fam, R, f10, G};
    PopPlan& pl = pop->plans[key];
    if (pl.built_for != pop->active_gen) {
        LDDMM_TRY(pop_sync(st));          // tables may be in use by queued work
        LDDMM_TRY(build_plan(pop, pl, fam, R, f10, G));
    }
    *out = &pl;
    return LDDMM_OK;
}

int upload_active(lddmm_population* pop, hipStream_t st) {
    if (!pop->active_dirty) return LDDMM_OK;
    LDDMM_TRY(pop_sync(st));
    LDDMM_TRY(dev_upload(pop->d_active, pop->active.data(), size_t(pop->P)));
    pop->active_dirty = false;
    return LDDMM_OK;
}

// Scratch: [fixed bytes][partials]; grows (and drops the graphs) when needed.
int pop_scratch(lddmm_population* pop, size_t bytes, hipStream_t st) {
    if (bytes <= pop->scratch_bytes) return LDDMM_OK;
    LDDMM_TRY(pop_sync(st));
    drop_graphs(pop);
    if (pop->scratch) LDDMM_HIP(hipFree(pop->scratch));
    pop->scratch = nullptr;
    size_t nb = bytes + bytes / 4 + (1 << 20);
    LDDMM_HIP(hipMalloc(reinterpret_cast<void**>(&pop->scratch), nb));
    pop->scratch_bytes = nb;
    return LDDMM_OK;
}

// Workgroups of the persistent launch.
template <typename T, class OP, int R>
int pop_grid() {
    DeviceInfo di;
    LDDMM_TRY(device_info(&di));
    int occ = blocks_per_cu(reinterpret_cast<const void*>(&k_pair_pop<T, OP, R>), kBlock, 0);
    return std::max(1, di.num_cu * std::max(occ, 1));
}

// A prepared launch of one op: plan + shape.
template <typename T>
struct PopLaunch {
    PopPlan* plan = nullptr;
    int R = 1, G = 1;
};

// `active` (device, one flag per pair) restricts the work to the active pairs;
// null runs every item (autotuning times the full population).
template <typename T, class OP>
int launch_pop(const typename OP::template Args<T>& args, const PopLaunch<T>& L, const double* c0s, T* phi, T* plo,
               const uint8_t* active, hipStream_t st) {
    if (L.R == 2)
        hipLaunchKernelGGL((k_pair_pop<T, OP, 2>), dim3(unsigned(L.G)), dim3(kBlock), 0, st, args, L.plan->d_items,
                           L.plan->d_bstart, L.plan->d_pairs, c0s, phi, plo, active);
    else
        hipLaunchKernelGGL((k_pair_pop<T, OP, 1>), dim3(unsigned(L.G)), dim3(kBlock), 0, st, args, L.plan->d_items,
                           L.plan->d_bstart, L.plan->d_pairs, c0s, phi, plo, active);
    LDDMM_CHECK_LAUNCH();
    return LDDMM_OK;
}

template <typename T, class OP>
int grid_for(int R, int* G) {
    *G = R == 2 ? pop_grid<T, OP, 2>() : pop_grid<T, OP, 1>();
    return *G > 0 ? LDDMM_OK : LDDMM_EHIP;
}

// Tuning key: (op, precision, multi-scale).
template <typename T>
std::vector<int64_t> tune_key(int op, bool ms) {
    return {op, int64_t(sizeof(T)), ms ? 1 : 0};
}

// Tuning choices shared by the populations of one process: a population with
// the same device and segment tables (FNV-1a hash of the vertex, face and
// target-face offsets) reuses the choice instead of timing the candidates
// again.  The choice only picks a launch shape; every shape gives the same
// sums' values up to the chunking of the float64 row sums.
std::mutex g_tune_mu;
std::map<std::vector<int64_t>, TuneChoice> g_tune_shared;

std::vector<int64_t> shared_tune_key(const lddmm_population* pop, const std::vector<int64_t>& key) {
    uint64_t h = 1469598103934665603ull;
    auto mix = [&h](int64_t v) {
        for (int b = 0; b < 8; ++b) {
            h ^= (uint64_t(v) >> (8 * b)) & 0xffu;
            h *= 1099511628211ull;
        }
    };
    for (const auto* tab : {&pop->voff, &pop->foff, &pop->tfoff}) {
        mix(int64_t(tab->size()));
        for (int64_t v : *tab) mix(v);
    }
    std::vector<int64_t> k = key;
    k.insert(k.end(), {int64_t(pop->device), int64_t(pop->P), pop->V, pop->F, pop->TF, int64_t(h)});
    return k;
}

// Partial buffers sit after the fixed region: phi at its start, plo at the
// middle of what remains, so both stay put while the plans change (a graph
// captured for one active set replays for any other that fits).
template <typename T>
struct PopPartials {
    T* phi;
    T* plo;
    int64_t cap;       // elements available to each of phi / plo
};

template <typename T>
PopPartials<T> pop_partials(const lddmm_population* pop, size_t fixed) {
    PopPartials<T> r;
    size_t avail = pop->scratch_bytes > fixed ? (pop->scratch_bytes - fixed) / sizeof(T) : 0;
    r.cap = int64_t(avail / 2) & ~int64_t(63);
    r.phi = reinterpret_cast<T*>(pop->scratch + fixed);
    r.plo = r.phi + r.cap;
    return r;
}

// Grows the scratch so that `elems` partial elements fit in each half.
template <typename T>
int pop_fit(lddmm_population* pop, size_t fixed, int64_t elems, hipStream_t st) {
    if (pop_partials<T>(pop, fixed).cap >= elems) return LDDMM_OK;
    return pop_scratch(pop, fixed + 2 * size_t(elems + 64) * sizeof(T) + 1024, st);
}

// Picks (R, items per workgroup) for op by timing every candidate once on the
// real inputs (the partials are overwritten by the real run), then returns
// the prepared launch for the current active set.  make_args() builds the
// op's arguments and prep(stream) enqueues what the launch reads (centres,
// records); both are called after the scratch has reached its final size.
template <typename T, class OP, class MakeArgs, class Prep>
int prepare_op(lddmm_population* pop, int op, int fam, bool ms, MakeArgs&& make_args, size_t fixed, Prep&& prep,
               hipStream_t st, PopLaunch<T>* out) {
    auto key = tune_key<T>(op, ms);
    auto it = pop->tune.find(key);
    const std::vector<int64_t> skey = shared_tune_key(pop, key);
    if (it == pop->tune.end()) {
        std::lock_guard<std::mutex> lk(g_tune_mu);
        auto g = g_tune_shared.find(skey);
        if (g != g_tune_shared.end()) {
            pop->tune[key] = g->second;
            it = pop->tune.find(key);
        }
    }
    if (it == pop->tune.end()) {
        struct Cand { PopLaunch<T> L; int f10; };
        std::vector<Cand> cands;
        int64_t need = 0;
        for (int R : kTuneR)
            for (int f10 : kTuneF) {
                Cand c;
                c.L.R = R;
                c.f10 = f10;
                LDDMM_TRY((grid_for<T, OP>(R, &c.L.G)));
                LDDMM_TRY(get_plan(pop, fam, R, f10, c.L.G, st, &c.L.plan));
                need = std::max(need, c.L.plan->rowchunks * OP::NCH);
                cands.push_back(c);
            }
        LDDMM_TRY(pop_fit<T>(pop, fixed, need, st));
        LDDMM_TRY(prep(st));
        const auto args = make_args();
        const PopPartials<T> pp = pop_partials<T>(pop, fixed);
        const double* c0s = reinterpret_cast<const double*>(pop->scratch);
        hipEvent_t e0, e1;
        LDDMM_HIP(hipEventCreate(&e0));
        LDDMM_HIP(hipEventCreate(&e1));
        float best = 1e30f;
        TuneChoice bc{cands[0].L.R, cands[0].f10};
        int rc = LDDMM_OK;
        for (const Cand& c : cands) {
            // one untimed launch (instruction cache, tables), one timed
            rc = launch_pop<T, OP>(args, c.L, c0s, pp.phi, pp.plo, nullptr, st);
            if (rc != LDDMM_OK) break;
            if (hipEventRecord(e0, st) != hipSuccess) { rc = LDDMM_EHIP; break; }
            rc = launch_pop<T, OP>(args, c.L, c0s, pp.phi, pp.plo, nullptr, st);
            if (rc != LDDMM_OK) break;
            if (hipEventRecord(e1, st) != hipSuccess || hipEventSynchronize(e1) != hipSuccess) {
                rc = LDDMM_EHIP;
                break;
            }
            float t_ms = 0.f;
            (void)hipEventElapsedTime(&t_ms, e0, e1);
            if (t_ms < best) { best = t_ms; bc = TuneChoice{c.L.R, c.f10}; }
        }
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        if (rc != LDDMM_OK) return rc;
        pop->tune[key] = bc;
        it = pop->tune.find(key);
        std::lock_guard<std::mutex> lk(g_tune_mu);
        g_tune_shared[skey] = bc;
    }
    PopLaunch<T> L;
    L.R = it->second.R;
    LDDMM_TRY((grid_for<T, OP>(L.R, &L.G)));
    LDDMM_TRY(get_plan(pop, fam, L.R, it->second.f10, L.G, st, &L.plan));
    LDDMM_TRY(pop_fit<T>(pop, fixed, L.plan->rowchunks * OP::NCH, st));
    *out = L;
    return LDDMM_OK;
}

// Fixed scratch layout (independent of the plans).
struct PopFixed {
    size_t c0 = 0, rowacc = 0, pbuf0 = 0, pbuf1 = 0, dat = 0, r1 = 0, z1 = 0, dz1 = 0, erow = 0, total = 0;
};

PopFixed fixed_layout(const lddmm_population* pop) {
    Carve cv;
    PopFixed f;
    f.c0 = cv.take(size_t(pop->P) * 4 * sizeof(double));
    f.rowacc = cv.take(size_t(pop->V) * sizeof(double));
    f.pbuf0 = cv.take(size_t(pop->V) * 3 * sizeof(double));
    f.pbuf1 = cv.take(size_t(pop->V) * 3 * sizeof(double));
    f.dat = cv.take(size_t(pop->V) * 3 * sizeof(double));
    f.r1 = cv.take(size_t(std::max<int64_t>(pop->F, 1)) * sizeof(FaceRec));
    f.z1 = cv.take(size_t(std::max<int64_t>(pop->F, 1)) * 3 * sizeof(double));
    f.dz1 = cv.take(size_t(std::max<int64_t>(pop->F, 1)) * 3 * sizeof(double));
    f.erow = cv.take(size_t(std::max<int64_t>(pop->F, 1)) * sizeof(double));
    f.total = cv.off + 256;
    return f;
}

unsigned nblk(int64_t n) { return unsigned((n + 255) / 256); }

// Graph cache: enqueue(stream) records the launches of one call; the key
// holds every pointer and scalar the launches use.
template <class Enq>
int run_graphed(lddmm_population* pop, const std::vector<uint64_t>& key, bool use_graph, Enq&& enqueue,
                hipStream_t st) {
    if (!use_graph) return enqueue(st);
    auto it = pop->graphs.find(key);
    if (it == pop->graphs.end()) {
        if (pop->graphs.size() >= 64) {     // bounded cache: start over
            LDDMM_TRY(pop_sync(st));
            drop_graphs(pop);
        }
        if (!pop->cap_stream) LDDMM_HIP(hipStreamCreateWithFlags(&pop->cap_stream, hipStreamNonBlocking));
        LDDMM_HIP(hipStreamBeginCapture(pop->cap_stream, hipStreamCaptureModeThreadLocal));
        int rc = enqueue(pop->cap_stream);
        hipGraph_t g = nullptr;
        hipError_t ee = hipStreamEndCapture(pop->cap_stream, &g);
        if (rc != LDDMM_OK) {
            if (g) (void)hipGraphDestroy(g);
            return rc;
        }
        if (ee != hipSuccess) return hip_fail(ee, "hipStreamEndCapture", __FILE__, __LINE__);
        GraphRec gr;
        hipError_t ei = hipGraphInstantiate(&gr.exec, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (ei != hipSuccess) return hip_fail(ei, "hipGraphInstantiate", __FILE__, __LINE__);
        it = pop->graphs.emplace(key, gr).first;
    }
    LDDMM_HIP(hipGraphLaunch(it->second.exec, st));
    return LDDMM_OK;
}

template <typename X>
void key_add(std::vector<uint64_t>& k, const X& v) {
    uint64_t w[(sizeof(X) + 7) / 8] = {};
    std::memcpy(w, &v, sizeof(X));
    for (uint64_t x : w) k.push_back(x);
}

void key_kernel(std::vector<uint64_t>& k, const lddmm_kernel* kd) {
    key_add(k, kd->kind);
    key_add(k, kd->nscale);
    for (int i = 0; i < kd->nscale; ++i) key_add(k, kd->sigma[i]);
    key_add(k, kd->kp);
    key_add(k, kd->prec);
    key_add(k, kd->order);
}

void key_ptr(std::vector<uint64_t>& k, const void* p) { k.push_back(uint64_t(reinterpret_cast<uintptr_t>(p))); }

template <typename T>
void key_launch(std::vector<uint64_t>& k, const PopLaunch<T>& L) {
    key_ptr(k, L.plan->d_items);
    key_ptr(k, L.plan->d_bstart);
    key_ptr(k, L.plan->d_pairs);
    k.push_back(uint64_t(L.R));
    k.push_back(uint64_t(L.G));
}

// ------------------------------------------------------------------ drivers

template <typename T, bool MS>
int pop_forward_t(const lddmm_kernel* kd, lddmm_population* pop, const double* at, int T_, double* xt, double r,
                  double* energies, hipStream_t st) {
    ScaleSet<T> sc = make_scales<T>(kd, nullptr);
    using OPK = OpApplyK<3, MS>;
    const PopFixed fx = fixed_layout(pop);
    LDDMM_TRY(pop_scratch(pop, fx.total, st));
    const int64_t V = pop->V;
    const size_t stride = size_t(V) * 3;
    auto prep = [&](hipStream_t s) -> int {
        hipLaunchKernelGGL(k_pop_center, dim3(unsigned(pop->P)), dim3(256), 0, s, xt, pop->d_voff, pop->d_active,
                           reinterpret_cast<double*>(pop->scratch + fx.c0));
        LDDMM_CHECK_LAUNCH();
        return LDDMM_OK;
    };
    auto mk = [&]() { return typename OPK::template Args<T>{xt, xt, at, nullptr, 3, 3, 0, sc}; };
    const bool tuned = pop->tune.count(tune_key<T>(kOpK, MS)) != 0;
    PopLaunch<T> L;
    LDDMM_TRY((prepare_op<T, OPK>(pop, kOpK, kFamVert, MS, mk, fx.total, prep, st, &L)));
    char* S = pop->scratch;
    const PopPartials<T> pp = pop_partials<T>(pop, fx.total);
    double* c0 = reinterpret_cast<double*>(S + fx.c0);
    double* rowacc = reinterpret_cast<double*>(S + fx.rowacc);
    const double dt = 1.0 / T_;
    auto enqueue = [&](hipStream_t s) -> int {
        hipLaunchKernelGGL(k_pop_zero, dim3(nblk(V)), dim3(256), 0, s, int(V), pop->d_vpair, pop->d_active, rowacc);
        LDDMM_CHECK_LAUNCH();
        // one centre per pair (from x_0) for the whole flow
        hipLaunchKernelGGL(k_pop_center, dim3(unsigned(pop->P)), dim3(256), 0, s, xt, pop->d_voff, pop->d_active, c0);
        LDDMM_CHECK_LAUNCH();
        for (int t = 0; t < T_; ++t) {
            const double* x = xt + size_t(t) * stride;
            const double* a = at + size_t(t) * stride;
            typename OPK::template Args<T> args{x, x, a, nullptr, 3, 3, 0, sc};
            LDDMM_TRY((launch_pop<T, OPK>(args, L, c0, pp.phi, pp.plo, pop->d_active, s)));
            hipLaunchKernelGGL((k_pop_shoot_epi<T>), dim3(nblk(V)), dim3(256), 0, s, pp.phi, pp.plo,
                               L.plan->d_pairs, pop->d_vpair, pop->d_active, int(V), x, a, dt,
                               xt + size_t(t + 1) * stride, rowacc, r * dt);
            LDDMM_CHECK_LAUNCH();
        }
        if (energies) {
            hipLaunchKernelGGL(k_pop_segsum, dim3(unsigned(pop->P)), dim3(256), 0, s, rowacc, pop->d_voff,
                               pop->d_active, 1.0, energies);
            LDDMM_CHECK_LAUNCH();
        }
        return LDDMM_OK;
    };
    std::vector<uint64_t> key = {1};
    key_kernel(key, kd);
    key_add(key, r);
    key_add(key, T_);
    for (const void* p : {(const void*)at, (const void*)xt, (const void*)energies, (const void*)S,
                          (const void*)pp.plo})
        key_ptr(key, p);
    key_launch(key, L);
    return run_graphed(pop, key, pop->graphs_on && tuned, enqueue, st);
}

template <typename T, bool MS>
int pop_adjoint_t(const lddmm_kernel* kd, lddmm_population* pop, const double* at, const double* xt, int T_,
                  const double* px1, double r, int euclidean, double* grad, hipStream_t st) {
    SingleScaleConsts ss;
    ScaleSet<T> sc = make_scales<T>(kd, &ss);
    const double fdiff = MS ? -1.0 : ss.diff;
    using OPD = OpDiffKT<3, true, MS>;
    const PopFixed fx = fixed_layout(pop);
    LDDMM_TRY(pop_scratch(pop, fx.total, st));
    const int64_t V = pop->V;
    const size_t stride = size_t(V) * 3;
    const double* xl = xt + size_t(T_ - 1) * stride;
    const double* al = at + size_t(T_ - 1) * stride;
    auto prep = [&](hipStream_t s) -> int {
        hipLaunchKernelGGL(k_pop_center, dim3(unsigned(pop->P)), dim3(256), 0, s, xl, pop->d_voff, pop->d_active,
                           reinterpret_cast<double*>(pop->scratch + fx.c0));
        LDDMM_CHECK_LAUNCH();
        return LDDMM_OK;
    };
    // the Euclidean gradient takes the fused step: one pair launch computes the
    // covector sum and K(x)(2 r a - p) (OpAdjoint), one reduce writes both
    using OPA = OpAdjoint<3, MS>;
    const bool tuned = euclidean ? pop->tune.count(tune_key<T>(kOpAdj, MS)) != 0
                                 : pop->tune.count(tune_key<T>(kOpDiffKT, MS)) != 0;
    PopLaunch<T> LD, LA;
    auto mkd = [&]() { return typename OPD::template Args<T>{xl, xl, px1, al, nullptr, 3, 3, 0, 2.0 * r, sc}; };
    if (euclidean) LDDMM_TRY((prepare_op<T, OPA>(pop, kOpAdj, kFamVert, MS, mkd, fx.total, prep, st, &LA)));
    else LDDMM_TRY((prepare_op<T, OPD>(pop, kOpDiffKT, kFamVert, MS, mkd, fx.total, prep, st, &LD)));
    char* S = pop->scratch;
    const PopPartials<T> pp = pop_partials<T>(pop, fx.total);
    double* c0 = reinterpret_cast<double*>(S + fx.c0);
    double* pb[2] = {reinterpret_cast<double*>(S + fx.pbuf0), reinterpret_cast<double*>(S + fx.pbuf1)};
    double* dat = reinterpret_cast<double*>(S + fx.dat);
    const double dt = 1.0 / T_;
    auto enqueue = [&](hipStream_t s) -> int {
        // p_T = px1
        hipLaunchKernelGGL(k_pop_axpby, dim3(nblk(V)), dim3(256), 0, s, int(V), pop->d_vpair, pop->d_active, 1.0,
                           px1, 0.0, nullptr, pb[T_ & 1]);
        LDDMM_CHECK_LAUNCH();
        hipLaunchKernelGGL(k_pop_center, dim3(unsigned(pop->P)), dim3(256), 0, s, xt, pop->d_voff, pop->d_active, c0);
        LDDMM_CHECK_LAUNCH();
        for (int t = T_ - 1; t >= 0; --t) {
            const double* x = xt + size_t(t) * stride;
            const double* a = at + size_t(t) * stride;
            double* pn = pb[(t + 1) & 1];
            double* pc = pb[t & 1];
            typename OPD::template Args<T> args{x, x, pn, a, nullptr, 3, 3, 0, 2.0 * r, sc};
            double* g = grad + size_t(t) * stride;
            if (euclidean) {
                LDDMM_TRY((launch_pop<T, OPA>(args, LA, c0, pp.phi, pp.plo, pop->d_active, s)));
                hipLaunchKernelGGL((k_pop_reduce_adj<T>), dim3(nblk(V)), dim3(256), 0, s, pp.phi, pp.plo,
                                   LA.plan->d_pairs, pop->d_vpair, pop->d_active, int(V), fdiff, dt, pn, pc, g);
                LDDMM_CHECK_LAUNCH();
                continue;
            }
            LDDMM_TRY((launch_pop<T, OPD>(args, LD, c0, pp.phi, pp.plo, pop->d_active, s)));
            // dat = 2 r a_t - p_{t+1}  (before p_t overwrites the other buffer)
            hipLaunchKernelGGL(k_pop_axpby, dim3(nblk(V)), dim3(256), 0, s, int(V), pop->d_vpair, pop->d_active,
                               2.0 * r, a, -1.0, pn, dat);
            LDDMM_CHECK_LAUNCH();
            // p_t = p_{t+1} + dt * fdiff * sum
            hipLaunchKernelGGL((k_pop_reduce3<T>), dim3(nblk(V)), dim3(256), 0, s, pp.phi, pp.plo,
                               LD.plan->d_pairs, pop->d_vpair, pop->d_active, int(V), 3, fdiff, dt, pn, pc);
            LDDMM_CHECK_LAUNCH();
            hipLaunchKernelGGL(k_pop_axpby, dim3(nblk(V)), dim3(256), 0, s, int(V), pop->d_vpair, pop->d_active, dt,
                               dat, 0.0, nullptr, g);
            LDDMM_CHECK_LAUNCH();
        }
        return LDDMM_OK;
    };
    std::vector<uint64_t> key = {2};
    key_kernel(key, kd);
    key_add(key, r);
    key_add(key, T_);
    key_add(key, euclidean);
    for (const void* p : {(const void*)at, (const void*)xt, (const void*)px1, (const void*)grad, (const void*)S,
                          (const void*)pp.plo})
        key_ptr(key, p);
    key_launch(key, euclidean ? LA : LD);
    return run_graphed(pop, key, pop->graphs_on && tuned, enqueue, st);
}

int ensure_target_records(lddmm_population* pop, hipStream_t st) {
    if (pop->trec_ready || pop->TF == 0) return LDDMM_OK;
    hipLaunchKernelGGL(k_face_features, dim3(nblk(pop->TF)), dim3(256), 0, st, pop->tvert, pop->tfaces,
                       int(pop->TF), pop->tvw, pop->d_trec, nullptr, nullptr, nullptr);
    LDDMM_CHECK_LAUNCH();
    pop->trec_ready = true;
    return LDDMM_OK;
}

template <typename T, bool MS>
int pop_varifold_t(const lddmm_kernel* kd, lddmm_population* pop, const double* vertices, double* energies,
                   double* grad, hipStream_t st) {
    SingleScaleConsts ss;
    ScaleSet<T> sc = make_scales<T>(kd, &ss);
    const double fD = MS ? -1.0 : ss.diff;
    const PopFixed fx = fixed_layout(pop);
    LDDMM_TRY(pop_scratch(pop, fx.total, st));
    LDDMM_TRY(ensure_target_records(pop, st));
    const int64_t F = pop->F;
    auto r1p = [&]() { return reinterpret_cast<FaceRec*>(pop->scratch + fx.r1); };
    auto c0p = [&]() { return reinterpret_cast<double*>(pop->scratch + fx.c0); };
    auto prep = [&](hipStream_t s) -> int {
        hipLaunchKernelGGL(k_face_features, dim3(nblk(F)), dim3(256), 0, s, vertices, pop->faces, int(F),
                           pop->vweights, r1p(), nullptr, nullptr, nullptr);
        LDDMM_CHECK_LAUNCH();
        hipLaunchKernelGGL(k_pop_center_faces, dim3(unsigned(pop->P)), dim3(256), 0, s, r1p(), pop->d_foff,
                           pop->d_trec, pop->d_tfoff, pop->d_active, c0p());
        LDDMM_CHECK_LAUNCH();
        return LDDMM_OK;
    };
    if (!grad) {
        using OPE = OpVarifoldE<MS>;
        auto mk = [&]() {
            typename OPE::template Args<T> A;
            A.f1 = r1p(); A.f2 = pop->d_trec; A.n1 = int(F); A.n2 = int(pop->TF); A.w2 = -2.0; A.c1 = 0;
            A.c0 = nullptr; A.sd = 1.0; A.sc = sc;
            return A;
        };
        PopLaunch<T> L;
        LDDMM_TRY((prepare_op<T, OPE>(pop, kOpVfE, kFamFace, MS, mk, fx.total, prep, st, &L)));
        const PopPartials<T> pp = pop_partials<T>(pop, fx.total);
        double* erow = reinterpret_cast<double*>(pop->scratch + fx.erow);
        LDDMM_TRY(prep(st));
        LDDMM_TRY((launch_pop<T, OPE>(mk(), L, c0p(), pp.phi, pp.plo, pop->d_active, st)));
        hipLaunchKernelGGL((k_pop_vf_erows<T>), dim3(nblk(F)), dim3(256), 0, st, pp.phi, pp.plo, L.plan->d_pairs,
                           pop->d_fpair, pop->d_active, int(F), r1p(), erow);
        LDDMM_CHECK_LAUNCH();
        if (energies) {
            hipLaunchKernelGGL(k_pop_segsum, dim3(unsigned(pop->P)), dim3(256), 0, st, erow, pop->d_foff,
                               pop->d_active, 1.0, energies);
            LDDMM_CHECK_LAUNCH();
        }
        return LDDMM_OK;
    }
    using OPG = OpVarifoldG<MS>;
    auto mk = [&]() {
        typename OPG::template Args<T> A;
        A.f1 = r1p(); A.f2 = pop->d_trec; A.n1 = int(F); A.n2 = int(pop->TF); A.w2 = -1.0; A.c1 = 0;
        A.c0 = nullptr; A.sd = 1.0; A.sc = sc;
        return A;
    };
    PopLaunch<T> L;
    LDDMM_TRY((prepare_op<T, OPG>(pop, kOpVfG, kFamFace, MS, mk, fx.total, prep, st, &L)));
    const PopPartials<T> pp = pop_partials<T>(pop, fx.total);
    double* z1 = reinterpret_cast<double*>(pop->scratch + fx.z1);
    double* dz1 = reinterpret_cast<double*>(pop->scratch + fx.dz1);
    double* erow = reinterpret_cast<double*>(pop->scratch + fx.erow);
    LDDMM_TRY(prep(st));
    LDDMM_TRY((launch_pop<T, OPG>(mk(), L, c0p(), pp.phi, pp.plo, pop->d_active, st)));
    hipLaunchKernelGGL((k_pop_vf_grad_epi<T>), dim3(nblk(F)), dim3(256), 0, st, pp.phi, pp.plo, L.plan->d_pairs,
                       pop->d_fpair, pop->d_active, int(F), r1p(), fD, z1, dz1, erow);
    LDDMM_CHECK_LAUNCH();
    if (energies) {
        hipLaunchKernelGGL(k_pop_segsum, dim3(unsigned(pop->P)), dim3(256), 0, st, erow, pop->d_foff, pop->d_active,
                           1.0, energies);
        LDDMM_CHECK_LAUNCH();
    }
    if (pop->V > 0) {
        hipLaunchKernelGGL(k_face_gather, dim3(nblk(pop->V)), dim3(256), 0, st, vertices, pop->faces, pop->adj_ptr,
                           pop->adj_idx, int(pop->V), dz1, z1, grad, pop->d_vpair, pop->d_active);
        LDDMM_CHECK_LAUNCH();
    }
    return LDDMM_OK;
}

int check_pop(const lddmm_population* pop) {
    LDDMM_REQUIRE(pop != nullptr, "null population");
    int dev = -1;
    LDDMM_HIP(hipGetDevice(&dev));
    LDDMM_REQUIRE(dev == pop->device, "population was created on device %d, current device is %d", pop->device, dev);
    return LDDMM_OK;
}

void free_pop(lddmm_population* pop) {
    if (!pop) return;
    drop_graphs(pop);
    for (auto& kv : pop->plans) {
        if (kv.second.d_items) (void)hipFree(kv.second.d_items);
        if (kv.second.d_bstart) (void)hipFree(kv.second.d_bstart);
        if (kv.second.d_pairs) (void)hipFree(kv.second.d_pairs);
    }
    for (void* p : {(void*)pop->d_voff, (void*)pop->d_foff, (void*)pop->d_tfoff, (void*)pop->d_vpair,
                    (void*)pop->d_fpair, (void*)pop->d_active, (void*)pop->d_trec, (void*)pop->scratch,
                    (void*)pop->d_segpart})
        if (p) (void)hipFree(p);
    if (pop->cap_stream) (void)hipStreamDestroy(pop->cap_stream);
    delete pop;
}

}  // namespace
}  // namespace lddmm

using namespace lddmm;

#define POP_DISPATCH(FN, ...)                                                                  \
    do {                                                                                       \
        const bool ms_ = general_path(kd);                                                       \
        if (kd->prec == LDDMM_PREC_F32C)                                                       \
            return ms_ ? FN<float, true>(__VA_ARGS__) : FN<float, false>(__VA_ARGS__);         \
        return ms_ ? FN<double, true>(__VA_ARGS__) : FN<double, false>(__VA_ARGS__);           \
    } while (0)

extern "C" int lddmm_population_create(const lddmm_population_desc* d, lddmm_population** out) {
    LDDMM_REQUIRE(out != nullptr, "population_create: null output");
    *out = nullptr;
    LDDMM_REQUIRE(d != nullptr, "population_create: null descriptor");
    LDDMM_REQUIRE(d->npairs >= 1, "population_create: npairs must be >= 1");
    LDDMM_REQUIRE(d->voff && d->foff && d->tfoff, "population_create: null offsets");
    const int P = d->npairs;
    for (int p = 0; p < P; ++p) {
        LDDMM_REQUIRE(d->voff[p + 1] >= d->voff[p] && d->foff[p + 1] >= d->foff[p] && d->tfoff[p + 1] >= d->tfoff[p],
                      "population_create: offsets must be non-decreasing (pair %d)", p);
    }
    LDDMM_REQUIRE(d->voff[0] == 0 && d->foff[0] == 0 && d->tfoff[0] == 0, "population_create: offsets start at 0");
    const int64_t V = d->voff[P], F = d->foff[P], TF = d->tfoff[P];
    LDDMM_REQUIRE(V <= kMaxPoints && F + TF < (int64_t(1) << 30), "population_create: sizes beyond int32");
    LDDMM_REQUIRE(F == 0 || d->faces, "population_create: null faces");
    LDDMM_REQUIRE(V == 0 || (d->adj_ptr && d->adj_idx), "population_create: adjacency required");
    LDDMM_REQUIRE(TF == 0 || (d->tvertices && d->tfaces), "population_create: null target arrays");
    lddmm_population* pop = new lddmm_population();
    LDDMM_HIP(hipGetDevice(&pop->device));
    pop->P = P;
    pop->V = V;
    pop->F = F;
    pop->TF = TF;
    pop->voff.assign(d->voff, d->voff + P + 1);
    pop->foff.assign(d->foff, d->foff + P + 1);
    pop->tfoff.assign(d->tfoff, d->tfoff + P + 1);
    pop->faces = d->faces;
    pop->vweights = d->vweights;
    pop->adj_ptr = d->adj_ptr;
    pop->adj_idx = d->adj_idx;
    pop->tvert = d->tvertices;
    pop->tfaces = d->tfaces;
    pop->tvw = d->tvweights;
    pop->active.assign(size_t(P), 1);
    std::vector<int32_t> vpair(size_t(std::max<int64_t>(V, 1)), 0), fpair(size_t(std::max<int64_t>(F, 1)), 0);
    for (int p = 0; p < P; ++p) {
        for (int64_t i = d->voff[p]; i < d->voff[p + 1]; ++i) vpair[size_t(i)] = p;
        for (int64_t i = d->foff[p]; i < d->foff[p + 1]; ++i) fpair[size_t(i)] = p;
    }
    int rc = LDDMM_OK;
    auto step = [&](int r) { if (rc == LDDMM_OK) rc = r; };
    step(dev_alloc(&pop->d_voff, size_t(P) + 1));
    step(dev_alloc(&pop->d_foff, size_t(P) + 1));
    step(dev_alloc(&pop->d_tfoff, size_t(P) + 1));
    step(dev_alloc(&pop->d_vpair, vpair.size()));
    step(dev_alloc(&pop->d_fpair, fpair.size()));
    step(dev_alloc(&pop->d_active, size_t(P)));
    step(dev_alloc(&pop->d_trec, size_t(std::max<int64_t>(TF, 1))));
    if (rc == LDDMM_OK) {
        step(dev_upload(pop->d_voff, d->voff, size_t(P) + 1));
        step(dev_upload(pop->d_foff, d->foff, size_t(P) + 1));
        step(dev_upload(pop->d_tfoff, d->tfoff, size_t(P) + 1));
        step(dev_upload(pop->d_vpair, vpair.data(), vpair.size()));
        step(dev_upload(pop->d_fpair, fpair.data(), fpair.size()));
        step(dev_upload(pop->d_active, pop->active.data(), size_t(P)));
    }
    if (rc != LDDMM_OK) {
        free_pop(pop);
        return rc;
    }
    pop->active_dirty = false;
    *out = pop;
    return LDDMM_OK;
}

extern "C" int lddmm_population_destroy(lddmm_population* pop) {
    if (!pop) return LDDMM_OK;
    (void)hipDeviceSynchronize();
    free_pop(pop);
    return LDDMM_OK;
}

extern "C" int lddmm_population_set_active(lddmm_population* pop, const uint8_t* active) {
    LDDMM_TRY(check_pop(pop));
    std::vector<uint8_t> a(size_t(pop->P), 1);
    if (active)
        for (int p = 0; p < pop->P; ++p) a[size_t(p)] = active[p] ? 1 : 0;
    if (a != pop->active) {
        pop->active.swap(a);
        pop->active_gen++;
        pop->active_dirty = true;
    }
    return LDDMM_OK;
}

extern "C" int lddmm_population_set_graphs(lddmm_population* pop, int32_t on) {
    LDDMM_TRY(check_pop(pop));
    pop->graphs_on = on != 0;
    return LDDMM_OK;
}

extern "C" int lddmm_population_shoot_forward(const lddmm_kernel* kd, lddmm_population* pop, const double* at,
                                              int32_t T, double* xt, double regweight, double* energies,
                                              void* stream) {
    LDDMM_TRY(validate_kernel(kd));
    LDDMM_TRY(check_pop(pop));
    LDDMM_REQUIRE(T >= 1, "population_shoot_forward: T must be >= 1");
    LDDMM_REQUIRE(pop->V == 0 || (at && xt), "population_shoot_forward: null arrays");
    hipStream_t st = static_cast<hipStream_t>(stream);
    LDDMM_TRY(upload_active(pop, st));
    if (pop->V == 0) return LDDMM_OK;
    POP_DISPATCH(pop_forward_t, kd, pop, at, T, xt, regweight, energies, st);
}

extern "C" int lddmm_population_shoot_adjoint(const lddmm_kernel* kd, lddmm_population* pop, const double* at,
                                              const double* xt, int32_t T, const double* px1, double regweight,
                                              int32_t euclidean, double* grad, void* stream) {
    LDDMM_TRY(validate_kernel(kd));
    LDDMM_TRY(check_pop(pop));
    LDDMM_REQUIRE(T >= 1, "population_shoot_adjoint: T must be >= 1");
    LDDMM_REQUIRE(pop->V == 0 || (at && xt && px1 && grad), "population_shoot_adjoint: null arrays");
    hipStream_t st = static_cast<hipStream_t>(stream);
    LDDMM_TRY(upload_active(pop, st));
    if (pop->V == 0) return LDDMM_OK;
    POP_DISPATCH(pop_adjoint_t, kd, pop, at, xt, T, px1, regweight, euclidean, grad, st);
}

extern "C" int lddmm_population_varifold(const lddmm_kernel* kd, lddmm_population* pop, const double* vertices,
                                         double* energies, double* grad, void* stream) {
    LDDMM_TRY(validate_kernel(kd, kValDataTerm));
    LDDMM_TRY(check_pop(pop));
    LDDMM_REQUIRE(energies || grad, "population_varifold: nothing requested");
    LDDMM_REQUIRE(pop->V == 0 || vertices, "population_varifold: null vertices");
    hipStream_t st = static_cast<hipStream_t>(stream);
    LDDMM_TRY(upload_active(pop, st));
    if (pop->F == 0) {
        // no template faces: zero energies / gradients of the active pairs
        std::vector<double> z(size_t(pop->P), 0.0);
        if (energies) {
            LDDMM_TRY(pop_sync(st));
            for (int p = 0; p < pop->P; ++p)
                if (pop->active[size_t(p)]) LDDMM_HIP(hipMemcpy(energies + p, &z[0], sizeof(double), hipMemcpyHostToDevice));
        }
        if (grad && pop->V > 0) {
            hipLaunchKernelGGL(k_pop_axpby, dim3(nblk(pop->V)), dim3(256), 0, st, int(pop->V), pop->d_vpair,
                               pop->d_active, 0.0, nullptr, 0.0, nullptr, grad);
            LDDMM_CHECK_LAUNCH();
        }
        return LDDMM_OK;
    }
    POP_DISPATCH(pop_varifold_t, kd, pop, vertices, energies, grad, st);
}

extern "C" int lddmm_population_seg_dot(lddmm_population* pop, const double* a, const double* b, int32_t T,
                                        double* out, void* stream) {
    LDDMM_TRY(check_pop(pop));
    LDDMM_REQUIRE(T >= 1 && out && (pop->V == 0 || (a && b)), "population_seg_dot: bad arguments");
    hipStream_t st = static_cast<hipStream_t>(stream);
    LDDMM_TRY(upload_active(pop, st));
    int64_t maxn = 0;
    for (int q = 0; q < pop->P; ++q) maxn = std::max(maxn, pop->voff[size_t(q) + 1] - pop->voff[size_t(q)]);
    const int nc = int(std::max<int64_t>(1, (maxn * 3 * T + kSegChunk - 1) / kSegChunk));
    const size_t need = size_t(pop->P) * size_t(nc);
    if (need > pop->segpart_cap) {
        LDDMM_TRY(pop_sync(st));
        if (pop->d_segpart) LDDMM_HIP(hipFree(pop->d_segpart));
        pop->d_segpart = nullptr;
        pop->segpart_cap = 0;
        LDDMM_TRY(dev_alloc(&pop->d_segpart, need));
        pop->segpart_cap = need;
    }
    hipLaunchKernelGGL(k_pop_segdot_chunks, dim3(unsigned(pop->P), unsigned(nc)), dim3(256), 0, st, a, b, int(T),
                       pop->V, pop->d_voff, pop->d_active, nc, pop->d_segpart);
    LDDMM_CHECK_LAUNCH();
    hipLaunchKernelGGL(k_pop_segdot_fold, dim3(unsigned((pop->P + 255) / 256)), dim3(256), 0, st, pop->d_segpart,
                       pop->P, int(T), pop->d_voff, pop->d_active, nc, out);
    LDDMM_CHECK_LAUNCH();
    return LDDMM_OK;
}

extern "C" int lddmm_population_seg_axpy(lddmm_population* pop, double* y, const double* alpha, const double* x,
                                         int32_t T, void* stream) {
    LDDMM_TRY(check_pop(pop));
    LDDMM_REQUIRE(T >= 1 && alpha && (pop->V == 0 || (x && y)), "population_seg_axpy: bad arguments");
    hipStream_t st = static_cast<hipStream_t>(stream);
    LDDMM_TRY(upload_active(pop, st));
    size_t tot = size_t(T) * size_t(pop->V) * 3;
    if (tot == 0) return LDDMM_OK;
    hipLaunchKernelGGL(k_pop_segaxpy, dim3(unsigned((tot + 255) / 256)), dim3(256), 0, st, y, alpha, x, int(T),
                       pop->V, pop->d_vpair, pop->d_active);
    LDDMM_CHECK_LAUNCH();
    return LDDMM_OK;
}

extern "C" int lddmm_population_plan_check(int32_t npairs, const int64_t* voff, const int64_t* foff,
                                           const int64_t* tfoff, const uint8_t* active, int32_t family, int32_t R,
                                           int32_t f10, int32_t G, int64_t* nitems) {
    LDDMM_REQUIRE(npairs >= 1 && voff && foff && tfoff && active, "population_plan_check: bad arguments");
    LDDMM_REQUIRE(family == 0 || family == 1, "population_plan_check: family must be 0 or 1");
    LDDMM_REQUIRE((R == 1 || R == 2) && f10 >= 1 && G >= 1, "population_plan_check: bad launch shape");
    PopShape sh{npairs, foff[npairs], voff, foff, tfoff, active};
    PopPlan pl;
    schedule_plan(sh, pl, family, R, f10, G);
    LDDMM_TRY(validate_plan(sh, pl, family, R, G));
    if (nitems) *nitems = int64_t(pl.items.size());
    return LDDMM_OK;
}

extern "C" int lddmm_population_tuning(const lddmm_population* pop, int32_t* out, int32_t cap, int32_t* count) {
    LDDMM_REQUIRE(pop != nullptr && count != nullptr, "population_tuning: null argument");
    int32_t n = 0;
    for (const auto& kv : pop->tune) {
        if (out && n < cap) {
            out[4 * n + 0] = int32_t(kv.first[0]);
            out[4 * n + 1] = int32_t(kv.first[1]);
            out[4 * n + 2] = kv.second.R;
            out[4 * n + 3] = kv.second.f10;
        }
        ++n;
    }
    *count = n;
    return LDDMM_OK;
}
