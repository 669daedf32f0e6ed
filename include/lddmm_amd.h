/*
 * lddmm_amd.h -- C ABI of the MI355X (gfx950) point-set hot path of py-lddmm.
 *
 * Every entry point takes DEVICE pointers (row-major float64 arrays, int32
 * face / adjacency arrays), plain sizes and a hipStream_t passed as void*.
 * Work is enqueued on that stream and the call returns without waiting,
 * except where a scalar result is returned through a host pointer (then the
 * call synchronises the stream).  Outputs are written, never accumulated.
 *
 * Return value: 0 on success, a negative LDDMM_E* code otherwise; the text of
 * the last error on the calling thread is given by lddmm_last_error().
 *
 * Precision (lddmm_kernel.prec):
 *   LDDMM_PREC_F64  -- float64 arithmetic throughout (the reference's numba
 *                      path is float64).
 *   LDDMM_PREC_F32C -- float32 pair arithmetic on centred coordinates, float32
 *                      accumulation compensated (two-sum) across source tiles
 *                      and chunks; results are returned as float64.
 *
 * Scratch memory: the library keeps one grow-only device workspace per GPU.
 * Calls that share a GPU must be issued on one stream (or be serialised by
 * the caller); calls on different GPUs are independent.
 *
 * Each declaration names the reference function it replaces
 * (paths relative to py-lddmm/).
 */
#ifndef LDDMM_AMD_H
#define LDDMM_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define LDDMM_MAX_SCALES 8

#define LDDMM_KERNEL_GAUSS     0
/* Laplacian (Matern) family: K = P_order(r) exp(-r), r = |x - y| / sigma,
 * order 0..4 (base/kernelFunctions_util.py:20-52, :92-95). */
#define LDDMM_KERNEL_LAPLACIAN 1

#define LDDMM_PREC_F64  0
#define LDDMM_PREC_F32C 1

#define LDDMM_OK           0
#define LDDMM_EINVAL      -1   /* bad argument (sizes, modes, null pointers) */
#define LDDMM_EUNSUPPORTED -2  /* kernel family / option not built */
#define LDDMM_EHIP        -3   /* HIP runtime error (launch, allocation) */

/* diffKT modes: base/kernelFunctions_util.py:785-791 */
#define LDDMM_DIFFKT_PLAIN 0
#define LDDMM_DIFFKT_SYM   1
#define LDDMM_DIFFKT_LDDMM 2

/* applyDiffK options: base/kernelFunctions_util.py:662-677 */
#define LDDMM_DIFFK_1      0
#define LDDMM_DIFFK_2      1
#define LDDMM_DIFFK_1AND2  2

/* applyDDiffK options: base/kernelFunctions_util.py:1100-1127 */
#define LDDMM_DDIFFK_11     0
#define LDDMM_DDIFFK_12     1
#define LDDMM_DDIFFK_11AND12 2

/* Affine part of a kernel (KernelSpec.setAffine, base/kernelFunctions.py:393-412):
 * 'affine' adds w1 (y - c)(x - c)^T a + w2 sum(a) to K(y, x) a, 'euclidean'
 * keeps the rotation part of the linear term (an orthonormal basis of
 * antisymmetric matrices) and the translation. */
#define LDDMM_AFFINE_NONE      0
#define LDDMM_AFFINE_FULL      1   /* affine = 'affine' */
#define LDDMM_AFFINE_EUCLIDEAN 2   /* affine = 'euclidean' */

/* Kernel description: mirrors KernelSpec (base/kernelFunctions.py:339-412):
 * name -> kind, sigma -> sigma[0..nscale), ms_exponent -> kp, pk_dtype -> prec,
 * order -> order (Laplacian only; the reference resets orders above 4 to 3,
 * this library rejects them), affine / w1 / w2 / center -> the fields of the
 * same names.  The fused data terms (surface and curve varifold, population
 * varifold) take LDDMM_KERNEL_GAUSS only.  The affine part is applied by the
 * shooting entry points (lddmm_shoot_forward / lddmm_shoot_adjoint, where the
 * reference's Kernel.applyK / applyDiffKT add it); every other entry point
 * restates a kernelFunctions_util function, which has none, and rejects a
 * description with affine != LDDMM_AFFINE_NONE. */
typedef struct lddmm_kernel {
    int32_t kind;                      /* LDDMM_KERNEL_GAUSS / LDDMM_KERNEL_LAPLACIAN */
    int32_t nscale;                    /* 1 .. LDDMM_MAX_SCALES */
    double  sigma[LDDMM_MAX_SCALES];
    double  kp;                        /* multi-scale exponent, reference default -1 */
    int32_t prec;                      /* LDDMM_PREC_F64 / LDDMM_PREC_F32C */
    int32_t order;                     /* Laplacian polynomial order 0..4; ignored for gauss */
    int32_t affine;                    /* LDDMM_AFFINE_* */
    int32_t reserved;
    double  w1, w2;                    /* linear and translation weights of the affine part */
    double  center[3];                 /* origin c of the linear part */
} lddmm_kernel;

const char* lddmm_last_error(void);
const char* lddmm_version(void);
/* Number of visible GPUs (0 when no HIP device is present). */
int lddmm_device_count(void);
/* Releases the calling device's workspace. */
int lddmm_release_workspace(void);

/* Instrumentation (no counterpart in the reference): when enabled, every
 * pair-sum kernel launch is bracketed by HIP events on its own stream.
 * lddmm_profile_read waits for the recorded events, writes the duration in ms
 * of the first `cap` launches (in launch order) to ms_each, the number of
 * launches since the last read to *launches, and clears the record. */
int lddmm_profile_enable(int32_t on);
int lddmm_profile_read(double* ms_each, int64_t cap, int64_t* launches);

/* Accumulation engine of the float32 (LDDMM_PREC_F32C) applyK and plain
 * applyDiffKT sums (no counterpart in the reference): LDDMM_ENGINE_VALU
 * (default) accumulates each pair on the vector ALUs; LDDMM_ENGINE_MFMA
 * computes the Gaussian values on the VALU and accumulates them against the
 * per-source weight columns on the matrix cores (v_mfma_f32_16x16x4f32).
 * Process-wide setting. */
#define LDDMM_ENGINE_VALU 0
#define LDDMM_ENGINE_MFMA 1
int lddmm_set_pair_engine(int32_t engine);

/* Engine of the float32 single-scale Gaussian varifold energy sums (no
 * counterpart in the reference): LDDMM_VARIFOLD_VALU (default) evaluates each
 * face pair on the vector ALUs; LDDMM_VARIFOLD_MFMA_COS takes the 16x16 blocks
 * of normal cosines n_i.n_j from v_mfma_f32_16x16x4f32; LDDMM_VARIFOLD_MFMA_COS_R2
 * also the squared centre distances (relative to an origin per block of 16
 * target faces).  The gradient and the float64 path stay on the VALU.
 * Process-wide setting. */
#define LDDMM_VARIFOLD_VALU 0
#define LDDMM_VARIFOLD_MFMA_COS 1
#define LDDMM_VARIFOLD_MFMA_COS_R2 2
int lddmm_set_varifold_engine(int32_t engine);

/* out[n,k] = sum_s w_s sum_j K_s(y_i, x_j) a[j,:]
 * Replaces ku.applyK / applyK_numba / applyK_pykeops
 * (base/kernelFunctions_util.py:443-466, :499-535), reached from
 * Kernel.applyK (base/kernelFunctions.py:452-475). dim in {2,3}, k >= 1. */
int lddmm_applyK(const lddmm_kernel* kd,
                 const double* y, int64_t n, const double* x, int64_t m, int32_t dim,
                 const double* a, int32_t k, double* out, void* stream);

/* lddmm_applyK without any allocation.  No counterpart in the reference: it
 * is for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  Same
 * arguments and result as lddmm_applyK; the centre and the chunk partial sums
 * live in one file-scope device array of LDDMM_NOALLOC_SCRATCH_BYTES per GPU
 * (held until the process ends) instead of the grow-only workspace, and the
 * call makes no hipMalloc / hipFree / allocating library call (host-side
 * caches of device facts and launch shapes are filled on a device's first
 * call).  Every launch writes all the scratch it reads, so nothing carries
 * over between calls.  Sizes: n * min(k, 4) * 8 bytes of partials per source
 * chunk must fit the array (n up to about 2.8M targets for k = 3); a plan
 * with more chunks than fit is given fewer, larger chunks (the same sums in
 * another summation order); larger n is rejected with LDDMM_EINVAL before
 * any launch.  All calls on one GPU share its array: calls that overlap on
 * one GPU (different streams or threads) race and must be serialised by the
 * caller.  Always uses the VALU engine; not instrumented by
 * lddmm_profile_enable. */
#define LDDMM_NOALLOC_SCRATCH_BYTES (64u << 20)
int lddmm_applyK_noalloc(const lddmm_kernel* kd,
                         const double* y, int64_t n, const double* x, int64_t m, int32_t dim,
                         const double* a, int32_t k, double* out, void* stream);

/* a[n,k] with K(x, x) a = b by conjugate gradients, all on the device.
 * Replaces Kernel.solve (base/kernelFunctions.py:676-680), which runs linearcg
 * (base/conjugateGradient.py:351-408) with Kernel.applyK as the operator.
 * Zero start; a = 0 when |b|^2 < 1e-10; stops when alpha mean|p| < 1e-10, when
 * the energy a.Ka - 2 b.a changes by less than 1e-20 of its previous value, or
 * after iter_max products.  sequence LDDMM_CG_REFERENCE reproduces linearcg as
 * it executes: its first direction is the residual array itself, which the
 * in-place residual update also moves, so the second direction is (1 + beta)
 * times the residual; LDDMM_CG_TEXTBOOK keeps them apart (textbook conjugate
 * gradients).  work: caller-provided device buffer of 4 n k + 8 doubles
 * (residual, direction, product, K a - 2 b, CG scalars).  Stream-ordered; the
 * host reads the stop flag every 16 products to skip the rest after a stop.
 * Affine kernel parts are not applied (affine must be LDDMM_AFFINE_NONE). */
#define LDDMM_CG_REFERENCE 0
#define LDDMM_CG_TEXTBOOK  1
int lddmm_kernel_solve(const lddmm_kernel* kd, const double* x, int64_t n, int32_t dim,
                       const double* b, int32_t k, int32_t iter_max, int32_t sequence, double* a,
                       double* work, void* stream);

/* out[n,dim] = sum_s s^(KP-1)/W sum_j -(y_i-x_j)/s K_s(y_i,x_j) c_ij with
 * c_ij = p_i.a_j (+ a_i.p_j for SYM, + a_i.p_j - 2 regweight a_i.a_j for LDDMM).
 * p is [n,k], a is [m,k]; SYM and LDDMM require n == m.
 * Replaces ku.applyDiffKT / applyDiffKT_numba / _pykeops
 * (base/kernelFunctions_util.py:765-845), reached from Kernel.applyDiffKT
 * (base/kernelFunctions.py:581-600). */
int lddmm_applyDiffKT(const lddmm_kernel* kd,
                      const double* y, int64_t n, const double* x, int64_t m, int32_t dim,
                      const double* p, const double* a, int32_t k,
                      int32_t mode, double regweight, double* out, void* stream);

/* out[n] = sum_s s^(KP-1)/W sum_j -((y_i-x_j)/s) . a_j K_s(y_i,x_j): the divergence
 * at y_i of the velocity field K(., x) a (a is [m,dim]).
 * Replaces ku.applyDiv / applyDiv_numba (base/kernelFunctions_util.py:596-620),
 * reached from Kernel.applyDivergence (base/kernelFunctions.py:638-663). */
int lddmm_applyDiv(const lddmm_kernel* kd, const double* y, int64_t n, const double* x, int64_t m, int32_t dim,
                   const double* a, double* out, void* stream);

/* out[n] = sum_s w_s sum_j K_s(y_i,x_j) (ay_i ax_j + (betay_i . betax_j)^2)
 * Replaces ku.applyktensor (base/kernelFunctions_util.py:1424-1506),
 * reached from Kernel.applyKTensor (base/kernelFunctions.py:509-517). */
int lddmm_applyktensor(const lddmm_kernel* kd,
                       const double* y, int64_t n, const double* x, int64_t m, int32_t dim,
                       const double* ay, const double* ax,
                       const double* betay, const double* betax, int32_t kb,
                       double* out, void* stream);

/* out[n,kb] = sum_s w_s sum_j K_s(y_i,x_j) beta[i,j,:]   (dense beta [n,m,kb])
 * Replaces ku.applykmat (base/kernelFunctions_util.py:1509-1540), reached from
 * Kernel.applyK(..., matrixWeights=True) (base/kernelFunctions.py:456,467). */
int lddmm_applykmat(const lddmm_kernel* kd,
                    const double* y, int64_t n, const double* x, int64_t m, int32_t dim,
                    const double* beta, int32_t kb, double* out, void* stream);

/* out[n,dim] = -sum_s s^(KP-1)/W sum_j K_s (y_i-x_j)/s beta[i,j]  (dense beta [n,m])
 * Replaces ku.applykdiffmat (base/kernelFunctions_util.py:1631-1669), reached
 * from Kernel.applyDiffKmat (base/kernelFunctions.py:565-570). */
int lddmm_applykdiffmat(const lddmm_kernel* kd,
                        const double* y, int64_t n, const double* x, int64_t m, int32_t dim,
                        const double* beta, double* out, void* stream);

/* First-derivative sums on one point set x[n,dim], a1 [n,dim], a2 [n,k]:
 * option 1AND2: out[i] = sum_s s^(KP-1)/W sum_l -K ((x_i-x_l)/s.(a1_i-a1_l)) a2_l
 * (options 1 and 2 as in the reference).  Replaces ku.applyDiffK
 * (base/kernelFunctions_util.py:646-680), reached from Kernel.applyDiffK1and2
 * (base/kernelFunctions.py:549-562) through applykdiff1and2 (:954-955). */
int lddmm_applyDiffK(const lddmm_kernel* kd,
                     const double* x, int64_t n, int32_t dim,
                     const double* a1, const double* a2, int32_t k,
                     int32_t option, double* out, void* stream);

/* Second-derivative sums on one point set: a1 [n,k], a2 [n,k], p [n,dim]:
 * out[i] = sum_s s^(KP-2)/W sum_l ((d.q) d - q) K_s (a1_i.a2_l), d=(x_i-x_l)/s,
 * q = p_i (11), -p_l (12), p_i - p_l (11AND12).  Replaces ku.applyDDiffK
 * (base/kernelFunctions_util.py:1084-1129), reached from
 * Kernel.applyDDiffK11and12 (base/kernelFunctions.py:665-674). */
int lddmm_applyDDiffK(const lddmm_kernel* kd,
                      const double* x, int64_t n, int32_t dim,
                      const double* a1, const double* a2, int32_t k, const double* p,
                      int32_t option, double* out, void* stream);

/* Triangulated surface on the device: vertices [nv,3] f64, faces [nf,3] int32,
 * vertex weights [nv] f64 (NULL = all ones, the reference default,
 * base/surfaces.py:238-240). */
typedef struct lddmm_surface {
    const double*  vertices;
    int64_t        nv;
    const int32_t* faces;
    int64_t        nf;
    const double*  vweights;
    /* vertex -> incident (face*3 + corner) adjacency, CSR, built by
     * lddmm_surface_adjacency; required only for gradients. */
    const int32_t* adj_ptr;            /* [nv+1] */
    const int32_t* adj_idx;            /* [3*nf] */
} lddmm_surface;

/* Host helper: builds the CSR vertex->(face,corner) adjacency of int32 faces
 * (host arrays).  adj_ptr has nv+1 entries, adj_idx 3*nf entries. */
int lddmm_surface_adjacency(const int32_t* faces_host, int64_t nf, int64_t nv,
                            int32_t* adj_ptr_host, int32_t* adj_idx_host);

/* Internal (regularity) costs of a velocity field on a surface
 * (base/surfaceDistances.py:631-872): h1 = normGrad (discrete Dirichlet
 * energy), elastic = elasticNorm, displacement = normDisplacement; a cost
 * list holds up to LDDMM_MAX_ICOSTS (kind, weight) entries, as
 * SurfaceMatching's internalCost option (base/surfaceMatching.py:184-204). */
#define LDDMM_IC_H1           0
#define LDDMM_IC_ELASTIC      1
#define LDDMM_IC_DISPLACEMENT 2
#define LDDMM_MAX_ICOSTS      4
typedef struct lddmm_internal_cost {
    int32_t ncost;                       /* 1 .. LDDMM_MAX_ICOSTS */
    int32_t kinds[LDDMM_MAX_ICOSTS];     /* LDDMM_IC_* */
    double  weights[LDDMM_MAX_ICOSTS];
} lddmm_internal_cost;

/* sum_k w_k C_k(s, phi) for the field phi [nv,3] at s's vertices:
 * energy (device f64[1], may be NULL), grad_phi / grad_x (device [nv,3], may
 * be NULL) receive sum_k w_k dC_k/dphi and dC_k/dx (the reference's
 * diffNormGrad / diffElasticNorm / diffNormDisplacement 'phi' and 'x'
 * entries); gradients need s's adjacency.  Replaces
 * SurfaceMatching.internalCost__ / internalCostGrad__
 * (base/surfaceMatching.py:156-173). */
int lddmm_surface_internal_cost(const lddmm_surface* s, const double* phi, const lddmm_internal_cost* ic,
                                double* energy, double* grad_phi, double* grad_x, void* stream);

/* Varifold data term between surfaces (orientation-free, f0(u) = 1 + u^2):
 *   tgt == NULL : *energy = |S1|^2                       (varifoldNorm0)
 *   tgt != NULL : *energy = |S1|^2 - 2 <S1, tgt>           (varifoldNormDef)
 * energy is a DEVICE pointer to one float64 (FP64 total reduction).
 * grad (may be NULL) receives d/dx of |S1 - tgt|^2 wrt S1's vertices,
 * [nv,3] f64 (varifoldNormGradient); requires S1's adjacency.
 * Replaces base/surfaceDistances.py:286-468 (varifoldNorm0/Def/Gradient). */
int lddmm_surface_varifold(const lddmm_kernel* kd,
                           const lddmm_surface* s1, const lddmm_surface* tgt,
                           double* energy, double* grad, void* stream);

/* Face features of a surface: centers [nf,3], surfel*face_weight [nf,3],
 * face weights [nf] (base/surfaces.py:326-335).  Any output may be NULL. */
int lddmm_surface_features(const lddmm_surface* s, double* centers, double* wsurfel,
                           double* face_weights, void* stream);

/* Scatter of per-face (dz1, z1) to vertices with the cross-product chain rule
 * px[f_v] += dz1_f - (x_{v+2} - x_{v+1}) x z1_f
 * (base/surfaceDistances.py:440-453 / :117-131).  Requires the adjacency. */
int lddmm_surface_face_gather(const lddmm_surface* s, const double* dz1, const double* z1,
                              double* px, void* stream);

/* Polygonal curve on the device: vertices [nv,dim] f64 (dim 2 or 3),
 * segments [ns,2] int32 (the reference's `faces`), vertex weights [nv] f64
 * (NULL = all ones; line weight = mean of the two, base/curves.py:117-125). */
typedef struct lddmm_curve {
    const double*  vertices;
    int64_t        nv;
    int32_t        dim;
    int32_t        reserved;
    const int32_t* segments;
    int64_t        ns;
    const double*  vweights;
    /* vertex -> incident (segment*2 + end) adjacency, CSR, built by
     * lddmm_curve_adjacency; required only for gradients. */
    const int32_t* adj_ptr;            /* [nv+1] */
    const int32_t* adj_idx;            /* [2*ns] */
} lddmm_curve;

/* Host helper: CSR vertex->(segment,end) adjacency of int32 segments. */
int lddmm_curve_adjacency(const int32_t* segments_host, int64_t ns, int64_t nv,
                          int32_t* adj_ptr_host, int32_t* adj_idx_host);

/* Curve varifold data term, f(u) = 1 + weight*u^2 (weight >= 0):
 *   tgt == NULL : *energy = |C1|^2                 (varifoldNorm0)
 *   tgt != NULL : *energy = |C1|^2 - 2 <C1, tgt>   (varifoldNormDef)
 * energy: DEVICE f64[1].  grad (may be NULL): d/dx |C1 - tgt|^2, [nv,dim] f64
 * (varifoldNormGradient, with_weights=False); requires C1's adjacency.  With
 * grad the energy uses the gradient form's segment weight |l| w instead of
 * |w l| (identical for unit weights).
 * Replaces base/curveDistances.py:186-336. */
int lddmm_curve_varifold(const lddmm_kernel* kd,
                         const lddmm_curve* c1, const lddmm_curve* tgt, double weight,
                         double* energy, double* grad, void* stream);

/* Segment features: centers [ns,dim], linel [ns,dim], line weights [ns]
 * (base/curves.py:117-125).  Any output may be NULL. */
int lddmm_curve_features(const lddmm_curve* c, double* centers, double* linel,
                         double* seg_weights, void* stream);

/* px[v] = sum over incident segments of dz1_s -/+ wz1_s (start / end vertex),
 * dz1 and wz1 [ns,dim] (base/curveDistances.py:97-103, :324-326). */
int lddmm_curve_seg_gather(const lddmm_curve* c, const double* dz1, const double* wz1,
                           double* px, void* stream);

/* Forward Euler landmark flow x_{t+1} = x_t + (1/T) K(x_t,x_t) a_t, t < T.
 * at [T,n,dim], xt [T+1,n,dim] (xt[0] is read, xt[1..T] written).
 * If energy != NULL (device f64[1]) it receives sum_t r dt a_t.K(x_t)a_t.
 * K includes the kernel's affine part (kd->affine).
 * Replaces base/pointEvolution.py:88-164 and the deformation energy of
 * base/surfaceMatching.py:533-548. */
int lddmm_shoot_forward(const lddmm_kernel* kd, const double* at, int32_t T, int64_t n,
                        int32_t dim, double* xt, double regweight, double* energy,
                        void* stream);

/* Discrete adjoint: p_T = px1; p_t = p_{t+1} + dt applyDiffKT(x_t,p_{t+1},a_t,lddmm,r);
 * grad_t = K(x_t)(2 r a_t - p_{t+1}) / T when euclidean != 0, else
 * (2 r a_t - p_{t+1}) / T.  pxt [T+1,n,dim] (may be NULL), grad [T,n,dim].
 * applyDiffKT and K include the kernel's affine part (Kernel.applyDiffKT,
 * base/kernelFunctions.py:601-625; Kernel.applyK, :476-505).
 * Replaces base/surfaceMatching.py:728-908 (hamiltonianCovector/Gradient)
 * and base/pointEvolution.py:167-236, with the /T of getGradient (:1086). */
int lddmm_shoot_adjoint(const lddmm_kernel* kd, const double* at, const double* xt,
                        int32_t T, int64_t n, int32_t dim, const double* px1,
                        double regweight, int32_t euclidean, double* pxt, double* grad,
                        void* stream);

/* The same two loops with internal costs on the template's triangulation
 * (faces and adjacency of `s`; its vertices are not read; dim 3, n = s->nv):
 * forward: energy += sum_t dt C(x_t, v_t), v_t = K(x_t) a_t
 *   (base/surfaceMatching.py:549-553);
 * adjoint: with Lv, DLv = dC/dphi, dC/dx at (x_t, v_t),
 *   p_t = p_{t+1} + dt (applyDiffKT(x_t, p_{t+1} - Lv, a_t, lddmm, r) - DLv),
 *   grad_t = K(x_t)(2 r a_t - p_{t+1} + Lv) / T (euclidean) or without K
 *   (base/surfaceMatching.py:787-804, :860-891). */
int lddmm_shoot_forward_ic(const lddmm_kernel* kd, const lddmm_surface* s, const lddmm_internal_cost* ic,
                           const double* at, int32_t T, double* xt, double regweight, double* energy,
                           void* stream);
int lddmm_shoot_adjoint_ic(const lddmm_kernel* kd, const lddmm_surface* s, const lddmm_internal_cost* ic,
                           const double* at, const double* xt, int32_t T, const double* px1, double regweight,
                           int32_t euclidean, double* pxt, double* grad, void* stream);

/* Geodesic (EPDiff) shooting: x_{k+1} = x_k + dt K a_k, a_{k+1} = a_k - dt applyDiffKT(x_k,a_k,a_k)
 * for k < T; xt, at [T+1,n,dim] with xt[0], at[0] read.
 * Replaces base/pointEvolution.py:385-472 (secondOrderEvolution). */
int lddmm_epdiff_forward(const lddmm_kernel* kd, int32_t T, int64_t n, int32_t dim,
                         double* xt, double* at, void* stream);

/* Adjoint of lddmm_epdiff_forward given (xt, at), px1, pa1; writes pxt, pat
 * [T+1,n,dim] and da0 = K(x0)a0 - pa_0 [n,dim].
 * Replaces base/pointEvolution.py:489-618 (secondOrderCovector/Gradient). */
int lddmm_epdiff_adjoint(const lddmm_kernel* kd, int32_t T, int64_t n, int32_t dim,
                         const double* xt, const double* at, const double* px1,
                         const double* pa1, double* pxt, double* pat, double* da0,
                         void* stream);

/* ------------------------------------------------------------------------
 * Point-cloud network primitives of the learned warm start (no counterpart
 * in the reference; they implement the momentum predictor that feeds
 * SurfaceMatching's initial control).  Float32 activations, BF16 weights
 * (uint16 bit patterns), FP32 accumulation; device pointers; a segment is
 * the contiguous point range [off[s], off[s+1]) of one template-target pair.
 * ---------------------------------------------------------------------- */
/* Y[P,N] = act(X[P,K] W + b + segbias[seg[p]]); W is [K,N] (or [N,K] when transW);
 * b, segbias, seg may be NULL; relu != 0 applies max(., 0). */
int lddmm_nn_linear(const float* X, int64_t P, int32_t K, const uint16_t* W, int32_t transW,
                    const float* b, const float* segbias, const int32_t* seg, int32_t N,
                    int32_t relu, float* Y, void* stream);
/* dZ = dY * (Y > 0) (ReLU backward on post-activation values). */
int lddmm_nn_relu_bwd(const float* Y, const float* dY, float* dZ, int64_t count, void* stream);
/* dW[K,N] = X^T dZ, reduced over point chunks in a fixed order (deterministic). */
int lddmm_nn_wgrad(const float* X, const float* dZ, int64_t P, int32_t K, int32_t N, float* dW,
                   void* stream);
/* out[s,n] = sum over segment s of dZ[p,n] (bias / segment-bias gradients). */
int lddmm_nn_seg_colsum(const float* dZ, const int32_t* off, int32_t nseg, int32_t N, float* out,
                        void* stream);
/* g[s,n] = max over segment s of H[p,n]; arg[s,n] = first point attaining it. */
int lddmm_nn_segmax(const float* H, const int32_t* off, int32_t nseg, int32_t N, float* g,
                    int32_t* arg, void* stream);
/* dH[arg[s,n], n] += dg[s,n]. */
int lddmm_nn_segmax_bwd(const float* dg, const int32_t* off, const int32_t* arg, int32_t nseg,
                        int32_t N, float* dH, void* stream);
/* Adam step on float32 master weights w (gradient g, moments m, v; bias
 * corrections for `step` >= 1; decoupled-free L2 term wd*w added to g);
 * wbf (may be NULL) receives the BF16 copy. */
int lddmm_nn_adam(float* w, uint16_t* wbf, const float* g, float* m, float* v, int64_t count,
                  float lr, float b1, float b2, float eps, int32_t step, float wd, void* stream);
int lddmm_nn_to_bf16(const float* w, uint16_t* wbf, int64_t count, void* stream);
/* Warm-start input feature (no counterpart in the reference): for each ray
 * i (origin orig[i,:], direction dir[i,:], float64 [n,3]) the signed
 * parameter t[i] of the intersection orig + t dir with the triangle mesh
 * (tv [nv,3] float64, tf [nf,3] int32; a face with an index outside [0, nv)
 * is treated as degenerate and never hit) that has the
 * smallest |t| over both directions, lowest face index on ties; hit[i] = 1,
 * or t[i] = 0 and hit[i] = 0 when the line meets no triangle.  With unit
 * template vertex normals as directions, t is the signed normal distance
 * from the template to the target surface.  Meshes of more than 256 faces
 * are split into face chunks whose per-ray hits sit in the library's
 * per-device workspace (calls sharing a GPU go on one stream). */
int lddmm_normal_raycast(const double* orig, const double* dir, int64_t n, const double* tv, int64_t nv,
                         const int32_t* tf, int64_t nf, double* t_out, int32_t* hit, void* stream);

/* ------------------------------------------------------------------------
 * L-BFGS direction on the device (the two-loop recursion of base/bfgs.py:188-207
 * with the step scalars kept in device memory; no host synchronisation).
 * lddmm_lbfgs_pair_scalars: for a stored pair (s, y) of n doubles, writes the
 * device scalars rho = 1/<y,s> and c = <s,y>/<y,y> (1 if below 1e-10).
 * lddmm_lbfgs_two_loop: q = H g for the m stored pairs (oldest first; host
 * arrays of device pointers; rho[k] device scalars; c_newest = c of pair m-1);
 * q may equal g.  Dot products are reduced in a fixed order (deterministic).
 * ---------------------------------------------------------------------- */
int lddmm_lbfgs_pair_scalars(const double* s, const double* y, int64_t n, double* rho, double* c,
                             void* stream);
int lddmm_lbfgs_two_loop(const double* g, int64_t n, int32_t m, const double* const* s,
                         const double* const* y, const double* const* rho, const double* c_newest,
                         double* q, void* stream);
/* The same product in the compact form (Byrd-Nocedal-Schnabel), for m <= 64
 * stored pairs, in four launches instead of 2m + 1 (base/bfgs.py:188-207
 * computes it by the two-loop recursion; the results agree to rounding).
 * lddmm_lbfgs_multidot: out[p * nt + t] = <probes[p], targets[t]> for
 * np <= 2 probes and nt <= 128 targets of n doubles (host arrays of device
 * pointers), reduced in a fixed order.  lddmm_lbfgs_compact: q = H g; s, y as
 * in lddmm_lbfgs_two_loop; cols[j] (device, 2 * col_len[j] doubles) holds the
 * dots of y_j with s_0..s_j then with y_0..y_j of the pairs stored up to j
 * (older pairs that have since left the window may lead each half;
 * col_len[j] >= j + 1); gdots = [<s_i, g>, <y_i, g>] (2m doubles); q may
 * equal g. */
int lddmm_lbfgs_multidot(const double* const* probes, int32_t np, const double* const* targets, int32_t nt,
                         int64_t n, double* out, void* stream);
int lddmm_lbfgs_compact(const double* g, int64_t n, int32_t m, const double* const* s, const double* const* y,
                        const double* const* cols, const int32_t* col_len, const double* gdots, double* q,
                        void* stream);

/* ------------------------------------------------------------------------
 * Population registration: P template-target surface pairs of uneven sizes
 * computed together.  No counterpart in the reference (it registers one pair
 * per SurfaceMatching); per pair, every result equals the single-pair entry
 * point's (lddmm_shoot_forward / lddmm_shoot_adjoint / lddmm_surface_varifold)
 * up to floating-point summation order.  Each Gaussian sum over all active
 * pairs is ONE persistent launch driven by a segment table of (pair, target
 * tile, source chunk) work items, scheduled longest-first over the
 * workgroups on the host; tile and chunk shapes come from an autotuner that
 * times the candidates on the first call.  The T-step shooting loops are
 * captured once into hipGraphs and replayed.
 * Template vertices of all pairs are concatenated ([V,3], pair p owns rows
 * [voff[p], voff[p+1])), and so are the template faces (global vertex
 * indices), the target vertices and the target faces.
 * ---------------------------------------------------------------------- */
typedef struct lddmm_population_desc {
    int32_t        npairs;
    int32_t        reserved;
    const int64_t* voff;        /* host [npairs+1] template vertex offsets */
    const int64_t* foff;        /* host [npairs+1] template face offsets */
    const int32_t* faces;       /* device [F,3] */
    const double*  vweights;    /* device [V] or NULL (ones) */
    const int32_t* adj_ptr;     /* device CSR vertex->(face,corner) over all */
    const int32_t* adj_idx;     /*   template vertices (lddmm_surface_adjacency) */
    const int64_t* tfoff;       /* host [npairs+1] target face offsets */
    const double*  tvertices;   /* device target vertices of all pairs */
    const int32_t* tfaces;      /* device [TF,3] */
    const double*  tvweights;   /* device or NULL */
} lddmm_population_desc;

typedef struct lddmm_population lddmm_population;   /* opaque; one device */

int lddmm_population_create(const lddmm_population_desc* desc, lddmm_population** out);
int lddmm_population_destroy(lddmm_population* pop);
/* Pairs computed by the following calls (active_host[p] != 0; NULL = all).
 * Rows of inactive pairs in every output are left untouched. */
int lddmm_population_set_active(lddmm_population* pop, const uint8_t* active_host);
/* 0 = eager launches, 1 = hipGraph capture/replay of the shooting loops (default). */
int lddmm_population_set_graphs(lddmm_population* pop, int32_t on);
/* Forward Euler flow of every active pair: at [T,V,3], xt [T+1,V,3] (xt[0]
 * read); energies (device f64[P], may be NULL) receive
 * sum_t r dt a_t.K(x_t)a_t per pair. */
int lddmm_population_shoot_forward(const lddmm_kernel* kd, lddmm_population* pop, const double* at,
                                   int32_t T, double* xt, double regweight, double* energies,
                                   void* stream);
/* Exact discrete adjoint from px1 [V,3]: grad [T,V,3] (Euclidean K(x_t)
 * correction when euclidean != 0), as lddmm_shoot_adjoint. */
int lddmm_population_shoot_adjoint(const lddmm_kernel* kd, lddmm_population* pop, const double* at,
                                   const double* xt, int32_t T, const double* px1, double regweight,
                                   int32_t euclidean, double* grad, void* stream);
/* Varifold data term of the deformed templates (vertices [V,3]) against the
 * targets: energies (device f64[P] or NULL) = |S_p|^2 - 2 <S_p, T_p>,
 * grad (device [V,3] or NULL) = d/dx |S_p - T_p|^2. */
int lddmm_population_varifold(const lddmm_kernel* kd, lddmm_population* pop, const double* vertices,
                              double* energies, double* grad, void* stream);
/* out[p] = sum over t < T and the rows of pair p of a.b ([T,V,3] arrays). */
int lddmm_population_seg_dot(lddmm_population* pop, const double* a, const double* b, int32_t T,
                             double* out, void* stream);
/* y[t,i,:] += alpha[pair(i)] * x[t,i,:] for the rows of active pairs (alpha device [P]). */
int lddmm_population_seg_axpy(lddmm_population* pop, double* y, const double* alpha, const double* x,
                              int32_t T, void* stream);
/* Host-only self-test of the segment-table scheduler: builds the work items
 * of one family (0 = vertex sums, 1 = varifold faces) for the given sizes,
 * active set and launch shape (R targets per lane, f10 = items per workgroup
 * x 10, G workgroups) and checks every item against the buffer bounds the
 * kernels use.  No device call. */
int lddmm_population_plan_check(int32_t npairs, const int64_t* voff, const int64_t* foff,
                                const int64_t* tfoff, const uint8_t* active, int32_t family,
                                int32_t R, int32_t f10, int32_t G, int64_t* nitems);
/* Autotuned launch shapes: up to cap records of 4 int32 (op, precision, R,
 * items per workgroup x 10); returns the number of records in *count. */
int lddmm_population_tuning(const lddmm_population* pop, int32_t* out, int32_t cap, int32_t* count);

#ifdef __cplusplus
}
#endif

#endif /* LDDMM_AMD_H */
