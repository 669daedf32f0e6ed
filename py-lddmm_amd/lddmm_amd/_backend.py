"""ctypes binding of liblddmm_amd.so (the HIP C ABI declared in include/lddmm_amd.h).

torch supplies device memory and the current HIP stream; every compute call
goes through the C ABI.  There is no CPU fallback: importing this module
without the built library, or calling a kernel without a GPU, raises.
"""
import ctypes
import os

import numpy as np
import torch  # loaded first so the library binds to torch's HIP runtime

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liblddmm_amd.so")

MAX_SCALES = 8
KERNEL_GAUSS = 0
KERNEL_LAPLACIAN = 1
PREC_F64 = 0
PREC_F32C = 1
OK, EINVAL, EUNSUPPORTED, EHIP = 0, -1, -2, -3

AFFINE_NONE, AFFINE_FULL, AFFINE_EUCLIDEAN = 0, 1, 2
DIFFKT_PLAIN, DIFFKT_SYM, DIFFKT_LDDMM = 0, 1, 2
CG_REFERENCE, CG_TEXTBOOK = 0, 1
DIFFK_1, DIFFK_2, DIFFK_1AND2 = 0, 1, 2
DDIFFK_11, DDIFFK_12, DDIFFK_11AND12 = 0, 1, 2


class LddmmKernel(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int32), ("nscale", ctypes.c_int32),
                ("sigma", ctypes.c_double * MAX_SCALES), ("kp", ctypes.c_double),
                ("prec", ctypes.c_int32), ("order", ctypes.c_int32),
                ("affine", ctypes.c_int32), ("reserved", ctypes.c_int32),
                ("w1", ctypes.c_double), ("w2", ctypes.c_double), ("center", ctypes.c_double * 3)]


class LddmmSurface(ctypes.Structure):
    _fields_ = [("vertices", ctypes.c_void_p), ("nv", ctypes.c_int64),
                ("faces", ctypes.c_void_p), ("nf", ctypes.c_int64),
                ("vweights", ctypes.c_void_p),
                ("adj_ptr", ctypes.c_void_p), ("adj_idx", ctypes.c_void_p)]


class LddmmCurve(ctypes.Structure):
    _fields_ = [("vertices", ctypes.c_void_p), ("nv", ctypes.c_int64),
                ("dim", ctypes.c_int32), ("reserved", ctypes.c_int32),
                ("segments", ctypes.c_void_p), ("ns", ctypes.c_int64),
                ("vweights", ctypes.c_void_p),
                ("adj_ptr", ctypes.c_void_p), ("adj_idx", ctypes.c_void_p)]


class LddmmPopulationDesc(ctypes.Structure):
    _fields_ = [("npairs", ctypes.c_int32), ("reserved", ctypes.c_int32),
                ("voff", ctypes.c_void_p), ("foff", ctypes.c_void_p),
                ("faces", ctypes.c_void_p), ("vweights", ctypes.c_void_p),
                ("adj_ptr", ctypes.c_void_p), ("adj_idx", ctypes.c_void_p),
                ("tfoff", ctypes.c_void_p), ("tvertices", ctypes.c_void_p),
                ("tfaces", ctypes.c_void_p), ("tvweights", ctypes.c_void_p)]


MAX_ICOSTS = 4
IC_KINDS = {'h1': 0, 'elastic': 1, 'displacement': 2}


class LddmmInternalCost(ctypes.Structure):
    _fields_ = [("ncost", ctypes.c_int32), ("kinds", ctypes.c_int32 * MAX_ICOSTS),
                ("weights", ctypes.c_double * MAX_ICOSTS)]


def internal_cost_desc(costs):
    """[(name, weight), ...] with name in 'h1', 'elastic', 'displacement'
    (SurfaceMatching's internalCost list, base/surfaceMatching.py:184-204)."""
    costs = list(costs)
    if not 1 <= len(costs) <= MAX_ICOSTS:
        raise ValueError(f"1..{MAX_ICOSTS} internal costs (got {len(costs)})")
    d = LddmmInternalCost()
    d.ncost = len(costs)
    for k, (name, w) in enumerate(costs):
        if name not in IC_KINDS:
            raise NotImplementedError(f"internal cost {name!r}: lddmm_amd builds {sorted(IC_KINDS)}")
        d.kinds[k] = IC_KINDS[name]
        d.weights[k] = float(w)
    return d


class LddmmError(RuntimeError):
    pass


_P = ctypes.c_void_p
_I64 = ctypes.c_int64
_I32 = ctypes.c_int32
_D = ctypes.c_double
_KP = ctypes.POINTER(LddmmKernel)
_SP = ctypes.POINTER(LddmmSurface)
_CP = ctypes.POINTER(LddmmCurve)
_ICP = ctypes.POINTER(LddmmInternalCost)

# name -> argtypes (restype is int for every compute entry point)
SIGNATURES = {
    "lddmm_applyK": [_KP, _P, _I64, _P, _I64, _I32, _P, _I32, _P, _P],
    "lddmm_applyK_noalloc": [_KP, _P, _I64, _P, _I64, _I32, _P, _I32, _P, _P],
    "lddmm_applyDiffKT": [_KP, _P, _I64, _P, _I64, _I32, _P, _P, _I32, _I32, _D, _P, _P],
    "lddmm_kernel_solve": [_KP, _P, _I64, _I32, _P, _I32, _I32, _I32, _P, _P, _P],
    "lddmm_applyDiv": [_KP, _P, _I64, _P, _I64, _I32, _P, _P, _P],
    "lddmm_applyktensor": [_KP, _P, _I64, _P, _I64, _I32, _P, _P, _P, _P, _I32, _P, _P],
    "lddmm_applykmat": [_KP, _P, _I64, _P, _I64, _I32, _P, _I32, _P, _P],
    "lddmm_applykdiffmat": [_KP, _P, _I64, _P, _I64, _I32, _P, _P, _P],
    "lddmm_applyDiffK": [_KP, _P, _I64, _I32, _P, _P, _I32, _I32, _P, _P],
    "lddmm_applyDDiffK": [_KP, _P, _I64, _I32, _P, _P, _I32, _P, _I32, _P, _P],
    "lddmm_surface_adjacency": [_P, _I64, _I64, _P, _P],
    "lddmm_surface_varifold": [_KP, _SP, _SP, _P, _P, _P],
    "lddmm_surface_features": [_SP, _P, _P, _P, _P],
    "lddmm_surface_face_gather": [_SP, _P, _P, _P, _P],
    "lddmm_curve_adjacency": [_P, _I64, _I64, _P, _P],
    "lddmm_curve_varifold": [_KP, _CP, _CP, _D, _P, _P, _P],
    "lddmm_curve_features": [_CP, _P, _P, _P, _P],
    "lddmm_curve_seg_gather": [_CP, _P, _P, _P, _P],
    "lddmm_shoot_forward": [_KP, _P, _I32, _I64, _I32, _P, _D, _P, _P],
    "lddmm_shoot_adjoint": [_KP, _P, _P, _I32, _I64, _I32, _P, _D, _I32, _P, _P, _P],
    "lddmm_surface_internal_cost": [_SP, _P, _ICP, _P, _P, _P, _P],
    "lddmm_shoot_forward_ic": [_KP, _SP, _ICP, _P, _I32, _P, _D, _P, _P],
    "lddmm_shoot_adjoint_ic": [_KP, _SP, _ICP, _P, _P, _I32, _P, _D, _I32, _P, _P, _P],
    "lddmm_epdiff_forward": [_KP, _I32, _I64, _I32, _P, _P, _P],
    "lddmm_epdiff_adjoint": [_KP, _I32, _I64, _I32, _P, _P, _P, _P, _P, _P, _P, _P],
    "lddmm_release_workspace": [],
    "lddmm_profile_enable": [_I32],
    "lddmm_profile_read": [_P, _I64, _P],
    "lddmm_device_count": [],
    "lddmm_set_pair_engine": [_I32],
    "lddmm_set_varifold_engine": [_I32],
    "lddmm_nn_linear": [_P, _I64, _I32, _P, _I32, _P, _P, _P, _I32, _I32, _P, _P],
    "lddmm_nn_relu_bwd": [_P, _P, _P, _I64, _P],
    "lddmm_nn_wgrad": [_P, _P, _I64, _I32, _I32, _P, _P],
    "lddmm_nn_seg_colsum": [_P, _P, _I32, _I32, _P, _P],
    "lddmm_nn_segmax": [_P, _P, _I32, _I32, _P, _P, _P],
    "lddmm_nn_segmax_bwd": [_P, _P, _P, _I32, _I32, _P, _P],
    "lddmm_nn_adam": [_P, _P, _P, _P, _P, _I64, ctypes.c_float, ctypes.c_float, ctypes.c_float,
                      ctypes.c_float, _I32, ctypes.c_float, _P],
    "lddmm_nn_to_bf16": [_P, _P, _I64, _P],
    "lddmm_normal_raycast": [_P, _P, _I64, _P, _I64, _P, _I64, _P, _P, _P],
    "lddmm_lbfgs_pair_scalars": [_P, _P, _I64, _P, _P, _P],
    "lddmm_lbfgs_two_loop": [_P, _I64, _I32, _P, _P, _P, _P, _P, _P],
    "lddmm_lbfgs_multidot": [_P, _I32, _P, _I32, _I64, _P, _P],
    "lddmm_lbfgs_compact": [_P, _I64, _I32, _P, _P, _P, _P, _P, _P, _P],
    "lddmm_population_create": [ctypes.POINTER(LddmmPopulationDesc), ctypes.POINTER(ctypes.c_void_p)],
    "lddmm_population_destroy": [_P],
    "lddmm_population_set_active": [_P, _P],
    "lddmm_population_set_graphs": [_P, _I32],
    "lddmm_population_shoot_forward": [_KP, _P, _P, _I32, _P, _D, _P, _P],
    "lddmm_population_shoot_adjoint": [_KP, _P, _P, _P, _I32, _P, _D, _I32, _P, _P],
    "lddmm_population_varifold": [_KP, _P, _P, _P, _P, _P],
    "lddmm_population_seg_dot": [_P, _P, _P, _I32, _P, _P],
    "lddmm_population_seg_axpy": [_P, _P, _P, _P, _I32, _P],
    "lddmm_population_tuning": [_P, _P, _I32, _P],
    "lddmm_population_plan_check": [_I32, _P, _P, _P, _P, _I32, _I32, _I32, _I32, _P],
}


def _load():
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"lddmm_amd: {LIB_PATH} is missing; build it with "
            "`python py-lddmm_amd/build_lib.py` (hipcc, gfx950)")
    lib = ctypes.CDLL(LIB_PATH, mode=ctypes.RTLD_GLOBAL)
    lib.lddmm_last_error.restype = ctypes.c_char_p
    lib.lddmm_last_error.argtypes = []
    lib.lddmm_version.restype = ctypes.c_char_p
    lib.lddmm_version.argtypes = []
    for name, args in SIGNATURES.items():
        f = getattr(lib, name)
        f.argtypes = args
        f.restype = ctypes.c_int
    return lib


LIB = _load()


def check(rc, what=""):
    if rc != OK:
        msg = LIB.lddmm_last_error().decode(errors="replace")
        kind = {EINVAL: "invalid argument", EUNSUPPORTED: "unsupported",
                EHIP: "HIP error"}.get(rc, f"error {rc}")
        if rc == EUNSUPPORTED:
            raise NotImplementedError(f"lddmm_amd {what}: {kind}: {msg}")
        raise LddmmError(f"lddmm_amd {what}: {kind}: {msg}")


_gpu_checked = False


def device():
    global _gpu_checked
    if not _gpu_checked:
        if not torch.cuda.is_available():
            raise LddmmError("lddmm_amd needs a HIP GPU (torch.cuda.is_available() is False)")
        _gpu_checked = True
    return torch.device("cuda", torch.cuda.current_device())


# torch's current stream of a device as a raw hipStream_t, without building a
# Stream object per call (the registration drivers make thousands of calls)
_raw_stream = getattr(torch._C, "_cuda_getCurrentRawStream", None)


def stream_handle():
    if _raw_stream is not None:
        return ctypes.c_void_p(_raw_stream(torch.cuda.current_device()))
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def ptr(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else ctypes.c_void_p(0)


_NP_DTYPE = {torch.float64: np.float64, torch.float32: np.float32, torch.int32: np.int32, torch.int64: np.int64}


def upload(arr):
    """Device copy of a contiguous numpy array on the current stream, without
    waiting for the GPU: the values are first copied into page-locked memory
    from torch's caching host allocator (a pageable copy would wait until every
    queued kernel has finished).  The caller may reuse `arr` at once."""
    t = torch.from_numpy(np.ascontiguousarray(arr))
    if t.numel() == 0:
        return t.to(device())
    return t.pin_memory().to(device(), non_blocking=True)


def to_dev(a, dtype=torch.float64):
    """Contiguous device tensor of `dtype` (numpy arrays are copied host->device)."""
    if isinstance(a, torch.Tensor):
        if (a.is_cuda and a.dtype == dtype and a.is_contiguous()
                and a.device.index == torch.cuda.current_device()):
            return a                     # what .to(...).contiguous() would return
        return a.to(device=device(), dtype=dtype).contiguous()
    npt = _NP_DTYPE.get(dtype) or torch.empty((), dtype=dtype).numpy().dtype
    return upload(np.ascontiguousarray(np.asarray(a), dtype=npt))


ENGINE_VALU, ENGINE_MFMA = 0, 1


def set_pair_engine(engine):
    """Accumulation engine of the float32 applyK / plain applyDiffKT sums."""
    code = {'valu': ENGINE_VALU, 'mfma': ENGINE_MFMA}.get(engine, engine)
    check(LIB.lddmm_set_pair_engine(int(code)), "set_pair_engine")


VARIFOLD_VALU, VARIFOLD_MFMA_COS, VARIFOLD_MFMA_COS_R2 = 0, 1, 2


def set_varifold_engine(engine):
    """Engine of the float32 single-scale varifold energy: 'valu' (default),
    'mfma_cos' (normal cosines on the matrix cores) or 'mfma_cos_r2' (cosines
    and squared distances)."""
    code = {'valu': VARIFOLD_VALU, 'mfma_cos': VARIFOLD_MFMA_COS,
            'mfma_cos_r2': VARIFOLD_MFMA_COS_R2}.get(engine, engine)
    check(LIB.lddmm_set_varifold_engine(int(code)), "set_varifold_engine")


def profile_enable(on=True):
    check(LIB.lddmm_profile_enable(1 if on else 0), "profile_enable")


def profile_read(cap=4096):
    """Durations (ms) of the pair-kernel launches recorded since the last read."""
    buf = (ctypes.c_double * cap)()
    n = ctypes.c_int64(0)
    check(LIB.lddmm_profile_read(buf, cap, ctypes.byref(n)), "profile_read")
    return [buf[i] for i in range(min(n.value, cap))]


def kernel_desc(sigma, kp=-1.0, prec='float64', kind=KERNEL_GAUSS, order=3):
    s = np.atleast_1d(np.asarray(sigma, dtype=np.float64)).ravel()
    if s.size < 1 or s.size > MAX_SCALES:
        raise ValueError(f"lddmm_amd supports 1..{MAX_SCALES} scales (got {s.size})")
    kd = LddmmKernel()
    kd.kind = kind
    kd.nscale = int(s.size)
    for i, v in enumerate(s):
        kd.sigma[i] = float(v)
    kd.kp = float(kp)
    kd.prec = precision_code(prec)
    if kind == KERNEL_LAPLACIAN:
        if int(order) != order or not 0 <= order <= 4:
            raise ValueError(f"laplacian kernel order must be an integer in 0..4 (got {order!r})")
        kd.order = int(order)
    return kd


def kernel_for(name, sigma, order, kp=-1.0, prec='float64'):
    """Descriptor of the reference kernel `name` ('gauss'/'gaussian' or
    'laplacian', matched as base/kernelFunctions_util.py:395-400 does)."""
    return kernel_desc(sigma, kp, prec, kernel_kind(name), order)


def kernel_of(K, prec=None, with_affine=False):
    """Descriptor of a Kernel object; with_affine carries its affine part
    (KernelSpec.setAffine: affine, w1, w2, center) for the shooting entry
    points, the only ones that apply it."""
    kd = kernel_for(K.name, K.sigma, K.order, K.ms_exponent, prec or K.pk_dtype)
    aff = getattr(K, 'affine', 'none')
    if aff in (None, 'none'):
        return kd
    if not with_affine:
        raise NotImplementedError(f"affine={aff!r} kernel terms are applied in shooting only")
    kinds = {'affine': AFFINE_FULL, 'euclidean': AFFINE_EUCLIDEAN}
    if aff not in kinds:
        raise NotImplementedError(f"affine={aff!r}: kernels take 'affine' or 'euclidean'")
    kd.affine = kinds[aff]
    kd.w1, kd.w2 = float(K.w1), float(K.w2)
    c = np.zeros(3)
    cc = np.asarray(K.center, dtype=np.float64).ravel()
    c[:cc.size] = cc[:3]
    for i in range(3):
        kd.center[i] = c[i]
    return kd


def kernel_solve(K, x, b, iter_max, sequence='reference'):
    """a with K(x, x) a = b by conjugate gradients on the device
    (lddmm_kernel_solve).  x [n, d] and b [n] or [n, k] float64 device
    tensors; K a Kernel without affine part.  Returns a with b's shape."""
    seq = {'reference': CG_REFERENCE, 'textbook': CG_TEXTBOOK}[sequence]
    x = to_dev(x)
    b2 = to_dev(b).reshape(b.shape[0], -1).contiguous()
    n, k = b2.shape
    if x.ndim != 2 or x.shape[0] != n:
        raise ValueError("kernel_solve: x must be [n, d] with n rows of b")
    a = torch.empty_like(b2)
    work = torch.empty(4 * n * k + 8, dtype=torch.float64, device=b2.device)
    check(LIB.lddmm_kernel_solve(kernel_of(K), ptr(x), n, x.shape[1], ptr(b2), k, int(iter_max), seq, ptr(a),
                                 ptr(work), stream_handle()), "kernel_solve")
    return a.reshape(b.shape)


def precision_code(prec):
    if prec in ('float64', 'double', np.float64, torch.float64, PREC_F64):
        return PREC_F64
    if prec in ('float32', 'float', np.float32, torch.float32, PREC_F32C):
        return PREC_F32C
    raise ValueError(f"unknown precision {prec!r} (use 'float64' or 'float32')")


def kernel_kind(name):
    if name is not None and 'gauss' in str(name):
        return KERNEL_GAUSS
    if name is not None and 'lap' in str(name):
        return KERNEL_LAPLACIAN
    raise NotImplementedError(
        f"lddmm_amd builds the gauss and laplacian kernel families (got name={name!r})")
