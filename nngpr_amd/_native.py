"""ctypes binding of libnngp_hip.so (the C ABI declared in include/nngp.h).

The library is loaded from the package directory.  There is no fallback: if it is
missing or cannot be loaded, every accelerated entry point raises.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libnngp_hip.so")

NNGP_MAX_DIM = 64
NNGP_MAX_NN = 1024
NNGP_MAX_K = 256

NNGP_OK = 0
NNGP_ERR_INVALID_ARGUMENT = -1
NNGP_ERR_UNSUPPORTED = -2
NNGP_ERR_NOT_POSITIVE_DEFINITE = -3
NNGP_ERR_WORKSPACE_TOO_SMALL = -4
NNGP_ERR_HIP = -5
NNGP_ERR_KNN_OVERFLOW = -6

PRECISION_FP64 = 0
PRECISION_FP32_REFINED = 1

# nngp_cond_output
COND_MEAN = 0
COND_STD = 1
COND_COV = 2


class NngpKernel(ctypes.Structure):
    """Mirror of `nngp_kernel` (include/nngp.h)."""
    _fields_ = [
        ("noise_level", ctypes.c_double),
        ("constant_value", ctypes.c_double),
        ("rbf_level", ctypes.c_double),
        ("n_length_scale", ctypes.c_int32),
        ("theta_noise_level", ctypes.c_int32),
        ("theta_constant_value", ctypes.c_int32),
        ("theta_rbf_level", ctypes.c_int32),
        ("theta_length_scale", ctypes.c_int32),
        ("n_theta", ctypes.c_int32),
        ("length_scale", ctypes.c_double * NNGP_MAX_DIM),
        ("covariance", ctypes.c_int32),
    ]


# nngp_covariance: stationary part of the covariance
COV_RBF = 0
COV_MATERN_12 = 1
COV_MATERN_32 = 2
COV_MATERN_52 = 3
_MATERN_FAMILY = {0.5: COV_MATERN_12, 1.5: COV_MATERN_32, 2.5: COV_MATERN_52}


def covariance_family(nu):
    """nngp_covariance of a Matern smoothness nu (None or inf: the RBF)."""
    if nu is None or np.isinf(nu):
        return COV_RBF
    fam = _MATERN_FAMILY.get(float(nu))
    if fam is None:
        raise NotImplementedError("Matern nu=%r is not accelerated (0.5, 1.5, 2.5 or inf)" % nu)
    return fam


class NngpStats(ctypes.Structure):
    _fields_ = [("knn_rows_recomputed", ctypes.c_int64),
                ("points_refactored_fp64", ctypes.c_int64)]


# (name, restype, argtypes) of every exported entry point.
_P = ctypes.c_void_p
_I64 = ctypes.c_int64
_I = ctypes.c_int
_D = ctypes.c_double
_SZ = ctypes.c_size_t
SIGNATURES = [
    ("nngp_status_string", ctypes.c_char_p, [_I]),
    ("nngp_last_error_detail", ctypes.c_char_p, []),
    ("nngp_abi_version", _I, []),
    ("nngp_knn_workspace_bytes", _SZ, [_I64, _I64, _I, _I, _I]),
    ("nngp_knn", _I, [_P, _I64, _P, _I64, _I, _P, _I, _I, _I, _I64, _I, _P, _P, _SZ, _P]),
    ("nngp_knn_rescale_state_bytes", _SZ, [_I64, _I64, _I, _I, _I]),
    ("nngp_knn_rescale_init", _I, [_P, _I64, _P, _I64, _I, _P, _I, _P, _I, _I, _P, _SZ, _P]),
    ("nngp_knn_rescale", _I, [_P, _I64, _P, _I64, _I, _P, _P, _I, _I, _I64, _P, _I, _I, _P, _P,
                              _SZ, _P]),
    ("nngp_fill_nn_kernel", _I, [_P, _I64, _I, _P, _I64, _I, ctypes.POINTER(NngpKernel), _P, _P, _P]),
    ("nngp_vecchia_workspace_bytes", _SZ, [_I64, _I, _I, _I, _I, _I, _I]),
    ("nngp_vecchia_loglik", _I, [_P, _P, _I64, _I, _I, _P, _I, _I64, _I64,
                                 ctypes.POINTER(NngpKernel), _D, _I, _I, _P, _P, _P, _P, _SZ, _P]),
    ("nngp_vecchia_loglik_noalloc", _I, [_P, _P, _I64, _I, _I, _P, _I, _I64, _I64,
                                         ctypes.POINTER(NngpKernel), _D, _I, _I, _P, _P, _P,
                                         _P]),
    ("nngp_predict_workspace_bytes", _SZ, [_I64, _I, _I, _I, _I]),
    ("nngp_predict", _I, [_P, _P, _I64, _I, _I, _P, _I64, _P, _I,
                          ctypes.POINTER(NngpKernel), _D, _I, _P, _P, _P, _SZ, _P]),
    ("nngp_predict_sequential_workspace_bytes", _SZ, [_I64, _I64, _I, _I, _I, _I]),
    ("nngp_predict_sequential", _I, [_P, _P, _I64, _P, _I64, _I, _I, _P, _I,
                                     ctypes.POINTER(NngpKernel), _D, _I, _P, _P, _P, _P, _SZ,
                                     _P]),
    ("nngp_sample_sequential_workspace_bytes", _SZ, [_I64, _I64, _I, _I, _I]),
    ("nngp_sample_sequential", _I, [_P, _P, _I64, _P, _I64, _I, _I, _P, _I,
                                    ctypes.POINTER(NngpKernel), _D, _I, _P, _P, _P, _SZ, _P]),
    ("nngp_read_stats", _I, [_P, ctypes.POINTER(NngpStats), _P]),
    ("nngp_model_create", _I, [_P, _P, _I64, _I, _I, ctypes.POINTER(NngpKernel), _D, _I, _P, _I, _I,
                               ctypes.POINTER(ctypes.c_void_p), _P]),
    ("nngp_model_destroy", _I, [_P]),
    ("nngp_model_predict_workspace_bytes", _SZ, [_P, _I64]),
    ("nngp_model_predict", _I, [_P, _P, _I64, _P, _P, _P, _SZ, _P]),
    ("nngp_knn_set_method", _I, [_I]),
    ("nngp_knn_get_method", _I, []),
    ("nngp_timing_enable", _I, [_I]),
    ("nngp_timing_read", _I, [_I, _I, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(_I64)]),
    ("nngp_timing_work", _I, [_I, _I, ctypes.POINTER(_I64)]),
]

TIMED_KNN_SCAN, TIMED_KNN_RERANK, TIMED_LOGLIK, TIMED_PREDICT = 0, 1, 2, 3
TIMED_KNN_GRID, TIMED_KNN_GRID_BUILD = 4, 5

KNN_AUTO, KNN_BRUTE_FORCE, KNN_BRUTE_FORCE_VALU, KNN_GRID_LANES, KNN_GRID_WAVES = 0, 1, 2, 3, 4


def knn_set_method(method):
    """Process-wide kNN search method (KNN_AUTO: grid search for d <= 3, k <= 64,
    one query per wave up to 100,000 queries and one per lane above;
    KNN_BRUTE_FORCE: always the scans; both screen d >= 4 on the matrix cores in
    FP32_REFINED; KNN_BRUTE_FORCE_VALU: the scans with the FP32 VALU screen;
    KNN_GRID_LANES / KNN_GRID_WAVES: AUTO with the one grid kernel for every
    query count).  Every method returns identical indices."""
    check(lib().nngp_knn_set_method(int(method)), "nngp_knn_set_method")


def knn_get_method():
    return int(lib().nngp_knn_get_method())


def timing_enable(on=True):
    lib().nngp_timing_enable(1 if on else 0)


def timing_read(kernel, reset=True):
    """(total milliseconds, launches) recorded for a timed kernel."""
    ms = ctypes.c_double()
    n = ctypes.c_int64()
    check(lib().nngp_timing_read(kernel, 1 if reset else 0, ctypes.byref(ms), ctypes.byref(n)),
          "nngp_timing_read")
    return ms.value, n.value


def timing_work(kernel, reset=True):
    """Work units measured inside a timed kernel since the last reset (the grid
    search: (query, ref) pairs keyed)."""
    u = ctypes.c_int64()
    check(lib().nngp_timing_work(kernel, 1 if reset else 0, ctypes.byref(u)), "nngp_timing_work")
    return u.value

_lib = None


class NativeLibraryError(RuntimeError):
    pass


def lib():
    """Load (once) and return the native library; raise if it is unavailable."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise NativeLibraryError(
            "libnngp_hip.so not found at %s; build it with `python -m nngpr_amd.build`" % LIB_PATH)
    try:
        handle = ctypes.CDLL(LIB_PATH)
    except OSError as e:
        raise NativeLibraryError("cannot load %s: %s" % (LIB_PATH, e))
    for name, res, args in SIGNATURES:
        fn = getattr(handle, name)
        fn.restype = res
        fn.argtypes = args
    _lib = handle
    return _lib


def status_text(rc):
    return lib().nngp_status_string(rc).decode()


def check(rc, what="nngp call"):
    """Map an nngp_status to the exception class the reference raises."""
    if rc == NNGP_OK:
        return
    msg = "%s failed: %s" % (what, status_text(rc))
    if rc == NNGP_ERR_HIP:
        msg += " (" + lib().nngp_last_error_detail().decode() + ")"
    if rc == NNGP_ERR_UNSUPPORTED:
        raise NotImplementedError(msg)
    if rc == NNGP_ERR_NOT_POSITIVE_DEFINITE:
        raise np.linalg.LinAlgError(msg)
    if rc == NNGP_ERR_INVALID_ARGUMENT:
        raise ValueError(msg)
    raise RuntimeError(msg)


def kernel_struct(params, mapping, n_theta):
    """Build an NngpKernel from decomposed parameters and their theta mapping."""
    k = NngpKernel()
    k.noise_level = float(params.get("noise_level", 0.0))
    k.constant_value = float(params.get("constant_value", 0.0))
    ls = params.get("length_scale")
    if ls is None:
        k.n_length_scale = 0
        k.rbf_level = 0.0
    else:
        lsa = np.asarray(ls, dtype=np.float64).reshape(-1)
        if lsa.size > NNGP_MAX_DIM:
            raise NotImplementedError("at most %d length scales are supported" % NNGP_MAX_DIM)
        k.n_length_scale = lsa.size
        for i, v in enumerate(lsa):
            k.length_scale[i] = float(v)
        k.rbf_level = float(params.get("rbf_level", 1.0))
        k.covariance = covariance_family(params.get("matern_nu"))

    def idx(name):
        v = mapping.get(name)
        if v is None:
            return -1
        return int(np.asarray(v).reshape(-1)[0])

    k.theta_noise_level = idx("noise_level")
    k.theta_constant_value = idx("constant_value")
    k.theta_rbf_level = idx("rbf_level")
    k.theta_length_scale = idx("length_scale")
    k.n_theta = int(n_theta)
    return k
