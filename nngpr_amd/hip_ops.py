"""Torch-level wrappers over the C ABI (device tensors in, device tensors out).

Every function runs the HIP kernels of libnngp_hip.so on the tensors' CUDA (HIP)
device and the current torch stream.  Inputs must already live on that device;
nothing here falls back to another implementation.
"""
import ctypes

import numpy as np
import torch

from . import _native as nat

PRECISION_FP64 = nat.PRECISION_FP64
PRECISION_FP32_REFINED = nat.PRECISION_FP32_REFINED
KNN_AUTO = nat.KNN_AUTO
KNN_BRUTE_FORCE = nat.KNN_BRUTE_FORCE
KNN_BRUTE_FORCE_VALU = nat.KNN_BRUTE_FORCE_VALU
KNN_GRID_LANES = nat.KNN_GRID_LANES
KNN_GRID_WAVES = nat.KNN_GRID_WAVES

_workspaces = {}


def _require_device(*tensors):
    dev = None
    for t in tensors:
        if t is None:
            continue
        if not t.is_cuda:
            raise ValueError("nngpr_amd HIP ops need device tensors, got a %s tensor" % t.device)
        if dev is None:
            dev = t.device
        elif t.device != dev:
            raise ValueError("all tensors must be on the same device")
    return dev


def _stream(dev):
    return ctypes.c_void_p(torch.cuda.current_stream(dev).cuda_stream)


def _p(t):
    return None if t is None else ctypes.c_void_p(t.data_ptr())


def workspace(dev, nbytes, slot="default"):
    """Per-(device, slot) scratch buffer, grown on demand (torch caching allocator)."""
    key = (str(dev), slot)
    buf = _workspaces.get(key)
    if buf is None or buf.numel() < nbytes:
        buf = torch.empty(max(int(nbytes), 256), dtype=torch.uint8, device=dev)
        _workspaces[key] = buf
    return buf


def clear_workspaces():
    _workspaces.clear()


def _f64(t):
    if t.dtype != torch.float64:
        t = t.to(torch.float64)
    return t.contiguous()


def _i32(t):
    if t.dtype != torch.int32:
        t = t.to(torch.int32)
    return t.contiguous()


def _scale_array(scale, d):
    if scale is None:
        return None
    a = np.ascontiguousarray(np.asarray(scale, dtype=np.float64).reshape(-1))
    if a.size not in (1, d):
        raise ValueError("scale must have 1 or d entries")
    return a


def knn(ref, query, scale, k, ref_cut=None, precision=PRECISION_FP32_REFINED, out=None,
        ws_slot="knn"):
    """k nearest refs of every query (see nngp_knn in include/nngp.h).

    Returns an int32 (n_query, k) device tensor, ascending (key, index), -1 padded.
    """
    ref = _f64(ref)
    query = _f64(query)
    dev = _require_device(ref, query)
    n_ref, d = ref.shape
    n_query = query.shape[0]
    if query.shape[1] != d:
        raise ValueError("ref and query must have the same number of features")
    if out is None:
        out = torch.empty((n_query, k), dtype=torch.int32, device=dev)
    sc = _scale_array(scale, d)
    L = nat.lib()
    nbytes = L.nngp_knn_workspace_bytes(n_ref, n_query, d, k, precision)
    ws = workspace(dev, nbytes, ws_slot)
    rc = L.nngp_knn(_p(ref), n_ref, _p(query), n_query, d,
                    sc.ctypes.data_as(ctypes.c_void_p) if sc is not None else None,
                    0 if sc is None else sc.size, k,
                    0 if ref_cut is None else 1, 0 if ref_cut is None else int(ref_cut),
                    precision, _p(out), _p(ws), ws.numel(), _stream(dev))
    nat.check(rc, "nngp_knn")
    return out


RESCALE_MAX_K = 72
# Largest k of the parallel searches (uniform grid, brute-force scans); larger
# lists are searched one row per thread, far slower.
KNN_FAST_MAX_K = 64


class KnnRescaler:
    """Exact neighbour lists at new length scales from one search's top-K lists
    (nngp_knn_rescale_init / nngp_knn_rescale in include/nngp.h).

    The search runs once with K = k + margin candidates at `scale`; `lists(s)`
    then returns exactly what `knn(ref, query, s, k, ref_cut)` would, re-ranking
    or copying the stored candidates and recomputing only the rows whose
    certificate fails (`last_rows_recomputed`).
    """

    def __init__(self, ref, query, scale, k, K, ref_cut=None,
                 precision=PRECISION_FP32_REFINED):
        if not (k <= K <= RESCALE_MAX_K):
            raise ValueError("need k <= K <= %d" % RESCALE_MAX_K)
        self.ref = _f64(ref)
        self.query = self.ref if query is ref else _f64(query)
        self.dev = _require_device(self.ref, self.query)
        self.d = self.ref.shape[1]
        self.k, self.K = int(k), int(K)
        self.ref_cut = ref_cut
        self.scale = _scale_array(scale, self.d)
        if self.scale is None:
            raise ValueError("rescaling needs a length scale")
        self.cand = knn(self.ref, self.query, self.scale, self.K, ref_cut=ref_cut,
                        precision=precision)
        L = nat.lib()
        nq, nr = self.query.shape[0], self.ref.shape[0]
        nbytes = L.nngp_knn_rescale_state_bytes(nr, nq, self.d, self.k, self.K)
        self.state = torch.empty(max(int(nbytes), 256), dtype=torch.uint8, device=self.dev)
        rc = L.nngp_knn_rescale_init(_p(self.ref), nr, _p(self.query), nq, self.d,
                                     self.scale.ctypes.data_as(ctypes.c_void_p), self.scale.size,
                                     _p(self.cand), self.K, self.k, _p(self.state),
                                     self.state.numel(), _stream(self.dev))
        nat.check(rc, "nngp_knn_rescale_init")
        self.last_rows_recomputed = 0

    def base_lists(self):
        """The exact top-k lists at the search's own scale."""
        return self.cand[:, :self.k].contiguous()

    def lists(self, scale, out=None):
        sc = _scale_array(scale, self.d)
        if sc is None or sc.size != self.scale.size:
            raise ValueError("scale must have as many entries as the base scale")
        nq, nr = self.query.shape[0], self.ref.shape[0]
        if out is None:
            out = torch.empty((nq, self.k), dtype=torch.int32, device=self.dev)
        cut = self.ref_cut
        rc = nat.lib().nngp_knn_rescale(
            _p(self.ref), nr, _p(self.query), nq, self.d,
            self.scale.ctypes.data_as(ctypes.c_void_p), sc.ctypes.data_as(ctypes.c_void_p),
            sc.size, 0 if cut is None else 1, 0 if cut is None else int(cut), _p(self.cand),
            self.K, self.k, _p(out), _p(self.state), self.state.numel(), _stream(self.dev))
        nat.check(rc, "nngp_knn_rescale")
        return out

    def rows_recomputed(self):
        """Rows of the last `lists` call recomputed by the exact per-row search
        (synchronises the stream)."""
        st = nat.NngpStats()
        nat.check(nat.lib().nngp_read_stats(_p(self.state), ctypes.byref(st), _stream(self.dev)),
                  "nngp_read_stats")
        return int(st.knn_rows_recomputed)


def stats(ws_slot, dev=None):
    """Counters of the last call that used workspace slot `ws_slot` ("knn",
    "vecchia", "predict", ...): kNN rows recomputed exactly, points re-factored in
    FP64 by the FP32 form.  None when the slot was never used."""
    dev = dev or torch.device("cuda", torch.cuda.current_device())
    buf = _workspaces.get((str(dev), ws_slot))
    if buf is None:
        return None
    st = nat.NngpStats()
    nat.lib().nngp_read_stats(_p(buf), ctypes.byref(st), _stream(dev))
    return {"knn_rows_recomputed": st.knn_rows_recomputed,
            "points_refactored_fp64": st.points_refactored_fp64}


def knn_stats(ws_slot="knn", dev=None):
    return stats(ws_slot, dev)


def fill_nn_kernel(x, nn, kstruct, eval_gradient=False):
    """Materialised neighbour blocks K (B, w, w) and gradients (B, w, w, n_theta)."""
    x = _f64(x)
    nn = _i32(nn)
    dev = _require_device(x, nn)
    B, w = nn.shape
    K = torch.empty((B, w, w), dtype=torch.float64, device=dev)
    G = torch.empty((B, w, w, kstruct.n_theta), dtype=torch.float64, device=dev) \
        if eval_gradient else None
    rc = nat.lib().nngp_fill_nn_kernel(_p(x), x.shape[0], x.shape[1], _p(nn), B, w,
                                       ctypes.byref(kstruct), _p(K), _p(G), _stream(dev))
    nat.check(rc, "nngp_fill_nn_kernel")
    if G is None:
        G = torch.empty(0, dtype=torch.float64, device=dev)
    return K, G


def _check_pd(ws, dev, what):
    st = nat.NngpStats()
    rc = nat.lib().nngp_read_stats(_p(ws), ctypes.byref(st), _stream(dev))
    if rc == nat.NNGP_ERR_NOT_POSITIVE_DEFINITE:
        raise np.linalg.LinAlgError(
            "%s: a neighbour covariance block is not positive definite" % what)
    if rc == nat.NNGP_ERR_INVALID_ARGUMENT:
        raise ValueError("%s: a neighbour index lies outside [-1, number of rows)" % what)
    nat.check(rc, "nngp_read_stats")
    return st


def check_status(ws_slot, what, dev=None):
    """Raise for the last call on workspace slot `ws_slot` as check_pd=True would
    (LinAlgError for a block that could not be factored, ValueError for an index
    out of range); for callers that launch with check_pd=False and read their
    results first."""
    dev = dev or torch.device("cuda", torch.cuda.current_device())
    buf = _workspaces.get((str(dev), ws_slot))
    if buf is not None:
        _check_pd(buf, dev, what)


def vecchia_loglik(x, y, nn, kstruct, alpha, precision=PRECISION_FP64, eval_gradient=False,
                   row0=0, row1=None, point_loglik=False, check_pd=True, ws_slot="vecchia"):
    """Vecchia log-likelihood of rows [row0, row1) and its theta gradient.

    x (n, d) and y (n,) or (n, ny) float64 device tensors; nn (n, m) int32 with the
    neighbours of every row among earlier rows.  Returns device tensors
    (loglik (1,), grad (n_theta,) or None, per-row loglik or None).
    """
    x = _f64(x)
    y = _f64(y)
    if y.ndim == 1:
        y = y[:, None]
    nn = _i32(nn)
    dev = _require_device(x, y, nn)
    n, d = x.shape
    ny = y.shape[1]
    m = nn.shape[1]
    row1 = n if row1 is None else row1
    nt = kstruct.n_theta
    ll = torch.empty(1, dtype=torch.float64, device=dev)
    grad = torch.empty(max(nt, 1), dtype=torch.float64, device=dev) if eval_gradient else None
    pts = torch.empty(row1 - row0, dtype=torch.float64, device=dev) if point_loglik else None
    L = nat.lib()
    nbytes = L.nngp_vecchia_workspace_bytes(row1 - row0, m, d, ny, nt, precision,
                                            1 if eval_gradient else 0)
    ws = workspace(dev, nbytes, ws_slot)
    rc = L.nngp_vecchia_loglik(_p(x), _p(y), n, d, ny, _p(nn), m, row0, row1,
                               ctypes.byref(kstruct), float(alpha), precision,
                               1 if eval_gradient else 0, _p(ll), _p(grad), _p(pts),
                               _p(ws), ws.numel(), _stream(dev))
    nat.check(rc, "nngp_vecchia_loglik")
    if check_pd:
        _check_pd(ws, dev, "log_marginal_likelihood")
    if grad is not None:
        grad = grad[:nt]
    return ll, grad, pts


def predict(x_train, y_train, x_query, nn, kstruct, alpha, precision=PRECISION_FP64,
            return_std=True, check_pd=True, ws_slot="predict", out_mean=None, out_std=None):
    """Only-train predictive mean (nq, ny) and std (nq,) of the query points."""
    x_train = _f64(x_train)
    y_train = _f64(y_train)
    if y_train.ndim == 1:
        y_train = y_train[:, None]
    x_query = _f64(x_query)
    nn = _i32(nn)
    dev = _require_device(x_train, y_train, x_query, nn)
    nt, d = x_train.shape
    ny = y_train.shape[1]
    nq = x_query.shape[0]
    m = nn.shape[1]
    mean = out_mean if out_mean is not None else torch.empty((nq, ny), dtype=torch.float64, device=dev)
    std = None
    if return_std:
        std = out_std if out_std is not None else torch.empty(nq, dtype=torch.float64, device=dev)
    L = nat.lib()
    nbytes = L.nngp_predict_workspace_bytes(nq, m, d, ny, precision)
    ws = workspace(dev, nbytes, ws_slot)
    rc = L.nngp_predict(_p(x_train), _p(y_train), nt, d, ny, _p(x_query), nq, _p(nn), m,
                        ctypes.byref(kstruct), float(alpha), precision, _p(mean), _p(std),
                        _p(ws), ws.numel(), _stream(dev))
    nat.check(rc, "nngp_predict")
    if check_pd:
        _check_pd(ws, dev, "predict")
    return mean, std


COND_MEAN = nat.COND_MEAN
COND_STD = nat.COND_STD
COND_COV = nat.COND_COV


def _sequential_inputs(x_train, y_train, x_query, nn):
    x_query = _f64(x_query)
    nn = _i32(nn)
    dev = _require_device(x_train, y_train, x_query, nn)
    nq, d = x_query.shape
    if x_train is None or x_train.shape[0] == 0:
        ny = 1 if y_train is None or y_train.ndim == 1 else max(1, y_train.shape[1])
        x_train = torch.empty((0, d), dtype=torch.float64, device=dev)
        y_train = torch.empty((0, ny), dtype=torch.float64, device=dev)
    else:
        x_train = _f64(x_train)
        y_train = _f64(y_train)
        if y_train.ndim == 1:
            y_train = y_train[:, None]
    if nn.ndim != 2 or nn.shape[0] != nq:
        raise ValueError("nn must have one row per query")
    if nn.shape[1] == 0:  # no neighbours at all: one empty slot per row
        nn = torch.full((nq, 1), -1, dtype=torch.int32, device=dev)
    return x_train, y_train, x_query, nn, dev


def predict_sequential(x_train, y_train, x_query, nn, kstruct, alpha, output=COND_STD,
                       ws_slot="sequential"):
    """Prediction with sequential conditioning (nngp_predict_sequential).

    nn (nq, m) indexes [x_train; x_query] (training rows, then earlier queries).
    Returns mean (nq, ny), std (nq,) or None, cov (nq, nq) or None.
    """
    x_train, y_train, x_query, nn, dev = _sequential_inputs(x_train, y_train, x_query, nn)
    nt, d = x_train.shape
    ny = y_train.shape[1]
    nq = x_query.shape[0]
    m = nn.shape[1]
    mean = torch.empty((nq, ny), dtype=torch.float64, device=dev)
    std = torch.empty(nq, dtype=torch.float64, device=dev) if output != COND_MEAN else None
    cov = torch.empty((nq, nq), dtype=torch.float64, device=dev) if output == COND_COV else None
    L = nat.lib()
    nbytes = L.nngp_predict_sequential_workspace_bytes(nt, nq, m, d, ny, output)
    ws = workspace(dev, nbytes, ws_slot)
    rc = L.nngp_predict_sequential(_p(x_train) if nt else None, _p(y_train) if nt else None, nt,
                                   _p(x_query), nq, d, ny, _p(nn), m, ctypes.byref(kstruct),
                                   float(alpha), int(output), _p(mean), _p(std), _p(cov), _p(ws),
                                   ws.numel(), _stream(dev))
    nat.check(rc, "nngp_predict_sequential")
    return mean, std, cov


def sample_sequential(x_train, y_train, x_query, nn, kstruct, alpha, normals,
                      ws_slot="sequential"):
    """Samples of the sequential process (nngp_sample_sequential); normals
    (nq, ny, n_samples) are standard normal draws.  Returns (nq, ny, n_samples)."""
    x_train, y_train, x_query, nn, dev = _sequential_inputs(x_train, y_train, x_query, nn)
    normals = _f64(normals)
    _require_device(normals, x_query)
    nt, d = x_train.shape
    ny = y_train.shape[1]
    nq = x_query.shape[0]
    m = nn.shape[1]
    if normals.ndim != 3 or normals.shape[0] != nq or normals.shape[1] != ny:
        raise ValueError("normals must have shape (n_query, n_targets, n_samples)")
    ns = normals.shape[2]
    out = torch.empty((nq, ny, ns), dtype=torch.float64, device=dev)
    L = nat.lib()
    nbytes = L.nngp_sample_sequential_workspace_bytes(nt, nq, m, d, ny)
    ws = workspace(dev, nbytes, ws_slot)
    rc = L.nngp_sample_sequential(_p(x_train) if nt else None, _p(y_train) if nt else None, nt,
                                  _p(x_query), nq, d, ny, _p(nn), m, ctypes.byref(kstruct),
                                  float(alpha), ns, _p(normals), _p(out), _p(ws), ws.numel(),
                                  _stream(dev))
    nat.check(rc, "nngp_sample_sequential")
    return out


class knn_method:
    """Context manager selecting the process-wide kNN search method:
    `with hip_ops.knn_method(hip_ops.KNN_BRUTE_FORCE): ...`."""

    def __init__(self, method):
        self.method = int(method)
        self.prev = None

    def __enter__(self):
        self.prev = nat.knn_get_method()
        nat.knn_set_method(self.method)
        return self

    def __exit__(self, *exc):
        nat.knn_set_method(self.prev)
        return False
