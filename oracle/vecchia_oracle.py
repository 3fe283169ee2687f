"""FP64 CPU restatement of the Vecchia / NNGP hot path of the reference.

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py): the HIP product path is checked
against these functions; it never calls them.

Kernel parameters are passed explicitly as the decomposition of
``WhiteKernel + ConstantKernel * RBF`` used by the reference's accelerated path
(cuda_kernel_accelerators.py:520-590):

    noise       w   (WhiteKernel noise_level, added on the diagonal by position)
    constant    C   (ConstantKernel summed with the RBF part)
    level       s   (ConstantKernel multiplying the RBF part)
    length      ls  (None: no RBF part; scalar: isotropic; (d,) array: ARD)
    nu              None or inf: RBF stationary part; 0.5, 1.5, 2.5: Matern of
                    that nu (batched_kernels.py:1003-1131)

and ``theta_map`` maps each of 'noise_level', 'constant_value', 'rbf_level',
'length_scale' to its index (or index array for ARD) in kernel.theta
(cuda_kernel_accelerators.py:592-649).
"""
import ctypes
import os

import numpy as np

_LIB = None


def _lib():
    global _LIB
    if _LIB is None:
        here = os.path.dirname(os.path.abspath(__file__))
        path = os.path.join(here, "liboracle_knn.so")
        if not os.path.exists(path):
            from . import build_oracle
            build_oracle.build(quiet=True)
        lib = ctypes.CDLL(path)
        lib.oracle_knn.restype = ctypes.c_int
        lib.oracle_knn.argtypes = [
            ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p, ctypes.c_int64, ctypes.c_int,
            ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int64,
            ctypes.c_void_p, ctypes.c_void_p]
        _LIB = lib
    return _LIB


def _ls_array(ls, d):
    if ls is None:
        return np.zeros(0)
    a = np.asarray(ls, dtype=np.float64).reshape(-1)
    if a.size not in (1, d):
        raise ValueError("length scale must be scalar or have d entries")
    return np.ascontiguousarray(a)


def knn(ref, query, ls, k, ref_cut=None, return_keys=False):
    """Exact FP64 kNN with the reference's key and ordering.

    Follows cuda_kernel_accelerators.py:196-381 (host) and :663-745 (kernel):
    scaled squared distance, top-k ascending, ties -> lower index, query i limited
    to refs [0, i + ref_cut) when ref_cut is not None, empty slots -1.
    """
    ref = np.ascontiguousarray(ref, dtype=np.float64)
    query = np.ascontiguousarray(query, dtype=np.float64)
    d = ref.shape[1]
    lsa = _ls_array(ls, d)
    out = np.empty((query.shape[0], k), dtype=np.int32)
    keys = np.empty((query.shape[0], k), dtype=np.float64) if return_keys else None
    rc = _lib().oracle_knn(
        ref.ctypes.data, ref.shape[0], query.ctypes.data, query.shape[0], d,
        lsa.ctypes.data if lsa.size else None, lsa.size, k,
        0 if ref_cut is None else 1, 0 if ref_cut is None else int(ref_cut),
        out.ctypes.data, keys.ctypes.data if keys is not None else None)
    if rc != 0:
        raise RuntimeError("oracle_knn failed: %d" % rc)
    return (out, keys) if return_keys else out


def knn_lexsort(ref, query, ls, k, ref_cut=None):
    """Independent small-size kNN: numpy keys (no FMA) + lexsort on (key, index).

    Used only to pin `knn` on inputs whose keys are exact or well separated.
    """
    ref = np.asarray(ref, dtype=np.float64)
    query = np.asarray(query, dtype=np.float64)
    d = ref.shape[1]
    lsa = _ls_array(ls, d)
    ur = ref / lsa if lsa.size else ref
    uq = query / lsa if lsa.size else query
    out = -np.ones((query.shape[0], k), dtype=np.int32)
    for i in range(query.shape[0]):
        lim = ref.shape[0] if ref_cut is None else min(max(i + ref_cut, 0), ref.shape[0])
        if lim == 0:
            continue
        key = ((ur[:lim] - uq[i]) ** 2).sum(-1)
        order = np.lexsort((np.arange(lim), key))[:k]
        out[i, :order.size] = order
    return out


def stationary(u, nu=None):
    """Stationary part E and its length-scale factor G at u = r^2 / 2, r the
    distance in length-scale units, with dE/dlog ls_f = G dx_f^2 / ls_f^2.

    RBF (nu None or inf; batched_kernels.py:844-920): E = exp(-u), G = E.
    Matern (batched_kernels.py:1069-1124), a = sqrt(2 nu) r:
      nu = 0.5  E = exp(-r)                  G = E / r (0 where r = 0, :1101-1104)
      nu = 1.5  E = (1 + a) exp(-a)          G = 3 exp(-a)
      nu = 2.5  E = (1 + a + a^2/3) exp(-a)  G = 5/3 (1 + a) exp(-a)
    """
    u = np.asarray(u, dtype=np.float64)
    if nu is None or np.isinf(nu):
        E = np.exp(-u)
        return E, E
    r = np.sqrt(2.0 * u)
    if nu == 0.5:
        E = np.exp(-r)
        G = np.where(r > 0, E / np.where(r > 0, r, 1.0), 0.0)
    elif nu == 1.5:
        a = np.sqrt(3.0) * r
        E = (1.0 + a) * np.exp(-a)
        G = 3.0 * np.exp(-a)
    elif nu == 2.5:
        a = np.sqrt(5.0) * r
        E = (1.0 + a + a * a / 3.0) * np.exp(-a)
        G = 5.0 / 3.0 * (1.0 + a) * np.exp(-a)
    else:
        raise NotImplementedError("Matern nu=%r" % nu)
    return E, G


def fill(xg, valid, w, C, s, ls, theta_map=None, n_theta=0, eval_gradient=False, nu=None):
    """Neighbour covariance blocks and their theta gradients.

    Restates the accelerated fill (cuda_kernel_accelerators.py:778-861):
    K_jl = s * E(u_jl) + C + w * [j == l] for valid j, l, u = sum_f dx_f^2 / (2 ls_f^2)
    and E the stationary part (`stationary`; the reference's accelerated fill has
    the RBF only) (white noise by position, :831); rows/cols of invalid slots are
    the identity with zero gradient (:787-811).  Gradients w.r.t. log-parameters
    (:833-857): d/dlog w = w [j == l], d/dlog C = C, d/dlog s = s E,
    d/dlog ls_f = s G dx_f^2/ls_f^2.

    xg: (B, M, d) gathered coordinates; valid: (B, M) bool.
    Returns K (B, M, M) and, if eval_gradient, Kg (B, M, M, n_theta).
    """
    xg = np.asarray(xg, dtype=np.float64)
    B, M, d = xg.shape
    diff = xg[:, :, None, :] - xg[:, None, :, :]          # (B, M, M, d)
    sq = diff * diff
    if ls is None:
        E = np.zeros((B, M, M))
        Gf = E
        half = np.zeros_like(sq)
    else:
        lsa = np.asarray(ls, dtype=np.float64).reshape(-1)
        half = sq / (2.0 * lsa * lsa) if lsa.size == d else sq / (2.0 * lsa[0] * lsa[0])
        E, Gf = stationary(half.sum(-1), nu)
    eye = np.eye(M, dtype=bool)[None]
    pair = valid[:, :, None] & valid[:, None, :]
    K = s * E + C + w * eye
    K = np.where(pair, K, eye.astype(np.float64))
    if not eval_gradient:
        return K, None
    Kg = np.zeros((B, M, M, n_theta))
    tm = theta_map or {}
    if tm.get("noise_level") is not None:
        Kg[..., tm["noise_level"]] = np.where(pair & eye, w, 0.0)
    if tm.get("constant_value") is not None:
        Kg[..., tm["constant_value"]] = np.where(pair, C, 0.0)
    if tm.get("rbf_level") is not None:
        Kg[..., tm["rbf_level"]] = np.where(pair, s * E, 0.0)
    if tm.get("length_scale") is not None and ls is not None:
        idx = np.asarray(tm["length_scale"]).reshape(-1)
        if idx.size == 1 and np.asarray(ls).size == 1:
            Kg[..., idx[0]] = np.where(pair, s * Gf * 2.0 * half.sum(-1), 0.0)
        else:
            for f, t in enumerate(idx):
                Kg[..., t] = np.where(pair, s * Gf * 2.0 * half[..., f], 0.0)
    return K, Kg


def _chol(A):
    return np.linalg.cholesky(A)


def _cho_solve(L, b):
    z = np.linalg.solve(L, b)
    return np.linalg.solve(np.swapaxes(L, -1, -2), z)


def vecchia_loglik(x, y, nn, w, C, s, ls, alpha, theta_map=None, n_theta=0,
                   eval_gradient=False, rows=None, nu=None):
    """Vecchia log-likelihood of the training set (and its theta gradient).

    Restates Nngpr._log_likelihood_worker (base_nngpr.py:530-614): for row i with
    neighbours nn[i] (earlier rows, -1 padded), idx = [i, nn[i]]; K from `fill`;
    alpha added on the diagonal of valid slots (:555-556, the point itself
    included); L = chol(K[1:, 1:]) (:564); mu = K_xn K^-1 y_N, sigma =
    sqrt(K_00 - K_xn K^-1 K_nx) (:578-580); l_i = sum over outputs of
    -(y - mu)^2/(2 sigma^2) - log(2 pi)/2 - log sigma (:581-584).  The gradient
    follows the reference's own formula through K_gradient and per-theta solves
    (:586-611).

    Returns (ll_total, grad or None, ll_per_row).
    """
    x = np.asarray(x, dtype=np.float64)
    y = np.asarray(y, dtype=np.float64)
    if y.ndim == 1:
        y = y[:, None]
    nn = np.asarray(nn)
    n = x.shape[0]
    rows = np.arange(n) if rows is None else np.asarray(rows)
    idx = np.concatenate([rows[:, None], nn[rows]], axis=1)
    valid = idx > -1
    xg = x[np.where(valid, idx, 0)]
    K, Kg = fill(xg, valid, w, C, s, ls, theta_map, n_theta, eval_gradient, nu)
    B, M = idx.shape
    K = K + alpha * (valid[:, :, None] & np.eye(M, dtype=bool)[None])
    L = _chol(K[:, 1:, 1:])
    yN = np.where(valid[:, 1:, None], y[np.where(valid[:, 1:], idx[:, 1:], 0)], 0.0)
    yi = y[rows]
    Kxn = K[:, 0:1, 1:]
    Kinv_y = _cho_solve(L, yN)
    Kinv_k = _cho_solve(L, np.swapaxes(Kxn, 1, 2))
    mu = (Kxn @ Kinv_y)[:, 0, :]
    sigma = np.sqrt(K[:, 0, 0][:, None] - (Kxn @ Kinv_k)[:, 0, :])
    ll = -0.5 * (yi - mu) ** 2 / sigma ** 2 - 0.5 * np.log(2 * np.pi) - np.log(sigma)
    ll_rows = ll.sum(-1)
    if not eval_gradient:
        return ll_rows.sum(), None, ll_rows
    # Reference gradient formula (base_nngpr.py:586-611), theta axis first.
    G = np.moveaxis(Kg, -1, 0)                               # (T, B, M, M)
    GNN = G[:, :, 1:, 1:]
    dKxn = G[:, :, 0:1, 1:]                                  # (T, B, 1, m)
    def dKinv(right):
        return -_cho_solve(L[None], GNN @ _cho_solve(L, right)[None])
    dmu = (dKxn @ Kinv_y[None])[:, :, 0, :] + (Kxn[None] @ dKinv(yN))[:, :, 0, :]
    dsig = 0.5 / sigma[None] * (G[:, :, 0, 0][..., None]
                                - 2 * (dKxn @ Kinv_k[None])[:, :, 0, :]
                                - (Kxn[None] @ dKinv(np.swapaxes(Kxn, 1, 2)))[:, :, 0, :])
    r = yi[None] - mu[None]
    g = (-1 / sigma[None] + r ** 2 / sigma[None] ** 3) * dsig + r / sigma[None] ** 2 * dmu
    return ll_rows.sum(), g.sum(axis=(1, 2)), ll_rows


def vecchia_adjoint_grad(x, y, nn, w, C, s, ls, alpha, theta_map, n_theta, rows=None, nu=None):
    """Closed-form adjoint of the Vecchia log-likelihood (no K_gradient solves).

    With Kt = K_NN + alpha I, k = K_N,self, c = K_self,self + alpha, a = Kt^-1 y_N,
    b = Kt^-1 k, mu = k.a, sigma^2 = c - k.b, r = y - mu,
    g_mu = r / sigma^2 and g_s = -1/(2 sigma^2) + r^2/(2 sigma^4) (summed over
    outputs where shared):
      dl/dtheta = g_s dc + dk.(sum_o g_mu a - 2 g_s b) + sum_jl dKt_jl (g_s b_j b_l - b_j (sum_o g_mu a)_l)
    which expands the reference's dmu/dtheta and dsigma/dtheta
    (base_nngpr.py:586-611).  This is the formula the HIP adjoint kernel uses;
    tests check it against `vecchia_loglik`'s reference formula.
    """
    x = np.asarray(x, dtype=np.float64)
    y = np.asarray(y, dtype=np.float64)
    if y.ndim == 1:
        y = y[:, None]
    n = x.shape[0]
    rows = np.arange(n) if rows is None else np.asarray(rows)
    idx = np.concatenate([rows[:, None], np.asarray(nn)[rows]], axis=1)
    valid = idx > -1
    xg = x[np.where(valid, idx, 0)]
    K, Kg = fill(xg, valid, w, C, s, ls, theta_map, n_theta, True, nu)
    B, M = idx.shape
    K = K + alpha * (valid[:, :, None] & np.eye(M, dtype=bool)[None])
    Kt = K[:, 1:, 1:]
    k = K[:, 1:, 0]
    c = K[:, 0, 0]
    L = _chol(Kt)
    yN = np.where(valid[:, 1:, None], y[np.where(valid[:, 1:], idx[:, 1:], 0)], 0.0)
    a = _cho_solve(L, yN)                       # (B, m, ny)
    b = _cho_solve(L, k[..., None])[..., 0]     # (B, m)
    mu = np.einsum("bj,bjo->bo", k, a)
    s2 = c - np.einsum("bj,bj->b", k, b)
    r = y[rows] - mu
    gmu = r / s2[:, None]
    gs = (-0.5 / s2[:, None] + 0.5 * r * r / s2[:, None] ** 2).sum(-1)
    at = np.einsum("bo,bjo->bj", gmu, a)
    W = gs[:, None, None] * b[:, :, None] * b[:, None, :] - b[:, :, None] * at[:, None, :]
    v = at - 2 * gs[:, None] * b
    dc = Kg[:, 0, 0, :]
    dk = Kg[:, 1:, 0, :]
    dK = Kg[:, 1:, 1:, :]
    g = gs[:, None] * dc + np.einsum("bj,bjt->bt", v, dk) + np.einsum("bjl,bjlt->bt", W, dK)
    return g.sum(0)


def predict(x_train, y_train, xq, nn, w, C, s, ls, alpha, nu=None):
    """Only-train prediction (mean, std) of query points.

    Restates Nngpr._predict_worker (base_nngpr.py:1101-1127) with the neighbour
    layout of _find_nn_indices_for_train_and_eval (:1454-1463): idx = [query,
    nbrs] on x_full = [x_train; xq]; alpha only on train neighbours (0 <= idx < nt,
    :1106-1107) so the query's own variance keeps white noise but no alpha;
    mean = K_xn K^-1 y_N, std = sqrt(K_00 - K_xn K^-1 K_nx) without clamping (:1126).
    Returns mean (nq, ny) and std (nq,).
    """
    x_train = np.asarray(x_train, dtype=np.float64)
    xq = np.asarray(xq, dtype=np.float64)
    y = np.asarray(y_train, dtype=np.float64)
    if y.ndim == 1:
        y = y[:, None]
    nn = np.asarray(nn)
    nq = xq.shape[0]
    nt = x_train.shape[0]
    x_full = np.concatenate([x_train, xq], axis=0)
    idx = np.concatenate([np.arange(nt, nt + nq)[:, None], nn], axis=1)
    valid = idx > -1
    xg = x_full[np.where(valid, idx, 0)]
    K, _ = fill(xg, valid, w, C, s, ls, nu=nu)
    M = idx.shape[1]
    train = (idx >= 0) & (idx < nt)
    K = K + alpha * (train[:, :, None] & np.eye(M, dtype=bool)[None])
    L = _chol(K[:, 1:, 1:])
    yN = np.where(valid[:, 1:, None], y[np.where(valid[:, 1:], idx[:, 1:], 0)], 0.0)
    Kxn = K[:, 0:1, 1:]
    mean = (Kxn @ _cho_solve(L, yN))[:, 0, :]
    with np.errstate(invalid="ignore"):
        std = np.sqrt(K[:, 0, 0] - (Kxn @ _cho_solve(L, np.swapaxes(Kxn, 1, 2)))[:, 0, 0])
    return mean, std


def vecchia_knn(x, ls, m):
    """Vecchia conditioning sets of a training set: row i's m nearest among rows < i
    (base_nngpr.py:530-535: ref = x[:i1], query = x[i0:i1], ref_cut_index = i0)."""
    return knn(x, x, ls, m, ref_cut=0)
