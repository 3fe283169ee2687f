"""Estimator core: scikit-learn style NNGP regressor on the HIP engine.

Keeps the public surface of the reference's Nngpr (base_nngpr.py:26-1464): the
constructor parameters and defaults, fit / log_marginal_likelihood / predict, the
sklearn parameter validation and the L-BFGS-B fit driver.  What runs underneath is
different: the training set lives in device memory, and every log-likelihood or
prediction is one kNN launch sequence plus one fused fill/factor/solve launch
sequence on one MI355X (nngpr_amd.hip_ops), instead of per-batch tensor algebra
distributed over worker processes.

Sequential conditioning (conditioning_method 'full' / 'full-double-nn'),
return_cov and sample_y run on the HIP engine too (nngp_predict_sequential /
nngp_sample_sequential).  Kernels outside White + Constant * (RBF or Matern
nu = 1/2, 3/2, 5/2) raise NotImplementedError.
"""
import multiprocessing
from numbers import Integral
from operator import itemgetter

import numpy as np
import torch
import xxhash
from sklearn import base as sk_base
from sklearn.gaussian_process import GaussianProcessRegressor
from sklearn.preprocessing._data import _handle_zeros_in_scale
from sklearn.utils import check_random_state
from sklearn.utils._param_validation import Interval, StrOptions
from sklearn.utils.validation import validate_data

from . import _native as nat
from . import batched_kernels
from . import hip_ops
from .array_module import TorchModule
from .kernel_params import KernelDecomposition


def _knn_lists(ref, query, scale, k, ref_cut=None, precision=hip_ops.PRECISION_FP32_REFINED):
    """k nearest refs of every query (hip_ops.knn).  Above the search's k limit the
    lists are accepted only when they hold every candidate ref of every query (the
    exact-GP settings, num_nn >= the number of rows): then each query lists all of
    its candidate refs in index order, the same set as the search would return
    (refs with a non-finite coordinate included)."""
    if k <= nat.NNGP_MAX_K:
        return hip_ops.knn(ref, query, scale, k, ref_cut=ref_cut, precision=precision)
    n_ref, nq = ref.shape[0], query.shape[0]
    dev = query.device
    avail = torch.full((nq,), n_ref, dtype=torch.int64, device=dev)
    if ref_cut is not None:
        avail = torch.clamp(torch.arange(nq, device=dev) + int(ref_cut), 0, n_ref)
    if nq and k < int(avail.max().item()):
        raise NotImplementedError("num_nn above %d is supported only when it covers every "
                                  "candidate row (%d here)" % (nat.NNGP_MAX_K, int(avail.max())))
    idx = torch.arange(k, device=dev)[None, :]
    return torch.where(idx < avail[:, None], idx, -1).to(torch.int32).contiguous()


def _fingerprint(a):
    """Content hash of a host array (shape, dtype and bytes)."""
    a = np.ascontiguousarray(a)
    h = xxhash.xxh3_64()
    h.update(repr((a.shape, a.dtype.str)).encode())
    try:
        h.update(memoryview(a).cast("B"))
    except (TypeError, ValueError):  # buffers memoryview cannot view as bytes
        h.update(a.tobytes())
    return h.intdigest()


def _precisions(allow_downcast_f32):
    """(kNN precision, fill/factor precision) for an allow_downcast_f32 setting.

    'no' -> FP64 everywhere; 'only-nn' (default) -> FP32-screened kNN (same exact
    neighbour sets) + FP64 factor; 'yes' -> FP32 factor with FP64 refinement too.
    """
    if allow_downcast_f32 == "no":
        return hip_ops.PRECISION_FP64, hip_ops.PRECISION_FP64
    if allow_downcast_f32 == "yes":
        return hip_ops.PRECISION_FP32_REFINED, hip_ops.PRECISION_FP32_REFINED
    return hip_ops.PRECISION_FP32_REFINED, hip_ops.PRECISION_FP64


class HipWorker:
    """Worker bound to one HIP device, exposing the reference's NngprWorker
    surface (base_nngpr.py:1558-1867) on top of the C ABI."""

    def __init__(self, device="cuda:0", random_gen_batch_byte_size=None,
                 nn_large_size_preferred=25000, batch_size=None):
        dev = torch.device(device)
        if dev.type != "cuda":
            raise ValueError("HipWorker needs a HIP device ('cuda' or 'cuda:N'), got %r" % device)
        if dev.index is None:
            dev = torch.device("cuda", 0)
        self.device = dev
        self.device_type = "cuda"
        self.xp = TorchModule(dev)
        self.random_gen_batch_byte_size = random_gen_batch_byte_size
        self.nn_large_size_preferred = nn_large_size_preferred
        self.batch_size = batch_size
        self._dec = KernelDecomposition()

    def __getstate__(self):
        st = dict(self.__dict__)
        st["device"] = str(self.device)
        st["_dec"] = None
        return st

    def __setstate__(self, st):
        self.__dict__.update(st)
        self.device = torch.device(st["device"])
        self._dec = KernelDecomposition()

    # --- NngprWorker surface -------------------------------------------------
    def get_array_module(self):
        return self.xp

    def clear_cache(self):
        hip_ops.clear_workspaces()
        if torch.cuda.is_available():
            torch.cuda.empty_cache()

    def decompose(self, kernel):
        dec = self._dec.update(kernel)
        if not dec.accelerated:
            raise NotImplementedError("Cannot accelerate this kernel: %s" % dec.error)
        return dec

    def kernel_struct(self, kernel):
        dec = self.decompose(kernel)
        return nat.kernel_struct(dec.params, dec.mapping, dec.n_theta), dec

    def knn_scale(self, kernel, nn_type):
        if nn_type != "kernel-space":
            return None
        dec = self.decompose(kernel)
        ls = dec.params.get("length_scale")
        return None if ls is None else np.asarray(ls, dtype=np.float64).reshape(-1)

    def find_nearest_neighbors(self, ref, query, nn_type, kernel, num_nn, ref_cut_index,
                               precision=hip_ops.PRECISION_FP32_REFINED):
        """base_nngpr.py:1613-1686 semantics on the HIP kNN (exact FP64 keys)."""
        ref = self.xp.from_numpy(ref)
        query = self.xp.from_numpy(query)
        return _knn_lists(ref, query, self.knn_scale(kernel, nn_type), int(num_nn),
                           ref_cut=ref_cut_index, precision=precision)

    def fill_nn_kernel(self, x, nn_indices, kernel, eval_gradient=False):
        """base_nngpr.py:1781-1826 semantics on the HIP fill kernel.  An index
        outside [-1, len(x)) raises IndexError, as the generic path's gather
        x[nn_indices] does."""
        ks, _ = self.kernel_struct(kernel)
        x = self.xp.from_numpy(x)
        nn = self.xp.from_numpy(nn_indices)
        if nn.numel() and bool(((nn < -1) | (nn >= x.shape[0])).any()):
            raise IndexError("neighbour index out of range for %d points" % x.shape[0])
        return hip_ops.fill_nn_kernel(x, nn, ks, eval_gradient=eval_gradient)

    def batched_chofactor(self, x, may_overwrite_x=False):
        return torch.linalg.cholesky(x, upper=False), True

    def batched_chosolve(self, L, y, lower, may_overwerite_b=False):
        return torch.cholesky_solve(y, L, upper=not lower)

    def batch_size_for_nn(self, num_nn, itemsize, kernel):
        if self.batch_size is not None:
            return self.batch_size
        return self.custom_batch_size_for_nn(num_nn, itemsize, kernel)

    def custom_batch_size_for_nn(self, num_nn, itemsize, kernel):
        return int(self.nn_large_size_preferred or 25000)

    def check_random_state(self, seed):
        if seed is None or isinstance(seed, Integral):
            g = torch.Generator(device=self.device)
            if seed is not None:
                g.manual_seed(int(seed))
            return g
        if not isinstance(seed, torch.Generator):
            raise ValueError("Seed is not a torch.Generator instance or an integer")
        return seed

    def random_normal(self, random_state, mu, sigma, shape):
        mu = self.xp.from_numpy(np.asarray(mu)).expand(shape)
        sigma = self.xp.from_numpy(np.asarray(sigma)).expand(shape)
        return torch.normal(mu, sigma, generator=random_state)


class Nngpr(GaussianProcessRegressor):
    """Nearest-neighbour GP regressor (Datta 2016 / Vecchia) on the HIP engine.

    Parameters are those of the reference (base_nngpr.py:127-191); see
    TorchNngpr / NumpyNngpr for the user-facing constructors.
    """

    _parameter_constraints: dict = {
        **GaussianProcessRegressor._parameter_constraints,
        "num_nn": [Interval(Integral, 1, None, closed="left")],
        "nn_type": [StrOptions({"kernel-space", "input-space"})],
        "batch_size": [Interval(Integral, 1, None, closed="left")],
        "allow_downcast_f32": [StrOptions({"yes", "no", "only-nn"})],
        "distribute_method": [StrOptions({"joblib", "multiprocessing",
                                          "multiprocessing-heterogenous", "sequential"})],
        "mp_start_method": [StrOptions(set(multiprocessing.get_all_start_methods()))],
        "random_gen_batch_byte_size": [Interval(Integral, 1, None, closed="left")],
    }

    def __init__(self, workers, kernel=None, *, alpha=1e-10, optimizer="fmin_l_bfgs_b",
                 n_restarts_optimizer=0, normalize_y=False, copy_X_train=True, random_state=None,
                 num_nn=None, nn_type=None, batch_size=None, allow_downcast_f32=None,
                 distribute_method=None, mp_start_method=None, random_gen_batch_byte_size=None,
                 distribute_environ=None):
        super().__init__(kernel=kernel, alpha=alpha, optimizer=optimizer,
                         n_restarts_optimizer=n_restarts_optimizer, normalize_y=normalize_y,
                         copy_X_train=copy_X_train, random_state=random_state)
        self.workers = workers
        self.num_nn = 32 if num_nn is None else num_nn
        self.nn_type = "kernel-space" if nn_type is None else nn_type
        self.allow_downcast_f32 = "only-nn" if allow_downcast_f32 is None else allow_downcast_f32
        fork = "fork" in multiprocessing.get_all_start_methods()
        self.distribute_method = (("multiprocessing" if fork else "joblib")
                                  if distribute_method is None else distribute_method)
        self.mp_start_method = (("fork" if fork else "spawn")
                                if mp_start_method is None else mp_start_method)
        self.random_gen_batch_byte_size = (1024 ** 3 if random_gen_batch_byte_size is None
                                           else random_gen_batch_byte_size)
        self.distribute_environ = distribute_environ
        if self.distribute_environ is None:
            self.distribute_environ = {"OPENBLAS_NUM_THREADS": "1", "MKL_NUM_THREADS": "1",
                                       "OMP_NUM_THREADS": "1"}
        self.batch_size = batch_size
        if self.batch_size is None:
            sizes = [w.nn_large_size_preferred for w in self.workers
                     if w.nn_large_size_preferred is not None]
            self.batch_size = max(1, min(sizes)) if sizes else 1000

    # ------------------------------------------------------------------ helpers
    @property
    def _worker(self):
        return self.workers[0]

    def get_default_kernel(self):
        """Fixed C(1) * RBF(1) (base_nngpr.py:1137-1151)."""
        return batched_kernels.ConstantKernel(1.0, constant_value_bounds="fixed") * \
            batched_kernels.RBF(1.0, length_scale_bounds="fixed")

    def __getstate__(self):
        """Pickled without the device-resident copies (rebuilt on first use)."""
        state = super().__getstate__()
        state.pop("_dev_cache", None)
        return state

    def _device_train(self):
        """Training arrays resident on the device (uploaded once per fit).

        With copy_X_train=False, X_train_ / y_train_ are the caller's arrays and
        may change in place after fit (scikit-learn's documented behaviour: the
        predictions follow them); outside fit's optimizer loop the cache is then
        keyed on a fingerprint of their contents too, so a change re-uploads the
        data and drops the cached neighbour lists."""
        key = (id(self.X_train_), id(self.y_train_))
        if not self.copy_X_train and not getattr(self, "_optimizing", False):
            key += (_fingerprint(self.X_train_), _fingerprint(self.y_train_))
        cache = getattr(self, "_dev_cache", None)
        if cache is None or cache[0] != key:
            xp = self._worker.get_array_module()
            x = xp.from_numpy(np.ascontiguousarray(self.X_train_, dtype=np.float64))
            y = self.y_train_
            if y.ndim == 1:
                y = y[:, None]
            y = xp.from_numpy(np.ascontiguousarray(y, dtype=np.float64))
            self._dev_cache = (key, x, y, {})
        return self._dev_cache[1], self._dev_cache[2], self._dev_cache[3]

    # extra candidates per row kept for reusing a search at other length scales
    _KNN_REUSE_MARGIN = 8
    # a base whose last rescale recomputed more than this fraction of the rows is
    # replaced by a new search at the next new scale
    _KNN_REBASE_FRACTION = 0.02

    def _train_knn(self, kernel, x_dev, knn_cache, precision):
        """Vecchia conditioning sets: each row's nearest earlier rows
        (base_nngpr.py:530-546, ref = x[:i1], query = x[i0:i1], ref_cut = i0).

        The reference searches again at every evaluation.  Here, in kernel space,
        one search keeps k + 8 exact candidates per row, and the lists at later
        length scales come from hip_ops.KnnRescaler: the same lists a new search
        would return (certified per row, uncertified rows searched exactly), for
        the cost of a copy under an isotropic change.  A base whose rescales
        recompute many rows is replaced by a search at the next new scale."""
        n = x_dev.shape[0]
        k = max(1, min(self.num_nn, n - 1))
        scale = self._worker.knn_scale(kernel, self.nn_type)
        key = (None if scale is None else tuple(scale.tolist()), k, self.nn_type)
        lists = knn_cache.setdefault("lists", {})
        nn = lists.get(key)
        if nn is not None:
            return nn
        lists.clear()
        K = k + self._KNN_REUSE_MARGIN
        # (reuse needs a search for K = k + margin candidates: only where that search
        # is as fast as the k-list search it replaces)
        if scale is None or not getattr(self, "_knn_reuse", True) or \
                K > min(hip_ops.RESCALE_MAX_K, hip_ops.KNN_FAST_MAX_K) or k > nat.NNGP_MAX_K:
            nn = _knn_lists(x_dev, x_dev, scale, k, ref_cut=0, precision=precision)
        else:
            base = knn_cache.get("rescaler")
            if base is not None and (base.k != k or base.scale.size != scale.size):
                base = None
            # the previous rescale's recomputed rows (read now: the stream has
            # synchronised since, for the log-likelihood value)
            if base is not None and knn_cache.pop("check", False) and \
                    base.rows_recomputed() > self._KNN_REBASE_FRACTION * n:
                base = None
            # A row stays certified when max_f(r_f)^2 times its m-th old key is below
            # min_f(r_f)^2 times its K-th (r_f = old / new length scale); with keys
            # growing like rank^(2/d) that fails for most rows once
            # max(r) / min(r) > (K / k)^(1/d), and every failing row is then searched
            # alone over all its refs: a new search is far cheaper.
            if base is not None and scale.size > 1:
                ratio = base.scale / scale
                if ratio.max() > ratio.min() * (K / k) ** (1.0 / x_dev.shape[1]):
                    base = None
            if base is None:
                base = hip_ops.KnnRescaler(x_dev, x_dev, scale, k, min(K, max(n - 1, k)),
                                           ref_cut=0, precision=precision)
                knn_cache["rescaler"] = base
                nn = base.base_lists()
            else:
                nn = base.lists(scale)
                knn_cache["check"] = True
        lists[key] = nn
        return nn

    # ---------------------------------------------------------------------- fit
    @sk_base._fit_context(prefer_skip_nested_validation=True)
    def fit(self, X, y):
        """Fit the regressor (base_nngpr.py:271-383): validation, normalize_y,
        L-BFGS-B over log-hyperparameters via scikit-learn's optimizer driver."""
        self.kernel_ = self.get_default_kernel() if self.kernel is None else sk_base.clone(self.kernel)
        self._rng = check_random_state(self.random_state)
        if self.kernel_.requires_vector_input:
            dtype, ensure_2d = "numeric", True
        else:
            dtype, ensure_2d = None, False
        X, y = validate_data(self, X, y, multi_output=True, y_numeric=True, ensure_2d=ensure_2d,
                             dtype=dtype)
        n_targets_seen = y.shape[1] if y.ndim > 1 else 1
        if self.n_targets is not None and n_targets_seen != self.n_targets:
            raise ValueError("The number of targets seen in `y` is different from the parameter "
                             "`n_targets`. Got %d != %d." % (n_targets_seen, self.n_targets))
        if self.normalize_y:
            self._y_train_mean = np.mean(y, axis=0)
            self._y_train_std = _handle_zeros_in_scale(np.std(y, axis=0), copy=False)
            y = (y - self._y_train_mean) / self._y_train_std
        else:
            shape = (y.shape[1],) if y.ndim == 2 else 1
            self._y_train_mean = np.zeros(shape=shape)
            self._y_train_std = np.ones(shape=shape)
        if np.iterable(self.alpha):
            if self.alpha.shape[0] == 1:
                self.alpha = self.alpha[0]
            else:
                raise ValueError("alpha must be a scalar or an array with only one element")
        self.X_train_ = np.copy(X) if self.copy_X_train else X
        self.y_train_ = np.copy(y) if self.copy_X_train else y
        # a new fit never reuses an earlier upload (the arrays may be the same
        # objects with new contents)
        self._dev_cache = None
        self._optimizing = True
        try:
            self._fit_hyperparameters()
        finally:
            self._optimizing = False
        for w in self.workers:
            w.clear_cache()
        return self

    def _fit_hyperparameters(self):
        """L-BFGS-B restarts of fit (base_nngpr.py:327-383) on the uploaded data."""
        if self.optimizer is not None and self.kernel_.n_dims > 0:
            def obj_func(theta, eval_gradient=True):
                if eval_gradient:
                    lml, grad = self.log_marginal_likelihood(theta, eval_gradient=True,
                                                             clone_kernel=False)
                    return -lml, -grad
                return -self.log_marginal_likelihood(theta, clone_kernel=False)

            optima = [self._constrained_optimization(obj_func, self.kernel_.theta,
                                                     self.kernel_.bounds)]
            if self.n_restarts_optimizer > 0:
                if not np.isfinite(self.kernel_.bounds).all():
                    raise ValueError("Multiple optimizer restarts (n_restarts_optimizer>0) "
                                     "requires that all bounds are finite.")
                bounds = self.kernel_.bounds
                for _ in range(self.n_restarts_optimizer):
                    theta0 = self._rng.uniform(bounds[:, 0], bounds[:, 1])
                    optima.append(self._constrained_optimization(obj_func, theta0, bounds))
            lml_values = list(map(itemgetter(1), optima))
            self.kernel_.theta = optima[np.argmin(lml_values)][0]
            self.kernel_._check_bounds_params()
            self.log_marginal_likelihood_value_ = -np.min(lml_values)
        else:
            self.log_marginal_likelihood_value_ = self.log_marginal_likelihood(
                self.kernel_.theta, clone_kernel=False)

    # ------------------------------------------------------------ log-likelihood
    def log_marginal_likelihood(self, theta=None, eval_gradient=False, clone_kernel=True):
        """Vecchia log-likelihood at theta (and its gradient), base_nngpr.py:385-451.

        One kNN pass (exact neighbour sets; cached while the search scale is
        unchanged) and one fused fill -> Cholesky -> mu/sigma -> loglik (+ adjoint
        gradient) pass over all training rows, reduced on the device.
        """
        if theta is None:
            if eval_gradient:
                raise ValueError("Gradient can only be evaluated for theta!=None")
            return self.log_marginal_likelihood_value_
        if clone_kernel:
            kernel = self.kernel_.clone_with_theta(theta)
        else:
            kernel = self.kernel_
            kernel.theta = theta
        x_dev, y_dev, cache = self._device_train()
        knn_prec, fac_prec = _precisions(self.allow_downcast_f32)
        ks, _ = self._worker.kernel_struct(kernel)
        nn = self._train_knn(kernel, x_dev, cache, knn_prec)
        ll, grad, _ = hip_ops.vecchia_loglik(x_dev, y_dev, nn, ks, float(self.alpha),
                                             precision=fac_prec, eval_gradient=eval_gradient,
                                             check_pd=False)
        # one transfer for the value and the gradient, then the call's status
        vals = (torch.cat([ll, grad]) if eval_gradient else ll).cpu().numpy()
        hip_ops.check_status("vecchia", "log_marginal_likelihood", x_dev.device)
        if eval_gradient:
            return float(vals[0]), vals[1:].astype(np.float64)
        return float(vals[0])

    # ------------------------------------------------------------------ predict
    def predict(self, X, return_std=False, return_cov=False, conditioning_method=None):
        """Only-train predictive mean (and std) of X (base_nngpr.py:763-894)."""
        conditioning_method = "only-train" if conditioning_method is None else conditioning_method
        if conditioning_method not in {"only-train", "full", "full-double-nn"}:
            raise ValueError("conditioning_method must be 'only-train', 'full' or 'full-double-nn'")
        if return_std and return_cov:
            raise RuntimeError("At most one of return_std or return_cov can be requested.")
        if self.kernel is None or self.kernel.requires_vector_input:
            dtype, ensure_2d = "numeric", True
        else:
            dtype, ensure_2d = None, False
        X = validate_data(self, X, ensure_2d=ensure_2d, dtype=dtype, reset=False)
        if return_cov or conditioning_method != "only-train":
            return self._predict_sequential(X, return_std, return_cov, conditioning_method)
        worker = self._worker
        xp = worker.get_array_module()
        if not hasattr(self, "X_train_"):
            kernel = self.kernel if self.kernel is not None else self.get_default_kernel()
            mean = np.zeros(X.shape[0])
            if return_std:
                if isinstance(kernel, batched_kernels.BatchedKernel):
                    diag = xp.to_numpy(kernel.diag(xp, xp.from_numpy(np.asarray(X, dtype=np.float64))))
                else:
                    diag = kernel.diag(X)
                return mean, np.sqrt(diag)
            return mean
        kernel = self.kernel_
        x_dev, y_dev, _ = self._device_train()
        knn_prec, fac_prec = _precisions(self.allow_downcast_f32)
        ks, _ = worker.kernel_struct(kernel)
        xq = xp.from_numpy(np.ascontiguousarray(X, dtype=np.float64))
        nt = x_dev.shape[0]
        k = max(1, min(nt, self.num_nn))
        nn = _knn_lists(x_dev, xq, worker.knn_scale(kernel, self.nn_type), k, ref_cut=None,
                         precision=knn_prec)
        mean_d, std_d = hip_ops.predict(x_dev, y_dev, xq, nn, ks, float(self.alpha),
                                        precision=fac_prec, return_std=return_std)
        mean = mean_d.cpu().numpy()
        ystd = np.asarray(self._y_train_std).astype(mean.dtype).reshape((1, -1))
        ymean = np.asarray(self._y_train_mean).astype(mean.dtype).reshape((1, -1))
        mean = ystd * mean + ymean
        if mean.shape[1] == 1:
            mean = mean[:, 0]
        if not return_std:
            return mean
        std = std_d.cpu().numpy()
        std = std.reshape((-1, 1)) * ystd
        if std.shape[1] == 1:
            std = std[:, 0]
        return mean, std

    # ------------------------------------------------- sequential conditioning
    def _conditioning_nn(self, kernel, x_train, xq, method, knn_prec):
        """Neighbour lists of the queries over [x_train; xq]
        (_find_nn_indices_for_train_and_eval, base_nngpr.py:1370-1464, without the
        leading self column): 'only-train' the nearest training rows; 'full' the
        nearest of the training rows and the earlier queries; 'full-double-nn' the
        nearest training rows, then the nearest earlier queries, valid entries first."""
        nt, nq = x_train.shape[0], xq.shape[0]
        dev = xq.device
        scale = self._worker.knn_scale(kernel, self.nn_type)
        empty = torch.empty((nq, 0), dtype=torch.int32, device=dev)
        if method == "only-train":
            k = min(nt, self.num_nn)
            return _knn_lists(x_train, xq, scale, k, precision=knn_prec) if k > 0 else empty
        if method == "full":
            k = min(self.num_nn, nt + nq)
            x_full = torch.cat([x_train, xq], dim=0)
            return _knn_lists(x_full, xq, scale, k, ref_cut=nt, precision=knn_prec)
        kq = min(self.num_nn, nq - 1)
        qn = _knn_lists(xq, xq, scale, kq, ref_cut=0, precision=knn_prec) if kq > 0 else empty
        qn = torch.where(qn >= 0, qn + nt, qn)
        if nt == 0:
            return qn
        kt = min(self.num_nn, nt)
        tr = _knn_lists(x_train, xq, scale, kt, precision=knn_prec)
        # valid training entries, then valid query entries, then -1 (the extra last
        # column takes the scatter of the empty query slots)
        merged = torch.full((nq, kt + kq + 1), -1, dtype=torch.int32, device=dev)
        merged[:, :kt] = tr
        if kq > 0:
            pos = (tr >= 0).sum(dim=1, keepdim=True) + torch.arange(kq, device=dev)[None, :]
            pos = torch.where(qn >= 0, pos, torch.full_like(pos, kt + kq))
            merged.scatter_(1, pos, qn)
        return merged[:, :kt + kq].contiguous()

    def _sequential_setup(self, X, method):
        worker = self._worker
        xp = worker.get_array_module()
        xq = xp.from_numpy(np.ascontiguousarray(X, dtype=np.float64))
        if hasattr(self, "X_train_"):
            kernel = self.kernel_
            x_dev, y_dev, _ = self._device_train()
        else:
            kernel = self.kernel if self.kernel is not None else self.get_default_kernel()
            x_dev = torch.empty((0, xq.shape[1]), dtype=torch.float64, device=xq.device)
            y_dev = torch.empty((0, 1), dtype=torch.float64, device=xq.device)
        knn_prec, _ = _precisions(self.allow_downcast_f32)
        ks, _ = worker.kernel_struct(kernel)
        nn = self._conditioning_nn(kernel, x_dev, xq, method, knn_prec)
        return x_dev, y_dev, xq, nn, ks

    def _y_scale(self, dtype):
        if not hasattr(self, "_y_train_std"):
            return None, None
        return (np.asarray(self._y_train_std).astype(dtype).reshape((1, -1)),
                np.asarray(self._y_train_mean).astype(dtype).reshape((1, -1)))

    def _predict_sequential(self, X, return_std, return_cov, method):
        """Sequential branch of predict (base_nngpr.py:896-1030) on the HIP engine:
        mean, the sparse-partial-covariance std, or the joint covariance."""
        x_dev, y_dev, xq, nn, ks = self._sequential_setup(X, method)
        output = hip_ops.COND_COV if return_cov else (hip_ops.COND_STD if return_std
                                                      else hip_ops.COND_MEAN)
        mean_d, std_d, cov_d = hip_ops.predict_sequential(x_dev, y_dev, xq, nn, ks,
                                                          float(self.alpha), output)
        mean = mean_d.cpu().numpy()
        ystd, ymean = self._y_scale(mean.dtype)
        if ystd is not None:
            mean = ystd * mean + ymean
        if mean.shape[1] == 1:
            mean = mean[:, 0]
        if return_std:
            std = std_d.cpu().numpy().reshape((-1, 1))
            if ystd is not None:
                std = std * ystd
            return mean, (std[:, 0] if std.shape[1] == 1 else std)
        if return_cov:
            cov = cov_d.cpu().numpy()
            if ystd is not None:
                cov = cov[:, :, None] * (ystd.reshape((1, 1, -1)) ** 2)
                if cov.shape[2] == 1:
                    cov = cov[:, :, 0]
            return mean, cov
        return mean

    def _generator(self, random_state, dev):
        """torch generator on the device seeded from random_state (None, int or a
        numpy RandomState, as check_random_state accepts)."""
        g = torch.Generator(device=dev)
        if random_state is None:
            g.seed()
        elif isinstance(random_state, (Integral, np.integer)):
            g.manual_seed(int(random_state))
        else:
            g.manual_seed(int(check_random_state(random_state).randint(0, 2 ** 31 - 1)))
        return g

    def sample_y(self, X, n_samples=1, random_state=0, conditioning_method=None):
        """Samples of the process at X (base_nngpr.py:616-761): 'only-train' draws
        independent normals around the only-train mean/std; the sequential methods
        draw each query given the samples already drawn at its earlier query
        neighbours.  Random draws come from a torch generator on the device seeded
        by random_state; the draws differ from the reference's generators."""
        conditioning_method = "only-train" if conditioning_method is None else conditioning_method
        if conditioning_method not in {"only-train", "full", "full-double-nn"}:
            raise ValueError("conditioning_method must be 'only-train', 'full' or 'full-double-nn'")
        if self.kernel is None or self.kernel.requires_vector_input:
            dtype, ensure_2d = "numeric", True
        else:
            dtype, ensure_2d = None, False
        X = validate_data(self, X, ensure_2d=ensure_2d, dtype=dtype, reset=False)
        nq = X.shape[0]
        dev = self._worker.get_array_module().torch_device
        gen = self._generator(random_state, dev)
        if conditioning_method == "only-train":
            mu, sigma = self.predict(X, return_std=True)
            mu = np.asarray(mu).reshape((nq, -1))
            sigma = np.asarray(sigma).reshape((nq, -1))
            z = torch.randn((nq, mu.shape[1], n_samples), generator=gen, dtype=torch.float64,
                            device=dev).cpu().numpy()
            y = mu[:, :, None] + sigma[:, :, None] * z
            return y[:, 0, :] if y.shape[1] == 1 else y
        x_dev, y_dev, xq, nn, ks = self._sequential_setup(X, conditioning_method)
        ny = y_dev.shape[1]
        z = torch.randn((nq, ny, n_samples), generator=gen, dtype=torch.float64, device=dev)
        y = hip_ops.sample_sequential(x_dev, y_dev, xq, nn, ks, float(self.alpha), z)
        y = y.cpu().numpy()
        ystd, ymean = self._y_scale(y.dtype)
        if ystd is not None:
            y = y * ystd[:, :, None] + ymean[:, :, None]
        return y[:, 0, :] if ny == 1 else y
