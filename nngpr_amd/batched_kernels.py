"""Batched GP kernels over any numpy-like array module (numpy, torch on a device).

The in-scope family of the reference's batched kernels (batched_kernels.py:10-297,
545-1131): WhiteKernel, ConstantKernel, RBF and Matern (isotropic or ARD), Sum,
Product.  Each
class subclasses its scikit-learn twin, so hyperparameters, theta (log
parameters), bounds and cloning behave exactly like scikit-learn; evaluation takes
the array module first and accepts inputs of shape (*batch, n, d):

    K = kernel(xp, X)                        # (*batch, n, n)
    K, K_grad = kernel(xp, X, eval_gradient=True)   # K_grad (*batch, n, n, n_theta)
    kernel.diag(xp, X)                       # (*batch, n)

The HIP hot path reads only the decomposed parameters of these kernels
(nngpr_amd.kernel_params); evaluation here serves the generic worker surface and
the scikit-learn compatible API.
"""
import numpy as np
from sklearn.base import clone
from sklearn.gaussian_process import kernels


class BatchedKernel(kernels.Kernel):
    """Base class; operators build batched Sum/Product trees."""

    def __add__(self, b):
        if not isinstance(b, BatchedKernel):
            return Sum(self, ConstantKernel(b))
        return Sum(self, b)

    def __radd__(self, b):
        if not isinstance(b, BatchedKernel):
            return Sum(ConstantKernel(b), self)
        return Sum(b, self)

    def __mul__(self, b):
        if not isinstance(b, BatchedKernel):
            return Product(self, ConstantKernel(b))
        return Product(self, b)

    def __rmul__(self, b):
        if not isinstance(b, BatchedKernel):
            return Product(ConstantKernel(b), self)
        return Product(b, self)

    def bind_to_array_module(self, xp):
        w = BatchedKernelWrapper(self)
        w.xp = xp
        return w


class BatchedKernelWrapper(kernels.Kernel):
    """A BatchedKernel bound to one array module, callable with the plain
    scikit-learn signature kernel(X, Y=None, eval_gradient=False)."""

    def __init__(self, base_kernel):
        self.base_kernel = base_kernel
        self.xp = None

    def __call__(self, X, Y=None, eval_gradient=False):
        return self.base_kernel(self.xp, X, Y, eval_gradient=eval_gradient)

    def diag(self, X):
        return self.base_kernel.diag(self.xp, X)

    def is_stationary(self):
        return self.base_kernel.is_stationary()

    @property
    def theta(self):
        return self.base_kernel.theta

    @theta.setter
    def theta(self, theta):
        self.base_kernel.theta = theta

    @property
    def bounds(self):
        return self.base_kernel.bounds

    @property
    def hyperparameters(self):
        return self.base_kernel.hyperparameters

    def clone_with_theta(self, theta):
        w = BatchedKernelWrapper(self.base_kernel.clone_with_theta(theta))
        w.xp = self.xp
        return w

    def get_params(self, deep=True):
        return self.base_kernel.get_params(deep=deep)

    def set_params(self, **params):
        self.base_kernel.set_params(**params)
        return self

    def __getattr__(self, name):
        if name in ("base_kernel", "xp"):
            raise AttributeError(name)
        return getattr(self.base_kernel, name)

    def __repr__(self):
        return repr(self.base_kernel)

    def __sklearn_clone__(self):
        w = BatchedKernelWrapper(clone(self.base_kernel))
        w.xp = self.xp
        return w


def _concat(xp, parts, dtype_like):
    return xp.concatenate(parts, axis=-1)


class Sum(BatchedKernel, kernels.Sum):
    def __call__(self, xp, X, Y=None, eval_gradient=False):
        if eval_gradient:
            K1, G1 = self.k1(xp, X, Y, eval_gradient=True)
            K2, G2 = self.k2(xp, X, Y, eval_gradient=True)
            return K1 + K2, xp.concatenate((G1, G2), axis=-1)
        return self.k1(xp, X, Y) + self.k2(xp, X, Y)

    def diag(self, xp, X):
        return self.k1.diag(xp, X) + self.k2.diag(xp, X)


class Product(BatchedKernel, kernels.Product):
    def __call__(self, xp, X, Y=None, eval_gradient=False):
        if eval_gradient:
            K1, G1 = self.k1(xp, X, Y, eval_gradient=True)
            K2, G2 = self.k2(xp, X, Y, eval_gradient=True)
            return K1 * K2, xp.concatenate((G1 * K2[..., None], G2 * K1[..., None]), axis=-1)
        return self.k1(xp, X, Y) * self.k2(xp, X, Y)

    def diag(self, xp, X):
        return self.k1.diag(xp, X) * self.k2.diag(xp, X)


class ConstantKernel(BatchedKernel, kernels.ConstantKernel):
    def __call__(self, xp, X, Y=None, eval_gradient=False):
        if Y is None:
            Y = X
        elif eval_gradient:
            raise ValueError("Gradient can only be evaluated when Y is None.")
        if tuple(X.shape[:-2]) != tuple(Y.shape[:-2]):
            raise ValueError("X and Y must have the same batch dimensions.")
        K = xp.full((*X.shape[:-1], Y.shape[-2]), self.constant_value, dtype=X.dtype)
        if not eval_gradient:
            return K
        n_dims = 0 if self.hyperparameter_constant_value.fixed else 1
        return K, xp.full((*X.shape[:-1], X.shape[-2], n_dims), self.constant_value, dtype=X.dtype)

    def diag(self, xp, X):
        return xp.full(X.shape[:-1], self.constant_value, dtype=X.dtype)


class WhiteKernel(BatchedKernel, kernels.WhiteKernel):
    def __call__(self, xp, X, Y=None, eval_gradient=False):
        if Y is not None and eval_gradient:
            raise ValueError("Gradient can only be evaluated when Y is None.")
        if Y is None:
            n = X.shape[-2]
            K = self.noise_level * xp.tile(xp.eye(n, dtype=X.dtype), (*X.shape[:-2], 1, 1))
            if not eval_gradient:
                return K
            if self.hyperparameter_noise_level.fixed:
                return K, xp.empty((*X.shape[:-1], n, 0), dtype=X.dtype)
            return K, xp.asarray(K[..., None], copy=True)
        if tuple(X.shape[:-2]) != tuple(Y.shape[:-2]):
            raise ValueError("X and Y must have the same batch dimensions.")
        return xp.zeros((*X.shape[:-1], Y.shape[-2]), dtype=X.dtype)

    def diag(self, xp, X):
        return xp.full(X.shape[:-1], self.noise_level, dtype=X.dtype)


def _check_length_scale(X, length_scale):
    ls = np.squeeze(length_scale).astype(float)
    if np.ndim(ls) > 1:
        raise ValueError("length_scale cannot be of dimension greater than 1")
    if np.ndim(ls) == 1 and X.shape[-1] != ls.shape[0]:
        raise ValueError("Anisotropic kernel must have the same number of dimensions as data "
                         "(%d!=%d)" % (ls.shape[0], X.shape[-1]))
    return ls


class RBF(BatchedKernel, kernels.RBF):
    def __call__(self, xp, X, Y=None, eval_gradient=False):
        ls = xp.asarray(_check_length_scale(X, self.length_scale), dtype=X.dtype)
        Xs = X / ls
        if Y is None:
            diff = Xs[..., :, None, :] - Xs[..., None, :, :]
            sq = diff * diff
            d2 = sq.sum(-1)
            K = xp.exp(-0.5 * d2)
        else:
            if eval_gradient:
                raise ValueError("Gradient can only be evaluated when Y is None.")
            Ys = Y / ls
            diff = Xs[..., :, None, :] - Ys[..., None, :, :]
            K = xp.exp(-0.5 * (diff * diff).sum(-1))
            return K
        if not eval_gradient:
            return K
        if self.hyperparameter_length_scale.fixed:
            return K, xp.empty((*X.shape[:-1], X.shape[-2], 0), dtype=X.dtype)
        if not self.anisotropic or np.ndim(_check_length_scale(X, self.length_scale)) == 0 \
                or np.size(self.length_scale) == 1:
            return K, (K * d2)[..., None]
        return K, K[..., None] * sq

    def diag(self, xp, X):
        return xp.ones(X.shape[:-1], dtype=X.dtype)


class Matern(RBF, kernels.Matern):
    """Matern kernel for nu in {0.5, 1.5, 2.5, inf} (batched_kernels.py:923-1131);
    any other nu raises NotImplementedError, as in the reference (:1081, :1120).

    With r the distance in length-scale units and a = sqrt(2 nu) r:
    nu = 0.5: exp(-r); 1.5: (1 + a) exp(-a); 2.5: (1 + a + a^2/3) exp(-a);
    inf: exp(-r^2 / 2).  d K / d log l_f = G * (dx_f / l_f)^2 with G = K / r (0 at
    r = 0), 3 exp(-a), 5/3 (1 + a) exp(-a) and K respectively.
    """

    def __call__(self, xp, X, Y=None, eval_gradient=False):
        ls = xp.asarray(_check_length_scale(X, self.length_scale), dtype=X.dtype)
        Xs = X / ls
        if Y is None:
            diff = Xs[..., :, None, :] - Xs[..., None, :, :]
        else:
            if eval_gradient:
                raise ValueError("Gradient can only be evaluated when Y is None.")
            diff = Xs[..., :, None, :] - (Y / ls)[..., None, :, :]
        sq = diff * diff
        d2 = sq.sum(-1)
        nu = self.nu
        if nu == 0.5:
            r = xp.sqrt(d2)
            K = xp.exp(-r)
        elif nu == 1.5:
            a = xp.sqrt(3.0 * d2)
            K = (1.0 + a) * xp.exp(-a)
        elif nu == 2.5:
            a = xp.sqrt(5.0 * d2)
            K = (1.0 + a + a * a / 3.0) * xp.exp(-a)
        elif np.isinf(nu):
            K = xp.exp(-0.5 * d2)
        else:
            raise NotImplementedError("Matern nu=%r is supported only for 0.5, 1.5, 2.5 and inf"
                                      % nu)
        if Y is not None or not eval_gradient:
            return K
        if self.hyperparameter_length_scale.fixed:
            return K, xp.empty((*X.shape[:-1], X.shape[-2], 0), dtype=X.dtype)
        if nu == 0.5:
            G = xp.where(r > 0, K / xp.where(r > 0, r, xp.ones_like(r)), xp.zeros_like(r))
        elif nu == 1.5:
            G = 3.0 * xp.exp(-a)
        elif nu == 2.5:
            G = 5.0 / 3.0 * (1.0 + a) * xp.exp(-a)
        else:
            G = K
        if not self.anisotropic or np.ndim(_check_length_scale(X, self.length_scale)) == 0 \
                or np.size(self.length_scale) == 1:
            return K, (G * d2)[..., None]
        return K, G[..., None] * sq
