"""Seeded synthetic workloads used by the benchmark and the tests.

Each generator returns a dict with x_train, y_train, x_test (when defined), the
kernel to use and the neighbour count.  Shapes and recipes:

  small_uniform      n=2,000 train / 500 test, d=3, x ~ U[0,1]^3,
                     y = sin(6 x0) + cos(4 x1) x2 + N(0, 0.1^2),
                     White(0.01) + C(1.0) * RBF(0.3) (all fixed), m=16
  road_mixture       n=434,874 train / 43,487 test, d=3: (x0, x1) from a
                     64-component isotropic Gaussian mixture in the unit square
                     (means U[0.05, 0.95]^2, std U[0.02, 0.1], Dirichlet(1) weights,
                     reflected into [0, 1]), x2 = 0.5 + 0.25 sin(2 pi x0) cos(2 pi x1)
                     + N(0, 0.05^2); y = standardised(sin(3 x0) + cos(5 x1) + x2)
                     + N(0, 0.1^2); White(0.01) + C(1.0) * RBF(0.1), m=32
  correlated_gauss   n=16,599, d=18, x ~ N(0, S) with S = Q diag(l) Q^T,
                     l log-uniform in [0.1, 10]; y = standardised sum of 200 random
                     cosine features + N(0, 0.05^2); White + C * RBF(ARD, 18), m=32
  uniform_8d         n=2,000,000, d=8, x ~ U[0,1]^8, y = sum_k sin(w_k . x) over 16
                     w_k ~ N(0, 9 I) + N(0, 0.1^2); White(0.01) + C(1.0) * RBF(0.3)
The road-mixture size can be reduced with `n_train`/`n_test` (same generator
stream, first rows).
"""
import numpy as np

from . import batched_kernels as bk


def small_uniform(seed=0):
    rng = np.random.default_rng(seed)
    xtr = rng.random((2000, 3))
    xte = rng.random((500, 3))
    ytr = np.sin(6 * xtr[:, 0]) + np.cos(4 * xtr[:, 1]) * xtr[:, 2] + rng.normal(0, 0.1, 2000)
    kernel = bk.WhiteKernel(0.01, noise_level_bounds="fixed") + \
        bk.ConstantKernel(1.0, constant_value_bounds="fixed") * bk.RBF(0.3, length_scale_bounds="fixed")
    return dict(x_train=xtr, y_train=ytr, x_test=xte, kernel=kernel, num_nn=16, alpha=1e-10)


def _mixture_xy(rng, n):
    k = 64
    means = rng.uniform(0.05, 0.95, size=(k, 2))
    stds = rng.uniform(0.02, 0.1, size=k)
    w = rng.dirichlet(np.ones(k))
    comp = rng.choice(k, size=n, p=w)
    xy = means[comp] + rng.normal(size=(n, 2)) * stds[comp, None]
    xy = np.abs(xy)
    xy = 1.0 - np.abs(1.0 - xy)
    xy = np.clip(xy, 0.0, 1.0)
    x2 = 0.5 + 0.25 * np.sin(2 * np.pi * xy[:, 0]) * np.cos(2 * np.pi * xy[:, 1]) + \
        rng.normal(0, 0.05, n)
    x = np.column_stack([xy, x2])
    f = np.sin(3 * x[:, 0]) + np.cos(5 * x[:, 1]) + x[:, 2]
    f = (f - f.mean()) / f.std()
    y = f + rng.normal(0, 0.1, n)
    return x, y


def road_mixture(seed=2, n_train=434_874, n_test=43_487):
    rng = np.random.default_rng(seed)
    x, y = _mixture_xy(rng, 434_874 + 43_487)
    perm = rng.permutation(x.shape[0])
    x, y = x[perm], y[perm]
    xtr, ytr = x[:434_874][:n_train], y[:434_874][:n_train]
    xte = x[434_874:][:n_test]
    kernel = bk.WhiteKernel(0.01) + bk.ConstantKernel(1.0) * bk.RBF(0.1)
    return dict(x_train=np.ascontiguousarray(xtr), y_train=np.ascontiguousarray(ytr),
                x_test=np.ascontiguousarray(xte), kernel=kernel, num_nn=32, alpha=1e-10)


def correlated_gauss(seed=3, n=16_599):
    rng = np.random.default_rng(seed)
    d = 18
    q, _ = np.linalg.qr(rng.normal(size=(d, d)))
    lam = np.exp(rng.uniform(np.log(0.1), np.log(10.0), size=d))
    cov = (q * lam) @ q.T
    x = rng.multivariate_normal(np.zeros(d), cov, size=n)
    wk = rng.normal(0, 0.5, size=(200, d))
    ak = rng.normal(size=200)
    bk_ = rng.uniform(0, 2 * np.pi, size=200)
    f = np.cos(x @ wk.T + bk_) @ ak
    f = (f - f.mean()) / f.std()
    y = f + rng.normal(0, 0.05, n)
    kernel = bk.WhiteKernel(0.1) + bk.ConstantKernel(1.0) * bk.RBF(np.full(d, 2.0))
    return dict(x_train=x, y_train=y, x_test=None, kernel=kernel, num_nn=32, alpha=1e-10)


def uniform_8d(seed=4, n=2_000_000, num_nn=32):
    rng = np.random.default_rng(seed)
    x = rng.random((n, 8))
    w = rng.normal(0, 3.0, size=(16, 8))
    y = np.sin(x @ w.T).sum(-1) + rng.normal(0, 0.1, n)
    kernel = bk.WhiteKernel(0.01) + bk.ConstantKernel(1.0) * bk.RBF(0.3)
    return dict(x_train=x, y_train=y, x_test=None, kernel=kernel, num_nn=num_nn, alpha=1e-10)
