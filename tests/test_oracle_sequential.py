"""Pins the sequential-conditioning oracle (oracle/sequential_oracle.py).

  exact sets      with neighbour lists holding every training row and every
                  earlier query ('full', 'full-double-nn'), the sequential mean,
                  std and joint covariance are the exact GP posterior: scikit-learn's
                  predict(return_std / return_cov); 'only-train' gives the exact
                  per-query variances on a diagonal covariance
  closed form     at small m the joint covariance equals (I - B)^-1 L (I - B)^-T and
                  the mean (I - B)^-1 t, with B the query-neighbour weights, L the
                  conditional variances and t the training part, assembled densely
                  from per-query solves; samples equal mean + (I - B)^-1 L^1/2 z
  sparse std      where every pair of a query's query neighbours is itself a
                  neighbour pair (points on a line), the sparse-partial-covariance
                  std equals the joint covariance's diagonal; dropping the
                  off-diagonal terms (a planted mistake) breaks that
"""
import numpy as np
import pytest
from sklearn import gaussian_process as skgp
from sklearn.gaussian_process import kernels as skk

from oracle import sequential_oracle as so
from oracle import vecchia_oracle as vo

W, C, S, LS, ALPHA = 0.05, 0.2, 1.3, 0.35, 1e-6


def sk_kernel(nu):
    st = skk.RBF(LS) if nu is None else skk.Matern(LS, nu=nu)
    return skk.ConstantKernel(S) * st + skk.ConstantKernel(C, "fixed") + skk.WhiteKernel(W)


def data(nt=40, nq=15, d=2, seed=0):
    rng = np.random.default_rng(seed)
    x = rng.random((nt, d))
    xq = rng.random((nq, d))
    y = np.column_stack([np.sin(5 * x[:, 0]), x[:, 1] ** 2]) + rng.normal(0, 0.05, (nt, 2))
    return x, y, xq


@pytest.mark.parametrize("nu", [None, 1.5])
@pytest.mark.parametrize("method", ["full", "full-double-nn"])
def test_exact_sets_equal_exact_gp(method, nu):
    x, y, xq = data()
    gpr = skgp.GaussianProcessRegressor(sk_kernel(nu), alpha=ALPHA, optimizer=None).fit(x, y)
    mean_ref, std_ref = gpr.predict(xq, return_std=True)
    _, cov_ref = gpr.predict(xq, return_cov=True)
    nn = so.conditioning_nn(x, xq, LS, 100, method)
    mean, std = so.predict_sequential(x, y, xq, nn, W, C, S, LS, ALPHA, "std", nu)
    np.testing.assert_allclose(mean, mean_ref, atol=1e-9)
    np.testing.assert_allclose(std, std_ref[:, 0] if std_ref.ndim > 1 else std_ref, atol=1e-9)
    mean2, cov = so.predict_sequential(x, y, xq, nn, W, C, S, LS, ALPHA, "cov", nu)
    np.testing.assert_allclose(mean2, mean_ref, atol=1e-9)
    np.testing.assert_allclose(cov, cov_ref[:, :, 0] if cov_ref.ndim > 2 else cov_ref, atol=1e-9)


def test_only_train_cov_is_diagonal_exact_variance():
    x, y, xq = data()
    gpr = skgp.GaussianProcessRegressor(sk_kernel(None), alpha=ALPHA, optimizer=None).fit(x, y[:, 0])
    mean_ref, cov_ref = gpr.predict(xq, return_cov=True)
    nn = so.conditioning_nn(x, xq, LS, 100, "only-train")
    mean, cov = so.predict_sequential(x, y[:, 0], xq, nn, W, C, S, LS, ALPHA, "cov")
    np.testing.assert_allclose(mean[:, 0], mean_ref, atol=1e-9)
    np.testing.assert_allclose(np.diag(cov), np.diag(cov_ref), atol=1e-9)
    assert np.all(cov[~np.eye(len(xq), dtype=bool)] == 0)


def dense_closed_form(x, y, xq, nn, nu=None):
    """(I - B), Lambda and the training part t from per-query dense solves."""
    nt, nq = x.shape[0], xq.shape[0]
    x_full = np.concatenate([x, xq])
    y = y.reshape((nt, -1))
    B = np.zeros((nq, nq))
    lam = np.zeros(nq)
    t = np.zeros((nq, y.shape[1]))
    for i in range(nq):
        row = nn[i][nn[i] >= 0]
        pts = np.concatenate([[nt + i], row])
        K, _ = vo.fill(x_full[pts][None], np.ones((1, pts.size), bool), W, C, S, LS, nu=nu)
        K = K[0]
        K[1:, 1:] += ALPHA * np.diag(row < nt)
        wv = np.linalg.solve(K[1:, 1:], K[1:, 0])
        lam[i] = max(0.0, K[0, 0] - K[1:, 0] @ wv)
        for p, j in enumerate(row):
            if j < nt:
                t[i] += wv[p] * y[j]
            else:
                B[i, j - nt] = wv[p]
    return np.eye(nq) - B, lam, t


@pytest.mark.parametrize("nu", [None, 0.5])
@pytest.mark.parametrize("method", ["full", "full-double-nn"])
def test_small_m_matches_closed_form(method, nu):
    x, y, xq = data(nt=60, nq=40, seed=3)
    nn = so.conditioning_nn(x, xq, LS, 5, method)
    IB, lam, t = dense_closed_form(x, y, xq, nn, nu)
    Minv = np.linalg.inv(IB)
    mean, cov = so.predict_sequential(x, y, xq, nn, W, C, S, LS, ALPHA, "cov", nu)
    np.testing.assert_allclose(mean, Minv @ t, atol=1e-10)
    np.testing.assert_allclose(cov, Minv @ np.diag(lam) @ Minv.T, atol=1e-10)
    z = np.random.default_rng(5).normal(size=(40, 2, 7))
    smp = so.sample_sequential(x, y, xq, nn, W, C, S, LS, ALPHA, z, nu)
    M = Minv @ np.diag(np.sqrt(lam))
    np.testing.assert_allclose(smp, (Minv @ t)[:, :, None] + np.einsum("ij,jos->ios", M, z),
                               atol=1e-10)


def line_case():
    nt, nq = 20, 30
    x = np.linspace(10.0, 11.0, nt)[:, None]
    xq = np.linspace(0.0, 1.0, nq)[:, None]
    y = np.cos(x[:, 0])
    return x, y, xq


def test_sparse_std_on_a_line_equals_joint_diagonal():
    x, y, xq = line_case()
    nn = so.conditioning_nn(x, xq, 0.3, 4, "full")
    assert np.all(nn[4:] >= 20)                    # the four previous queries
    _, std = so.predict_sequential(x, y, xq, nn, W, C, S, 0.3, ALPHA, "std")
    _, cov = so.predict_sequential(x, y, xq, nn, W, C, S, 0.3, ALPHA, "cov")
    np.testing.assert_allclose(std, np.sqrt(np.diag(cov)), rtol=1e-10)
    _, bad = so.predict_sequential(x, y, xq, nn, W, C, S, 0.3, ALPHA, "std", _offdiag=False)
    assert not np.allclose(bad, std, rtol=1e-6)


def test_neighbour_lists():
    x, _, xq = data(nt=30, nq=12)
    full = so.conditioning_nn(x, xq, LS, 6, "full")
    dbl = so.conditioning_nn(x, xq, LS, 6, "full-double-nn")
    for i in range(12):
        assert np.all((full[i] < 30 + i))          # training rows and earlier queries only
        tr, qq = dbl[i][:6], dbl[i][6:]
        assert np.all(tr < 30) and np.all((qq == -1) | ((qq >= 30) & (qq < 30 + i)))
        assert np.sum(qq >= 0) == min(6, i)
    # prior: no training rows
    nn0 = so.conditioning_nn(np.zeros((0, 2)), xq, LS, 4, "full-double-nn")
    assert nn0.shape == (12, 4) and np.all(nn0[0] == -1)
