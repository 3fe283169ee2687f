"""Pins the FP64 oracle before anything is compared against it.

Each oracle function is checked against something it does not share code with:
  knn             numpy lexsort of (key, index) on exact-key / well-separated inputs;
                  the reference's own ref_cut cases (None, 0, 2, -1)
  fill            scikit-learn's kernels on the gathered points (values and gradients),
                  RBF and Matern nu = 1/2, 3/2, 5/2 (iso and ARD, duplicate points)
  vecchia_loglik  scikit-learn's exact GaussianProcessRegressor when m >= n - 1
                  (the reference's own known-answer test, nngpr_base_tester.py:174-192);
                  dense conditional Gaussians by brute force when m < n;
                  central finite differences of its own value for the gradient
  adjoint grad    the reference gradient formula
  predict         scikit-learn's exact predictive mean/std when m >= n
                  (nngpr_base_tester.py:228-257)
Planted mistakes (wrong sign in the white derivative, dropped alpha, transposed
neighbour order) are shown to fail these checks.
"""
import numpy as np
import pytest
from sklearn import gaussian_process as skgp
from sklearn.gaussian_process import kernels as skk

from oracle import vecchia_oracle as vo


def ref_test_data():
    # Data recipe of the reference's estimator tests (nngpr_base_tester.py:120-133).
    rng = np.random.RandomState(234058)
    x = rng.rand(150, 4)
    x_eval = rng.rand(50, 4)
    beta = rng.rand(4, 3)
    beta[:, -1] += 1
    y = x @ beta + rng.rand(150, 3)
    y[:, 0] *= 10
    return x, x_eval, y


# kernel C(c) * RBF(l) + White(w): theta = log([c, l, w]) in sklearn order
def sk_kernel(c=1.0, l=1.0, w=1.0):
    return skk.ConstantKernel(c) * skk.RBF(l) + skk.WhiteKernel(w)


THETA_MAP = {"rbf_level": 0, "length_scale": 1, "noise_level": 2}


# ----------------------------------------------------------------------------- knn
@pytest.mark.parametrize("cut", [None, 0, 2, -1])
@pytest.mark.parametrize("ls", [None, 0.37, "ard"])
def test_knn_matches_lexsort(cut, ls):
    rng = np.random.default_rng(11)
    ref = rng.random((300, 3))
    qry = rng.random((120, 3)) if cut is None else ref[:120]
    if ls == "ard":
        ls = np.array([0.3, 0.9, 2.0])
    a = vo.knn(ref, qry, ls, 9, ref_cut=cut)
    b = vo.knn_lexsort(ref, qry, ls, 9, ref_cut=cut)
    np.testing.assert_array_equal(a, b)


def test_knn_ties_keep_lower_index():
    # Integer grid with duplicates: every key is exact, many keys tie.
    rng = np.random.default_rng(5)
    pts = rng.integers(0, 4, size=(200, 2)).astype(np.float64)
    a = vo.knn(pts, pts, None, 12, ref_cut=0)
    b = vo.knn_lexsort(pts, pts, None, 12, ref_cut=0)
    np.testing.assert_array_equal(a, b)
    # all-equal points: row i takes refs 0..min(i,k)-1 in order (reference test shape)
    ones = np.ones((10, 3))
    c = vo.knn(ones, ones, 1.0, 4, ref_cut=0)
    for i in range(10):
        exp = list(range(min(i, 4))) + [-1] * (4 - min(i, 4))
        assert list(c[i]) == exp


def test_knn_vecchia_padding_and_earlier_only():
    rng = np.random.default_rng(1)
    x = rng.random((50, 2))
    nn = vo.vecchia_knn(x, 0.5, 8)
    for i in range(50):
        valid = nn[i][nn[i] >= 0]
        assert len(valid) == min(i, 8)
        assert np.all(valid < i)
        assert np.all(nn[i][len(valid):] == -1)


def test_knn_planted_order_mistake_fails():
    rng = np.random.default_rng(3)
    ref = rng.random((100, 3))
    good = vo.knn_lexsort(ref, ref[:20], None, 5)
    bad = good[:, ::-1]
    assert not np.array_equal(vo.knn(ref, ref[:20], None, 5), bad)


# ---------------------------------------------------------------------------- fill
def test_fill_matches_sklearn():
    rng = np.random.default_rng(2)
    xg = rng.random((6, 7, 3))
    valid = np.ones((6, 7), dtype=bool)
    valid[2, 5] = False
    c, l, w = 1.7, 0.6, 0.3
    K, G = vo.fill(xg, valid, w, 0.0, c, l, THETA_MAP, 3, eval_gradient=True)
    k = sk_kernel(c, l, w)
    for b in range(6):
        Kr, Gr = k(xg[b], eval_gradient=True)
        v = valid[b]
        np.testing.assert_allclose(K[b][np.ix_(v, v)], Kr[np.ix_(v, v)], rtol=1e-13, atol=1e-15)
        np.testing.assert_allclose(G[b][np.ix_(v, v)], Gr[np.ix_(v, v)], rtol=1e-12, atol=1e-15)
        inv = ~v
        if inv.any():
            j = np.where(inv)[0][0]
            assert K[b][j, j] == 1.0 and np.all(K[b][j, v] == 0) and np.all(G[b][j] == 0)


def test_fill_ard_matches_sklearn():
    rng = np.random.default_rng(4)
    xg = rng.random((3, 5, 4))
    ls = np.array([0.5, 1.5, 0.8, 2.5])
    K, G = vo.fill(xg, np.ones((3, 5), bool), 0.2, 0.0, 1.3, ls,
                   {"rbf_level": 0, "length_scale": np.arange(1, 5), "noise_level": 5}, 6, True)
    k = skk.ConstantKernel(1.3) * skk.RBF(ls) + skk.WhiteKernel(0.2)
    for b in range(3):
        Kr, Gr = k(xg[b], eval_gradient=True)
        np.testing.assert_allclose(K[b], Kr, rtol=1e-13)
        np.testing.assert_allclose(G[b], Gr, rtol=1e-12, atol=1e-15)


MATERN_NU = [0.5, 1.5, 2.5]


def matern_case(nu, ard, d=3):
    ls = np.linspace(0.4, 1.1, d) if ard else 0.45
    k = skk.ConstantKernel(1.3) * skk.Matern(ls, nu=nu) + skk.WhiteKernel(0.2)
    tm = {"rbf_level": 0, "length_scale": np.arange(1, d + 1) if ard else 1,
          "noise_level": d + 1 if ard else 2}
    return k, ls, tm, (d + 2 if ard else 3)


@pytest.mark.parametrize("nu", MATERN_NU)
@pytest.mark.parametrize("ard", [False, True])
def test_fill_matern_matches_sklearn(nu, ard):
    """Matern values and log-parameter gradients against scikit-learn's Matern,
    with a duplicated point (r = 0: the nu = 1/2 derivative is 0 there)."""
    rng = np.random.default_rng(21)
    xg = rng.random((4, 6, 3))
    xg[1, 4] = xg[1, 2]
    k, ls, tm, nt = matern_case(nu, ard)
    K, G = vo.fill(xg, np.ones((4, 6), bool), 0.2, 0.0, 1.3, ls, tm, nt, True, nu=nu)
    for b in range(4):
        Kr, Gr = k(xg[b], eval_gradient=True)
        np.testing.assert_allclose(K[b], Kr, rtol=1e-13, atol=1e-15)
        np.testing.assert_allclose(G[b], Gr, rtol=1e-12, atol=1e-15)


def test_fill_matern_planted_rbf_derivative_fails(monkeypatch):
    """Taking the RBF's length-scale factor (G = E) for a Matern is caught."""
    rng = np.random.default_rng(22)
    xg = rng.random((2, 5, 3))
    k, ls, tm, nt = matern_case(1.5, True)
    real = vo.stationary
    monkeypatch.setattr(vo, "stationary", lambda u, nu=None: (real(u, nu)[0],) * 2)
    _, G = vo.fill(xg, np.ones((2, 5), bool), 0.2, 0.0, 1.3, ls, tm, nt, True, nu=1.5)
    _, Gr = k(xg[0], eval_gradient=True)
    assert not np.allclose(G[0], Gr, rtol=1e-6)


@pytest.mark.parametrize("nu", MATERN_NU)
def test_exact_gp_matern_loglik_and_grad(nu):
    x, _, y = ref_test_data()
    y = y[:, :2]
    k, ls, tm, nt = matern_case(nu, True, d=4)
    gpr = skgp.GaussianProcessRegressor(kernel=k, alpha=1e-10)
    gpr.X_train_, gpr.y_train_, gpr.kernel_ = x, y, k
    ll_ref, g_ref = gpr.log_marginal_likelihood(k.theta, eval_gradient=True)
    nn = vo.vecchia_knn(x, ls, x.shape[0] - 1)
    ll, g, _ = vo.vecchia_loglik(x, y, nn, 0.2, 0.0, 1.3, ls, 1e-10, tm, nt, eval_gradient=True,
                                 nu=nu)
    assert abs(ll / ll_ref - 1) < 1e-10
    np.testing.assert_allclose(g, g_ref, rtol=1e-8, atol=1e-8)
    g_adj = vo.vecchia_adjoint_grad(x, y, nn, 0.2, 0.0, 1.3, ls, 1e-10, tm, nt, nu=nu)
    np.testing.assert_allclose(g_adj, g_ref, rtol=1e-8, atol=1e-8)


@pytest.mark.parametrize("nu", MATERN_NU)
def test_exact_gp_matern_predict(nu):
    x, x_eval, y = ref_test_data()
    k = skk.ConstantKernel(0.8) * skk.Matern(0.6, nu=nu) + skk.WhiteKernel(0.3)
    gpr = skgp.GaussianProcessRegressor(kernel=k, optimizer=None, alpha=1e-10).fit(x, y[:, 1])
    mean_ref, std_ref = gpr.predict(x_eval, return_std=True)
    nn = vo.knn(x, x_eval, 0.6, x.shape[0])
    mean, std = vo.predict(x, y[:, 1], x_eval, nn, 0.3, 0.0, 0.8, 0.6, 1e-10, nu=nu)
    np.testing.assert_allclose(mean[:, 0], mean_ref, atol=1e-9)
    np.testing.assert_allclose(std, std_ref, atol=1e-9)


@pytest.mark.parametrize("nu", MATERN_NU)
def test_vecchia_matern_grad_central_differences(nu):
    rng = np.random.default_rng(23)
    x = rng.random((70, 2))
    y = np.sin(5 * x[:, 0]) + rng.normal(0, 0.1, 70)
    theta = np.log([0.9, 0.3, 0.02])
    nn = vo.vecchia_knn(x, np.exp(theta[1]), 6)

    def f(t):
        c, l, w = np.exp(t)
        return vo.vecchia_loglik(x, y, nn, w, 0.0, c, l, 1e-8, nu=nu)[0]

    c, l, w = np.exp(theta)
    _, g, _ = vo.vecchia_loglik(x, y, nn, w, 0.0, c, l, 1e-8, THETA_MAP, 3, eval_gradient=True,
                                nu=nu)
    h = 1e-5
    fd = np.array([(f(theta + h * e) - f(theta - h * e)) / (2 * h) for e in np.eye(3)])
    np.testing.assert_allclose(g, fd, rtol=1e-6, atol=1e-6)


# ------------------------------------------------------------ exact GP known answers
@pytest.mark.parametrize("y_dim", [1, 3])
def test_exact_gp_loglik_and_grad(y_dim):
    x, _, y = ref_test_data()
    y = y[:, :y_dim]
    theta = np.full(3, 0.5)
    gpr = skgp.GaussianProcessRegressor(kernel=sk_kernel(), alpha=1e-10)
    gpr.X_train_, gpr.y_train_, gpr.kernel_ = x, y, sk_kernel()
    ll_ref, g_ref = gpr.log_marginal_likelihood(theta, eval_gradient=True)
    c, l, w = np.exp(theta)
    nn = vo.vecchia_knn(x, l, x.shape[0] - 1)
    ll, g, _ = vo.vecchia_loglik(x, y, nn, w, 0.0, c, l, 1e-10, THETA_MAP, 3, eval_gradient=True)
    assert abs(ll / ll_ref - 1) < 1e-10
    np.testing.assert_allclose(g, g_ref, rtol=1e-9)
    # planted mistake: drop alpha -> still tiny; flip the noise derivative sign -> fails
    gbad = g.copy()
    gbad[2] = -gbad[2]
    assert not np.allclose(gbad, g_ref, rtol=1e-6)


def test_exact_gp_predict():
    x, x_eval, y = ref_test_data()
    y = y[:, 0]
    kern = sk_kernel(1.0, 1.0, 1.0)
    gpr = skgp.GaussianProcessRegressor(kernel=kern, optimizer=None, alpha=1e-10).fit(x, y)
    mean_ref, std_ref = gpr.predict(x_eval, return_std=True)
    nn = vo.knn(x, x_eval, 1.0, x.shape[0])
    mean, std = vo.predict(x, y, x_eval, nn, 1.0, 0.0, 1.0, 1.0, 1e-10)
    np.testing.assert_allclose(mean[:, 0], mean_ref, atol=1e-10)
    np.testing.assert_allclose(std, std_ref, atol=1e-10)


def test_exact_gp_predict_alpha_only_on_training_rows():
    """alpha = 1e-3 is large enough to tell the reference's rule apart: alpha is
    added to the training neighbours' variances only, never to the query's own
    K_00 (base_nngpr.py:1106-1107), which is what scikit-learn's exact predictive
    std does too.  Adding it to K_00 moves std by ~alpha / (2 std) and fails."""
    x, x_eval, y = ref_test_data()
    y = y[:, 1]
    alpha = 1e-3
    kern = sk_kernel(1.3, 0.8, 0.05)
    gpr = skgp.GaussianProcessRegressor(kernel=kern, optimizer=None, alpha=alpha).fit(x, y)
    mean_ref, std_ref = gpr.predict(x_eval, return_std=True)
    nn = vo.knn(x, x_eval, 0.8, x.shape[0])
    mean, std = vo.predict(x, y, x_eval, nn, 0.05, 0.0, 1.3, 0.8, alpha)
    np.testing.assert_allclose(mean[:, 0], mean_ref, rtol=1e-9, atol=1e-10)
    np.testing.assert_allclose(std, std_ref, rtol=1e-9, atol=1e-10)
    # planted mistake: alpha on the query's own variance as well
    std_bad = np.sqrt(std ** 2 + alpha)
    assert not np.allclose(std_bad, std_ref, rtol=1e-6, atol=1e-8)


# ------------------------------------------- numpy_port (the timed CPU baseline)
def test_numpy_port_knn_matches_oracle():
    """The restated generic kNN (dense -k blocks + argsort) returns the oracle's
    neighbour lists on tie-free data, with and without the Vecchia cut."""
    from oracle import numpy_port as npp
    rng = np.random.default_rng(31)
    x = rng.random((400, 3))
    q = rng.random((90, 3))
    np.testing.assert_array_equal(npp.generic_knn(x, q, 0.0, 1.0, 0.3, 7, None),
                                  vo.knn(x, q, 0.3, 7))
    for i0, i1 in [(0, 50), (120, 400)]:
        got = npp.generic_knn(x[:i1], x[i0:i1], 0.2, 1.4, 0.3, 7, i0)
        np.testing.assert_array_equal(got, vo.vecchia_knn(x, 0.3, 7)[i0:i1])


@pytest.mark.parametrize("i0,i1", [(0, 40), (40, 300), (250, 400)])
def test_numpy_port_loglik_matches_oracle(i0, i1):
    """One task of the restated CPU worker (rows [i0, i1), neighbours among the
    rows before each) equals the oracle's per-row log-likelihood summed over the
    same rows, and its reference-formula gradient the oracle's gradient."""
    from oracle import numpy_port as npp
    rng = np.random.default_rng(32)
    x = rng.random((400, 3))
    y = np.sin(5 * x[:, 0]) + x[:, 1] * x[:, 2] + rng.normal(0, 0.1, 400)
    w, C, s, ls, k, alpha = 0.02, 0.0, 1.1, 0.25, 9, 1e-8
    tmap = {"noise_level": 0, "rbf_level": 1, "length_scale": 2}
    ll, g, _, _ = npp.loglik_batch(x, y, i0, i1, w, C, s, ls, k, alpha, tmap, 3)
    nn = vo.vecchia_knn(x, ls, k)
    rows = np.arange(i0, i1)
    _, _, pts = vo.vecchia_loglik(x, y, nn, w, C, s, ls, alpha, tmap, 3)
    np.testing.assert_allclose(ll, pts[rows].sum(), rtol=1e-10)
    # the oracle's gradient of the same rows: its adjoint restricted to them
    g_o = vo.vecchia_adjoint_grad(x, y, nn, w, C, s, ls, alpha, tmap, 3, rows=rows)
    np.testing.assert_allclose(g, g_o, rtol=1e-9, atol=1e-10)


# ----------------------------------------------------------- Vecchia, m < n, brute
def test_vecchia_matches_dense_conditionals():
    rng = np.random.default_rng(9)
    x = rng.random((60, 3))
    y = rng.normal(size=60)
    c, l, w, alpha = 1.2, 0.4, 0.05, 1e-6
    nn = vo.vecchia_knn(x, l, 5)
    _, _, llp = vo.vecchia_loglik(x, y, nn, w, 0.0, c, l, alpha)
    Kfull = sk_kernel(c, l, w)(x) + alpha * np.eye(60)
    for i in range(60):
        S = nn[i][nn[i] >= 0]
        if len(S):
            A = Kfull[np.ix_(S, S)]
            kx = Kfull[i, S]
            mu = kx @ np.linalg.solve(A, y[S])
            s2 = Kfull[i, i] - kx @ np.linalg.solve(A, kx)
        else:
            mu, s2 = 0.0, Kfull[i, i]
        ref = -0.5 * (y[i] - mu) ** 2 / s2 - 0.5 * np.log(2 * np.pi * s2)
        assert abs(llp[i] - ref) < 1e-10 * max(1, abs(ref))


def test_vecchia_grad_central_differences():
    rng = np.random.default_rng(10)
    x = rng.random((80, 2))
    y = np.sin(5 * x[:, 0]) + rng.normal(0, 0.1, 80)
    theta = np.log([0.9, 0.3, 0.02])
    nn = vo.vecchia_knn(x, np.exp(theta[1]), 6)

    def f(t):
        c, l, w = np.exp(t)
        return vo.vecchia_loglik(x, y, nn, w, 0.0, c, l, 1e-8)[0]

    c, l, w = np.exp(theta)
    _, g, _ = vo.vecchia_loglik(x, y, nn, w, 0.0, c, l, 1e-8, THETA_MAP, 3, eval_gradient=True)
    h = 1e-5
    fd = np.array([(f(theta + h * e) - f(theta - h * e)) / (2 * h) for e in np.eye(3)])
    np.testing.assert_allclose(g, fd, rtol=1e-6, atol=1e-6)


@pytest.mark.parametrize("ard", [False, True])
def test_adjoint_equals_reference_formula(ard):
    rng = np.random.default_rng(12)
    x = rng.random((70, 3))
    y = rng.normal(size=(70, 2))
    if ard:
        ls = np.array([0.3, 0.5, 0.9])
        tm = {"noise_level": 0, "constant_value": 1, "rbf_level": 2, "length_scale": np.arange(3, 6)}
        nt = 6
    else:
        ls = 0.4
        tm = {"noise_level": 0, "constant_value": 1, "rbf_level": 2, "length_scale": 3}
        nt = 4
    nn = vo.vecchia_knn(x, ls, 7)
    _, g_ref, _ = vo.vecchia_loglik(x, y, nn, 0.05, 0.3, 1.1, ls, 1e-8, tm, nt, eval_gradient=True)
    g_adj = vo.vecchia_adjoint_grad(x, y, nn, 0.05, 0.3, 1.1, ls, 1e-8, tm, nt)
    np.testing.assert_allclose(g_adj, g_ref, rtol=1e-10, atol=1e-10)


def test_golden_small_uniform_reproduced():
    import os
    from nngpr_amd import datasets
    g = np.load(os.path.join(os.path.dirname(__file__), "golden", "small_uniform.npz"))
    ds = datasets.small_uniform()
    nn = vo.vecchia_knn(ds["x_train"], 0.3, 16)
    np.testing.assert_array_equal(nn, g["nn_train"])
    ll, _, _ = vo.vecchia_loglik(ds["x_train"], ds["y_train"], nn, 0.01, 0.0, 1.0, 0.3, 1e-10)
    assert abs(ll - float(g["loglik"])) < 1e-9 * abs(ll)
