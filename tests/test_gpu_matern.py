"""Matern stationary part (nu = 1/2, 3/2, 5/2) on every fused path and entry
point vs the FP64 oracle (itself pinned against scikit-learn's Matern and exact
GP) and vs scikit-learn directly where the Vecchia factorisation is exact.

Paths: the segment kernel (m <= 64, d <= 32: 16-, 32- and 64-row blocks, 3-, 8-
and 32-feature layouts), the register kernel (d > 32) and the generic kernel
(m > 64).  Tolerances as for the RBF: FP64 form 1e-10 relative per point and
1e-9 on gradients; FP32 + refinement form |dl|/n <= 1e-6, per point and
gradients 1e-5.
"""
import functools

import numpy as np
import pytest
import torch
from sklearn import gaussian_process as skgp
from sklearn.gaussian_process import kernels as skk

from nngpr_amd import TorchNngpr, hip_ops
from nngpr_amd import _native as nat
from nngpr_amd import batched_kernels as bk
from nngpr_amd.array_module import NumpyModule
from nngpr_amd.kernel_params import get_kernel_params, get_kernel_theta_mapping
from nngpr_amd.serving import ResidentModel
from oracle import vecchia_oracle as vo

pytestmark = pytest.mark.gpu
DEV = "cuda:0"
NUS = [0.5, 1.5, 2.5]


def t(a, dtype=torch.float64):
    return torch.from_numpy(np.ascontiguousarray(a)).to(DEV, dtype)


def kstruct(w, C, s, ls, nu, theta_map, n_theta):
    params = {"noise_level": w, "constant_value": C, "length_scale": ls, "rbf_level": s,
              "matern_nu": nu}
    return nat.kernel_struct(params, theta_map, n_theta)


# (n, d, m, ard): the path each case takes
CASES = {
    "seg16_d3": (600, 3, 12, False),
    "seg32_d3_ard": (600, 3, 30, True),
    "seg64_d8_ard": (400, 6, 50, True),
    "seg32_d32_ard": (400, 18, 20, True),
    "reg_d40_ard": (300, 40, 10, True),
    "reg_d40_iso": (300, 40, 10, False),
    "generic_m80": (400, 2, 80, False),
    "generic_m70_ard": (300, 3, 70, True),
}


@functools.lru_cache(maxsize=None)
def case_data(name, nu):
    n, d, m, ard = CASES[name]
    rng = np.random.default_rng(10 * list(CASES).index(name) + int(2 * nu))
    x = rng.normal(size=(n, d))
    y = np.column_stack([np.sin(x[:, 0]) + 0.3 * x[:, -1], np.cos(x[:, 1])])
    y = y + rng.normal(0, 0.05, y.shape)
    scale = np.sqrt(d / 3.0)
    if ard:
        ls = rng.uniform(0.8, 2.5, size=d) * scale
        tm = {"noise_level": 0, "constant_value": 1, "rbf_level": 2,
              "length_scale": np.arange(3, 3 + d)}
        nt = 3 + d
    else:
        ls = 1.2 * scale
        tm = {"noise_level": 0, "constant_value": 1, "rbf_level": 2, "length_scale": 3}
        nt = 4
    w, C, s = 0.03, 0.2, 1.1
    nn = vo.vecchia_knn(x, ls, m)
    ll_o, _, pts_o = vo.vecchia_loglik(x, y, nn, w, C, s, ls, 1e-10, nu=nu)
    g_o = vo.vecchia_adjoint_grad(x, y, nn, w, C, s, ls, 1e-10, tm, nt, nu=nu)
    return x, y, nn, (w, C, s, ls, tm, nt), (ll_o, g_o, pts_o)


@pytest.mark.parametrize("prec", [0, 1])
@pytest.mark.parametrize("nu", NUS)
@pytest.mark.parametrize("case", list(CASES))
def test_loglik_grad_vs_oracle(case, nu, prec):
    x, y, nn, (w, C, s, ls, tm, nt), (ll_o, g_o, pts_o) = case_data(case, nu)
    ks = kstruct(w, C, s, ls, nu, tm, nt)
    ll, g, pts = hip_ops.vecchia_loglik(t(x), t(y), t(nn, torch.int32), ks, 1e-10, precision=prec,
                                        eval_gradient=True, point_loglik=True)
    n = x.shape[0]
    assert abs(ll.item() - ll_o) / n <= (1e-12 if prec == 0 else 1e-6)
    tol = 1e-10 if prec == 0 else 1e-5
    np.testing.assert_allclose(pts.cpu().numpy(), pts_o, rtol=tol, atol=tol)
    gt = 1e-9 if prec == 0 else 1e-5
    np.testing.assert_allclose(g.cpu().numpy(), g_o, rtol=gt, atol=gt * max(1.0, np.abs(g_o).max()))


@pytest.mark.parametrize("prec", [0, 1])
@pytest.mark.parametrize("nu", NUS)
@pytest.mark.parametrize("m", [20, 80])
def test_predict_vs_oracle(m, nu, prec):
    rng = np.random.default_rng(int(10 * nu) + m)
    x = rng.random((1500, 2))
    y = rng.normal(size=(1500, 2))
    xq = rng.random((300, 2))
    nn = vo.knn(x, xq, 0.2, m)
    mean_o, std_o = vo.predict(x, y, xq, nn, 0.1, 0.5, 1.5, 0.2, 1e-8, nu=nu)
    ks = kstruct(0.1, 0.5, 1.5, 0.2, nu, {}, 0)
    mean, std = hip_ops.predict(t(x), t(y), t(xq), t(nn, torch.int32), ks, 1e-8, precision=prec)
    tol = 1e-10 if prec == 0 else 1e-5
    np.testing.assert_allclose(mean.cpu().numpy(), mean_o, rtol=tol, atol=tol)
    np.testing.assert_allclose(std.cpu().numpy(), std_o, rtol=tol)


@pytest.mark.parametrize("nu", NUS)
@pytest.mark.parametrize("ard", [False, True])
def test_fill_kernel_vs_oracle(nu, ard):
    rng = np.random.default_rng(5)
    x = rng.random((300, 3))
    x[7] = x[3]  # a duplicated point: r = 0 off the diagonal
    nn = rng.integers(-1, 300, size=(50, 9)).astype(np.int32)
    nn[0, :3] = [3, 7, 11]
    ls = np.array([0.4, 0.7, 1.2]) if ard else 0.5
    tm = {"noise_level": 0, "constant_value": 1, "rbf_level": 2,
          "length_scale": np.arange(3, 6) if ard else 3}
    nt = 6 if ard else 4
    ks = kstruct(0.1, 0.3, 1.4, ls, nu, tm, nt)
    K, G = hip_ops.fill_nn_kernel(t(x), t(nn, torch.int32), ks, eval_gradient=True)
    valid = nn >= 0
    Ko, Go = vo.fill(x[np.where(valid, nn, 0)], valid, 0.1, 0.3, 1.4, ls, tm, nt, True, nu=nu)
    np.testing.assert_allclose(K.cpu().numpy(), Ko, rtol=1e-13, atol=1e-15)
    np.testing.assert_allclose(G.cpu().numpy(), Go, rtol=1e-12, atol=1e-14)


def ref_data():
    rng = np.random.RandomState(234058)
    x = rng.rand(150, 4)
    x_eval = rng.rand(50, 4)
    beta = rng.rand(4, 3)
    beta[:, -1] += 1
    y = x @ beta + rng.rand(150, 3)
    y[:, 0] *= 10
    return x, x_eval, y


@pytest.mark.parametrize("nu", NUS)
@pytest.mark.parametrize("ard", [False, True])
def test_estimator_exact_gp_equivalence(nu, ard):
    """num_nn >= n: the estimator's log-likelihood, gradient and predictions equal
    scikit-learn's exact GP with the same Matern kernel
    (the reference's known-answer tests, nngpr_base_tester.py:174-257)."""
    x, x_eval, y = ref_data()
    ls = [0.8, 1.1, 0.6, 1.4] if ard else 1.0

    def kern():
        return bk.ConstantKernel(1.0) * bk.Matern(ls, nu=nu) + bk.WhiteKernel(0.5)

    theta = kern().theta + 0.3
    ref = skgp.GaussianProcessRegressor(kernel=kern().bind_to_array_module(NumpyModule()),
                                        normalize_y=True)
    ref.X_train_, ref.y_train_, ref.kernel_ = x, y, kern().bind_to_array_module(NumpyModule())
    ll_ref, g_ref = ref.log_marginal_likelihood(theta, eval_gradient=True)
    gpr = TorchNngpr(DEV, kernel=kern(), normalize_y=True, num_nn=x.shape[0])
    gpr.X_train_, gpr.y_train_, gpr.kernel_ = x, y, kern()
    ll, g = gpr.log_marginal_likelihood(theta, eval_gradient=True)
    assert abs(ll / ll_ref - 1) < 1e-10
    np.testing.assert_allclose(g, g_ref, rtol=1e-8, atol=1e-8 * np.abs(g_ref).max())

    k_fixed = kern().clone_with_theta(theta)
    for p in ("k1__k1__constant_value_bounds", "k1__k2__length_scale_bounds",
              "k2__noise_level_bounds"):
        k_fixed.set_params(**{p: "fixed"})
    ref = skgp.GaussianProcessRegressor(kernel=k_fixed.bind_to_array_module(NumpyModule()),
                                        normalize_y=True).fit(x, y[:, 1])
    mean_ref, std_ref = ref.predict(x_eval, return_std=True)
    gpr = TorchNngpr(DEV, kernel=k_fixed, normalize_y=True, num_nn=x.shape[0]).fit(x, y[:, 1])
    mean, std = gpr.predict(x_eval, return_std=True)
    assert np.max(np.abs(mean - mean_ref)) < 1e-9 * max(1.0, np.abs(mean_ref).max())
    assert np.max(np.abs(std - std_ref)) < 1e-9 * max(1.0, np.abs(std_ref).max())


@pytest.mark.parametrize("nu", NUS)
def test_gradient_central_differences_on_gpu(nu):
    rng = np.random.default_rng(31)
    x = rng.random((400, 3))
    y = np.cos(3 * x[:, 0]) * x[:, 2] + rng.normal(0, 0.1, 400)
    kern = bk.WhiteKernel(0.05) + bk.ConstantKernel(0.8) * bk.Matern([0.3, 0.5, 0.4], nu=nu)
    nn = vo.vecchia_knn(x, np.array([0.3, 0.5, 0.4]), 10)
    nnt = t(nn, torch.int32)

    def f(theta, grad=False):
        k = kern.clone_with_theta(theta)
        ks = nat.kernel_struct(get_kernel_params(k), get_kernel_theta_mapping(k), len(theta))
        ll, g, _ = hip_ops.vecchia_loglik(t(x), t(y), nnt, ks, 1e-10, eval_gradient=grad)
        return (ll.item(), g.cpu().numpy()) if grad else ll.item()

    th = kern.theta
    _, g = f(th, True)
    h = 1e-5
    fd = np.array([(f(th + h * e) - f(th - h * e)) / (2 * h) for e in np.eye(len(th))])
    np.testing.assert_allclose(g, fd, rtol=1e-6, atol=1e-6)


@pytest.mark.parametrize("prec", [0, 1])
def test_resident_model_matern(prec):
    rng = np.random.default_rng(17)
    x = rng.random((2000, 3))
    y = np.sin(5 * x[:, 0]) + rng.normal(0, 0.05, 2000)
    k = bk.WhiteKernel(0.02) + bk.ConstantKernel(1.3) * bk.Matern(0.25, nu=1.5)
    model = ResidentModel(x, y, k, alpha=1e-10, num_nn=16, precision=prec)
    try:
        xq = rng.random((500, 3))
        mean, std = model.predict(xq)
    finally:
        model.close()
    nn = vo.knn(x, xq, 0.25, 16)
    mo, so = vo.predict(x, y, xq, nn, 0.02, 0.0, 1.3, 0.25, 1e-10, nu=1.5)
    tol = 1e-10 if prec == 0 else 1e-5
    np.testing.assert_allclose(mean[:, 0], mo[:, 0], rtol=tol, atol=tol)
    np.testing.assert_allclose(std[:, 0], so, rtol=tol)
