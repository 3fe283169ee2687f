"""HIP fused Vecchia kernels vs the FP64 oracle and scikit-learn's exact GP.

Tolerances: FP64 form 1e-10 relative per point (1e-9 on gradients); FP32 +
refinement form |dl|/n <= 1e-6 and mean/std relative error <= 1e-5 (the
targets stated for that form).
"""
import os

import numpy as np
import pytest
import torch
from sklearn import gaussian_process as skgp
from sklearn.gaussian_process import kernels as skk

from nngpr_amd import _native as nat
from nngpr_amd import datasets, hip_ops
from nngpr_amd import batched_kernels as bk
from nngpr_amd.kernel_params import get_kernel_params, get_kernel_theta_mapping
from oracle import vecchia_oracle as vo

pytestmark = pytest.mark.gpu
DEV = "cuda:0"
GOLD = os.path.join(os.path.dirname(__file__), "golden", "small_uniform.npz")
SU_MAP = {"noise_level": 0, "rbf_level": 1, "length_scale": 2}


def t(a, dtype=torch.float64):
    return torch.from_numpy(np.ascontiguousarray(a)).to(DEV, dtype)


def kstruct(w, C, s, ls, theta_map, n_theta):
    params = {"noise_level": w, "constant_value": C}
    if ls is not None:
        params["length_scale"] = ls
        params["rbf_level"] = s
    return nat.kernel_struct(params, theta_map, n_theta)


def tol(prec):
    return 1e-10 if prec == 0 else 1e-5


@pytest.mark.parametrize("prec", [0, 1])
def test_small_uniform_loglik_golden(prec):
    g = np.load(GOLD)
    ds = datasets.small_uniform()
    x, y = ds["x_train"], ds["y_train"]
    ks = kstruct(0.01, 0.0, 1.0, 0.3, SU_MAP, 3)
    ll, grad, pts = hip_ops.vecchia_loglik(t(x), t(y), t(g["nn_train"], torch.int32), ks, 1e-10,
                                           precision=prec, eval_gradient=True, point_loglik=True)
    n = x.shape[0]
    assert abs(ll.item() - float(g["loglik"])) / n <= (1e-12 if prec == 0 else 1e-6)
    np.testing.assert_allclose(pts.cpu().numpy(), g["loglik_points"], rtol=tol(prec), atol=tol(prec))
    np.testing.assert_allclose(grad.cpu().numpy(), g["grad"], rtol=1e-9 if prec == 0 else 1e-5)


@pytest.mark.parametrize("prec", [0, 1])
def test_small_uniform_predict_golden(prec):
    g = np.load(GOLD)
    ds = datasets.small_uniform()
    ks = kstruct(0.01, 0.0, 1.0, 0.3, {}, 0)
    mean, std = hip_ops.predict(t(ds["x_train"]), t(ds["y_train"]), t(ds["x_test"]),
                                t(g["nn_test"], torch.int32), ks, 1e-10, precision=prec)
    np.testing.assert_allclose(mean.cpu().numpy()[:, 0], g["mean"], rtol=tol(prec), atol=tol(prec))
    np.testing.assert_allclose(std.cpu().numpy(), g["std"], rtol=tol(prec))


@pytest.mark.parametrize("prec", [0, 1])
@pytest.mark.parametrize("m", [1, 5, 16, 32, 64, 90])
def test_loglik_grad_vs_oracle(prec, m):
    rng = np.random.default_rng(m)
    x = rng.random((700, 3))
    y = np.column_stack([np.sin(4 * x[:, 0]), x[:, 1] ** 2]) + rng.normal(0, 0.1, (700, 2))
    tm = {"noise_level": 0, "constant_value": 1, "rbf_level": 2, "length_scale": 3}
    w, C, s, ls = 0.05, 0.4, 1.2, 0.25
    nn = vo.vecchia_knn(x, ls, m)
    ll_o, g_o, pts_o = vo.vecchia_loglik(x, y, nn, w, C, s, ls, 1e-10, tm, 4, eval_gradient=True)
    ks = kstruct(w, C, s, ls, tm, 4)
    ll, g, pts = hip_ops.vecchia_loglik(t(x), t(y), t(nn, torch.int32), ks, 1e-10, precision=prec,
                                        eval_gradient=True, point_loglik=True)
    assert abs(ll.item() - ll_o) / x.shape[0] <= (1e-12 if prec == 0 else 1e-6)
    np.testing.assert_allclose(pts.cpu().numpy(), pts_o, rtol=tol(prec), atol=tol(prec))
    np.testing.assert_allclose(g.cpu().numpy(), g_o, rtol=1e-9 if prec == 0 else 1e-5,
                               atol=1e-9 if prec == 0 else 1e-5)


@pytest.mark.parametrize("prec", [0, 1])
@pytest.mark.parametrize("d,m", [(6, 12), (18, 32), (32, 40)])
def test_ard_grad_vs_oracle(prec, d, m):
    rng = np.random.default_rng(3 + d)
    x = rng.normal(size=(500, d))
    y = np.sin(x[:, 0]) + 0.3 * x[:, 3] + rng.normal(0, 0.05, 500)
    ls = rng.uniform(0.8, 3.0, size=d) * np.sqrt(d / 6.0)
    tm = {"noise_level": 0, "rbf_level": 1, "length_scale": np.arange(2, 2 + d)}
    nn = vo.vecchia_knn(x, ls, m)
    ll_o, g_o, _ = vo.vecchia_loglik(x, y, nn, 0.02, 0.0, 0.9, ls, 1e-10, tm, 2 + d,
                                     eval_gradient=True)
    ll, g, _ = hip_ops.vecchia_loglik(t(x), t(y), t(nn, torch.int32),
                                      kstruct(0.02, 0.0, 0.9, ls, tm, 2 + d),
                                      1e-10, precision=prec, eval_gradient=True)
    assert abs(ll.item() - ll_o) / 500 <= (1e-12 if prec == 0 else 1e-6)
    np.testing.assert_allclose(g.cpu().numpy(), g_o, rtol=1e-9 if prec == 0 else 1e-5, atol=1e-8)


@pytest.mark.parametrize("prec", [0, 1])
def test_predict_vs_oracle_multioutput(prec):
    rng = np.random.default_rng(8)
    x = rng.random((1500, 2))
    y = rng.normal(size=(1500, 3))
    xq = rng.random((400, 2))
    nn = vo.knn(x, xq, 0.2, 20)
    mean_o, std_o = vo.predict(x, y, xq, nn, 0.1, 0.5, 1.5, 0.2, 1e-8)
    mean, std = hip_ops.predict(t(x), t(y), t(xq), t(nn, torch.int32),
                                kstruct(0.1, 0.5, 1.5, 0.2, {}, 0), 1e-8, precision=prec)
    np.testing.assert_allclose(mean.cpu().numpy(), mean_o, rtol=tol(prec), atol=tol(prec))
    np.testing.assert_allclose(std.cpu().numpy(), std_o, rtol=tol(prec))


def test_exact_gp_equivalence_m_ge_n_minus_1():
    # The reference's known-answer test (nngpr_base_tester.py:174-192) on the HIP path.
    rng = np.random.RandomState(234058)
    x = rng.rand(150, 4)
    rng.rand(50, 4)
    beta = rng.rand(4, 3)
    beta[:, -1] += 1
    y = x @ beta + rng.rand(150, 3)
    y[:, 0] *= 10
    kern = skk.ConstantKernel(1) * skk.RBF(1) + skk.WhiteKernel(1)
    theta = np.full(3, 0.5)
    gpr = skgp.GaussianProcessRegressor(kernel=kern, alpha=1e-10)
    gpr.X_train_, gpr.y_train_, gpr.kernel_ = x, y, kern
    ll_ref, g_ref = gpr.log_marginal_likelihood(theta, eval_gradient=True)
    k2 = kern.clone_with_theta(theta)
    params, mapping = get_kernel_params(k2), get_kernel_theta_mapping(k2)
    ks = nat.kernel_struct(params, mapping, 3)
    nn = hip_ops.knn(t(x), t(x), np.exp(0.5), 149, ref_cut=0)
    ll, g, _ = hip_ops.vecchia_loglik(t(x), t(y), nn, ks, 1e-10, eval_gradient=True)
    assert abs(ll.item() / ll_ref - 1) < 1e-10
    np.testing.assert_allclose(g.cpu().numpy(), g_ref, rtol=1e-9)


def test_fill_kernel_vs_oracle():
    rng = np.random.default_rng(4)
    x = rng.random((300, 3))
    nn = rng.integers(-1, 300, size=(50, 9)).astype(np.int32)
    tm = {"noise_level": 0, "constant_value": 1, "rbf_level": 2, "length_scale": 3}
    ks = kstruct(0.1, 0.3, 1.4, 0.5, tm, 4)
    K, G = hip_ops.fill_nn_kernel(t(x), t(nn, torch.int32), ks, eval_gradient=True)
    valid = nn >= 0
    Ko, Go = vo.fill(x[np.where(valid, nn, 0)], valid, 0.1, 0.3, 1.4, 0.5, tm, 4, True)
    np.testing.assert_allclose(K.cpu().numpy(), Ko, rtol=1e-13, atol=1e-15)
    np.testing.assert_allclose(G.cpu().numpy(), Go, rtol=1e-12, atol=1e-15)


def test_fill_exponential_accuracy():
    """The covariance entries' exp(-u) (the library's own FP64 routine) against
    numpy's over u in [0, 760]: within 2 ulp in the normal range, the subnormal
    tail within 2 of its units, and 0 from u = 746 on.  One feature, pairs (0, r)
    with r^2 / (2 l^2) = u, s = 1 and no other kernel term, so each off-diagonal
    entry is exactly exp(-u) of the kernel's own u."""
    rng = np.random.default_rng(21)
    ls = 0.5
    u = np.concatenate([np.linspace(0.0, 760.0, 4001), rng.uniform(0.0, 50.0, 4000),
                        rng.uniform(0.0, 1e-3, 500), rng.uniform(700.0, 750.0, 500)])
    r = np.sqrt(2.0 * u) * ls
    x = np.concatenate([np.zeros(1), r])[:, None]
    nn = np.stack([np.zeros(u.size, np.int32), np.arange(1, u.size + 1, dtype=np.int32)], 1)
    tm = {"noise_level": -1, "constant_value": -1, "rbf_level": 0, "length_scale": 1}
    ks = kstruct(0.0, 0.0, 1.0, ls, tm, 2)
    K, _ = hip_ops.fill_nn_kernel(t(x), t(nn, torch.int32), ks, eval_gradient=False)
    got = K.cpu().numpy()[:, 0, 1]
    uk = (r * r) * (1.0 / (2.0 * ls * ls))      # the kernel's own u, as it forms it
    exp = np.exp(-uk)
    normal = exp >= np.finfo(np.float64).tiny
    ulp = np.spacing(exp[normal])
    assert np.max(np.abs(got[normal] - exp[normal]) / ulp) <= 2.0
    assert np.max(np.abs(got[~normal] - exp[~normal])) <= 2 * 5e-324
    assert np.all(got[uk >= 746.0] == 0.0)


def near_duplicate_set(seed=12):
    """300 points in the unit square, then a copy of each moved by about 1e-3 (rows
    300-599, so every copy conditions on its original).  With l = 0.05, White(1e-6)
    and alpha = 0 a block holding both copies has a pivot near 4e-4 against a
    diagonal of 1 (pivot ratio ~2.5e3 > 1e3), which FP64 still factors to ~1e-12;
    blocks of the first 300 rows are well conditioned."""
    rng = np.random.default_rng(seed)
    base = rng.random((300, 2))
    twin = base + rng.normal(0, 1e-3 / np.sqrt(2), base.shape)
    x = np.concatenate([base, twin], axis=0)
    y = np.sin(4 * x[:, 0]) + np.cos(3 * x[:, 1]) + rng.normal(0, 0.05, 600)
    return x, y


@pytest.mark.parametrize("what", ["loglik", "predict"])
def test_fp32_pivot_ratio_refactor_matches_oracle(what):
    """The FP32 form sends blocks whose pivot ratio exceeds 1e3 (or that fail to
    factor) to an FP64 re-factor in the same call: the count is positive and the
    results equal the FP64 form and the oracle."""
    x, y = near_duplicate_set()
    w, s, ls, alpha, m = 1e-6, 1.0, 0.05, 0.0, 12
    tm = {"noise_level": 0, "rbf_level": 1, "length_scale": 2}
    ks = kstruct(w, 0.0, s, ls, tm, 3)
    if what == "loglik":
        nn = vo.vecchia_knn(x, ls, m)
        ll_o, g_o, pts_o = vo.vecchia_loglik(x, y, nn, w, 0.0, s, ls, alpha, tm, 3,
                                             eval_gradient=True)
        out = {}
        for prec in (0, 1):
            ll, g, pts = hip_ops.vecchia_loglik(t(x), t(y), t(nn, torch.int32), ks, alpha,
                                                precision=prec, eval_gradient=True,
                                                point_loglik=True)
            out[prec] = (ll.item(), g.cpu().numpy(), pts.cpu().numpy(),
                         hip_ops.stats("vecchia")["points_refactored_fp64"])
        assert out[0][3] == 0
        assert 0 < out[1][3] < x.shape[0]
        for prec, rt in ((0, 1e-8), (1, 1e-6)):
            assert abs(out[prec][0] - ll_o) / x.shape[0] <= rt
            np.testing.assert_allclose(out[prec][2], pts_o, rtol=10 * rt, atol=10 * rt)
            np.testing.assert_allclose(out[prec][1], g_o, rtol=10 * rt, atol=10 * rt)
    else:
        xq = x[::7] + 3e-5
        nn = vo.knn(x, xq, ls, m)
        mean_o, std_o = vo.predict(x, y, xq, nn, w, 0.0, s, ls, alpha)
        for prec in (0, 1):
            mean, std = hip_ops.predict(t(x), t(y), t(xq), t(nn, torch.int32),
                                        kstruct(w, 0.0, s, ls, {}, 0), alpha, precision=prec)
            nref = hip_ops.stats("predict")["points_refactored_fp64"]
            assert (nref == 0) if prec == 0 else (nref > 0)
            np.testing.assert_allclose(mean.cpu().numpy(), mean_o, rtol=1e-6, atol=1e-7)
            np.testing.assert_allclose(std.cpu().numpy(), std_o, rtol=1e-5)


@pytest.mark.parametrize("prec", [0, 1])
def test_neighbour_index_out_of_range_rejected(prec):
    """An index >= n (or below -1) is never gathered: the fused entry points report
    it through the workspace status (ValueError), the plug-in fill raises
    IndexError, and the raw fill entry point returns NaN for that slot."""
    rng = np.random.default_rng(5)
    x = rng.random((200, 3))
    y = rng.normal(size=200)
    nn = vo.vecchia_knn(x, 0.3, 8)
    ks = kstruct(0.05, 0.0, 1.0, 0.3, {"noise_level": 0, "rbf_level": 1, "length_scale": 2}, 3)
    for bad in (200, 10 ** 9, -7):
        nb = nn.copy()
        nb[150, 3] = bad
        with pytest.raises(ValueError, match="outside"):
            hip_ops.vecchia_loglik(t(x), t(y), t(nb, torch.int32), ks, 1e-10, precision=prec,
                                   eval_gradient=True)
        with pytest.raises(ValueError, match="outside"):
            hip_ops.predict(t(x), t(y), t(x[:20]), t(nb[140:160], torch.int32),
                            kstruct(0.05, 0.0, 1.0, 0.3, {}, 0), 1e-10, precision=prec)
    # the same call with valid indices still succeeds afterwards
    ll, _, _ = hip_ops.vecchia_loglik(t(x), t(y), t(nn, torch.int32), ks, 1e-10, precision=prec)
    ll_o, _, _ = vo.vecchia_loglik(x, y, nn, 0.05, 0.0, 1.0, 0.3, 1e-10)
    assert abs(ll.item() - ll_o) / 200 <= (1e-12 if prec == 0 else 1e-6)
    nb = nn[:10].copy()
    nb[4, 2] = 200
    K, _ = hip_ops.fill_nn_kernel(t(x), t(nb, torch.int32), ks)
    K = K.cpu().numpy()
    assert np.isnan(K[4, 2]).all() and np.isnan(K[4, :, 2]).all()
    assert np.isfinite(np.delete(np.delete(K[4], 2, 0), 2, 1)).all()
    assert np.isfinite(K[[0, 1, 2, 3, 5]]).all()
    from nngpr_amd.base_nngpr import HipWorker
    from nngpr_amd import batched_kernels as bkm
    kern = bkm.WhiteKernel(0.05) + bkm.RBF(0.3)
    with pytest.raises(IndexError):
        HipWorker("cuda:0").fill_nn_kernel(x, nb, kern)


@pytest.mark.parametrize("prec", [0, 1])
def test_not_positive_definite_raises(prec):
    x = np.repeat(np.random.default_rng(1).random((20, 2)), 3, axis=0)
    y = np.zeros(60)
    nn = vo.vecchia_knn(x, 1.0, 4)
    ks = kstruct(0.0, 0.0, 1.0, 1.0, {}, 0)
    with pytest.raises(np.linalg.LinAlgError):
        hip_ops.vecchia_loglik(t(x), t(y), t(nn, torch.int32), ks, 0.0, precision=prec)


def test_gradient_central_differences_on_gpu():
    rng = np.random.default_rng(21)
    x = rng.random((400, 3))
    y = np.cos(3 * x[:, 0]) * x[:, 2] + rng.normal(0, 0.1, 400)
    kern = bk.WhiteKernel(0.05) + bk.ConstantKernel(0.8) * bk.RBF(0.3)
    nn = vo.vecchia_knn(x, 0.3, 10)
    nnt = t(nn, torch.int32)

    def f(theta, grad=False):
        k = kern.clone_with_theta(theta)
        ks = nat.kernel_struct(get_kernel_params(k), get_kernel_theta_mapping(k), 3)
        ll, g, _ = hip_ops.vecchia_loglik(t(x), t(y), nnt, ks, 1e-10, eval_gradient=grad)
        return (ll.item(), g.cpu().numpy()) if grad else ll.item()

    th = kern.theta
    _, g = f(th, True)
    h = 1e-5
    fd = np.array([(f(th + h * e) - f(th - h * e)) / (2 * h) for e in np.eye(3)])
    np.testing.assert_allclose(g, fd, rtol=1e-6, atol=1e-6)
