"""Sequential conditioning on the HIP engine (nngp_predict_sequential /
nngp_sample_sequential and the estimator's predict(conditioning_method=...,
return_cov=...) and sample_y) vs the sequential oracle (pinned against
scikit-learn's exact GP and the closed form in test_oracle_sequential.py).

Tolerances: mean, std and covariance 1e-10 (absolute on values of order 1); the
sparse-covariance std and the samples 1e-9.
"""
import numpy as np
import pytest
import torch
from sklearn import gaussian_process as skgp
from sklearn.gaussian_process import kernels as skk

from nngpr_amd import TorchNngpr, hip_ops
from nngpr_amd import _native as nat
from nngpr_amd import batched_kernels as bk
from oracle import sequential_oracle as so

pytestmark = pytest.mark.gpu
DEV = "cuda:0"
W, C, S, LS, ALPHA = 0.05, 0.2, 1.3, 0.35, 1e-6
OUT = {"mean": hip_ops.COND_MEAN, "std": hip_ops.COND_STD, "cov": hip_ops.COND_COV}


def t(a, dtype=torch.float64):
    return torch.from_numpy(np.ascontiguousarray(a)).to(DEV, dtype)


def kstruct(nu=None):
    p = {"noise_level": W, "constant_value": C, "rbf_level": S, "length_scale": LS}
    if nu is not None:
        p["matern_nu"] = nu
    return nat.kernel_struct(p, {}, 0)


def data(nt, nq, d, seed):
    rng = np.random.default_rng(seed)
    x = rng.random((nt, d))
    xq = rng.random((nq, d))
    y = np.column_stack([np.sin(5 * x[:, 0]), x[:, 1] ** 2]) + rng.normal(0, 0.05, (nt, 2))
    return x, y, xq


CASES = [  # (nt, nq, d, num_nn, method)
    (300, 120, 2, 8, "full"),
    (300, 120, 3, 6, "full-double-nn"),
    (0, 90, 2, 5, "full"),
    (0, 90, 2, 5, "full-double-nn"),
    (250, 80, 2, 10, "only-train"),
    (200, 60, 2, 70, "full"),          # m > 64 per row: the generic weights kernel too
]


@pytest.mark.parametrize("nu", [None, 2.5])
@pytest.mark.parametrize("output", ["mean", "std", "cov"])
@pytest.mark.parametrize("case", CASES, ids=lambda c: "%d-%d-%d-%d-%s" % c)
def test_predict_sequential_vs_oracle(case, output, nu):
    nt, nq, d, k, method = case
    if method == "only-train" and output == "std":
        pytest.skip("only-train std is the parallel predict path")
    x, y, xq = data(nt, nq, d, nt + nq + d)
    nn = so.conditioning_nn(x, xq, LS, k, method)
    exp = so.predict_sequential(x, y, xq, nn, W, C, S, LS, ALPHA, output, nu)
    mean, std, cov = hip_ops.predict_sequential(t(x), t(y), t(xq), t(nn, torch.int32), kstruct(nu),
                                                ALPHA, OUT[output])
    if output == "mean":
        np.testing.assert_allclose(mean.cpu().numpy(), exp, atol=1e-10)
        return
    np.testing.assert_allclose(mean.cpu().numpy(), exp[0], atol=1e-10)
    if output == "std":
        np.testing.assert_allclose(std.cpu().numpy(), exp[1], atol=1e-9)
    else:
        np.testing.assert_allclose(cov.cpu().numpy(), exp[1], atol=1e-10)
        np.testing.assert_allclose(std.cpu().numpy() ** 2, np.diag(exp[1]), atol=1e-10)


@pytest.mark.parametrize("case", CASES[:4], ids=lambda c: "%d-%d-%d-%d-%s" % c)
def test_sample_sequential_vs_oracle(case):
    nt, nq, d, k, method = case
    x, y, xq = data(nt, nq, d, 7 + nq)
    nn = so.conditioning_nn(x, xq, LS, k, method)
    ny = y.shape[1] if nt else 1
    z = np.random.default_rng(3).normal(size=(nq, ny, 5))
    exp = so.sample_sequential(x, y if nt else np.zeros((0, 1)), xq, nn, W, C, S, LS, ALPHA, z)
    got = hip_ops.sample_sequential(t(x) if nt else None, t(y) if nt else None, t(xq),
                                    t(nn, torch.int32), kstruct(), ALPHA, t(z))
    np.testing.assert_allclose(got.cpu().numpy(), exp, atol=1e-9)


def test_neighbour_lists_match_oracle():
    x, _, xq = data(500, 200, 2, 1)
    gpr = TorchNngpr(DEV, kernel=bk.ConstantKernel(S) * bk.RBF(LS), num_nn=7)
    xd, xqd = t(x), t(xq)
    for method in ("only-train", "full", "full-double-nn"):
        got = gpr._conditioning_nn(gpr.kernel, xd, xqd, method, hip_ops.PRECISION_FP64)
        np.testing.assert_array_equal(got.cpu().numpy(), so.conditioning_nn(x, xq, LS, 7, method))


def sk_kern(fixed=True):
    b = "fixed" if fixed else (1e-5, 1e5)
    return skk.ConstantKernel(S, b) * skk.RBF(LS, b) + skk.WhiteKernel(W, b)


def bk_kern():
    return bk.ConstantKernel(S, "fixed") * bk.RBF(LS, "fixed") + bk.WhiteKernel(W, "fixed")


@pytest.mark.parametrize("method", ["full", "full-double-nn"])
def test_estimator_exact_sets_equal_sklearn(method):
    """num_nn >= n_train + n_query: the sequential predictions are the exact GP
    posterior (mean, std, joint covariance), with normalize_y."""
    x, y, xq = data(80, 30, 2, 11)
    ref = skgp.GaussianProcessRegressor(sk_kern(), alpha=ALPHA, normalize_y=True,
                                        optimizer=None).fit(x, y)
    gpr = TorchNngpr(DEV, kernel=bk_kern(), alpha=ALPHA, normalize_y=True, num_nn=200).fit(x, y)
    m_ref, s_ref = ref.predict(xq, return_std=True)
    m, s = gpr.predict(xq, return_std=True, conditioning_method=method)
    np.testing.assert_allclose(m, m_ref, atol=1e-8)
    np.testing.assert_allclose(s, s_ref, atol=1e-8)
    _, c_ref = ref.predict(xq, return_cov=True)
    m2, c = gpr.predict(xq, return_cov=True, conditioning_method=method)
    np.testing.assert_allclose(m2, m_ref, atol=1e-8)
    np.testing.assert_allclose(c, c_ref, atol=1e-8)
    assert gpr.predict(xq, conditioning_method=method).shape == m_ref.shape


def test_estimator_small_m_vs_oracle():
    x, y, xq = data(400, 150, 2, 12)
    gpr = TorchNngpr(DEV, kernel=bk_kern(), alpha=ALPHA, num_nn=6).fit(x, y[:, 0])
    nn = so.conditioning_nn(x, xq, LS, 6, "full")
    m_o, s_o = so.predict_sequential(x, y[:, 0], xq, nn, W, 0.0, S, LS, ALPHA, "std")
    m, s = gpr.predict(xq, return_std=True, conditioning_method="full")
    np.testing.assert_allclose(m, m_o[:, 0], atol=1e-10)
    np.testing.assert_allclose(s, s_o, atol=1e-9)
    nn = so.conditioning_nn(x, xq, LS, 6, "only-train")
    _, c_o = so.predict_sequential(x, y[:, 0], xq, nn, W, 0.0, S, LS, ALPHA, "cov")
    _, c = gpr.predict(xq, return_cov=True)
    np.testing.assert_allclose(c, c_o, atol=1e-10)


def test_prior_sequential():
    """Unfitted estimator: the sequential prior over the queries alone."""
    _, _, xq = data(0, 40, 2, 13)
    gpr = TorchNngpr(DEV, kernel=bk_kern(), num_nn=5)
    nn = so.conditioning_nn(np.zeros((0, 2)), xq, LS, 5, "full")
    _, c_o = so.predict_sequential(np.zeros((0, 2)), np.zeros((0, 1)), xq, nn, W, 0.0, S, LS,
                                   1e-10, "cov")
    m, c = gpr.predict(xq, return_cov=True, conditioning_method="full")
    np.testing.assert_allclose(m, 0.0, atol=1e-12)
    np.testing.assert_allclose(c, c_o, atol=1e-10)


@pytest.mark.parametrize("method", ["only-train", "full", "full-double-nn"])
def test_sample_y_moments_and_seed(method):
    """Sample moments match the predictive mean / std (and, for the sequential
    methods, the joint covariance); a seed repeats the draws."""
    x, y, xq = data(150, 12, 2, 14)
    gpr = TorchNngpr(DEV, kernel=bk_kern(), alpha=ALPHA, num_nn=200, normalize_y=True).fit(x, y)
    ns = 20000
    smp = gpr.sample_y(xq, n_samples=ns, random_state=1, conditioning_method=method)
    assert smp.shape == (12, 2, ns)
    again = gpr.sample_y(xq, n_samples=ns, random_state=1, conditioning_method=method)
    np.testing.assert_array_equal(smp, again)
    mean, sd = gpr.predict(xq, return_std=True, conditioning_method=method)
    assert np.all(np.abs(smp.mean(-1) - mean) <= 5 * sd / np.sqrt(ns))
    if method != "only-train":
        _, cov = gpr.predict(xq, return_cov=True, conditioning_method=method)
        c0 = cov[:, :, 0]
        emp = np.cov(smp[:, 0, :])
        scale = np.sqrt(np.outer(np.diag(c0), np.diag(c0)))
        assert np.all(np.abs(emp - c0) <= 6 * scale / np.sqrt(ns))
    one = gpr.sample_y(xq[:3], n_samples=2, random_state=0, conditioning_method=method)
    assert one.shape == (3, 2, 2)


def test_invalid_neighbour_lists_rejected():
    x, y, xq = data(50, 10, 2, 15)
    nn = so.conditioning_nn(x, xq, LS, 4, "full")
    bad = nn.copy()
    bad[3, 0] = 50 + 3          # the query itself: not an earlier query
    with pytest.raises(ValueError):
        hip_ops.predict_sequential(t(x), t(y), t(xq), t(bad, torch.int32), kstruct(), ALPHA)
    bad[3, 0] = 50 + 10         # past the end of [x_train; x_query]
    with pytest.raises(ValueError):
        hip_ops.predict_sequential(t(x), t(y), t(xq), t(bad, torch.int32), kstruct(), ALPHA)
