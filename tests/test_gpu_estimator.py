"""The drop-in facades (TorchNngpr, NumpyNngpr) against scikit-learn's exact GP,
restating the reference's estimator known-answer tests
(nngpr_base_tester.py:174-257) on the HIP engine: with num_nn >= n the Vecchia
factorisation is exact, so log-likelihood, gradient, fitted theta and
predictions must match scikit-learn."""
import numpy as np
import pytest
from sklearn import gaussian_process as skgp

from nngpr_amd import NumpyNngpr, TorchNngpr
from nngpr_amd import batched_kernels as bk
from nngpr_amd.array_module import NumpyModule

pytestmark = pytest.mark.gpu


def data():
    rng = np.random.RandomState(234058)
    x = rng.rand(150, 4)
    x_eval = rng.rand(50, 4)
    beta = rng.rand(4, 3)
    beta[:, -1] += 1
    y = x @ beta + rng.rand(150, 3)
    y[:, 0] *= 10
    return x, x_eval, y


def kern(fixed=False):
    if fixed:
        return bk.ConstantKernel(1, constant_value_bounds="fixed") * bk.RBF(
            length_scale=1, length_scale_bounds="fixed") + bk.WhiteKernel(1, noise_level_bounds="fixed")
    return bk.ConstantKernel(1) * bk.RBF(length_scale=1) + bk.WhiteKernel(1)


def make(cls, **kw):
    if cls is TorchNngpr:
        return TorchNngpr("cuda:0", **kw)
    return NumpyNngpr(**kw)


@pytest.mark.parametrize("cls", [TorchNngpr, NumpyNngpr])
@pytest.mark.parametrize("y_dim", [1, 100])
def test_same_lkl(cls, y_dim):
    x, _, y = data()
    if y_dim == 100:
        y = np.tile(y, (1, 34))[:, :100]
    theta = np.full(3, 0.5)
    ref = skgp.GaussianProcessRegressor(kernel=kern().bind_to_array_module(NumpyModule()),
                                        normalize_y=True)
    ref.X_train_, ref.y_train_, ref.kernel_ = x, y[:, :y_dim], kern().bind_to_array_module(NumpyModule())
    ll_ref, g_ref = ref.log_marginal_likelihood(theta, eval_gradient=True)
    gpr = make(cls, kernel=kern(), normalize_y=True, num_nn=x.shape[0])
    gpr.X_train_, gpr.y_train_, gpr.kernel_ = x, y[:, :y_dim], kern()
    ll, g = gpr.log_marginal_likelihood(theta, eval_gradient=True)
    assert abs(ll / ll_ref - 1) < 1e-10
    assert np.max(np.abs(g / g_ref - 1)) < 1e-9


@pytest.mark.parametrize("y_dim", [1, 3])
def test_same_fit_results(y_dim):
    x, _, y = data()
    ref = skgp.GaussianProcessRegressor(kernel=kern().bind_to_array_module(NumpyModule()),
                                        normalize_y=True).fit(x, y[:, :y_dim])
    gpr = make(TorchNngpr, kernel=kern(), normalize_y=True, num_nn=x.shape[0]).fit(x, y[:, :y_dim])
    assert np.max(np.abs(ref.kernel_.theta / gpr.kernel_.theta - 1)) < 1e-6


def test_same_fit_results_with_restarts():
    """n_restarts_optimizer: the same restart draws as scikit-learn (theta0 uniform
    in the log bounds from random_state), so the best optimum matches."""
    x, _, y = data()
    ref = skgp.GaussianProcessRegressor(kernel=kern().bind_to_array_module(NumpyModule()),
                                        normalize_y=True, n_restarts_optimizer=2,
                                        random_state=3).fit(x, y[:, 0])
    gpr = make(TorchNngpr, kernel=kern(), normalize_y=True, num_nn=x.shape[0],
               n_restarts_optimizer=2, random_state=3).fit(x, y[:, 0])
    assert abs(gpr.log_marginal_likelihood_value_ / ref.log_marginal_likelihood_value_ - 1) < 1e-8
    np.testing.assert_allclose(gpr.kernel_.theta, ref.kernel_.theta, rtol=1e-5, atol=1e-6)


@pytest.mark.parametrize("cls", [TorchNngpr, NumpyNngpr])
@pytest.mark.parametrize("use_eval", [True, False])
@pytest.mark.parametrize("y_dim", [1, 3])
def test_predict_same_results(cls, use_eval, y_dim):
    x, x_eval, y = data()
    xq = x_eval if use_eval else x
    ref = skgp.GaussianProcessRegressor(kernel=kern(True).bind_to_array_module(NumpyModule()),
                                        normalize_y=True).fit(x, y[:, :y_dim])
    mean_ref, std_ref = ref.predict(xq, return_std=True)
    gpr = make(cls, kernel=kern(True), normalize_y=True, num_nn=x.shape[0]).fit(x, y[:, :y_dim])
    mean, std = gpr.predict(xq, return_std=True)
    assert mean.shape == mean_ref.shape and std.shape == std_ref.shape
    assert np.max(np.abs(mean - mean_ref)) < 1e-10 * max(1.0, np.abs(mean_ref).max())
    assert np.max(np.abs(std - std_ref)) < 1e-10 * max(1.0, np.abs(std_ref).max())


def test_prior_predict():
    x, _, _ = data()
    ref = skgp.GaussianProcessRegressor(kernel=kern(True).bind_to_array_module(NumpyModule()))
    mean_ref, std_ref = ref.predict(x, return_std=True)
    gpr = make(TorchNngpr, kernel=kern(True), num_nn=x.shape[0])
    mean, std = gpr.predict(x, return_std=True)
    np.testing.assert_allclose(mean, mean_ref, atol=1e-12)
    np.testing.assert_allclose(std, std_ref, atol=1e-12)


def test_small_m_fit_predict_runs_and_improves():
    rng = np.random.default_rng(0)
    x = rng.random((3000, 3))
    y = np.sin(6 * x[:, 0]) + np.cos(4 * x[:, 1]) * x[:, 2] + rng.normal(0, 0.1, 3000)
    k = bk.WhiteKernel(0.5) + bk.ConstantKernel(0.5) * bk.RBF(1.0)
    gpr = TorchNngpr("cuda", kernel=k, num_nn=16).fit(x, y)
    assert gpr.log_marginal_likelihood_value_ > gpr.log_marginal_likelihood(k.theta)
    xq = rng.random((200, 3))
    mean, std = gpr.predict(xq, return_std=True)
    truth = np.sin(6 * xq[:, 0]) + np.cos(4 * xq[:, 1]) * xq[:, 2]
    assert np.sqrt(np.mean((mean - truth) ** 2)) < 0.1
    assert np.all(std > 0)


def test_unsupported_paths_raise():
    x, x_eval, y = data()
    gpr = make(TorchNngpr, kernel=kern(True), num_nn=10).fit(x, y[:, 0])
    with pytest.raises(RuntimeError):
        gpr.predict(x_eval, return_std=True, return_cov=True)
    with pytest.raises(ValueError):
        gpr.predict(x_eval, conditioning_method="nearest")
    m = TorchNngpr("cuda", kernel=bk.RBF(1.0) + bk.RBF(2.0), num_nn=5, optimizer=None)
    with pytest.raises(NotImplementedError):
        m.fit(x, y[:, 0])


def ydim(y, y_dim):
    return np.tile(y, (1, 34))[:, :y_dim] if y_dim > 3 else y[:, :y_dim]


@pytest.mark.parametrize("method", ["only-train", "full", "full-double-nn"])
@pytest.mark.parametrize("use_eval", [True, False])
@pytest.mark.parametrize("y_dim", [1, 100])
def test_predict_same_results_conditioning(method, use_eval, y_dim):
    """nngpr_base_tester.py:228-257: with enough neighbours every conditioning
    method gives scikit-learn's posterior mean and std, and the sequential ones its
    joint covariance (1e-10 absolute)."""
    x, x_eval, y = data()
    xq = x_eval if use_eval else x
    yy = ydim(y, y_dim)
    ref = skgp.GaussianProcessRegressor(kernel=kern(True).bind_to_array_module(NumpyModule()),
                                        normalize_y=True).fit(x, yy)
    mean_ref, std_ref = ref.predict(xq, return_std=True)
    _, cov_ref = ref.predict(xq, return_cov=True)
    num_nn = x.shape[0] + xq.shape[0] if method == "full" else x.shape[0]
    gpr = make(TorchNngpr, kernel=kern(True), normalize_y=True, num_nn=num_nn).fit(x, yy)
    _, cov = gpr.predict(xq, return_cov=True, conditioning_method=method)
    mean, std = gpr.predict(xq, return_std=True, conditioning_method=method)
    assert mean.shape == mean_ref.shape and std.shape == std_ref.shape
    assert cov.shape == cov_ref.shape
    assert np.max(np.abs(mean - mean_ref)) < 1e-10 * max(1.0, np.abs(mean_ref).max())
    assert np.max(np.abs(std - std_ref)) < 1e-10 * max(1.0, np.abs(std_ref).max())
    if method != "only-train":
        assert np.max(np.abs(cov - cov_ref)) < 1e-10 * max(1.0, np.abs(cov_ref).max())


@pytest.mark.parametrize("method", ["only-train", "full", "full-double-nn"])
def test_prior_predict_same_results_conditioning(method):
    """nngpr_base_tester.py:207-226: unfitted, every conditioning method gives
    scikit-learn's prior mean and std, and the sequential ones its covariance."""
    x, _, _ = data()
    ref = skgp.GaussianProcessRegressor(kernel=kern().bind_to_array_module(NumpyModule()),
                                        normalize_y=True)
    mean_ref, std_ref = ref.predict(x, return_std=True)
    _, cov_ref = ref.predict(x, return_cov=True)
    gpr = make(TorchNngpr, kernel=kern(True), normalize_y=True, num_nn=x.shape[0])
    _, cov = gpr.predict(x, return_cov=True, conditioning_method=method)
    mean, std = gpr.predict(x, return_std=True, conditioning_method=method)
    assert np.max(np.abs(mean - mean_ref)) < 1e-10
    assert np.max(np.abs(std - std_ref)) < 1e-10
    if method != "only-train":
        assert np.max(np.abs(cov - cov_ref)) < 1e-10


@pytest.mark.parametrize("nn_type", ["kernel-space", "input-space"])
def test_runs(nn_type):
    """nngpr_base_tester.py:148-172: the public calls run, before and after fit."""
    x, x_eval, y = data()
    gpr = make(TorchNngpr, kernel=kern(), num_nn=8, nn_type=nn_type, batch_size=7)
    gpr.predict(x_eval, return_std=True)
    gpr.predict(x_eval, return_cov=True)
    gpr.sample_y(x_eval, n_samples=2)
    gpr.fit(x, y)
    gpr.predict(x)
    gpr.predict(x, return_std=True)
    gpr.predict(x_eval, return_std=True)
    gpr.predict(x, return_cov=True)
    gpr.predict(x_eval, return_cov=True)
    s = gpr.sample_y(x_eval, n_samples=2)
    assert s.shape == (x_eval.shape[0], y.shape[1], 2)


@pytest.mark.parametrize("method", ["only-train", "full", "full-double-nn"])
def test_prior_sample(method):
    """nngpr_base_tester.py:259-280: prior samples agree with the predictions as
    closely as scikit-learn's own samples agree with its predictions (2x margin)."""
    x, _, _ = data()
    n_samples = 10000
    gpr = make(TorchNngpr, kernel=kern(True), normalize_y=True, num_nn=x.shape[0])
    samples = gpr.sample_y(x, n_samples=n_samples, random_state=1, conditioning_method=method)
    mean_pred, std_pred = gpr.predict(x, conditioning_method=method, return_std=True)
    ref = skgp.GaussianProcessRegressor(kernel=kern(True).bind_to_array_module(NumpyModule()),
                                        normalize_y=True)
    samples_ref = ref.sample_y(x, n_samples=n_samples, random_state=1)
    mean_ref, std_ref = ref.predict(x, return_std=True)
    mean_delta = np.max(np.abs(np.mean(samples_ref, axis=1) - mean_ref))
    assert np.max(np.abs(np.mean(samples, axis=1) - mean_pred)) < mean_delta * 2
    std_delta = np.max(np.abs(np.std(samples_ref, axis=1) - std_ref))
    assert np.max(np.abs(np.std(samples, axis=1) - std_pred)) < std_delta * 2


@pytest.mark.parametrize("method", ["only-train", "full", "full-double-nn"])
@pytest.mark.parametrize("use_eval", [False, True])
@pytest.mark.parametrize("y_dim", [1, 100])
def test_sample(method, use_eval, y_dim):
    """nngpr_base_tester.py:282-310: posterior samples agree with the predictions
    as closely as scikit-learn's samples agree with its own (2x margin)."""
    x, x_eval, y = data()
    xq = x_eval if use_eval else x
    yy = ydim(y, y_dim)
    n_samples = 1000
    gpr = make(TorchNngpr, kernel=kern(True), normalize_y=True, num_nn=x.shape[0]).fit(x, yy)
    samples = gpr.sample_y(xq, n_samples=n_samples, random_state=1, conditioning_method=method)
    mean_pred, std_pred = gpr.predict(xq, conditioning_method=method, return_std=True)
    ref = skgp.GaussianProcessRegressor(kernel=kern(True).bind_to_array_module(NumpyModule()),
                                        normalize_y=True).fit(x, yy)
    samples_ref = ref.sample_y(xq, n_samples=n_samples, random_state=1)
    mean_ref, std_ref = ref.predict(xq, return_std=True)
    assert samples.shape == samples_ref.shape
    mean_delta = np.max(np.abs(np.mean(samples_ref, axis=-1) - mean_ref))
    assert np.max(np.abs(np.mean(samples, axis=-1) - mean_pred)) < mean_delta * 2
    std_delta = np.max(np.abs(np.std(samples_ref, axis=-1) - std_ref))
    assert np.max(np.abs(np.std(samples, axis=-1) - std_pred)) < std_delta * 2


def test_large_sample_size():
    """nngpr_base_tester.py:312-330: 2 GB of only-train samples."""
    x, _, _ = data()
    gpr = make(TorchNngpr, kernel=kern(True), normalize_y=True, num_nn=32)
    n_samples = int(2 * 1024 ** 3 / x.shape[0] / 8)
    samples = gpr.sample_y(x, n_samples=n_samples, random_state=1, conditioning_method="only-train")
    assert samples.shape == (x.shape[0], n_samples)
    del samples
    xl = np.concatenate([x] * int(n_samples / x.shape[0] / 10), axis=0)
    n_samples = int(2 * 1024 ** 3 / xl.shape[0] / 8)
    samples = gpr.sample_y(xl, n_samples=n_samples, random_state=2, conditioning_method="only-train")
    assert samples.shape == (xl.shape[0], n_samples)
