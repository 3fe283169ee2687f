"""Exact kNN lists at new length scales from one search (nngp_knn_rescale).

The lists must equal, bit for bit, a fresh nngp_knn search at the new scale:
isotropic and ARD changes, small and large steps, every Vecchia cut the
reference tests (None, 0, 2, -1), exact ties and duplicated points, non-finite
coordinates, the full road-mixture training set; and a fit that reuses one search
across its evaluations follows exactly the same optimizer path as a fit that
searches at every evaluation (the reference's behaviour, base_nngpr.py:530-548).
"""
import numpy as np
import pytest
import torch

from nngpr_amd import TorchNngpr, datasets, hip_ops
from nngpr_amd import batched_kernels as bk
from oracle import vecchia_oracle as vo

pytestmark = pytest.mark.gpu
DEV = "cuda:0"


def t(a):
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64)).to(DEV)


def check_scales(x, q, s0, scales, k, cut, K=None):
    """Rescaled lists equal fresh searches at every scale; returns the rows
    recomputed per scale."""
    xt = t(x)
    qt = xt if q is None else t(q)
    K = K or k + 8
    r = hip_ops.KnnRescaler(xt, qt, s0, k, K, ref_cut=cut)
    fresh0 = hip_ops.knn(xt, qt, s0, k, ref_cut=cut)
    assert torch.equal(r.base_lists(), fresh0)
    counts = []
    for s in scales:
        got = r.lists(s)
        counts.append(r.rows_recomputed())
        want = hip_ops.knn(xt, qt, s, k, ref_cut=cut)
        assert torch.equal(got, want), "scale %r" % (s,)
    return counts


@pytest.mark.parametrize("cut", [None, 0, 2, -1])
def test_isotropic_changes(cut):
    rng = np.random.default_rng(1)
    x = rng.random((20_000, 3))
    q = None if cut is not None else rng.random((3_000, 3))
    scales = [0.1, 0.05, 0.2, 0.1 * (1 + 1e-12), 0.37, 1e-3, 10.0, 1e4]
    counts = check_scales(x, q, 0.1, scales, 16, cut)
    # an isotropic change keeps every row with separated keys: few recomputed rows
    assert max(counts) < 0.01 * x.shape[0]


@pytest.mark.parametrize("cut", [None, 0])
@pytest.mark.parametrize("d", [6, 18])
def test_ard_changes(cut, d):
    rng = np.random.default_rng(2 + d)
    x = rng.normal(size=(6_000, d))
    q = None if cut is not None else rng.normal(size=(1_000, d))
    s0 = rng.uniform(0.5, 2.0, d)
    scales = [s0, s0 * (1 + 1e-4 * rng.normal(size=d)), s0 * (1 + 1e-2 * rng.normal(size=d)),
              s0 * rng.uniform(0.5, 2.0, d), np.full(d, 1.0)]
    check_scales(x, q, s0, scales, 32, cut)


def test_ties_duplicates_and_non_finite():
    rng = np.random.default_rng(3)
    g = np.stack(np.meshgrid(np.arange(12), np.arange(12), np.arange(12)), -1).reshape(-1, 3)
    x = np.concatenate([g.astype(np.float64), g[:300] + 0.0, rng.random((500, 3)) * 11])
    x = x[rng.permutation(x.shape[0])]
    x[[5, 77, 901]] = np.nan
    x[1000, 1] = np.inf
    for cut in (None, 0):
        check_scales(x, None if cut == 0 else x[::3], 1.0, [1.0, 0.5, 3.0, 1.0 + 1e-9, 0.01],
                     20, cut)
    check_scales(x, None, np.array([1.0, 2.0, 0.5]),
                 [np.array([1.0, 2.0, 0.5]), np.array([2.0, 4.0, 1.0]), np.array([1.0, 1.0, 1.0])],
                 20, 0)


def test_small_sets_and_complete_lists():
    """Rows with fewer refs than K list every ref: exact at any scale."""
    rng = np.random.default_rng(4)
    for n, k in ((1, 1), (5, 3), (30, 16), (60, 32)):
        x = rng.random((n, 2))
        check_scales(x, None, 0.3, [0.3, 0.1, 2.0], k, 0, K=min(k + 8, 72))


def test_full_road_mixture_against_oracle_rows():
    ds = datasets.road_mixture()
    x = ds["x_train"]
    xt = t(x)
    r = hip_ops.KnnRescaler(xt, xt, 0.1, 32, 40, ref_cut=0)
    for s in (0.13, 0.071):
        got = r.lists(s)
        assert r.rows_recomputed() < 0.01 * x.shape[0]
        assert torch.equal(got, hip_ops.knn(xt, xt, s, 32, ref_cut=0))
        g = got.cpu().numpy()
        for row in np.random.default_rng(5).integers(1, x.shape[0], 20):
            np.testing.assert_array_equal(g[row], vo.knn(x[:row], x[row:row + 1], s, 32)[0])


@pytest.mark.parametrize("ard", [False, True])
def test_fit_reuse_follows_the_same_path(ard):
    """L-BFGS-B from the same start with the search reused across evaluations and
    with a search at every evaluation: identical theta, log-likelihood and
    neighbour lists."""
    ds = datasets.road_mixture(n_train=30_000, n_test=10)
    x, y = ds["x_train"], ds["y_train"]
    ls = np.array([0.1, 0.12, 0.09]) if ard else 0.1
    kern = bk.WhiteKernel(0.01) + bk.ConstantKernel(1.0) * bk.RBF(ls)
    fits = []
    for reuse in (True, False):
        est = TorchNngpr("cuda:0", kernel=kern, num_nn=16, alpha=1e-10)
        est._knn_reuse = reuse
        est.fit(x, y)
        fits.append(est)
    np.testing.assert_array_equal(fits[0].kernel_.theta, fits[1].kernel_.theta)
    assert fits[0].log_marginal_likelihood_value_ == fits[1].log_marginal_likelihood_value_


def test_fit_reuse_policy():
    """When a search is reused and when a new one runs: lists of k + 8 above the
    parallel searches' limit are never reused (every evaluation searches k
    lists), an anisotropic length-scale change beyond (K / k)^(1/d) starts a new
    search, a smaller one rescales the stored candidates; the log-likelihood is
    the same as with a search at every evaluation in every case."""
    ds = datasets.road_mixture(n_train=20_000, n_test=10)
    x, y = ds["x_train"], ds["y_train"]
    kern = bk.WhiteKernel(0.01) + bk.ConstantKernel(1.0) * bk.RBF(np.array([0.1, 0.1, 0.1]))
    ref = TorchNngpr("cuda:0", kernel=kern, num_nn=16, alpha=1e-10, optimizer=None)
    ref._knn_reuse = False
    ref.fit(x, y)
    est = TorchNngpr("cuda:0", kernel=kern, num_nn=16, alpha=1e-10, optimizer=None)
    est.fit(x, y)
    cache = est._dev_cache[3]
    base = cache.get("rescaler")
    assert base is not None
    th0 = est.kernel_.theta.copy()
    th_small = th0.copy()
    th_small[2:5] += np.log([1.0, 1.01, 1.02])   # spread 1.02 < (24 / 16)^(1/3)
    th_big = th0.copy()
    th_big[2:5] += np.log([1.0, 1.3, 0.8])       # spread 1.63
    for th, same_base in ((th_small, True), (th_big, False)):
        ll = est.log_marginal_likelihood(th)
        assert (cache.get("rescaler") is base) == same_base
        base = cache.get("rescaler")
        assert ll == ref.log_marginal_likelihood(th)
    est64 = TorchNngpr("cuda:0", kernel=kern, num_nn=64, alpha=1e-10, optimizer=None)
    est64.fit(x, y)
    est64.log_marginal_likelihood(th_small)
    assert est64._dev_cache[3].get("rescaler") is None
