"""HipWorker, the NngprWorker-compatible plug-in surface, restating the
reference's WorkerTester (nngpr_base_tester.py:12-115) on the HIP device: the
array module, find_nearest_neighbors for both neighbour types and every
ref_cut_index the reference tests (None, 0, 2, -1) on all-equal points,
batched_chofactor / batched_chosolve against numpy's solve, fill_nn_kernel with
and without gradients, and random_normal.  Beyond the reference's type checks,
the values are compared with the oracle (neighbour lists, covariance blocks and
their gradients)."""
import itertools

import numpy as np
import pytest
import torch

from nngpr_amd import batched_kernels as bk
from nngpr_amd.base_nngpr import HipWorker
from oracle import vecchia_oracle as vo

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def worker():
    return HipWorker("cuda:0")


def test_empty(worker):
    xp = worker.get_array_module()
    for shape, dtype in itertools.product([2, (5, 6)], [np.float32, np.float64]):
        a = xp.empty(shape, dtype=dtype)
        assert a.is_cuda and tuple(a.shape) == ((2,) if shape == 2 else shape)


def test_get_array_module(worker):
    xp = worker.get_array_module()
    x1 = xp.empty((5, 2, 3), dtype=np.float32)
    x2 = xp.empty((10, 2, 3), dtype=np.float32)
    assert xp.moveaxis(x1, 0, 2).shape == (2, 3, 5)
    assert xp.swapaxes(x2, 1, 2).shape == (10, 3, 2)
    xp.log(x1 + 2)
    xp.sqrt(x2)
    xp.arange(10)
    xp.where(x1 > 2)
    xp.sum(x1)
    assert xp.tile(x1, (3, 1, 1)).shape == (15, 2, 3)
    x2 = xp.ones((10, 3, 4, 4), dtype=np.float32)
    assert torch.equal(x2 @ x2, torch.full((10, 3, 4, 4), 4.0, device=x2.device))


def test_to_float32(worker):
    xp = worker.get_array_module()
    x1 = xp.ones((5, 2, 3), dtype=np.float32)
    x2 = xp.ones((10, 2, 3), dtype=np.float64)
    assert xp.to_float32(x1) is x1
    assert xp.to_float32(x2) is not x2 and xp.to_float32(x2).dtype == torch.float32


def test_to_and_from_numpy(worker):
    xp = worker.get_array_module()
    assert type(xp.to_numpy(xp.empty(3, dtype=np.float32))) is np.ndarray
    assert type(xp.from_numpy(np.ones(6))) is type(xp.empty(3, dtype=np.float32))


@pytest.mark.parametrize("nn_type", ["kernel-space", "input-space"])
@pytest.mark.parametrize("cut", [None, 0, 2, -1])
def test_find_nearest_neighbors_all_equal_points(worker, nn_type, cut):
    """nngpr_base_tester.py:70-78: ref = ones (10, 3), query = ones + 2, k = 2.
    Every key is equal, so the lists are the lowest admissible indices."""
    xp = worker.get_array_module()
    ref = xp.ones((10, 3))
    query = xp.ones((4, 3)) + 2
    k = bk.RBF(length_scale=1.0)
    res = worker.find_nearest_neighbors(ref, query, nn_type, k, 2, ref_cut_index=cut)
    assert type(res) is type(ref)
    np.testing.assert_array_equal(res.cpu().numpy(),
                                  vo.knn(np.ones((10, 3)), np.full((4, 3), 3.0), None, 2, ref_cut=cut))
    lim = [10 if cut is None else min(max(i + cut, 0), 10) for i in range(4)]
    expect = np.array([[j if j < l else -1 for j in range(2)] for l in lim])
    np.testing.assert_array_equal(res.cpu().numpy(), expect)


@pytest.mark.parametrize("nn_type", ["kernel-space", "input-space"])
@pytest.mark.parametrize("cut", [None, 0, 2, -1])
def test_find_nearest_neighbors_values(worker, nn_type, cut):
    rng = np.random.default_rng(17)
    ref = rng.random((500, 4))
    query = ref[:120] if cut is not None else rng.random((120, 4))
    ls = np.array([0.3, 1.2, 0.7, 2.0])
    k = bk.ConstantKernel(1.5) * bk.RBF(ls) + bk.WhiteKernel(0.1)
    res = worker.find_nearest_neighbors(ref, query, nn_type, k, 9, ref_cut_index=cut)
    scale = ls if nn_type == "kernel-space" else None
    np.testing.assert_array_equal(res.cpu().numpy(), vo.knn(ref, query, scale, 9, ref_cut=cut))


def test_chofactor_and_chosolve(worker):
    """nngpr_base_tester.py:80-92."""
    rng = np.random.default_rng(3)
    x = rng.random((3, 5, 5))
    x = x @ np.swapaxes(x, -1, -2) + np.eye(5)[None]
    y = rng.random((3, 5, 8))
    xp = worker.get_array_module()
    L, lower = worker.batched_chofactor(xp.from_numpy(x))
    res = worker.batched_chosolve(L, xp.from_numpy(y), lower)
    assert type(res) is type(xp.from_numpy(x))
    np.testing.assert_allclose(xp.to_numpy(res), np.linalg.solve(x, y), rtol=1e-10, atol=1e-12)


@pytest.mark.parametrize("eval_gradient", [False, True])
def test_fill_nn_kernel(worker, eval_gradient):
    """nngpr_base_tester.py:94-103 (float32 zeros, random indices, RBF(1)), then
    the values on random points against the oracle's fill."""
    xp = worker.get_array_module()
    x = xp.zeros((10, 5), dtype=np.float32)
    nn = np.random.default_rng(4).integers(0, 10, (3, 4)).astype(np.int32)
    k = bk.RBF(length_scale=1.0)
    res = worker.fill_nn_kernel(xp.from_numpy(x), xp.from_numpy(nn), k, eval_gradient=eval_gradient)
    assert type(res[0]) is type(x) and type(res[1]) is type(x)
    np.testing.assert_allclose(res[0].cpu().numpy(), np.ones((3, 4, 4)), rtol=1e-15)
    rng = np.random.default_rng(5)
    xr = rng.random((60, 3))
    nn = rng.integers(-1, 60, (25, 7)).astype(np.int32)
    k = bk.WhiteKernel(0.2) + bk.ConstantKernel(0.7) * bk.RBF(np.array([0.4, 0.9, 0.6]))
    K, G = worker.fill_nn_kernel(xr, nn, k, eval_gradient=eval_gradient)
    valid = nn >= 0
    tm = {"noise_level": 0, "rbf_level": 1, "length_scale": np.arange(2, 5)}
    Ko, Go = vo.fill(xr[np.where(valid, nn, 0)], valid, 0.2, 0.0, 0.7, np.array([0.4, 0.9, 0.6]),
                     tm, 5, eval_gradient)
    np.testing.assert_allclose(K.cpu().numpy(), Ko, rtol=1e-13, atol=1e-15)
    if eval_gradient:
        np.testing.assert_allclose(G.cpu().numpy(), Go, rtol=1e-12, atol=1e-15)


def test_random_normal(worker):
    """nngpr_base_tester.py:105-115, plus the moments of a large draw."""
    mu = np.ones((3, 1, 2, 1))
    sigma = np.ones((1, 5, 2, 1))
    shape = (3, 5, 2, 10)
    xp = worker.get_array_module()
    for random_state in [None, 0, 324]:
        s = worker.check_random_state(random_state)
        res = worker.random_normal(s, mu, sigma, shape)
        assert type(res) is type(xp.from_numpy(mu)) and tuple(res.shape) == shape
    g = worker.check_random_state(7)
    z = worker.random_normal(g, np.full((1,), 2.0), np.full((1,), 3.0), (400_000,)).cpu().numpy()
    assert abs(z.mean() - 2.0) < 0.02 and abs(z.std() - 3.0) < 0.02
    a = worker.random_normal(worker.check_random_state(9), mu, sigma, shape)
    b = worker.random_normal(worker.check_random_state(9), mu, sigma, shape)
    assert torch.equal(a, b)
