"""HIP kNN vs the FP64 oracle: neighbour indices must be bit-identical.

Covers both search forms (exact FP64 scan; FP32 screen + FP64 re-rank with
certificate), Vecchia cuts like the reference's tests (None, 0, 2, -1),
isotropic / ARD / raw coordinates, ties and duplicates, empty inputs, k > 64 and
d > 8 (per-row path), the stored golden vectors, and a slice of the 434,874-point
road-mixture workload.
"""
import os

import numpy as np
import pytest
import torch

from nngpr_amd import datasets, hip_ops
from oracle import vecchia_oracle as vo

pytestmark = pytest.mark.gpu
DEV = "cuda:0"
GOLD = os.path.join(os.path.dirname(__file__), "golden", "small_uniform.npz")


def t(a):
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64)).to(DEV)


def hip_knn(ref, qry, ls, k, cut, prec):
    out = hip_ops.knn(t(ref), t(qry), ls, k, ref_cut=cut, precision=prec)
    return out.cpu().numpy()


@pytest.mark.parametrize("prec", [0, 1])
def test_golden_small_uniform(prec):
    g = np.load(GOLD)
    ds = datasets.small_uniform()
    x, xq = ds["x_train"], ds["x_test"]
    np.testing.assert_array_equal(hip_knn(x, x, 0.3, 16, 0, prec), g["nn_train"])
    np.testing.assert_array_equal(hip_knn(x, xq, 0.3, 16, None, prec), g["nn_test"])


@pytest.mark.parametrize("prec", [0, 1])
@pytest.mark.parametrize("cut", [None, 0, 2, -1])
@pytest.mark.parametrize("d,ls", [(1, 0.2), (2, None), (3, "ard"), (5, 0.7), (8, "ard")])
@pytest.mark.parametrize("k", [1, 7, 33, 64])
def test_against_oracle(prec, cut, d, ls, k):
    rng = np.random.default_rng(100 + d * 7 + k)
    ref = rng.normal(size=(2100, d))
    qry = ref[:700] if cut is not None else rng.normal(size=(700, d))
    if ls == "ard":
        ls = rng.uniform(0.2, 2.0, size=d)
    exp = vo.knn(ref, qry, ls, k, ref_cut=cut)
    got = hip_knn(ref, qry, ls, k, cut, prec)
    np.testing.assert_array_equal(got, exp)


@pytest.mark.parametrize("prec", [0, 1])
@pytest.mark.parametrize("d,k", [(18, 32), (3, 100), (12, 5)])
def test_per_row_path(prec, d, k):
    rng = np.random.default_rng(7)
    x = rng.normal(size=(900, d))
    exp = vo.knn(x, x, 1.3, k, ref_cut=0)
    np.testing.assert_array_equal(hip_knn(x, x, 1.3, k, 0, prec), exp)


@pytest.mark.parametrize("prec", [0, 1])
def test_ties_and_duplicates(prec):
    rng = np.random.default_rng(5)
    pts = rng.integers(0, 5, size=(3000, 3)).astype(np.float64)
    for cut in (0, None):
        exp = vo.knn(pts, pts, None, 24, ref_cut=cut)
        np.testing.assert_array_equal(hip_knn(pts, pts, None, 24, cut, prec), exp)
    ones = np.ones((70, 3))
    exp = vo.knn(ones, ones, 1.0, 16, ref_cut=0)
    np.testing.assert_array_equal(hip_knn(ones, ones, 1.0, 16, 0, prec), exp)


@pytest.mark.parametrize("prec", [0, 1])
def test_empty_inputs(prec):
    x = np.random.default_rng(0).random((10, 3))
    out = hip_ops.knn(t(x), t(np.zeros((0, 3))), None, 4, precision=prec)
    assert out.shape == (0, 4)
    out = hip_ops.knn(t(np.zeros((0, 3))), t(x), None, 4, precision=prec).cpu().numpy()
    assert np.all(out == -1)


@pytest.mark.parametrize("prec", [0, 1])
def test_road_mixture_slice(prec):
    ds = datasets.road_mixture(n_train=60_000, n_test=4_000)
    x, xq = ds["x_train"], ds["x_test"]
    exp = vo.knn(x, x, 0.1, 32, ref_cut=0)
    np.testing.assert_array_equal(hip_knn(x, x, 0.1, 32, 0, prec), exp)
    exp = vo.knn(x, xq, 0.1, 32)
    np.testing.assert_array_equal(hip_knn(x, xq, 0.1, 32, None, prec), exp)
