"""The batched kernels against their scikit-learn twins.

Restates the reference's test_batched_kernels.py:11-138 (perform_test_on_kernel
over Constant, White, RBF (isotropic and ARD), Matern nu in {0.5, 1.5, 2.5, inf}
and every Sum / Product of the four): values and gradients equal scikit-learn's,
batched inputs (*batch, n, d) give the broadcast result, float32 inputs stay
float32, the diagonal, and the two-argument form K(X, Y).  Run over the numpy
module and torch on the CPU here, and torch on the GPU under the gpu marker.
"""
import itertools

import numpy as np
import pytest
import torch
from sklearn.gaussian_process import kernels as skk

from nngpr_amd import batched_kernels as bk
from nngpr_amd.array_module import NumpyModule, TorchModule

RNG = np.random.default_rng(290835)
X = RNG.random((10, 3))
Y = RNG.random((5, 3))
BATCH = (2, 3, 4)


def modules():
    mods = [pytest.param(NumpyModule(), id="numpy"), pytest.param(TorchModule("cpu"), id="torch-cpu")]
    mods.append(pytest.param("cuda", id="torch-gpu", marks=pytest.mark.gpu))
    return mods


@pytest.fixture(params=modules())
def xp(request):
    return TorchModule("cuda:0") if request.param == "cuda" else request.param


def all_kernels():
    return [(bk.ConstantKernel(3.1), skk.ConstantKernel(3.1)),
            (bk.WhiteKernel(3.2), skk.WhiteKernel(3.2)),
            (bk.RBF(3.3), skk.RBF(3.3)),
            (bk.Matern(1.5), skk.Matern(1.5))]


def close(a, b):
    np.testing.assert_allclose(a, b, rtol=1e-12, atol=1e-14)


def check_kernel(xp, k_in, k_ref):
    """test_batched_kernels.py:31-114 (perform_test_on_kernel)."""
    for eval_grad in (False, True):
        k = k_in.bind_to_array_module(xp)
        x, y = xp.from_numpy(X), xp.from_numpy(Y)
        res = k(x, eval_gradient=eval_grad)
        res_ref = k_ref(X, eval_gradient=eval_grad)
        grad = grad_ref = None
        if eval_grad:
            assert type(res[1]) is type(x)
            close(xp.to_numpy(res[1]), res_ref[1])
            grad, grad_ref = res[1], res_ref[1]
            res, res_ref = res[0], res_ref[0]
        assert type(res) is type(x)
        res = xp.to_numpy(res)
        close(res, res_ref)
        # batched input
        new_x = xp.tile(x, (*BATCH, *(1 for _ in range(X.ndim))))
        new_res = k(new_x, eval_gradient=eval_grad)
        if eval_grad:
            assert tuple(new_res[1].shape) == (*BATCH, *grad.shape)
            close(xp.to_numpy(new_res[1]), np.broadcast_to(grad_ref, (*BATCH, *grad_ref.shape)))
            new_res = new_res[0]
        assert tuple(new_res.shape) == (*BATCH, *res.shape)
        close(xp.to_numpy(new_res), np.broadcast_to(res_ref, (*BATCH, *res_ref.shape)))
        # float32 in, float32 out
        tmp = k(xp.from_numpy(X.astype("float32")), eval_gradient=eval_grad)
        if eval_grad:
            assert type(tmp[1]) is type(x)
            assert xp.to_numpy(tmp[0]).dtype == np.float32
            assert xp.to_numpy(tmp[1]).dtype == np.float32
            continue
        assert xp.to_numpy(tmp).dtype == np.float32
        # the diagonal
        diag_ref = k_ref.diag(X)
        diag = k.diag(x)
        assert type(diag) is type(x)
        diag = xp.to_numpy(diag)
        assert diag.shape == diag_ref.shape
        close(diag, diag_ref)
        diag_b = xp.to_numpy(k.diag(new_x))
        assert diag_b.shape == (*BATCH, *diag_ref.shape)
        close(diag_b, np.broadcast_to(diag_ref, (*BATCH, *diag_ref.shape)))
        assert xp.to_numpy(k.diag(xp.from_numpy(X.astype("float32")))).dtype == np.float32
        # K(X, Y)
        res = k(x, y)
        res_ref = k_ref(X, Y)
        assert type(res) is type(x)
        close(xp.to_numpy(res), res_ref)
        new_y = xp.tile(y, (*BATCH, *(1 for _ in range(Y.ndim))))
        res = k(new_x, new_y)
        assert tuple(res.shape) == (*BATCH, *res_ref.shape)
        close(xp.to_numpy(res), np.broadcast_to(res_ref, (*BATCH, *res_ref.shape)))
        res = xp.to_numpy(k(xp.from_numpy(X.astype("float32")), xp.from_numpy(Y.astype("float32"))))
        assert res.dtype == np.float32


def test_constant(xp):
    check_kernel(xp, bk.ConstantKernel(3), skk.ConstantKernel(3))


def test_white(xp):
    check_kernel(xp, bk.WhiteKernel(3.3), skk.WhiteKernel(3.3))


@pytest.mark.parametrize("ard", [False, True])
def test_rbf(xp, ard):
    ls = 6.5 * np.ones(X.shape[1]) if ard else 3.4
    check_kernel(xp, bk.RBF(ls), skk.RBF(ls))


@pytest.mark.parametrize("nu", [0.5, 1.5, 2.5, np.inf])
@pytest.mark.parametrize("ard", [False, True])
def test_matern(xp, nu, ard):
    ls = 6.7 * np.ones(X.shape[1]) if ard else 3.5
    check_kernel(xp, bk.Matern(ls, nu=nu), skk.Matern(ls, nu=nu))


@pytest.mark.parametrize("i,j", list(itertools.product(range(4), range(4))))
def test_sum(xp, i, j):
    ks = all_kernels()
    check_kernel(xp, ks[i][0] + ks[j][0], ks[i][1] + ks[j][1])


@pytest.mark.parametrize("i,j", list(itertools.product(range(4), range(4))))
def test_product(xp, i, j):
    ks = all_kernels()
    check_kernel(xp, ks[i][0] * ks[j][0], ks[i][1] * ks[j][1])


def test_torch_module_keeps_device_and_dtype():
    xp = TorchModule("cpu")
    a = xp.from_numpy(np.ones((3, 2), dtype=np.float32))
    assert isinstance(a, torch.Tensor) and a.dtype == torch.float32
    assert xp.to_float32(xp.from_numpy(np.ones(3))).dtype == torch.float32
