"""torch.autograd access to the Vecchia log-likelihood and its adjoint gradient.

`vecchia_log_likelihood(theta, x, y, nn, kernel, alpha)` returns the Vecchia
log-likelihood of the training rows as a 0-d tensor that is differentiable with
respect to `theta`, the log-hyperparameters of `kernel` (scikit-learn theta order,
as in GaussianProcessRegressor.log_marginal_likelihood).  Forward runs one fused
fill -> Cholesky -> mu/sigma -> loglik pass; when theta requires a gradient that
same pass also runs the hand-written adjoint (which recomputes the neighbour
factors in registers instead of storing them) and keeps the nθ-vector dl/dtheta,
so backward only scales it by the incoming gradient.

This is the gradient of base_nngpr.py:586-614 (same quantity, computed by the
closed-form adjoint rather than K_gradient solves).
"""
import numpy as np
import torch

from . import _native as nat
from . import hip_ops
from .kernel_params import get_kernel_params, get_kernel_theta_mapping


def _kernel_struct(kernel, theta_np):
    k = kernel.clone_with_theta(theta_np)
    return nat.kernel_struct(get_kernel_params(k), get_kernel_theta_mapping(k), len(theta_np))


class VecchiaLogLikelihood(torch.autograd.Function):
    """Autograd node: theta (1-D float64 tensor) -> Vecchia log-likelihood."""

    @staticmethod
    def forward(ctx, theta, x, y, nn, kernel, alpha, precision):
        th = theta.detach().cpu().numpy().astype(np.float64)
        ks = _kernel_struct(kernel, th)
        # one fused pass: the adjoint gradient comes with the value when theta
        # needs it, and backward only scales it
        need = bool(ctx.needs_input_grad[0])
        ll, g, _ = hip_ops.vecchia_loglik(x, y, nn, ks, alpha, precision=precision,
                                          eval_gradient=need)
        ctx.grad = g.to(theta.device, theta.dtype) if need else None
        return ll[0].to(theta.device)

    @staticmethod
    def backward(ctx, gout):
        return (gout * ctx.grad, None, None, None, None, None, None)


def vecchia_log_likelihood(theta, x, y, nn, kernel, alpha=1e-10,
                           precision=hip_ops.PRECISION_FP64):
    """Differentiable Vecchia log-likelihood (see module docstring).

    theta: 1-D float64 tensor of log-hyperparameters (any device);
    x (n, d), y (n,) or (n, ny): float64 device tensors; nn (n, m) int32 device
    tensor of each row's neighbours among earlier rows (hip_ops.knn with
    ref_cut=0); kernel: a White + Constant * RBF kernel (scikit-learn or
    nngpr_amd.batched_kernels) whose theta layout `theta` follows.
    """
    if not isinstance(theta, torch.Tensor):
        theta = torch.as_tensor(np.asarray(theta, dtype=np.float64))
    return VecchiaLogLikelihood.apply(theta, x, y, nn, kernel, float(alpha), int(precision))


def loglik_and_grad(theta, x, y, nn, kernel, alpha=1e-10, precision=hip_ops.PRECISION_FP64):
    """(loglik, dloglik/dtheta) as floats/ndarray from one fused pass: the form an
    L-BFGS objective wants."""
    th = np.asarray(theta, dtype=np.float64)
    ks = _kernel_struct(kernel, th)
    ll, g, _ = hip_ops.vecchia_loglik(x, y, nn, ks, alpha, precision=precision,
                                      eval_gradient=True)
    return float(ll.item()), g.cpu().numpy()
