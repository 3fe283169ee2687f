"""nngp_vecchia_loglik_noalloc (the log-likelihood entry point that allocates
nothing, with its scratch in a fixed device array) against nngp_vecchia_loglik
(caller workspace), both called through the C ABI: identical log-likelihood,
gradient and per-row values on two calls in a row, the second smaller; sizes
beyond the fixed scratch rejected before any launch; failures marked NaN."""
import ctypes

import numpy as np
import pytest
import torch

from nngpr_amd import _native as nat
from nngpr_amd import batched_kernels as bk
from nngpr_amd import datasets, hip_ops
from nngpr_amd.kernel_params import get_kernel_params, get_kernel_theta_mapping
from oracle import vecchia_oracle as vo

pytestmark = pytest.mark.gpu
DEV = "cuda:0"
SCRATCH = 32 << 20


def t(a, dtype=torch.float64):
    return torch.from_numpy(np.ascontiguousarray(a)).to(DEV, dtype)


def _p(x):
    return None if x is None else ctypes.c_void_p(x.data_ptr())


def noalloc(x, y, nn, ks, alpha, prec, row0=0, row1=None, grad=True):
    """(rc, loglik, grad, per-row loglik) of nngp_vecchia_loglik_noalloc."""
    n, d = x.shape
    ny = 1 if y.ndim == 1 else y.shape[1]
    row1 = n if row1 is None else row1
    ll = torch.full((1,), -1.0, dtype=torch.float64, device=DEV)
    g = torch.full((max(ks.n_theta, 1),), -1.0, dtype=torch.float64, device=DEV)
    pts = torch.full((max(row1 - row0, 1),), -1.0, dtype=torch.float64, device=DEV)
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    rc = nat.lib().nngp_vecchia_loglik_noalloc(
        _p(x), _p(y), n, d, ny, _p(nn), nn.shape[1], row0, row1, ctypes.byref(ks), float(alpha),
        prec, 1 if grad else 0, _p(ll), _p(g) if grad else None, _p(pts), stream)
    torch.cuda.synchronize()
    return rc, ll.item(), g.cpu().numpy()[:ks.n_theta], pts.cpu().numpy()[:row1 - row0]


def with_workspace(x, y, nn, ks, alpha, prec, row0=0, row1=None):
    ll, g, pts = hip_ops.vecchia_loglik(x, y, nn, ks, alpha, precision=prec, eval_gradient=True,
                                        row0=row0, row1=row1, point_loglik=True)
    return ll.item(), g.cpu().numpy(), pts.cpu().numpy()


def workload(n, seed):
    ds = datasets.road_mixture(n_train=n, n_test=1) if seed == 0 else None
    if ds is None:
        rng = np.random.default_rng(seed)
        x = rng.random((n, 3))
        y = np.sin(5 * x[:, 0]) + x[:, 1] * x[:, 2] + rng.normal(0, 0.1, n)
        kern = bk.WhiteKernel(0.02) + bk.ConstantKernel(1.3) * bk.RBF(0.2)
    else:
        x, y, kern = ds["x_train"], ds["y_train"], ds["kernel"]
    p, mp = get_kernel_params(kern), get_kernel_theta_mapping(kern)
    ks = nat.kernel_struct(p, mp, len(kern.theta))
    ls = float(np.asarray(p["length_scale"]).reshape(-1)[0])
    xt, yt = t(x), t(y)
    nn = hip_ops.knn(xt, xt, ls, 32, ref_cut=0)
    return xt, yt, nn, ks


@pytest.mark.parametrize("prec", [0, 1])
def test_noalloc_equals_workspace_entry_point(prec):
    """A 434,874-row call, then a 20,000-row call on other data, then a row range:
    each equal to the workspace entry point's result bit for bit (same kernels,
    same reduction order)."""
    cases = [(workload(434_874, 0), 0, None), (workload(20_000, 5), 0, None),
             (workload(20_000, 5), 3_000, 11_000)]
    for (xt, yt, nn, ks), r0, r1 in cases:
        rc, ll, g, pts = noalloc(xt, yt, nn, ks, 1e-10, prec, r0, r1)
        assert rc == 0, nat.status_text(rc)
        ll_w, g_w, pts_w = with_workspace(xt, yt, nn, ks, 1e-10, prec, r0, r1)
        assert ll == ll_w
        np.testing.assert_array_equal(g, g_w)
        np.testing.assert_array_equal(pts, pts_w)
        assert np.isfinite(ll)


def test_noalloc_rejects_sizes_beyond_scratch():
    """1.2 M rows with 3 theta need 38 MB of per-row values: rejected on the host
    with WORKSPACE_TOO_SMALL before any launch, the outputs untouched."""
    n = 1_200_000
    x = torch.zeros((n, 3), dtype=torch.float64, device=DEV)
    y = torch.zeros(n, dtype=torch.float64, device=DEV)
    nn = torch.full((n, 32), -1, dtype=torch.int32, device=DEV)
    ks = nat.kernel_struct({"noise_level": 0.1, "length_scale": 0.3, "rbf_level": 1.0},
                           {"noise_level": 0, "rbf_level": 1, "length_scale": 2}, 3)
    need = nat.lib().nngp_vecchia_workspace_bytes(n, 32, 3, 1, 3, 0, 1)
    assert need > SCRATCH
    rc, ll, g, _ = noalloc(x, y, nn, ks, 1e-10, 0)
    assert rc == nat.NNGP_ERR_WORKSPACE_TOO_SMALL
    assert b"scratch" in nat.lib().nngp_last_error_detail()
    assert ll == -1.0 and (g == -1.0).all()


@pytest.mark.parametrize("prec", [0, 1])
def test_noalloc_marks_failures_nan_then_recovers(prec):
    """An out-of-range neighbour index, then a block that cannot be factored, give
    NaN; the next valid call on the same scratch is exact again."""
    xt, yt, nn, ks = workload(5_000, 6)
    bad = nn.clone()
    bad[1234, 5] = 5_000
    rc, ll, g, _ = noalloc(xt, yt, bad, ks, 1e-10, prec)
    assert rc == 0 and np.isnan(ll) and np.isnan(g).all()
    xd = np.repeat(np.random.default_rng(1).random((20, 2)), 3, axis=0)
    nnd = vo.vecchia_knn(xd, 1.0, 4)
    ksd = nat.kernel_struct({"length_scale": 1.0, "rbf_level": 1.0}, {}, 0)
    rc, ll, _, _ = noalloc(t(xd), t(np.zeros(60)), t(nnd, torch.int32), ksd, 0.0, prec, grad=False)
    assert rc == 0 and np.isnan(ll)
    rc, ll, g, pts = noalloc(xt, yt, nn, ks, 1e-10, prec)
    ll_w, g_w, pts_w = with_workspace(xt, yt, nn, ks, 1e-10, prec)
    assert rc == 0 and ll == ll_w
    np.testing.assert_array_equal(g, g_w)
