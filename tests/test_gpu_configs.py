"""The benchmark workloads at their full sizes on the GPU.

  road_mixture     n = 434,874 train / 43,487 test, d = 3, m = 32
  uniform_8d       n = 2,000,000, d = 8, m = 32 and 64
  correlated_gauss n = 16,599, d = 18 (18 ARD length scales, 20 theta), m = 32
  serving          the resident road-mixture model (FP32 + refinement form),
                   request sizes log-uniform in [1, 65,536]

At these sizes the oracle runs on sampled rows (each Vecchia row against its own
prefix of refs, each query against every training row) and the rest is checked
through size-independent properties: identical neighbour lists across precision
forms and search methods, the FP32 + refinement form within its stated tolerance
of the FP64 form (|dl|/n <= 1e-6, gradient and mean/std relative error <= 1e-5),
the log-likelihood additive over row ranges, and each served result identical
alone and in the mix.
"""
import numpy as np
import pytest
import torch

from nngpr_amd import _native as nat
from nngpr_amd import datasets, hip_ops
from nngpr_amd.kernel_params import get_kernel_params, get_kernel_theta_mapping
from nngpr_amd.serving import NngpServer, ResidentModel
from oracle import vecchia_oracle as vo

pytestmark = pytest.mark.gpu
DEV = "cuda:0"
METHODS = [hip_ops.KNN_AUTO, hip_ops.KNN_BRUTE_FORCE, hip_ops.KNN_GRID_WAVES]


def t(a, dtype=torch.float64):
    return torch.from_numpy(np.ascontiguousarray(a)).to(DEV, dtype)


def kernel_args(kernel):
    """(kernel struct, oracle arguments w, C, s, ls, theta map, n_theta)."""
    p, mp = get_kernel_params(kernel), get_kernel_theta_mapping(kernel)
    ks = nat.kernel_struct(p, mp, len(kernel.theta))
    ls = np.asarray(p["length_scale"], dtype=np.float64).reshape(-1)
    ls = float(ls[0]) if ls.size == 1 else ls
    return ks, (p.get("noise_level", 0.0), p.get("constant_value", 0.0), p.get("rbf_level", 0.0),
                ls, mp, len(kernel.theta))


def vecchia_rows_vs_oracle(x, nn, rows, ls, k):
    """Each sampled Vecchia row's list equals the oracle's over its prefix."""
    for r in rows:
        e = vo.knn(x[:r], x[r:r + 1], ls, k)[0] if r > 0 else -np.ones(k, np.int32)
        np.testing.assert_array_equal(nn[r], e, err_msg="row %d" % r)


def sample_rows(n, count, seed):
    rng = np.random.default_rng(seed)
    return np.unique(np.concatenate([[0, 1, 2, 31, 32, 33, 64, 65, n - 1],
                                     rng.integers(0, n, count)]))


# ------------------------------------------------------------ road mixture, full
@pytest.fixture(scope="module")
def road():
    ds = datasets.road_mixture()
    ks, oargs = kernel_args(ds["kernel"])
    xt, yt, qt = t(ds["x_train"]), t(ds["y_train"]), t(ds["x_test"])
    return ds, ks, oargs, xt, yt, qt


def test_road_mixture_knn_identical_across_forms_and_methods(road):
    ds, ks, oargs, xt, yt, qt = road
    x = ds["x_train"]
    vecchia, query = [], []
    for m in METHODS:
        with hip_ops.knn_method(m):
            for prec in (0, 1):
                vecchia.append(hip_ops.knn(xt, xt, 0.1, 32, ref_cut=0, precision=prec).cpu().numpy())
                query.append(hip_ops.knn(xt, qt, 0.1, 32, precision=prec).cpu().numpy())
    for v in vecchia[1:]:
        np.testing.assert_array_equal(v, vecchia[0])
    for q in query[1:]:
        np.testing.assert_array_equal(q, query[0])
    vecchia_rows_vs_oracle(x, vecchia[0], sample_rows(x.shape[0], 40, 1), 0.1, 32)
    qs = sample_rows(ds["x_test"].shape[0], 40, 2)
    np.testing.assert_array_equal(query[0][qs], vo.knn(x, ds["x_test"][qs], 0.1, 32))


def test_road_mixture_fp32_form_within_tolerance_of_fp64(road):
    ds, ks, oargs, xt, yt, qt = road
    x, y = ds["x_train"], ds["y_train"]
    n = x.shape[0]
    nn = hip_ops.knn(xt, xt, 0.1, 32, ref_cut=0)
    res = {}
    for prec in (0, 1):
        ll, g, pts = hip_ops.vecchia_loglik(xt, yt, nn, ks, ds["alpha"], precision=prec,
                                            eval_gradient=True, point_loglik=True)
        res[prec] = (ll.item(), g.cpu().numpy(), pts.cpu().numpy())
    assert abs(res[1][0] - res[0][0]) / n <= 1e-6
    np.testing.assert_allclose(res[1][1], res[0][1], rtol=1e-5)
    # sampled rows against the oracle
    rows = sample_rows(n, 300, 3)
    w, C, s, ls, mp, nt = oargs
    _, _, pts_o = vo.vecchia_loglik(x, y, nn.cpu().numpy(), w, C, s, ls, ds["alpha"], rows=rows)
    np.testing.assert_allclose(res[0][2][rows], pts_o, rtol=1e-10, atol=1e-10)
    np.testing.assert_allclose(res[1][2][rows], pts_o, rtol=1e-5, atol=1e-5)
    # predictions: FP32 form vs FP64 form, and sampled queries vs the oracle
    nq = hip_ops.knn(xt, qt, 0.1, 32)
    out = {p: [v.cpu().numpy() for v in hip_ops.predict(xt, yt, qt, nq, ks, ds["alpha"],
                                                         precision=p)] for p in (0, 1)}
    np.testing.assert_allclose(out[1][0], out[0][0], rtol=1e-5, atol=1e-5)
    np.testing.assert_allclose(out[1][1], out[0][1], rtol=1e-5)
    qs = sample_rows(qt.shape[0], 300, 4)
    mo, so = vo.predict(x, y, ds["x_test"][qs], nq.cpu().numpy()[qs], w, C, s, ls, ds["alpha"])
    np.testing.assert_allclose(out[0][0][qs], mo, rtol=1e-10, atol=1e-10)
    np.testing.assert_allclose(out[0][1][qs], so, rtol=1e-10)


@pytest.mark.parametrize("prec", [0, 1])
def test_road_mixture_loglik_additive_over_row_ranges(road, prec):
    """The reference sums per-task (rows [i0, i1)) results on the host
    (base_nngpr.py:440-449); three ranges summed equal one call over all rows."""
    ds, ks, oargs, xt, yt, qt = road
    n = xt.shape[0]
    nn = hip_ops.knn(xt, xt, 0.1, 32, ref_cut=0)
    ll, g, pts = hip_ops.vecchia_loglik(xt, yt, nn, ks, ds["alpha"], precision=prec,
                                        eval_gradient=True, point_loglik=True)
    cuts = [0, 1000, 250_001, n]
    parts = [hip_ops.vecchia_loglik(xt, yt, nn, ks, ds["alpha"], precision=prec,
                                    eval_gradient=True, row0=a, row1=b, point_loglik=True)
             for a, b in zip(cuts[:-1], cuts[1:])]
    ll_sum = sum(p[0].item() for p in parts)
    g_sum = sum(p[1].cpu().numpy() for p in parts)
    assert abs(ll_sum - ll.item()) <= 1e-12 * abs(ll.item())
    np.testing.assert_allclose(g_sum, g.cpu().numpy(), rtol=1e-12, atol=1e-9)
    # each row's value does not depend on the range it was computed in
    np.testing.assert_array_equal(np.concatenate([p[2].cpu().numpy() for p in parts]),
                                  pts.cpu().numpy())


def test_road_mixture_serving_mixed_equals_alone(road):
    """Requests of log-uniform sizes in [1, 65,536] served concurrently by the
    resident road-mixture model (FP32 + refinement): each result equals the same
    request run alone, and sampled queries equal the oracle within 1e-5."""
    ds, ks, oargs, xt, yt, qt = road
    model = ResidentModel(ds["x_train"], ds["y_train"], ds["kernel"], alpha=ds["alpha"],
                          num_nn=32, precision=hip_ops.PRECISION_FP32_REFINED)
    try:
        rng = np.random.default_rng(11)
        sizes = np.floor(np.exp(rng.uniform(0.0, np.log(65536.0), 14))).astype(int)
        sizes[:2] = [1, 65536]
        pool = ds["x_test"]  # held-out draws of the same mixture
        reqs = [pool[rng.integers(0, pool.shape[0], s)] for s in sizes]
        with NngpServer(model, max_batch_points=65536, max_slots=4) as srv:
            futs = [srv.submit(r) for r in reqs]
            results = [f.result(timeout=100) for f in futs]
        w, C, s, ls, mp, nt = oargs
        for r, (mean, std) in zip(reqs, results):
            m_alone, s_alone = model.predict(r)
            np.testing.assert_array_equal(mean, m_alone)
            np.testing.assert_array_equal(std, s_alone)
        r, (mean, std) = reqs[1], results[1]
        qs = sample_rows(r.shape[0], 200, 12)
        nn_o = vo.knn(ds["x_train"], r[qs], 0.1, 32)
        mo, so = vo.predict(ds["x_train"], ds["y_train"], r[qs], nn_o, w, C, s, ls, ds["alpha"])
        np.testing.assert_allclose(mean[qs], mo, rtol=1e-5, atol=1e-5)
        np.testing.assert_allclose(std[qs, 0], so, rtol=1e-5)
    finally:
        model.close()


# ----------------------------------------------------------- uniform 8-d, 2M rows
@pytest.fixture(scope="module")
def uniform8():
    ds = datasets.uniform_8d()
    return ds, t(ds["x_train"]), t(ds["y_train"])


@pytest.mark.parametrize("m", [32, 64])
def test_uniform_8d_knn_and_fp32_form(uniform8, m):
    ds, xt, yt = uniform8
    x, y = ds["x_train"], ds["y_train"]
    n = x.shape[0]
    ks, (w, C, s, ls, mp, nt) = kernel_args(ds["kernel"])
    nn = {p: hip_ops.knn(xt, xt, ls, m, ref_cut=0, precision=p) for p in (0, 1)}
    np.testing.assert_array_equal(nn[0].cpu().numpy(), nn[1].cpu().numpy())
    nn_h = nn[0].cpu().numpy()
    vecchia_rows_vs_oracle(x, nn_h, sample_rows(n, 16, 5), ls, m)
    del nn[1]
    res = {}
    for prec in (0, 1):
        ll, g, pts = hip_ops.vecchia_loglik(xt, yt, nn[0], ks, ds["alpha"], precision=prec,
                                            eval_gradient=True, point_loglik=True)
        res[prec] = (ll.item(), g.cpu().numpy(), pts.cpu().numpy())
    assert abs(res[1][0] - res[0][0]) / n <= 1e-6
    np.testing.assert_allclose(res[1][1], res[0][1], rtol=1e-5)
    rows = sample_rows(n, 200, 6)
    _, _, pts_o = vo.vecchia_loglik(x, y, nn_h, w, C, s, ls, ds["alpha"], rows=rows)
    np.testing.assert_allclose(res[0][2][rows], pts_o, rtol=1e-10, atol=1e-10)
    np.testing.assert_allclose(res[1][2][rows], pts_o, rtol=1e-5, atol=1e-5)


# ------------------------------------------------------ correlated gauss, d = 18
def test_correlated_gauss_full_size():
    """18 ARD length scales (20 theta) at n = 16,599: neighbour lists identical in
    both forms and equal to the oracle on sampled rows; the FP32 form within
    tolerance of the FP64 form; the gradient of a row range (rows [6000, 9000),
    row0 > 0) equal to the oracle's reference-formula gradient of those rows."""
    ds = datasets.correlated_gauss()
    x, y = ds["x_train"], ds["y_train"]
    n = x.shape[0]
    ks, (w, C, s, ls, mp, nt) = kernel_args(ds["kernel"])
    xt, yt = t(x), t(y)
    nn = {p: hip_ops.knn(xt, xt, ls, 32, ref_cut=0, precision=p) for p in (0, 1)}
    np.testing.assert_array_equal(nn[0].cpu().numpy(), nn[1].cpu().numpy())
    nn_h = nn[0].cpu().numpy()
    vecchia_rows_vs_oracle(x, nn_h, sample_rows(n, 60, 7), ls, 32)
    res = {}
    for prec in (0, 1):
        ll, g, _ = hip_ops.vecchia_loglik(xt, yt, nn[0], ks, ds["alpha"], precision=prec,
                                          eval_gradient=True)
        res[prec] = (ll.item(), g.cpu().numpy())
    assert abs(res[1][0] - res[0][0]) / n <= 1e-6
    np.testing.assert_allclose(res[1][1], res[0][1], rtol=1e-5, atol=1e-6)
    r0, r1 = 6000, 9000
    ll_r, g_r, _ = hip_ops.vecchia_loglik(xt, yt, nn[0], ks, ds["alpha"], eval_gradient=True,
                                          row0=r0, row1=r1)
    ll_o, g_o, _ = vo.vecchia_loglik(x, y, nn_h, w, C, s, ls, ds["alpha"], mp, nt,
                                     eval_gradient=True, rows=np.arange(r0, r1))
    assert abs(ll_r.item() - ll_o) <= 1e-10 * abs(ll_o)
    np.testing.assert_allclose(g_r.cpu().numpy(), g_o, rtol=1e-9, atol=1e-9)
