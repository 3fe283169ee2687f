"""Serving loop: graph replays equal eager runs, every request's result is the
same alone and in a mix of concurrent requests, a small HBM budget queues work
instead of failing, queued requests can be cancelled, and results agree with the
estimator's predict and with the FP64 oracle."""
import numpy as np
import pytest
import torch

from nngpr_amd import TorchNngpr, datasets, hip_ops
from nngpr_amd import batched_kernels as bk
from nngpr_amd.serving import NngpServer, ResidentModel
from oracle import vecchia_oracle as vo

pytestmark = pytest.mark.gpu
DEV = "cuda:0"


@pytest.fixture(scope="module")
def small():
    ds = datasets.small_uniform()
    models = {p: ResidentModel(ds["x_train"], ds["y_train"], ds["kernel"], alpha=ds["alpha"],
                               num_nn=16, precision=p) for p in (0, 1)}
    yield ds, models
    for m in models.values():
        m.close()


@pytest.mark.parametrize("prec", [0, 1])
def test_resident_predict_matches_oracle(small, prec):
    ds, models = small
    rng = np.random.default_rng(3)
    xq = rng.random((700, 3))
    mean, std = models[prec].predict(xq)
    nn = vo.knn(ds["x_train"], xq, 0.3, 16)
    mo, so = vo.predict(ds["x_train"], ds["y_train"], xq, nn, 0.01, 0.0, 1.0, 0.3, ds["alpha"])
    tol = 1e-10 if prec == 0 else 1e-5
    np.testing.assert_allclose(mean[:, 0], mo[:, 0], rtol=tol, atol=tol)
    np.testing.assert_allclose(std[:, 0], so, rtol=tol)


def test_graph_replays_match_eager(small):
    ds, models = small
    model = models[1]
    b = 1024
    rng = np.random.default_rng(4)
    q = torch.from_numpy(rng.random((b, 3))).to(DEV)
    ws = torch.empty(model.workspace_bytes(b), dtype=torch.uint8, device=DEV)
    me, se = (torch.empty((b, 1), dtype=torch.float64, device=DEV),
              torch.empty(b, dtype=torch.float64, device=DEV))
    model.predict_into(q, me, se, ws)
    torch.cuda.synchronize()
    mg = torch.zeros_like(me)
    sg = torch.zeros_like(se)
    wsg = torch.empty_like(ws)
    s = torch.cuda.Stream()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g, stream=s):
        model.predict_into(q, mg, sg, wsg)
    other = torch.from_numpy(rng.random((3000, 3))).to(DEV)
    for rep in range(4):
        mg.zero_()
        sg.zero_()
        g.replay()
        torch.cuda.synchronize()
        assert torch.equal(mg, me) and torch.equal(sg, se), rep
        # eager work between replays (another prediction on the default stream)
        model.predict(other.cpu().numpy())
        torch.cuda.synchronize()


@pytest.mark.parametrize("graphs", [True, False])
def test_mixed_requests_equal_alone(small, graphs):
    ds, models = small
    model = models[1]
    rng = np.random.default_rng(5)
    sizes = [1, 7, 300, 1024, 1500, 2500, 33, 4000, 256, 999]
    reqs = [rng.random((s, 3)) for s in sizes]
    with NngpServer(model, max_batch_points=1024, max_slots=3, use_graphs=graphs) as srv:
        futs = [srv.submit(r) for r in reqs]
        results = [f.result(timeout=60) for f in futs]
        assert len(srv.latencies_ms()) == len(reqs)
    for r, (mean, std) in zip(reqs, results):
        m_alone, s_alone = model.predict(r)
        np.testing.assert_array_equal(mean, m_alone)
        np.testing.assert_array_equal(std, s_alone)


def test_small_budget_queues_instead_of_failing(small):
    ds, models = small
    model = models[0]
    rng = np.random.default_rng(6)
    reqs = [rng.random((1000, 3)) for _ in range(5)]
    budget = model.workspace_bytes(256) + 256 * 8 * 12
    with NngpServer(model, hbm_budget_bytes=budget, max_batch_points=65536) as srv:
        assert len(srv.slots) == 1 and srv.slot_capacity == 256
        futs = [srv.submit(r) for r in reqs]
        out = [f.result(timeout=60) for f in futs]
    for r, (mean, std) in zip(reqs, out):
        m, s = model.predict(r)
        np.testing.assert_array_equal(mean, m)


def test_cancel_queued_requests(small):
    ds, models = small
    model = models[1]
    rng = np.random.default_rng(7)
    budget = model.workspace_bytes(256) + 256 * 8 * 12
    with NngpServer(model, hbm_budget_bytes=budget, use_graphs=False) as srv:
        futs = [srv.submit(rng.random((2000, 3))) for _ in range(8)]
        cancelled = [f.cancel() for f in futs[4:]]
        done = [f.result(timeout=60) for f in futs[:4]]
    assert len(done) == 4
    for f, c in zip(futs[4:], cancelled):
        assert f.cancelled() == c


def test_from_estimator_matches_predict():
    ds = datasets.small_uniform()
    k = bk.WhiteKernel(0.01) + bk.ConstantKernel(1.0) * bk.RBF(0.3)
    est = TorchNngpr("cuda:0", kernel=k, num_nn=16, optimizer=None, normalize_y=True,
                     allow_downcast_f32="no").fit(ds["x_train"], ds["y_train"])
    model = ResidentModel.from_estimator(est)
    mean_e, std_e = est.predict(ds["x_test"], return_std=True)
    with NngpServer(model, max_batch_points=4096) as srv:
        mean, std = srv.submit(ds["x_test"]).result(timeout=60)
    np.testing.assert_allclose(mean[:, 0], mean_e, rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(std[:, 0], std_e, rtol=1e-12)
    model.close()


def test_pickled_estimator_round_trip_and_serving(tmp_path):
    """A fitted estimator survives pickling (without its device copies) and loads
    into the resident serving model."""
    import pickle
    ds = datasets.small_uniform()
    k = bk.WhiteKernel(0.01) + bk.ConstantKernel(1.0) * bk.RBF(0.3)
    est = TorchNngpr("cuda:0", kernel=k, num_nn=16, optimizer=None, normalize_y=True,
                     allow_downcast_f32="no").fit(ds["x_train"], ds["y_train"])
    mean_e, std_e = est.predict(ds["x_test"], return_std=True)
    blob = pickle.dumps(est)
    assert b"_dev_cache" not in blob
    est2 = pickle.loads(blob)
    mean2, std2 = est2.predict(ds["x_test"], return_std=True)
    np.testing.assert_array_equal(mean2, mean_e)
    np.testing.assert_array_equal(std2, std_e)
    path = tmp_path / "model.pkl"
    path.write_bytes(blob)
    model = ResidentModel.from_pickle(str(path))
    try:
        mean, std = model.predict(ds["x_test"])
    finally:
        model.close()
    np.testing.assert_allclose(mean[:, 0], mean_e, rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(std[:, 0], std_e, rtol=1e-12)
