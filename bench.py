"""Benchmark of the Vecchia hot path on one MI355X.

Workload: road-mixture regression, n_train=434,874, n_test=43,487, d=3, m=32,
White(0.01) + C(1.0) * RBF(0.1) (nngpr_amd.datasets.road_mixture, seeded).

A step is one log-likelihood + gradient evaluation over all training rows, as
the reference computes it on every optimizer call (base_nngpr.py:530-614): the
Vecchia kNN (each row's 32 nearest earlier rows) followed by the fused
fill -> Cholesky -> mu/sigma -> loglik -> adjoint gradient pass.  `value` is the
reference's default precision mode (allow_downcast_f32='only-nn': kNN screened
in FP32 with exact FP64 keys, FP64 factor).  The line also reports the 'yes'
(FP32 factor + FP64 refinement) and 'no' (FP64 everywhere) modes, kNN-cached
evaluation rates, predictions/s on the test set, the FP32-vs-FP64 errors, the
fused kernel's roofline position and the restated NumpyNngpr CPU path timed on
this host.

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_FP32_VALU_TFLOPS = 157.3
PEAK_FP64_VALU_TFLOPS = 78.6
PEAK_BF16_MFMA_TFLOPS = 2500.0   # dense bf16 matrix peak (no sparsity)
PEAK_HBM_GBS = 8000.0
_T_START = time.perf_counter()


def progress(msg):
    """One line on stderr per finished phase (the run's own progress record)."""
    print("bench: %s (%.1f s)" % (msg, time.perf_counter() - _T_START), file=sys.stderr,
          flush=True)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--n-train", type=int, default=434_874)
    ap.add_argument("--n-test", type=int, default=43_487)
    ap.add_argument("--skip-extras", action="store_true",
                    help="only the headline timed region (used for profiling runs)")
    ap.add_argument("--extras", default=None,
                    help="comma-separated subset of the extras to run: modes, serving, "
                         "fit_d18, uniform_8d, matern, sequential, fit_predict_restarts "
                         "(default: all)")
    ap.add_argument("--knn-method", choices=("auto", "brute"), default="auto",
                    help="kNN search of the headline step: auto (grid search for d <= 3) or "
                         "brute (the FP32 screen scan + FP64 re-rank; used for profiling "
                         "the scan kernel)")
    ap.add_argument("--copy-probe-bytes", type=int, default=0,
                    help="run only a plain device-to-device copy of this many bytes, 5 times "
                         "(calibrates the HBM byte counters against a known traffic)")
    return ap.parse_args()


TRAFFIC_PROFILE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles",
                               "loglik_road_mixture_hbm_traffic.md")


def measured_traffic(kernel):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 FETCH_SIZE /
    WRITE_SIZE summary (calibrated against a plain copy); None when absent."""
    try:
        text = open(TRAFFIC_PROFILE).read()
    except OSError:
        return None
    for line in text.splitlines():
        parts = line.split()
        if len(parts) == 4 and parts[0] == "traffic_bytes_per_launch" and parts[1] == kernel:
            return int(parts[3])
    return None


def c3_fit(steps=50):
    """correlated_gauss (n=16,599, d=18, White + C * RBF with 18 length scales, 20
    theta, m=32): loglik+grad evaluations/s with the kNN recomputed (the ARD
    length-scale ratios change the conditioning sets) and an L-BFGS-B fit of
    `steps` iterations (FP64 factor, reference default precision)."""
    import numpy as np
    import scipy.optimize
    import torch
    from nngpr_amd import TorchNngpr, datasets
    ds = datasets.correlated_gauss()
    counter = {"evals": 0}

    def opt(obj, theta0, bounds):
        def f(th):
            counter["evals"] += 1
            return obj(th)
        r = scipy.optimize.minimize(f, theta0, method="L-BFGS-B", jac=True, bounds=bounds,
                                    options={"maxiter": steps})
        return r.x, r.fun

    est = TorchNngpr("cuda:0", kernel=ds["kernel"], num_nn=ds["num_nn"], optimizer=None)
    est.fit(ds["x_train"], ds["y_train"])
    theta = est.kernel_.theta
    est.log_marginal_likelihood(theta, eval_gradient=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    reps = 30
    for r in range(reps):
        est.log_marginal_likelihood(theta + 1e-3 * r, eval_gradient=True)
    torch.cuda.synchronize()
    t_eval = (time.perf_counter() - t0) / reps
    est2 = TorchNngpr("cuda:0", kernel=ds["kernel"], num_nn=ds["num_nn"], optimizer=opt)
    t0 = time.perf_counter()
    est2.fit(ds["x_train"], ds["y_train"])
    t_fit = time.perf_counter() - t0
    return {"workload": "correlated_gauss: n=16,599 d=18 m=32, White+C*RBF(ARD, 18 length "
                        "scales), FP64 factor",
            "evals_per_s_knn_per_eval": 1.0 / t_eval,
            "fit_seconds": t_fit, "fit_lbfgs_max_iter": steps,
            "fit_objective_evals": counter["evals"],
            "fit_loglik": float(est2.log_marginal_likelihood_value_)}


# The reference's published seconds per restart at m=32 on 3droad (n_train subsets,
# RTX 4070 Super; read from its benchmark plot, +-15 %, benchmark.ipynb:459).
REFERENCE_SECONDS_PER_RESTART = {
    (200_000, 32): {"TorchNngpr (RTX 4070 Super, 3droad)": 53.0,
                    "CupyNngpr (RTX 4070 Super, 3droad)": 69.0},
    (200_000, 64): {"TorchNngpr (RTX 4070 Super, 3droad)": 307.0,
                    "CupyNngpr (RTX 4070 Super, 3droad)": 272.0},
}


def fit_predict_restarts(ds, cases=((200_000, 32), (434_874, 32), (200_000, 64)), restarts=3):
    """The reference's published benchmark protocol (benchmark.ipynb:115, 186-231,
    427-431) on the road-mixture data: fit(x_train[:n]) with
    ConstantKernel(0.1) * RBF(0.1, ARD over the 3 features) + WhiteKernel(0.1),
    alpha=1e-8, num_nn=m (32, 64), n_restarts_optimizer=3 (1 + 3 L-BFGS-B runs), then
    predict the test set; the wall time divided by the 3 restarts is the quantity the
    reference plots (seconds per restart)."""
    import numpy as np
    from nngpr_amd import TorchNngpr
    from nngpr_amd import batched_kernels as bk
    out = []
    for n, m in cases:
        kern = bk.ConstantKernel(0.1) * bk.RBF(np.full(3, 0.1)) + bk.WhiteKernel(0.1)
        est = TorchNngpr("cuda:0", kernel=kern, num_nn=m, alpha=1e-8,
                         n_restarts_optimizer=restarts, random_state=0)
        calls = {"n": 0}
        lml = est.log_marginal_likelihood

        def counted(*a, **k):
            calls["n"] += 1
            return lml(*a, **k)
        est.log_marginal_likelihood = counted
        t0 = time.perf_counter()
        est.fit(ds["x_train"][:n], ds["y_train"][:n])
        t_fit = time.perf_counter() - t0
        t1 = time.perf_counter()
        est.predict(ds["x_test"], return_std=True)
        t_pred = time.perf_counter() - t1
        ref = REFERENCE_SECONDS_PER_RESTART.get((n, m))
        out.append({"n_train": n, "num_nn": m, "n_test": int(ds["x_test"].shape[0]),
                    "seconds_per_restart": (t_fit + t_pred) / restarts,
                    "reference_seconds_per_restart": ref,
                    "fit_seconds": t_fit, "predict_seconds": t_pred,
                    "objective_evals": calls["n"],
                    "fit_loglik": float(est.log_marginal_likelihood_value_)})
    return out


def fit_loop(ds, maxiter=60):
    """L-BFGS-B fit of White + C * RBF (isotropic, 3 theta, all free) on the
    road-mixture training set, from the workload's kernel: objective evaluations
    per second through the estimator (validation, kernel decomposition, kNN, fused
    pass, host transfers), with one kNN search reused across the evaluations
    (exact rescale of k + 8 candidates) and with a search at every evaluation as
    the reference does (base_nngpr.py:530-548).  Both fits must follow the same
    path: theta and the log-likelihood are compared bit for bit."""
    import numpy as np
    import scipy.optimize
    from nngpr_amd import TorchNngpr
    out = {"workload": "road_mixture n=%d, White(0.01)+C(1.0)*RBF(0.1), m=32, L-BFGS-B "
                       "maxiter=%d, reference default precision" % (ds["x_train"].shape[0], maxiter)}
    fits = {}
    for reuse in (True, False):
        counter = {"evals": 0, "seconds": 0.0}

        def opt(obj, theta0, bounds, counter=counter):
            def f(th):
                counter["evals"] += 1
                t = time.perf_counter()
                r = obj(th)
                counter["seconds"] += time.perf_counter() - t
                return r
            r = scipy.optimize.minimize(f, theta0, method="L-BFGS-B", jac=True, bounds=bounds,
                                        options={"maxiter": maxiter})
            return r.x, r.fun

        est = TorchNngpr("cuda:0", kernel=ds["kernel"], num_nn=32, alpha=ds["alpha"],
                         optimizer=opt)
        est._knn_reuse = reuse
        est.fit(ds["x_train"][:2000], ds["y_train"][:2000])  # warm-up (compiles, allocator)
        counter["evals"], counter["seconds"] = 0, 0.0
        t0 = time.perf_counter()
        est.fit(ds["x_train"], ds["y_train"])
        t = time.perf_counter() - t0
        fits[reuse] = est
        key = "knn_reused" if reuse else "knn_per_eval"
        # evals_per_s: objective calls over the time spent inside them (the first
        # includes the one search); fit_seconds also holds validation, the upload
        # and the optimizer's own work
        out[key] = {"fit_seconds": t, "objective_evals": counter["evals"],
                    "objective_seconds": counter["seconds"],
                    "evals_per_s": counter["evals"] / counter["seconds"],
                    "fit_loglik": float(est.log_marginal_likelihood_value_)}
    out["identical_path"] = bool(
        np.array_equal(fits[True].kernel_.theta, fits[False].kernel_.theta) and
        fits[True].log_marginal_likelihood_value_ == fits[False].log_marginal_likelihood_value_)
    return out


def c4_eval(num_nn):
    """uniform_8d (n=2,000,000, d=8): one loglik+grad evaluation in the FP32 +
    FP64-refinement form, kNN (brute-force scans, d > 3) and fused pass timed
    apart."""
    import numpy as np
    import torch
    from nngpr_amd import _native as nat
    from nngpr_amd import datasets, hip_ops
    from nngpr_amd.kernel_params import get_kernel_params, get_kernel_theta_mapping
    ds = datasets.uniform_8d(num_nn=num_nn)
    dev = torch.device("cuda:0")
    x = torch.from_numpy(ds["x_train"]).to(dev)
    y = torch.from_numpy(ds["y_train"]).to(dev)
    kern = ds["kernel"]
    params, mapping = get_kernel_params(kern), get_kernel_theta_mapping(kern)
    ks = nat.kernel_struct(params, mapping, len(kern.theta))
    ls = float(np.asarray(params["length_scale"]).reshape(-1)[0])
    F32 = hip_ops.PRECISION_FP32_REFINED
    n, d = x.shape
    nn = hip_ops.knn(x, x, ls, num_nn, ref_cut=0, precision=F32)   # warm-up
    torch.cuda.synchronize()
    nat.timing_enable(True)
    for k in (nat.TIMED_KNN_SCAN, nat.TIMED_KNN_RERANK):
        nat.timing_read(k, reset=True)
    t0 = time.perf_counter()
    nn = hip_ops.knn(x, x, ls, num_nn, ref_cut=0, precision=F32)
    torch.cuda.synchronize()
    t_knn = time.perf_counter() - t0
    screen_ms = nat.timing_read(nat.TIMED_KNN_SCAN)[0]
    rerank_ms = nat.timing_read(nat.TIMED_KNN_RERANK)[0]
    nat.timing_enable(False)
    recomputed = hip_ops.knn_stats()["knn_rows_recomputed"]
    # Matrix-core screen: the Vecchia self-search scores, for query tile t (32 rows),
    # ref tiles 0..t (its rows' limits end inside tile t): n_t (n_t + 1) / 2 tiles of
    # 32 x 32 pairs, each NK = ceil((3 d + 2) / 16) bf16 products of 32 x 32 x 16.
    n_t = (n + 31) // 32
    tiles = n_t * (n_t + 1) / 2.0
    nk = -(-(3 * d + 2) // 16)
    mfma_flop = tiles * nk * 2.0 * 32 * 32 * 16
    ach = mfma_flop / (screen_ms * 1e-3) / 1e12
    roofline_screen = {"kernel": "mfma_screen_kernel (bf16 hi/lo split products, FP32 accumulation)",
                       "bound": "mfma", "achieved": ach, "peak": PEAK_BF16_MFMA_TFLOPS,
                       "unit": "TFLOP/s", "frac": ach / PEAK_BF16_MFMA_TFLOPS,
                       "avg_ms": screen_ms, "tiles": tiles, "products_per_tile": nk,
                       "algorithmic_pairs": n * (n - 1) / 2.0,
                       "fp32_equivalent_tflops": 3.0 * d * n * (n - 1) / 2.0 / (screen_ms * 1e-3) / 1e12}
    t_valu = None
    if num_nn == 32:
        # the same search with the FP32 VALU screen (identical indices)
        with hip_ops.knn_method(hip_ops.KNN_BRUTE_FORCE_VALU):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            nn_v = hip_ops.knn(x, x, ls, num_nn, ref_cut=0, precision=F32)
            torch.cuda.synchronize()
            t_valu = time.perf_counter() - t0
        assert torch.equal(nn_v, nn), "screens disagree"
        del nn_v
    hip_ops.vecchia_loglik(x, y, nn, ks, ds["alpha"], precision=F32, eval_gradient=True,
                           check_pd=False)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    reps = 3
    for _ in range(reps):
        hip_ops.vecchia_loglik(x, y, nn, ks, ds["alpha"], precision=F32, eval_gradient=True,
                               check_pd=False)
    torch.cuda.synchronize()
    t_fused = (time.perf_counter() - t0) / reps
    # minimum HBM bytes of the fused pass: neighbour indices, gathered coordinates
    # and targets of every neighbour and the point, per-point outputs
    min_bytes = n * (num_nn * 4 + (num_nn + 1) * (d + 1) * 8 + (1 + len(kern.theta)) * 8)
    out = {"workload": "uniform_8d: n=2,000,000 d=8 m=%d, White+C*RBF(0.3), FP32 factor + "
                       "FP64 refinement" % num_nn,
           "knn_ms": t_knn * 1e3, "knn_screen_ms": screen_ms, "knn_rerank_ms": rerank_ms,
           "knn_rows_recomputed": recomputed,
           "knn_ms_valu_screen": t_valu * 1e3 if t_valu is not None else None,
           "fused_loglik_grad_ms": t_fused * 1e3,
           "evals_per_s_knn_per_eval": 1.0 / (t_knn + t_fused),
           "fused_min_hbm_gb_per_s": min_bytes / t_fused / 1e9,
           "roofline_mfma_screen": roofline_screen}
    del x, y, nn
    torch.cuda.empty_cache()
    return out


def copy_probe(nbytes):
    """Five launches of torch's copy kernel over nbytes (reads nbytes, writes nbytes)."""
    import torch
    n = nbytes // 4
    src = torch.ones(n, dtype=torch.float32, device="cuda:0")
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    for _ in range(5):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_probe_bytes": n * 4, "launches": 5}))


def cpu_baseline(ds):
    """Restated NumpyNngpr path on a bounded sample of the same evaluation."""
    from oracle import numpy_port
    x, y = ds["x_train"], ds["y_train"]
    workers = len(os.sched_getaffinity(0))
    res = numpy_port.time_sample(x, y, 0.01, 0.0, 1.0, 0.1, 32, ds["alpha"],
                                 rows_per_task=24, n_tasks=2 * workers, workers=workers)
    est = res["est_seconds_per_eval"]
    return {
        "value": 1.0 / est,
        "unit": "loglik+grad evals/s",
        "cores": workers,
        "kind": "port",
        "sample": ("%d tasks x %d rows at row positions spread over [0, n), timed in %d worker "
                   "processes (%.1f s wall); full evaluation extrapolated with a per-row cost "
                   "a + b*i fitted to the sample (%.0f CPU-s per evaluation)"
                   % (res["tasks"], res["rows_per_task"], workers, res["sample_wall_s"],
                      res["cpu_seconds_per_eval_single_core"])),
    }


def cpu_sequential(ds, n_sample=2000):
    """The reference's sequential 'full' conditioning loop (restated, one query at a
    time) on the first n_sample test queries, extrapolated per query."""
    from oracle import sequential_oracle as so
    x, y, xq = ds["x_train"], ds["y_train"], ds["x_test"][:n_sample]
    nn = so.conditioning_nn(x, xq, 0.1, ds["num_nn"], "full")
    t0 = time.perf_counter()
    so.predict_sequential(x, y, xq, nn, 0.01, 0.0, 1.0, 0.1, ds["alpha"], "std")
    per_q = (time.perf_counter() - t0) / n_sample
    return {"queries_per_s": 1.0 / per_q, "cores": 1, "kind": "port",
            "sample": "first %d test queries, conditioning_method='full', mean + std "
                      "(kNN not timed)" % n_sample}


def sequential_extras(x, y, xq, ks, ls, m, alpha, steps):
    """Sequential conditioning on the test set: 'full' mean + std over all test
    queries (kNN over [train; test] with the cut + nngp_predict_sequential), the
    joint covariance of the first 4,096 queries, and 16 samples per query."""
    import numpy as np
    import torch
    from nngpr_amd import hip_ops
    nt = x.shape[0]
    F32 = hip_ops.PRECISION_FP32_REFINED

    def nn_full(q):
        return hip_ops.knn(torch.cat([x, q]), q, ls, m, ref_cut=nt, precision=F32)

    def run(q, output):
        return hip_ops.predict_sequential(x, y, q, nn_full(q), ks, alpha, output)

    def timed_call(fn):
        fn()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            fn()
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) / steps

    nq = xq.shape[0]
    nn = nn_full(xq).cpu().numpy()
    level = np.zeros(nq, dtype=np.int64)
    for i in range(nq):
        q = nn[i][nn[i] >= nt] - nt
        level[i] = level[q].max() + 1 if q.size else 0
    t_std = timed_call(lambda: run(xq, hip_ops.COND_STD))
    t_mean = timed_call(lambda: run(xq, hip_ops.COND_MEAN))
    xc = xq[:4096]
    t_cov = timed_call(lambda: run(xc, hip_ops.COND_COV))
    z = torch.randn((nq, 1, 16), dtype=torch.float64, device=x.device)
    t_smp = timed_call(lambda: hip_ops.sample_sequential(x, y, xq, nn_full(xq), ks, alpha, z))
    return {"queries": int(nq), "dependency_levels": int(level.max() + 1),
            "full_mean_std_ms": t_std * 1e3, "full_mean_std_queries_per_s": nq / t_std,
            "full_mean_ms": t_mean * 1e3,
            "full_cov_4096_ms": t_cov * 1e3,
            "full_sample_16_ms": t_smp * 1e3}


def serve_sweep(ds, rates, seconds=3.0, max_requests=300):
    """Poisson request streams against the resident model: batch sizes
    log-uniform in [1, 65536], queries drawn from the test points; returns p50/p99
    latency (submit -> result) and points/s per arrival rate."""
    import numpy as np
    from nngpr_amd import hip_ops
    from nngpr_amd.serving import NngpServer, ResidentModel
    model = ResidentModel(ds["x_train"], ds["y_train"], ds["kernel"], alpha=ds["alpha"],
                          num_nn=ds["num_nn"], precision=hip_ops.PRECISION_FP32_REFINED)
    pool = ds["x_test"]
    rng = np.random.default_rng(5)
    out = {}
    with NngpServer(model, max_batch_points=65536, max_slots=3) as srv:
        for f in [srv.submit(pool[:5000]) for _ in range(3)]:
            f.result()
        for lam in rates:
            n = int(min(max_requests, lam * seconds))
            gaps = rng.exponential(1.0 / lam, size=n)
            sizes = np.maximum(1, np.floor(np.exp(rng.uniform(0, np.log(65536), size=n)))).astype(int)
            starts = rng.integers(0, pool.shape[0], size=n)
            srv.completed.clear()
            futs = []
            t0 = time.perf_counter()
            t_next = t0
            for i in range(n):
                t_next += gaps[i]
                dt = t_next - time.perf_counter()
                if dt > 0:
                    time.sleep(dt)
                idx = (starts[i] + np.arange(sizes[i])) % pool.shape[0]
                futs.append(srv.submit(pool[idx]))
            for f in futs:
                f.result()
            t1 = max(r.t_done for r in srv.completed)
            lat = srv.latencies_ms()
            out[str(lam)] = {"requests": n, "points": int(sizes.sum()),
                             "p50_ms": float(np.percentile(lat, 50)),
                             "p99_ms": float(np.percentile(lat, 99)),
                             "points_per_s": float(sizes.sum() / (t1 - t0))}
        out["slots"] = len(srv.slots)
        out["slot_capacity"] = srv.slot_capacity
    model.close()
    return out


def main():
    args = parse()
    if args.copy_probe_bytes:
        copy_probe(args.copy_probe_bytes)
        return
    if args.gpus != 1:
        print("bench.py runs on exactly one GPU: this build has no multi-GPU path "
              "(--gpus %d requested)" % args.gpus, file=sys.stderr)
        sys.exit(2)
    import numpy as np

    from nngpr_amd import datasets
    ds = datasets.road_mixture(n_train=args.n_train, n_test=args.n_test)
    cpu = None
    cpu_seq = None
    ALL_EXTRAS = {"modes", "serving", "fit_d18", "uniform_8d", "matern", "sequential",
                  "fit_predict_restarts", "fit_loop"}
    want = set() if args.skip_extras else (
        ALL_EXTRAS if args.extras is None else set(args.extras.split(",")))
    if want - ALL_EXTRAS:
        raise SystemExit("unknown extras: %s" % ", ".join(sorted(want - ALL_EXTRAS)))
    if not args.no_cpu:
        cpu = cpu_baseline(ds)  # before the GPU is touched (fork-based pool)
        progress("cpu baseline done")
        if "sequential" in want:
            cpu_seq = cpu_sequential(ds)
            progress("cpu sequential baseline done")

    import torch
    from nngpr_amd import _native as nat
    from nngpr_amd import hip_ops
    from nngpr_amd.kernel_params import get_kernel_params, get_kernel_theta_mapping

    if args.knn_method == "brute":
        nat.knn_set_method(nat.KNN_BRUTE_FORCE)
    dev = torch.device("cuda:0")
    x = torch.from_numpy(ds["x_train"]).to(dev)
    y = torch.from_numpy(ds["y_train"]).to(dev)
    xq = torch.from_numpy(ds["x_test"]).to(dev)
    kern = ds["kernel"]
    params, mapping = get_kernel_params(kern), get_kernel_theta_mapping(kern)
    ks = nat.kernel_struct(params, mapping, len(kern.theta))
    ls = float(np.asarray(params["length_scale"]).reshape(-1)[0])
    m = ds["num_nn"]
    alpha = ds["alpha"]
    n, d = x.shape
    F32, F64 = hip_ops.PRECISION_FP32_REFINED, hip_ops.PRECISION_FP64

    # Precision modes of the reference (allow_downcast_f32): kNN form, factor form.
    MODES = {"only-nn": (F32, F64), "yes": (F32, F32), "no": (F64, F64)}

    def evaluate(mode, nn=None):
        knn_prec, prec = MODES[mode]
        if nn is None:
            nn = hip_ops.knn(x, x, ls, m, ref_cut=0, precision=knn_prec)
        ll, g, _ = hip_ops.vecchia_loglik(x, y, nn, ks, alpha, precision=prec,
                                          eval_gradient=True, check_pd=False)
        return ll, g, nn

    def predict(mode, nn=None):
        knn_prec, prec = MODES[mode]
        if nn is None:
            nn = hip_ops.knn(x, xq, ls, m, precision=knn_prec)
        return hip_ops.predict(x, y, xq, nn, ks, alpha, precision=prec, check_pd=False)

    def timed(fn, steps, warmup):
        for _ in range(warmup):
            fn()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            fn()
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) / steps

    TIMED = {"knn_prefix_scan": nat.TIMED_KNN_SCAN, "knn_rerank": nat.TIMED_KNN_RERANK,
             "knn_grid_build": nat.TIMED_KNN_GRID_BUILD, "knn_grid_search": nat.TIMED_KNN_GRID,
             "fused_loglik_grad": nat.TIMED_LOGLIK}

    def stage_times(fn, steps):
        """Per-step milliseconds of each timed kernel group over `steps` calls of fn
        (HIP events on the launch stream), and the wall time per step."""
        nat.timing_enable(True)
        for k in TIMED.values():
            nat.timing_read(k, reset=True)
            nat.timing_work(k, reset=True)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            fn()
        torch.cuda.synchronize()
        wall = (time.perf_counter() - t0) / steps
        nat.timing_enable(False)
        out = {}
        for name, k in TIMED.items():
            ms, cnt = nat.timing_read(k)
            out[name] = ms / steps
            out[name + "_launches"] = cnt / steps
        # (query, ref) pairs the grid search keyed per step (its measured work)
        out["knn_grid_search_pairs_work"] = nat.timing_work(nat.TIMED_KNN_GRID) / steps
        return wall, out

    def stage_ms(st):
        return {k: v for k, v in st.items() if not k.endswith("_launches") and
                not k.endswith("_work")}

    # ---------------- headline: the reference's default mode ('only-nn') -------
    HEAD = "only-nn"
    for _ in range(args.warmup):
        evaluate(HEAD)
    torch.cuda.synchronize()
    step_s, stages = stage_times(lambda: evaluate(HEAD), args.steps)
    progress("headline timed region done: %.3f ms per step" % (step_s * 1e3))
    fused_avg = stages["fused_loglik_grad"] / max(stages["fused_loglik_grad_launches"], 1)

    # Roofline of the fused fill -> Cholesky -> solve -> loglik -> adjoint pass
    # (FP64 VALU).  Algorithmic flop per point (m neighbours, d features,
    # E = (m+1)(m+2)/2 covariance entries): Cholesky 2m^3/3, two triangular solves
    # 4m^2, fill E(3d+2), adjoint E(d+3) + 2m^2.
    full_size = (n, int(xq.shape[0])) == (434_874, 43_487)
    E = (m + 1) * (m + 2) / 2.0
    fused_flop_pt = 2.0 * m ** 3 / 3.0 + 4.0 * m * m + E * (3 * d + 2) + E * (d + 3) + 2.0 * m * m
    fused_flop = fused_flop_pt * n
    achieved = fused_flop / (fused_avg * 1e-3) / 1e12
    roofline = {"kernel": "vecchia_seg_kernel<double,32,3> fused loglik+grad (FP64)", "bound": "valu",
                "achieved": achieved, "peak": PEAK_FP64_VALU_TFLOPS, "unit": "TFLOP/s",
                "frac": achieved / PEAK_FP64_VALU_TFLOPS,
                # the committed counter summary covers the default workload only
                "traffic": measured_traffic("vecchia_seg_kernel<double,32,3>") if full_size else None,
                "traffic_source": "profiles/loglik_road_mixture_hbm_traffic.md" if full_size else None,
                "avg_ms": fused_avg, "algorithmic_flop": fused_flop,
                "algorithmic_flop_per_point": fused_flop_pt}
    # Roofline of the grid kNN search (FP64 keys on the VALU): the work is the
    # (query, ref) pairs its lanes keyed, counted inside the kernel, at 3d flop
    # per pair (sub, mul, add per feature), over the timed span of the search
    # launch (its two small ordering kernels included).
    grid_ms = stages["knn_grid_search"]
    grid_pairs = stages["knn_grid_search_pairs_work"]
    roofline_grid = None
    if grid_ms > 0 and grid_pairs > 0:
        grid_flop = 3.0 * d * grid_pairs
        ach_grid = grid_flop / (grid_ms * 1e-3) / 1e12
        roofline_grid = {
            "kernel": "lane_knn_kernel<3,32> grid kNN search (FP64 keys)", "bound": "valu",
            "achieved": ach_grid, "peak": PEAK_FP64_VALU_TFLOPS, "unit": "TFLOP/s",
            "frac": ach_grid / PEAK_FP64_VALU_TFLOPS,
            "traffic": measured_traffic("lane_knn_kernel<3,32>") if full_size else None,
            "traffic_source": "profiles/loglik_road_mixture_hbm_traffic.md" if full_size else None,
            "avg_ms": grid_ms, "examined_pairs": grid_pairs, "algorithmic_flop": grid_flop,
            "pairs_per_query": grid_pairs / max(n - 64, 1)}
    # the headline roofline names the kernel that takes longer; both are kept
    if roofline_grid is not None and grid_ms > fused_avg:
        roofline, roofline_other = roofline_grid, ("roofline_fused_loglik", roofline)
    else:
        roofline_other = ("roofline_knn_grid_search", roofline_grid)

    extras = {"stage_ms": stage_ms(stages), roofline_other[0]: roofline_other[1]}
    if "modes" in want:
        nq = xq.shape[0]
        res = {}
        for mode in MODES:
            t_eval = step_s if mode == HEAD else timed(lambda: evaluate(mode), args.steps, 1)
            ll, g, nn = evaluate(mode)
            t_cached = timed(lambda: evaluate(mode, nn), args.steps, 1)
            t_pred = timed(lambda: predict(mode), args.steps, 1)
            mean, std = predict(mode)
            res[mode] = dict(t_eval=t_eval, t_cached=t_cached, t_pred=t_pred, ll=ll.item(),
                             g=g.cpu().numpy(), nn=nn, mean=mean.cpu().numpy()[:, 0],
                             std=std.cpu().numpy())
        _, stages32 = stage_times(lambda: evaluate("yes"), max(2, args.steps // 2))
        # the same headline step with the brute-force kNN scans (FP32 screen +
        # FP64 re-rank) instead of the grid search
        with hip_ops.knn_method(hip_ops.KNN_BRUTE_FORCE):
            evaluate(HEAD)
            tbf, stages_bf = stage_times(lambda: evaluate(HEAD), max(2, args.steps // 2))
        scan_ms = stages_bf["knn_prefix_scan"]
        pairs = n * (n - 1) / 2.0
        knn_flops = 3.0 * d * pairs  # sub, mul, add per feature per pair
        ach_scan = knn_flops / (scan_ms * 1e-3) / 1e12
        roofline_scan = {"kernel": "knn brute-force screen scan (FP32 VALU, |r|^2 - 2 q.r)",
                         "bound": "valu", "achieved": ach_scan, "peak": PEAK_FP32_VALU_TFLOPS,
                         "unit": "TFLOP/s", "frac": ach_scan / PEAK_FP32_VALU_TFLOPS,
                         "avg_ms": scan_ms, "algorithmic_flop": knn_flops}
        r32, r64 = res["yes"], res["no"]
        rel_m = np.abs(r32["mean"] - r64["mean"]) / np.maximum(np.abs(r64["mean"]), 1e-300)
        rel_s = np.abs(r32["std"] - r64["std"]) / np.abs(r64["std"])
        extras.update({
            "evals_per_s": {k: 1.0 / v["t_eval"] for k, v in res.items()},
            "evals_per_s_knn_cached": {k: 1.0 / v["t_cached"] for k, v in res.items()},
            "preds_per_s": {k: nq / v["t_pred"] for k, v in res.items()},
            "speedup_fp32_vs_fp64": {
                "evals": r64["t_eval"] / r32["t_eval"],
                "evals_knn_cached": r64["t_cached"] / r32["t_cached"],
                "preds": r64["t_pred"] / r32["t_pred"]},
            "fp32_vs_fp64_error": {
                "abs_dloglik_over_n": abs(r32["ll"] - r64["ll"]) / n,
                "grad_max_rel": float(np.max(np.abs(r32["g"] - r64["g"]) / np.abs(r64["g"]))),
                "mean_rel_max": float(rel_m.max()), "mean_rel_mean": float(rel_m.mean()),
                "std_rel_max": float(rel_s.max()), "std_rel_mean": float(rel_s.mean()),
            },
            "default_vs_fp64_loglik_identical": res[HEAD]["ll"] == r64["ll"],
            "knn_indices_identical_across_modes": bool(
                torch.equal(res["yes"]["nn"], res["no"]["nn"]) and
                torch.equal(res[HEAD]["nn"], res["no"]["nn"])),
            "stage_ms_fp32_form": stage_ms(stages32),
            "evals_per_s_brute_force_knn": 1.0 / tbf,
            "stage_ms_brute_force_knn": stage_ms(stages_bf),
            "roofline_knn_brute_force_scan": roofline_scan,
            "loglik": r64["ll"],
        })
        progress("precision modes done")

    serving = None
    if "serving" in want:
        serving = serve_sweep(ds, (50, 100, 200, 500, 1000, 2000))
        progress("serving sweep done")
    if "fit_loop" in want:
        extras["fit_loop_road_mixture"] = fit_loop(ds)
        progress("fit loop done")
    if "fit_d18" in want:
        extras["correlated_gauss_d18"] = c3_fit()
        progress("18-feature fit done")
    if "fit_predict_restarts" in want:
        extras["fit_predict_3_restarts"] = fit_predict_restarts(ds)
        progress("fit + predict restarts done")
    if "uniform_8d" in want:
        extras["uniform_8d_2m"] = [c4_eval(32)]
        progress("2M-row 8-feature m=32 done")
        extras["uniform_8d_2m"].append(c4_eval(64))
        progress("2M-row 8-feature m=64 done")
    if "matern" in want:
        # the same evaluation with a Matern stationary part (same parameters)
        matern = {}
        for nu in (0.5, 1.5, 2.5):
            km = nat.kernel_struct(dict(params, matern_nu=nu), mapping, len(kern.theta))

            def ev_m(km=km):
                nn_m = hip_ops.knn(x, x, ls, m, ref_cut=0, precision=F32)
                return hip_ops.vecchia_loglik(x, y, nn_m, km, alpha, precision=F64,
                                              eval_gradient=True, check_pd=False)
            ev_m()
            t_m, st_m = stage_times(ev_m, max(2, args.steps // 2))
            matern["nu=%g" % nu] = {"evals_per_s": 1.0 / t_m,
                                    "fused_ms": st_m["fused_loglik_grad"] /
                                    max(st_m["fused_loglik_grad_launches"], 1),
                                    "loglik": float(ev_m()[0].item())}
        extras["matern"] = matern
        progress("matern done")
    if "sequential" in want:
        seq = sequential_extras(x, y, xq, ks, ls, m, alpha, max(2, args.steps // 2))
        if cpu_seq is not None:
            seq["cpu_reference_loop"] = cpu_seq
            seq["speedup_vs_cpu_loop"] = seq["full_mean_std_queries_per_s"] / cpu_seq["queries_per_s"]
        extras["sequential_conditioning"] = seq
        progress("sequential conditioning done")

    line = {
        "metric": "loglik+grad evals/s (reference default precision 'only-nn': FP64 factor, "
                  "kNN recomputed per eval)",
        "value": 1.0 / step_s,
        "unit": "evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": step_s * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp64",
        "data": "synthetic",
        "config": {"workload": "road_mixture: n_train=%d n_test=%d d=3 m=32, "
                               "White(0.01)+C(1.0)*RBF(0.1)" % (n, xq.shape[0]),
                   "n_train": n, "n_test": int(xq.shape[0]), "d": d, "m": m},
        "roofline": roofline,
        "cpu_baseline": cpu,
        **extras,
        "serving": serving,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
