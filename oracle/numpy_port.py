"""Restatement of the reference's CPU path, used as the timed CPU baseline.

TEST / BENCHMARK INFRASTRUCTURE ONLY (see oracle/__init__.py): bench.py times it on
the host cores beside the GPU numbers; the product never calls it.

What it restates (NumpyNngpr over NumpyWorker, FP64 throughout):
  * Nngpr._log_likelihood_worker (base_nngpr.py:453-614) over row batches of
    2,400 (NngprWorker.custom_batch_size_for_nn, base_nngpr.py:1867);
  * the generic kNN of NngprWorker.find_nearest_neighbors (base_nngpr.py:1613-1686):
    dense blocks of -k(query, ref) capped at 0.5 GiB, Vecchia mask, full-row argsort;
  * the generic fill (base_nngpr.py:1781-1826 with the batched kernels);
  * NumpyWorker.batched_chofactor / batched_chosolve (numpy_nngpr.py:140-170):
    per-matrix scipy.linalg.cho_factor / cho_solve loops;
  * work spread over worker processes with BLAS threads = 1
    (numpy_nngpr.py:82, base_nngpr.py:173-178).
"""
import os
import time

import numpy as np
from scipy import linalg as sl


def _kernel_dists(q, ref, C, s, ls):
    # -k(q, ref) for White + C + s * RBF with Y given (White(X, Y) = 0).
    qs, rs = q / ls, ref / ls
    d2 = (qs * qs).sum(-1)[:, None] + (rs * rs).sum(-1)[None, :] - 2.0 * qs @ rs.T
    return -(C + s * np.exp(-0.5 * np.maximum(d2, 0.0)))


def generic_knn(ref, query, C, s, ls, k, ref_cut):
    nn = -np.ones((query.shape[0], k), dtype=np.int64)
    i_q_max = max(0, 1 - ref_cut) if ref_cut is not None else 0
    while i_q_max < query.shape[0]:
        i_q_min = i_q_max
        bs = max(1, int(1024 ** 3 * 0.5 / (ref.itemsize * ref.shape[0])))
        i_q_max = min(query.shape[0], i_q_min + bs)
        this_ref = ref[:i_q_max + ref_cut] if ref_cut is not None else ref
        dists = _kernel_dists(query[i_q_min:i_q_max], this_ref, C, s, ls)
        if ref_cut is not None:
            unmask = (np.arange(i_q_min, i_q_max)[:, None] + ref_cut) <= np.arange(this_ref.shape[0])[None, :]
            dists[unmask] = np.inf
        idx = np.argsort(dists, -1)[..., :k]
        if ref_cut is not None:
            if unmask.shape[1] >= idx.shape[1]:
                idx[unmask[:, :idx.shape[1]]] = -1
        nn[i_q_min:i_q_max, :idx.shape[1]] = idx
    return nn


def _fill(xg, valid, w, C, s, ls, nt, tmap):
    diff = xg[:, :, None, :] - xg[:, None, :, :]
    d2 = (diff * diff).sum(-1) / (ls * ls)
    E = np.exp(-0.5 * d2)
    M = xg.shape[1]
    eye = np.eye(M)[None]
    K = s * E + C + w * eye
    G = np.zeros(K.shape + (nt,))
    G[..., tmap["noise_level"]] = w * eye
    G[..., tmap["rbf_level"]] = s * E
    G[..., tmap["length_scale"]] = s * E * d2
    neg = ~valid
    b, j = np.where(neg)
    K[b, j, :] = 0
    K[b, :, j] = 0
    G[b, j, :] = 0
    G[b, :, j] = 0
    K[b, j, j] = 1
    return K, G


def _cho_factor(A):
    out = np.empty_like(A)
    for i in range(A.shape[0]):
        out[i] = sl.cho_factor(A[i], lower=True)[0]
    return out


def _cho_solve(L, y):
    shape = np.broadcast_shapes(L.shape[:-2], y.shape[:-2])
    out = np.empty(shape + y.shape[-2:])
    Lb = np.broadcast_to(L, shape + L.shape[-2:])
    yb = np.broadcast_to(y, shape + y.shape[-2:])
    for idx in np.ndindex(shape):
        out[idx] = sl.cho_solve((Lb[idx], True), yb[idx])
    return out


def loglik_batch(x, y, i0, i1, w, C, s, ls, k, alpha, tmap, nt, eval_gradient=True):
    """Loglik (+grad) of training rows [i0, i1): one task of the reference worker."""
    t0 = time.perf_counter()
    kk = min(k, max(i1 - 1, 1))
    nn = generic_knn(x[:i1], x[i0:i1], C, s, ls, kk, i0)
    t_knn = time.perf_counter() - t0
    idx = np.concatenate([np.arange(i0, i1)[:, None], nn], axis=1)
    valid = idx > -1
    K, Kg = _fill(x[np.where(valid, idx, 0)], valid, w, C, s, ls, nt, tmap)
    n = K.shape[0]
    K[np.where(valid)[0], np.where(valid)[1], np.where(valid)[1]] += alpha
    Kg = np.moveaxis(Kg, -1, 0)
    L = _cho_factor(K[:, 1:, 1:])
    yN = np.where(valid[:, 1:], y[np.where(valid[:, 1:], idx[:, 1:], 0)], 0.0)[..., None]
    yi = y[i0:i1][:, None]
    Kxn = K[:, 0:1, 1:]
    Kinv_y = _cho_solve(L, yN)
    Kinv_k = _cho_solve(L, np.swapaxes(Kxn, 1, 2))
    mu = (Kxn @ Kinv_y).reshape(n, 1)
    sigma = np.sqrt(K[:, 0, 0].reshape(n, 1) - (Kxn @ Kinv_k).reshape(n, 1))
    ll = (-0.5 * (yi - mu) ** 2 / sigma ** 2 - 0.5 * np.log(2 * np.pi) - np.log(sigma)).sum()
    g = np.zeros(nt)
    if eval_gradient:
        GNN = Kg[:, :, 1:, 1:]
        dKxn = Kg[:, :, 0:1, 1:]
        def dKinv(right):
            return -_cho_solve(L[None], GNN @ _cho_solve(L, right)[None])
        dmu = (dKxn @ Kinv_y[None]).reshape(nt, n, 1) + (Kxn[None] @ dKinv(yN)).reshape(nt, n, 1)
        dsig = 0.5 / sigma[None] * (Kg[:, :, 0, 0].reshape(nt, n, 1)
                                    - 2 * (dKxn @ Kinv_k[None]).reshape(nt, n, 1)
                                    - (Kxn[None] @ dKinv(np.swapaxes(Kxn, 1, 2))).reshape(nt, n, 1))
        r = yi[None] - mu[None]
        gg = (-1 / sigma[None] + r ** 2 / sigma[None] ** 3) * dsig + r / sigma[None] ** 2 * dmu
        g = gg.sum(axis=(1, 2))
    return ll, g, t_knn, time.perf_counter() - t0


def _worker(args):
    os.environ.setdefault("OMP_NUM_THREADS", "1")
    return loglik_batch(*args)


def time_sample(x, y, w, C, s, ls, k, alpha, rows_per_task=64, n_tasks=None, workers=None):
    """Times tasks of the reference worker on rows sampled uniformly over [k+1, n)
    and extrapolates one full evaluation.

    Row i costs a_knn(i) (distance block over i refs + argsort, linear-log in i)
    plus a roughly constant fill/factor/solve part; every sampled task of
    `rows_per_task` rows ending at a uniformly spaced row measures both.  A
    least-squares line a + b i through the per-row costs is summed over all n
    rows to estimate one full evaluation.  Returns a dict with the estimate and
    the sample description.
    """
    import multiprocessing as mp
    n = x.shape[0]
    workers = workers or len(os.sched_getaffinity(0))
    n_tasks = n_tasks or workers
    tmap = {"noise_level": 0, "rbf_level": 1, "length_scale": 2}
    ends = np.linspace(max(rows_per_task + k + 1, n // n_tasks), n, n_tasks).astype(int)
    tasks = [(x, y, int(e - rows_per_task), int(e), w, C, s, ls, k, alpha, tmap, 3, True)
             for e in ends]
    t0 = time.perf_counter()
    with mp.get_context("fork").Pool(workers) as pool:
        res = pool.map(_worker, tasks)
    wall = time.perf_counter() - t0
    per_row = np.array([r[3] / rows_per_task for r in res])
    centres = ends - rows_per_task / 2.0
    # per-row cost a + b * i (kNN block + argsort grow with i; factor/solve constant)
    b, a = np.polyfit(centres, per_row, 1) if len(per_row) > 1 else (0.0, per_row[0])
    a = max(a, 0.0)
    total_cpu_s = float(a * n + b * n * (n - 1) / 2.0)
    return {"cpu_seconds_per_eval_single_core": total_cpu_s,
            "est_seconds_per_eval": total_cpu_s / workers,
            "sample_wall_s": wall, "workers": workers, "tasks": len(tasks),
            "rows_per_task": rows_per_task, "sampled_rows": int(len(tasks) * rows_per_task)}
