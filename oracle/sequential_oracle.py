"""FP64 CPU restatement of the reference's sequential conditioning.

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py): the HIP product path is checked
against these functions; it never calls them.

Follows base_nngpr.py:896-1030 (predict with conditioning_method 'full' /
'full-double-nn' and return_cov), :680-752 (sample_y's sequential branch) and
:1370-1464 (the neighbour lists), one query at a time in index order, as the
reference does.  Kernel parameters as in vecchia_oracle (w, C, s, ls, nu).
"""
import numpy as np

from . import vecchia_oracle as vo


def conditioning_nn(x_train, xq, ls, num_nn, method):
    """Neighbour lists of the queries over [x_train; xq], without the leading
    self column (_find_nn_indices_for_train_and_eval, base_nngpr.py:1370-1464)."""
    x_train = np.asarray(x_train, dtype=np.float64)
    xq = np.asarray(xq, dtype=np.float64)
    nt, nq = x_train.shape[0], xq.shape[0]
    empty = np.zeros((nq, 0), dtype=np.int32)
    if method == "only-train":
        k = min(nt, num_nn)
        return vo.knn(x_train, xq, ls, k) if k > 0 else empty
    if method == "full":
        x_full = np.concatenate([x_train, xq], axis=0)
        return vo.knn(x_full, xq, ls, min(num_nn, nt + nq), ref_cut=nt)
    kq = min(num_nn, nq - 1)
    qn = vo.knn(xq, xq, ls, kq, ref_cut=0) if kq > 0 else empty
    qn = np.where(qn >= 0, qn + nt, qn)
    if nt == 0:
        return qn
    tr = vo.knn(x_train, xq, ls, min(num_nn, nt))
    out = -np.ones((nq, tr.shape[1] + qn.shape[1]), dtype=np.int32)
    for i in range(nq):          # valid training entries, then valid query entries (:1442-1455)
        v = np.concatenate([tr[i][tr[i] >= 0], qn[i][qn[i] >= 0]])
        out[i, :v.size] = v
    return out


def _blocks(x_train, xq, nn, w, C, s, ls, alpha, nu):
    """Per query: K over [query, sorted neighbours] with alpha on training
    neighbours (:897-898, :946-950), and the sorted neighbour row."""
    x_train = np.asarray(x_train, dtype=np.float64)
    xq = np.asarray(xq, dtype=np.float64)
    nt, nq = x_train.shape[0], xq.shape[0]
    x_full = np.concatenate([x_train, xq], axis=0)
    nn = np.asarray(nn).reshape((nq, -1))
    if nn.shape[1] == 0:
        nn = -np.ones((nq, 1), dtype=np.int32)
    idx = np.concatenate([np.arange(nt, nt + nq)[:, None], np.sort(nn, axis=1)], axis=1)
    valid = idx > -1
    K, _ = vo.fill(x_full[np.where(valid, idx, 0)], valid, w, C, s, ls, nu=nu)
    M = idx.shape[1]
    train = (idx >= 0) & (idx < nt)
    K = K + alpha * (train[:, :, None] & np.eye(M, dtype=bool)[None])
    return K, idx[:, 1:], nt


def _targets(y_train):
    y = np.asarray(y_train, dtype=np.float64)
    return y[:, None] if y.ndim == 1 else y


def _solver(Kn):
    L = np.linalg.cholesky(Kn)

    def solve(b):
        return np.linalg.solve(L.T, np.linalg.solve(L, b))
    return solve


def predict_sequential(x_train, y_train, xq, nn, w, C, s, ls, alpha, output="std", nu=None,
                       _offdiag=True):
    """Mean (nq, ny) and std (nq,) [output 'std'] or joint covariance (nq, nq)
    [output 'cov'] or the mean alone [output 'mean'] of the sequential branch of
    predict (base_nngpr.py:896-1030), for the normalised targets.

    'std' keeps the reference's sparse partial covariance: the covariance of two
    query neighbours is the one stored when the later had the earlier among its
    neighbours (:918-931, :986-988, :1003-1008), else 0.  (_offdiag=False drops
    those off-diagonal covariances: a planted mistake for the tests.)
    """
    K, rows, nt = _blocks(x_train, xq, nn, w, C, s, ls, alpha, nu)
    y = _targets(y_train)
    ny = y.shape[1]
    nq, m = rows.shape
    mean = np.full((nt + nq, ny), np.nan)
    mean[:nt] = y
    var = np.full(nq, np.nan)
    pc = {}
    cov = np.zeros((nq, nq)) if output == "cov" else None
    for i in range(nq):
        row = rows[i]
        valid = row > -1
        solve = _solver(K[i, 1:, 1:])
        kxn = K[i, 0, 1:]
        this_y = np.where(valid[:, None], mean[np.where(valid, row, 0)], 0.0)
        mean[nt + i] = kxn @ solve(this_y)
        if output == "mean":
            continue
        nontrain = row >= nt
        pos = np.where(nontrain)[0]
        q = row[nontrain] - nt
        this_cov = np.zeros((m, m))
        for a in range(len(pos) if _offdiag else 0):
            for b in range(a):
                this_cov[pos[a], pos[b]] = cov[q[a], q[b]] if cov is not None \
                    else pc.get((q[a], q[b]), 0.0)
        this_cov = this_cov + this_cov.T
        this_cov[pos, pos] = var[q]
        std2_raw = max(0.0, K[i, 0, 0] - kxn @ solve(kxn))
        var[i] = max(0.0, std2_raw + kxn @ solve(this_cov @ solve(kxn)))
        if cov is not None:
            if valid.any():
                A = (kxn @ np.linalg.inv(K[i, 1:, 1:]))[nontrain]
                cov[i, :i] = A @ cov[q, :i]
                cov[:i, i] = cov[i, :i]
            cov[i, i] = var[i]
        else:
            cw = (kxn @ solve(this_cov))[nontrain]
            for qq, v in zip(q, cw):
                pc[(i, qq)] = v
    mean = mean[nt:]
    if output == "cov":
        return mean, cov
    if output == "std":
        return mean, np.sqrt(var)
    return mean


def sample_sequential(x_train, y_train, xq, nn, w, C, s, ls, alpha, z, nu=None):
    """Samples (nq, ny, ns) of the sequential branch of sample_y
    (base_nngpr.py:703-752) with the standard normals z (nq, ny, ns) standing for
    the random draws: y_i = K_xn K^-1 y_N + sigma_i z_i, y_N holding training
    targets and the samples of earlier query neighbours, sigma_i =
    max(0, sqrt(K_00 - K_xn K^-1 K_nx)) (0 where the root is not real)."""
    K, rows, nt = _blocks(x_train, xq, nn, w, C, s, ls, alpha, nu)
    y = _targets(y_train)
    z = np.asarray(z, dtype=np.float64)
    nq, ny, ns = z.shape
    out = np.full((nq, ny, ns), np.nan)
    for i in range(nq):
        row = rows[i]
        y_nn = np.zeros((row.size, ny, ns))
        tr = (row >= 0) & (row < nt)
        y_nn[tr] = y[row[tr]][:, :, None]
        qm = row >= nt
        y_nn[qm] = out[row[qm] - nt]
        kxn = K[i, 0, 1:].copy()
        kxn[row < 0] = 0.0
        solve = _solver(K[i, 1:, 1:])
        wv = solve(kxn)
        mu = np.einsum("i,ijk->jk", wv, y_nn)
        v = K[i, 0, 0] - kxn @ wv
        sigma = np.sqrt(v) if v > 0 else 0.0
        out[i] = mu + sigma * z[i]
    return out
