"""HIP kNN vs the FP64 oracle: neighbour indices must be bit-identical.

Covers every search form (uniform-grid search for d <= 3, one query per lane
and one query per wave; brute-force exact FP64 scan; FP32 screen + FP64 re-rank with certificate), Vecchia cuts like the
reference's tests (None, 0, 2, -1), isotropic / ARD / raw coordinates, ties and
duplicates, non-finite coordinates, clustered and degenerate point sets, queries
far outside the refs, empty inputs, k > 64 and d > 8 (per-row path), the stored
golden vectors, and slices of the 434,874-point road-mixture workload.
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


# AUTO runs the grid search one query per wave at these sizes; GRID_LANES forces
# the lane-per-query grid kernel, GRID_WAVES the wave-per-query one at every size
METHODS = [hip_ops.KNN_AUTO, hip_ops.KNN_BRUTE_FORCE, hip_ops.KNN_BRUTE_FORCE_VALU,
           hip_ops.KNN_GRID_LANES, hip_ops.KNN_GRID_WAVES]


@pytest.fixture(params=METHODS, ids=["auto", "brute", "valu", "lanes", "waves"])
def method(request):
    with hip_ops.knn_method(request.param):
        yield request.param


def hip_knn(ref, qry, ls, k, cut, prec):
    out = hip_ops.knn(t(ref), t(qry), ls, k, ref_cut=cut, precision=prec)
    return out.cpu().numpy()


def hip_knn_self(x, ls, k, cut, prec):
    """The query set is the ref tensor itself (the Vecchia self-search)."""
    xt = t(x)
    return hip_ops.knn(xt, xt, ls, k, ref_cut=cut, precision=prec).cpu().numpy()


@pytest.mark.parametrize("prec", [0, 1])
def test_golden_small_uniform(prec, method):
    g = np.load(GOLD)
    ds = datasets.small_uniform()
    x, xq = ds["x_train"], ds["x_test"]
    np.testing.assert_array_equal(hip_knn(x, x, 0.3, 16, 0, prec), g["nn_train"])
    np.testing.assert_array_equal(hip_knn(x, xq, 0.3, 16, None, prec), g["nn_test"])


@pytest.mark.parametrize("prec", [0, 1])
@pytest.mark.parametrize("cut", [None, 0, 2, -1])
@pytest.mark.parametrize("d,ls", [(1, 0.2), (2, None), (3, "ard"), (5, 0.7), (8, "ard"),
                                  (12, "ard"), (18, 0.9), (32, "ard")])
@pytest.mark.parametrize("k", [1, 7, 33, 64])
def test_against_oracle(prec, cut, d, ls, k, method):
    rng = np.random.default_rng(100 + d * 7 + k)
    ref = rng.normal(size=(2100, d))
    qry = ref[:700] if cut is not None else rng.normal(size=(700, d))
    if ls == "ard":
        ls = rng.uniform(0.2, 2.0, size=d)
    exp = vo.knn(ref, qry, ls, k, ref_cut=cut)
    got = hip_knn(ref, qry, ls, k, cut, prec)
    np.testing.assert_array_equal(got, exp)


@pytest.mark.parametrize("prec", [0, 1])
@pytest.mark.parametrize("d,k", [(18, 32), (3, 100), (12, 5), (33, 8), (40, 70)])
def test_per_row_path(prec, d, k):
    rng = np.random.default_rng(7)
    x = rng.normal(size=(900, d))
    exp = vo.knn(x, x, 1.3, k, ref_cut=0)
    np.testing.assert_array_equal(hip_knn(x, x, 1.3, k, 0, prec), exp)


@pytest.mark.parametrize("prec", [0, 1])
def test_ties_and_duplicates(prec, method):
    rng = np.random.default_rng(5)
    pts = rng.integers(0, 5, size=(3000, 3)).astype(np.float64)
    for cut in (0, None):
        exp = vo.knn(pts, pts, None, 24, ref_cut=cut)
        np.testing.assert_array_equal(hip_knn(pts, pts, None, 24, cut, prec), exp)
        np.testing.assert_array_equal(hip_knn_self(pts, None, 24, cut, prec), exp)
    ones = np.ones((70, 3))
    exp = vo.knn(ones, ones, 1.0, 16, ref_cut=0)
    np.testing.assert_array_equal(hip_knn(ones, ones, 1.0, 16, 0, prec), exp)


@pytest.mark.parametrize("prec", [0, 1])
def test_empty_inputs(prec, method):
    x = np.random.default_rng(0).random((10, 3))
    out = hip_ops.knn(t(x), t(np.zeros((0, 3))), None, 4, precision=prec)
    assert out.shape == (0, 4)
    out = hip_ops.knn(t(np.zeros((0, 3))), t(x), None, 4, precision=prec).cpu().numpy()
    assert np.all(out == -1)


@pytest.mark.parametrize("prec", [0, 1])
def test_road_mixture_slice(prec, method):
    ds = datasets.road_mixture(n_train=60_000, n_test=4_000)
    x, xq = ds["x_train"], ds["x_test"]
    exp = vo.knn(x, x, 0.1, 32, ref_cut=0)
    np.testing.assert_array_equal(hip_knn(x, x, 0.1, 32, 0, prec), exp)
    np.testing.assert_array_equal(hip_knn_self(x, 0.1, 32, 0, prec), exp)
    exp = vo.knn(x, xq, 0.1, 32)
    np.testing.assert_array_equal(hip_knn(x, xq, 0.1, 32, None, prec), exp)


@pytest.mark.parametrize("d", [1, 2, 3, 5, 9])
@pytest.mark.parametrize("cut", [None, 0, 5])
def test_non_finite_coordinates(d, cut, method):
    """Refs with a NaN or infinite coordinate are never neighbours; a query with one
    gets -1 only (the scans' admission never takes a NaN or infinite key)."""
    rng = np.random.default_rng(40 + d)
    x = np.round(rng.normal(size=(3000, d)), 2)
    x[[3, 50, 777], 0] = np.nan
    x[[9, 1500], d - 1] = np.inf
    x[2222, 0] = -np.inf
    q = x if cut is not None else np.concatenate([x[:50], rng.normal(size=(50, d))])
    exp = vo.knn(x, q, None, 17, ref_cut=cut)
    np.testing.assert_array_equal(hip_knn(x, q, None, 17, cut, 1), exp)
    if cut is not None:
        np.testing.assert_array_equal(hip_knn_self(x, None, 17, cut, 0), exp)


@pytest.mark.parametrize("shape", ["cluster", "line", "plane", "point", "far_queries", "spread"])
def test_point_set_shapes(shape, method):
    """Clustered, lower-dimensional and degenerate point sets, and queries far
    outside the refs' bounding box."""
    rng = np.random.default_rng(11)
    n = 20_000
    if shape == "cluster":
        x = np.concatenate([rng.normal(size=(n // 2, 3)) * 1e-4, rng.normal(size=(n // 2, 3)) * 10])
        rng.shuffle(x)
    elif shape == "line":
        s_ = rng.random(n)
        x = np.stack([s_, 2 * s_ + 1, np.full(n, 0.5)], 1)
    elif shape == "plane":
        x = np.concatenate([rng.random((n, 2)), np.zeros((n, 1))], 1)
    elif shape == "point":
        x = np.full((n // 10, 3), 2.5)
    elif shape == "spread":
        x = rng.random((n, 3)) * np.array([1e6, 1.0, 1e-6]) + np.array([3e7, 0.0, -1.0])
    else:
        x = rng.random((n, 3))
    q = rng.normal(size=(500, 3)) * 50 if shape == "far_queries" else x[:700] + 1e-3
    for cut, qq in ((0, x), (None, q)):
        exp = vo.knn(x, qq, None, 32, ref_cut=cut)
        got = hip_knn_self(x, None, 32, 0, 1) if cut == 0 else hip_knn(x, qq, None, 32, None, 1)
        np.testing.assert_array_equal(got, exp)


@pytest.mark.parametrize("k", [1, 16, 33, 64])
def test_k_beyond_refs_and_small_sets(k, method):
    rng = np.random.default_rng(k)
    for n in (1, 5, 40, 130):
        x = rng.random((n, 3))
        q = rng.random((9, 3))
        np.testing.assert_array_equal(hip_knn(x, q, 0.5, k, None, 0), vo.knn(x, q, 0.5, k))
        np.testing.assert_array_equal(hip_knn_self(x, 0.5, k, 0, 1), vo.knn(x, x, 0.5, k, ref_cut=0))


def test_methods_agree_on_road_mixture_prefix():
    """Grid search and brute-force scans on the first 150,000 road-mixture rows
    (self-search, Vecchia cut 0, m = 32): identical indices."""
    ds = datasets.road_mixture()
    x = ds["x_train"][:150_000]
    outs = []
    for m in METHODS:
        with hip_ops.knn_method(m):
            outs.append(hip_knn_self(x, 0.1, 32, 0, 1))
    for o in outs[1:]:
        np.testing.assert_array_equal(outs[0], o)
    # no equal keys at the k-th place here: the lane kernel queues no query
    with hip_ops.knn_method(hip_ops.KNN_GRID_LANES):
        hip_knn_self(x, 0.1, 32, 0, 1)
        assert hip_ops.knn_stats()["knn_rows_recomputed"] == 0
    # a sample of rows against the oracle on each row's own prefix of refs
    rows = np.concatenate([[0, 1, 31, 32, 33], np.random.default_rng(0).choice(x.shape[0], 40)])
    for r in rows:
        e = vo.knn(x[:r], x[r:r + 1], 0.1, 32)[0] if r > 0 else -np.ones(32, np.int32)
        np.testing.assert_array_equal(outs[0][r], e)


@pytest.mark.parametrize("k", [16, 32, 64])
@pytest.mark.parametrize("cut", [0, None])
def test_lattice_ties_lane_kernel(k, cut):
    """150,000 points on an integer lattice of 60^3 sites (most keys tie): the
    lane-per-query kernel, whose lists of 32 and 64 order keys alone and queue
    every query with an equal key at the k-th place for the exact one-query-per-
    wave search, gives the same indices as that search for every query, and the
    oracle's on a sample of rows (each against its own prefix for cut 0)."""
    rng = np.random.default_rng(11 + k)
    x = rng.integers(0, 60, size=(150_000, 3)).astype(np.float64)
    xt = t(x)
    outs = []
    for m in (hip_ops.KNN_GRID_LANES, hip_ops.KNN_GRID_WAVES):
        with hip_ops.knn_method(m):
            outs.append(hip_ops.knn(xt, xt, None, k, ref_cut=cut).cpu().numpy())
            if m == hip_ops.KNN_GRID_LANES:
                queued = hip_ops.knn_stats()["knn_rows_recomputed"]
    np.testing.assert_array_equal(outs[0], outs[1])
    # lists of 16 keep (key, index) order in the lane kernel and queue nothing
    assert (queued > 10_000) if k >= 32 else (queued == 0), queued
    rows = np.concatenate([[0, 1, k, k + 1], rng.choice(x.shape[0], 12)])
    for r in rows:
        if cut == 0:
            e = vo.knn(x[:r], x[r:r + 1], None, k)[0] if r > 0 else -np.ones(k, np.int32)
        else:
            e = vo.knn(x, x[r:r + 1], None, k)[0]
        np.testing.assert_array_equal(outs[0][r], e)


@pytest.mark.parametrize("d", [1, 2, 3])
def test_grid_one_far_point_large(d):
    """100,000 uniform points and one far one (the lane-per-query kernel's size):
    the far point no longer coarsens the grid (the search took 22 ms with it and
    0.65 ms without it at d = 3); exact on the far row, the rows after it and a
    sample of rows, each against its own prefix."""
    import time
    rng = np.random.default_rng(40 + d)
    n, k = 100_001, 32
    x = rng.random((n, d))
    x[n // 3] = 1000.0
    xt = t(x)
    with hip_ops.knn_method(hip_ops.KNN_AUTO):
        got = hip_ops.knn(xt, xt, 0.1, k, ref_cut=0).cpu().numpy()
        rows = np.concatenate([[n // 3, n // 3 + 1, n - 1], rng.choice(n, 30)])
        for r in rows:
            np.testing.assert_array_equal(got[r], vo.knn(x[:r], x[r:r + 1], 0.1, k)[0])
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(5):
            hip_ops.knn(xt, xt, 0.1, k, ref_cut=0)
        torch.cuda.synchronize()
        assert (time.perf_counter() - t0) / 5 < 0.006


@pytest.mark.parametrize("ls", [[1e5, 1e5, 1e-5], [731.42, 302.17, 0.012312]])
def test_strongly_anisotropic_scales(ls):
    """Length scales that put the scaled points on (nearly) a line: the grid is one
    cell wide in x and the search walks only each ring's boundary rows.  Exact
    against the oracle on a prefix, and bounded in time at 50,000 rows (the full
    ring walk took seconds there)."""
    import time
    ls = np.array(ls)
    x = np.random.default_rng(7).random((50_000, 3))
    sub = x[:4000]
    np.testing.assert_array_equal(hip_knn_self(sub, ls, 32, 0, 0), vo.knn(sub, sub, ls, 32, ref_cut=0))
    np.testing.assert_array_equal(hip_knn(x, sub[:500], ls, 32, None, 1),
                                  vo.knn(x, sub[:500], ls, 32))
    xt = t(x)
    hip_ops.knn(xt, xt, ls, 32, ref_cut=0, precision=1)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    hip_ops.knn(xt, xt, ls, 32, ref_cut=0, precision=1)
    torch.cuda.synchronize()
    assert time.perf_counter() - t0 < 0.5


@pytest.mark.parametrize("screen", [hip_ops.KNN_BRUTE_FORCE, hip_ops.KNN_BRUTE_FORCE_VALU],
                         ids=["mfma", "valu"])
@pytest.mark.parametrize("cut", [None, 0])
@pytest.mark.parametrize("case", ["one_far_point", "far_cluster", "heavy_tails"])
def test_screen_with_outliers(case, cut, screen):
    """Screens of the brute-force scans (d = 5: the matrix-core bf16x3 screen and
    the FP32 VALU screen) on point sets with far points: the indices stay exact,
    and the per-row certificate (centred on the mean, bounded with the row's own
    norm) still certifies almost every row, so far points do not send the whole
    search to the exact per-row recomputation."""
    rng = np.random.default_rng(["one_far_point", "far_cluster", "heavy_tails"].index(case))
    n, d, k = 6000, 5, 16
    x = rng.random((n, d))
    if case == "one_far_point":
        x[n // 2] = 1000.0
    elif case == "far_cluster":
        x[n // 3: n // 3 + n // 100] = x[n // 3: n // 3 + n // 100] * 0.01 + 100.0
    else:
        x = rng.standard_t(2.0, size=(n, d))
    with hip_ops.knn_method(screen):
        got = hip_knn_self(x, 0.1, k, cut, 1)
        rec = hip_ops.knn_stats()["knn_rows_recomputed"]
    np.testing.assert_array_equal(got, vo.knn(x, x, 0.1, k, ref_cut=cut))
    # rows whose query is itself far from every other point keep a wide margin
    # requirement; the rest certify.  The matrix-core screen's error bound is ~40x
    # the FP32 screen's (bf16 hi/lo split), so a few more rows near the shifted
    # centre miss their margin: at most 2.5 % of the rows.
    limit = {"one_far_point": 2, "far_cluster": n // 100 + 2, "heavy_tails": n // 50}[case]
    if screen == hip_ops.KNN_BRUTE_FORCE:
        limit = max(limit, n // 40)
    assert rec <= limit, rec


@pytest.mark.parametrize("grid", [hip_ops.KNN_GRID_WAVES, hip_ops.KNN_GRID_LANES],
                         ids=["waves", "lanes"])
@pytest.mark.parametrize("cut", [None, 0])
@pytest.mark.parametrize("case", ["one_far_point", "few_far_points", "far_cluster", "heavy_tails"])
def test_grid_with_outliers(case, cut, grid):
    """Uniform-grid search (d = 3) when a few points lie far out (the grid then
    covers a box sampled from the refs and keeps the far refs out of its cells;
    every search keys them, and far queries are searched one per wave): indices stay
    exact for the far rows, the rows around them, separate queries far outside on
    every side and queries next to the far refs, with both grid kernels."""
    rng = np.random.default_rng(10 + ["one_far_point", "few_far_points", "far_cluster",
                                      "heavy_tails"].index(case))
    n, k = 5000, 16
    x = rng.random((n, 3))
    if case == "one_far_point":
        x[n // 2] = [1000.0, -1000.0, 3.0]
    elif case == "few_far_points":
        far = rng.choice(n, 6, replace=False)
        x[far] = rng.normal(0.0, 50.0, (6, 3))
        x[far[0], 1] = 0.5                      # far in x and z only
        x[far[1]] = x[far[2]] + [1e-3, 0, 0]    # two far points next to each other
    elif case == "far_cluster":
        x[n // 3: n // 3 + 50] = x[n // 3: n // 3 + 50] * 0.01 + 100.0
    else:
        x = rng.standard_t(1.5, size=(n, 3))
    with hip_ops.knn_method(grid):
        got = hip_knn_self(x, 0.1, k, cut, 1)
        np.testing.assert_array_equal(got, vo.knn(x, x, 0.1, k, ref_cut=cut))
        if cut is None:
            far = np.nonzero(np.abs(x - 0.5).max(axis=1) > 2.0)[0][:20]
            xq = np.concatenate([rng.random((200, 3)), rng.normal(0.0, 300.0, (100, 3)),
                                 x[rng.choice(n, 50)] + 1e-4, x[far] + 1e-3])
            np.testing.assert_array_equal(hip_knn(x, xq, 0.1, k, None, 0),
                                          vo.knn(x, xq, 0.1, k))


@pytest.mark.parametrize("screen", [hip_ops.KNN_BRUTE_FORCE, hip_ops.KNN_BRUTE_FORCE_VALU],
                         ids=["mfma", "valu"])
@pytest.mark.parametrize("prec", [0, 1])
@pytest.mark.parametrize("cut", [None, 0, 2, -1])
@pytest.mark.parametrize("k", [1, 7, 32, 64])
def test_ref_stripes_small_inputs(prec, cut, k, screen):
    """Few query blocks (n = 9,000, d = 18): the scans split the refs into stripes and
    merge the stripes' lists; the indices equal the oracle's, including the first
    rows of the Vecchia cut (whose limits end inside the first stripe) and
    duplicated points across stripes (ties go to the lower index).  k = 1 and 7
    take the smallest lists (16 / 24 screen candidates), k = 64 the widest (72
    screen candidates, 64 exact).  The matrix-core screen splits each stripe over
    the two halves of its query waves (two lists per stripe)."""
    if prec == 0 and screen == hip_ops.KNN_BRUTE_FORCE_VALU:
        pytest.skip("FP64 runs the same exact scan under both screen methods")
    rng = np.random.default_rng(41 + prec)
    n = 9000
    x = rng.normal(size=(n, 18))
    x[7000:7010] = x[10:20]        # exact duplicates in another stripe
    ls = rng.uniform(0.5, 2.0, 18)
    with hip_ops.knn_method(screen):
        got = hip_knn_self(x, ls, k, cut, prec)
    np.testing.assert_array_equal(got, vo.knn(x, x, ls, k, ref_cut=cut))


@pytest.mark.parametrize("cut", [0, None])
@pytest.mark.parametrize("k", [32, 64])
def test_mfma_screen_uniform_8d(cut, k):
    """The matrix-core screen on the uniform 8-feature distribution with length
    scale 0.3 (n = 20,000): indices equal the oracle's, and the bf16x3 error bound
    (~1.3e-4 of (|p| + max |r|)^2) leaves all but a few rows certified by the
    candidate margin (K - k >= 8), so the exact per-row recompute stays rare."""
    rng = np.random.default_rng(8 + k)
    x = rng.random((20_000, 8))
    with hip_ops.knn_method(hip_ops.KNN_AUTO):
        got = hip_knn_self(x, 0.3, k, cut, 1)
        rec = hip_ops.knn_stats()["knn_rows_recomputed"]
    np.testing.assert_array_equal(got, vo.knn(x, x, 0.3, k, ref_cut=cut))
    assert rec <= 200, rec
