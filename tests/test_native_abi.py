"""The C ABI library loads and exports every function include/nngp.h declares
(no compute calls: no GPU needed)."""
import ctypes
import os
import re

import numpy as np
import pytest

from nngpr_amd import _native as nat

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    text = open(os.path.join(ROOT, "include", "nngp.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"\b(nngp_[a-z0-9_]+)\s*\(", text)
    return sorted(set(names))


def test_library_exports_every_declared_symbol():
    lib = nat.lib()
    names = declared_functions()
    assert len(names) >= 10
    for name in names:
        assert hasattr(lib, name), name
    bound = {n for n, _, _ in nat.SIGNATURES}
    assert set(names) == bound


def test_abi_version_and_status_strings():
    lib = nat.lib()
    assert lib.nngp_abi_version() == 2
    assert nat.status_text(0) == "success"
    assert "positive definite" in nat.status_text(nat.NNGP_ERR_NOT_POSITIVE_DEFINITE)


def test_kernel_struct_layout():
    # double x3, int32 x6, double[64], int32 covariance, pad to 8
    assert ctypes.sizeof(nat.NngpKernel) == 3 * 8 + 6 * 4 + 64 * 8 + 8
    assert nat.NngpKernel.length_scale.offset == 48
    assert nat.NngpKernel.covariance.offset == 48 + 64 * 8


@pytest.mark.parametrize("nu,fam", [(0.5, nat.COV_MATERN_12), (1.5, nat.COV_MATERN_32),
                                    (2.5, nat.COV_MATERN_52), (np.inf, nat.COV_RBF)])
@pytest.mark.parametrize("ard", [False, True])
def test_matern_decomposition(nu, fam, ard):
    """C * Matern + White reduces like C * RBF + White, with the family kept aside
    (nu is not a hyperparameter: it maps to no theta)."""
    from sklearn.gaussian_process import kernels as skk
    from nngpr_amd import batched_kernels as bk
    from nngpr_amd.kernel_params import get_kernel_params, get_kernel_theta_mapping
    ls = np.array([0.3, 0.7]) if ard else 0.4
    for mod in (skk, bk):
        k = mod.ConstantKernel(1.3) * mod.Matern(ls, nu=nu) + mod.WhiteKernel(0.2)
        p, mp = get_kernel_params(k), get_kernel_theta_mapping(k)
        assert ("matern_nu" in p) == (not np.isinf(nu))
        assert "matern_nu" not in mp
        assert mp["rbf_level"] == 0 and mp["noise_level"] == (3 if ard else 2)
        np.testing.assert_array_equal(np.atleast_1d(mp["length_scale"]),
                                      [1, 2] if ard else [1])
        ks = nat.kernel_struct(p, mp, len(k.theta))
        assert ks.covariance == fam and ks.n_length_scale == (2 if ard else 1)
        assert ks.rbf_level == pytest.approx(1.3) and ks.noise_level == pytest.approx(0.2)


def test_matern_other_nu_not_accelerated():
    from sklearn.gaussian_process import kernels as skk
    from nngpr_amd.kernel_params import can_accelerate_kernel
    assert not can_accelerate_kernel(skk.Matern(0.4, nu=1.0) + skk.WhiteKernel(0.1))
    assert can_accelerate_kernel(skk.Matern(0.4, nu=2.5) + skk.WhiteKernel(0.1))
    # two stationary parts cannot be summed
    assert not can_accelerate_kernel(skk.Matern(0.4, nu=0.5) + skk.RBF(0.2))


def test_decomposition_cache_keeps_mapping_while_structure_holds(monkeypatch):
    """The cached decomposition re-reads the parameter values at every new theta
    but maps theta again only when the kernel's structure changes; each cached
    result equals a fresh decomposition."""
    from sklearn.gaussian_process import kernels as skk
    from nngpr_amd import kernel_params as kp
    calls = []
    orig = kp.get_kernel_theta_mapping
    monkeypatch.setattr(kp, "get_kernel_theta_mapping", lambda k: calls.append(1) or orig(k))

    def same_as_fresh(dec, k):
        p, mp = kp.get_kernel_params(k), orig(k)
        assert dec.params.keys() == p.keys() and dec.mapping.keys() == mp.keys()
        for a, b in ((dec.params, p), (dec.mapping, mp)):
            for name in a:
                np.testing.assert_array_equal(np.asarray(a[name]), np.asarray(b[name]))

    k = skk.WhiteKernel(0.01) + skk.ConstantKernel(1.0) * skk.RBF([0.1, 0.2])
    dec = kp.KernelDecomposition()
    dec.update(k)
    for step in range(3):  # an optimizer moving theta in place
        k.theta = k.theta + 0.1 * (step + 1)
        same_as_fresh(dec.update(k), k)
    assert len(calls) == 1
    k.set_params(k2__k1__constant_value_bounds="fixed")  # one theta fewer
    same_as_fresh(dec.update(k), k)
    assert len(calls) == 2
    k.k1, k.k2 = k.k2, k.k1  # same theta count, new order
    same_as_fresh(dec.update(k), k)
    assert len(calls) == 3
    other = skk.WhiteKernel(0.01) + skk.ConstantKernel(1.0) * skk.RBF([0.1, 0.2])
    same_as_fresh(dec.update(other), other)  # another kernel object
    assert len(calls) == 4


def test_workspace_queries_are_host_only():
    lib = nat.lib()
    assert lib.nngp_knn_workspace_bytes(1000, 1000, 3, 16, 1) > 0
    assert lib.nngp_vecchia_workspace_bytes(1000, 16, 3, 1, 3, 0, 1) > 0
    assert lib.nngp_predict_workspace_bytes(100, 16, 3, 1, 0) > 0


def test_error_mapping():
    with pytest.raises(NotImplementedError):
        nat.check(nat.NNGP_ERR_UNSUPPORTED)
    import numpy as np
    with pytest.raises(np.linalg.LinAlgError):
        nat.check(nat.NNGP_ERR_NOT_POSITIVE_DEFINITE)


def test_knn_method_setting():
    """The search-method switch is host state only (no GPU call)."""
    assert nat.knn_get_method() == nat.KNN_AUTO
    nat.knn_set_method(nat.KNN_BRUTE_FORCE)
    assert nat.knn_get_method() == nat.KNN_BRUTE_FORCE
    for m in (nat.KNN_GRID_LANES, nat.KNN_GRID_WAVES, nat.KNN_AUTO):
        nat.knn_set_method(m)
        assert nat.knn_get_method() == m
    with pytest.raises(ValueError):
        nat.knn_set_method(7)
    assert nat.knn_get_method() == nat.KNN_AUTO
