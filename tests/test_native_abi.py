"""The C ABI library loads and exports every function include/nngp.h declares
(no compute calls: no GPU needed)."""
import ctypes
import os
import re

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
    assert lib.nngp_abi_version() == 1
    assert nat.status_text(0) == "success"
    assert "positive definite" in nat.status_text(nat.NNGP_ERR_NOT_POSITIVE_DEFINITE)


def test_kernel_struct_layout():
    # double x3, int32 x6, pad to 8, double[64]
    assert ctypes.sizeof(nat.NngpKernel) == 3 * 8 + 6 * 4 + 64 * 8
    assert nat.NngpKernel.length_scale.offset == 48


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
    nat.knn_set_method(nat.KNN_AUTO)
    assert nat.knn_get_method() == nat.KNN_AUTO
    with pytest.raises(ValueError):
        nat.knn_set_method(7)
    assert nat.knn_get_method() == nat.KNN_AUTO
