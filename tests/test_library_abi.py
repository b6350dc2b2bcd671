"""The C-ABI library loads on a host without a GPU and exports every
function declared in include/bem_hip.h; calls fail cleanly (no fallback)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared(header):
    text = open(header).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(bem_\w+)\s*\(", text)))


def test_exports_match_header(bem):
    names = declared(os.path.join(ROOT, "include", "bem_hip.h"))
    assert len(names) >= 25
    L = ctypes.CDLL(bem.LIB_PATH)
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    assert set(names) == set(bem.EXPORTS)


def test_no_device_fails_loudly(bem):
    import numpy as np
    try:
        import torch
        if torch.cuda.is_available():
            pytest.skip("a GPU is present")
    except ImportError:
        pass
    with pytest.raises(bem.BemError):
        bem.Context(0)
    tri = bem.meshes.icosphere(1, 1.0)
    with pytest.raises(bem.BemError):
        bem.Solver(tri, bem.meshes.robin_bc(len(tri)), leaf=8, eta=2.0)


def test_null_arguments_rejected(bem):
    L = bem.lib()
    assert L.bem_destroy(None) == bem.BEM_ERR_ARG
    assert L.bem_hmat_destroy(None) == bem.BEM_ERR_ARG
    assert L.bem_solver_destroy(None) == bem.BEM_ERR_ARG
    assert L.bem_create(None, 0) == bem.BEM_ERR_ARG


def test_non_finite_vertices_refused_before_any_device_call(bem):
    """bem_solver_create refuses a mesh with a non-finite vertex coordinate
    (BEM_ERR_ARG) before it builds the cluster tree or touches a device, so
    the refusal is the same with or without a GPU."""
    import numpy as np
    tri = bem.meshes.icosphere(2, 0.5)
    bc = bem.meshes.robin_bc(len(tri), 0.0, 1.0, 1.0)
    for bad in (np.nan, np.inf, -np.inf):
        t = tri.copy()
        t[37, 4] = bad
        with pytest.raises(bem.BemError) as e:
            bem.Solver(t, bc, leaf=16, eta=2.0)
        assert e.value.code == bem.BEM_ERR_ARG
