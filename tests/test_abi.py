"""The C-ABI library loads without a GPU and exports every entry point that
include/bemuse_amd.h declares; host-only entry points work on the CPU."""
import ctypes
import os
import re

import numpy as np
import pytest

import bemuse_amd as bm

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "bemuse_amd.h")).read()
    return sorted(set(re.findall(r"^\s*(?:int|void|const char\*)\s+(bem_\w+)\s*\(", text, re.M)))


def test_header_symbols_exported():
    L = bm.lib()
    names = declared_symbols()
    assert len(names) >= 20
    for n in names:
        assert hasattr(L, n), n
    assert set(names) == set(bm.exported_symbols())


def test_mesh_counts_and_errors():
    m = bm.Mesh.half_cylinder(10.0, 20.0, 8, 24, 32)
    nn, ne = m.size()
    assert ne == 1024 and nn == 1 + 32 * 32
    nodes, elems = m.arrays()
    assert (elems[:, 3] < 0).sum() == 32
    with pytest.raises(bm.BemError):
        bm.Mesh.half_cylinder(-1.0, 20.0, 8, 24, 32)
    with pytest.raises(bm.BemError):
        bm.Mesh.from_arrays(nodes, np.array([[0, 1, 99999, 2]]))


def test_semisub_mesh_size():
    m = bm.Mesh.semisub3()
    assert 5900 <= m.n_panels <= 6100


def test_random_body_deterministic():
    a = bm.Mesh.random_body(20261203, 5)
    b = bm.Mesh.random_body(20261203, 5)
    na, ea = a.arrays(); nb, eb = b.arrays()
    assert np.array_equal(na, nb) and np.array_equal(ea, eb)
    assert abs(a.params["n"] - a.params["target"]) < 0.25 * a.params["target"]


def test_context_without_gpu_fails_cleanly():
    import torch
    if torch.cuda.is_available():
        pytest.skip("GPU present")
    with pytest.raises(bm.BemError):
        bm.Context(0)
