"""Failure paths on the device, reached from the host without faulting or
hanging any kernel:

  - refinement that does not reach its target: the mixed modes re-solve the
    frequency with the FP64 LU (fell_back_fp64 = 1). The refinement limit set
    to 0 correction steps makes every mixed frequency take that path, whose
    results must then equal FP64 mode bit for bit;
  - a numerically singular matrix (an exactly zero column: every elimination
    step leaves it zero, so its pivot is exactly 0 in either precision):
    BEM_ERR_NUMERIC from bem_dense_solve, where Eigen's PartialPivLU
    (Solver_Base.h:85) would return NaN silently;
  - non-finite entries are rejected on the host (BEM_ERR_ARG) before any GPU
    work.
"""
import numpy as np
import pytest

import bemuse_amd as bm

pytestmark = pytest.mark.gpu


@pytest.fixture
def no_refinement(ctx):
    ctx.set_refine_limit(0)
    try:
        yield ctx
    finally:
        ctx.set_refine_limit(4)


@pytest.fixture(scope="module")
def cyl(ctx):
    body = bm.Body(ctx, bm.Mesh.half_cylinder(5.0, 8.0, 4, 8, 16))
    body.assemble_rankine()
    yield body
    body.close()


def test_refine_limit_forces_fp64_fallback(cyl, no_refinement):
    betas = [0.0, 0.6]
    g64 = cyl.solve(0.9, betas=betas, mode=bm.MODE_FP64)
    gmx = cyl.solve(0.9, betas=betas, mode=bm.MODE_MIXED)
    assert gmx["info"]["fell_back_fp64"] == 1 and gmx["info"]["refine_iters"] == 0
    for k in ("A", "B", "F", "FK", "SC"):
        assert np.array_equal(gmx[k], g64[k]), k
    omegas = np.linspace(0.3, 1.8, 5)
    s64 = cyl.solve_sweep(omegas, betas=betas, mode=bm.MODE_FP64, lanes=3)
    smx = cyl.solve_sweep(omegas, betas=betas, mode=bm.MODE_MIXED, lanes=3)
    assert all(i["fell_back_fp64"] == 1 for i in smx["info"])
    for k in ("A", "B", "F"):
        assert np.array_equal(smx[k], s64[k]), k


def test_refine_limit_restored(cyl, ctx):
    g = cyl.solve(0.9, betas=[0.0], mode=bm.MODE_MIXED)
    assert g["info"]["fell_back_fp64"] == 0 and g["info"]["residual"] < 1e-13
    with pytest.raises(bm.BemError) as e:
        ctx.set_refine_limit(5)
    assert e.value.code == bm.ERR_ARG


def test_dense_solve_fallback_matches_fp64(ctx, no_refinement):
    rng = np.random.default_rng(5)
    n = 200
    A = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n)) + 3 * np.eye(n)
    B = rng.standard_normal((n, 4)) + 1j * rng.standard_normal((n, 4))
    X64, i64 = ctx.dense_solve(A, B, 64)
    X32, i32 = ctx.dense_solve(A, B, 32)
    assert i32["fell_back_fp64"] == 1 and i64["fell_back_fp64"] == 0
    assert np.array_equal(X32, X64)
    assert np.abs(X64 - np.linalg.solve(A, B)).max() < 1e-12 * np.abs(X64).max()


@pytest.mark.parametrize("precision", [64, 32])
@pytest.mark.parametrize("n,col", [(130, 70), (64, 0), (300, 299)])
def test_singular_matrix_reports_numeric(ctx, precision, n, col):
    rng = np.random.default_rng(n + col)
    A = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    A[:, col] = 0.0
    B = rng.standard_normal((n, 2)) + 1j * rng.standard_normal((n, 2))
    with pytest.raises(bm.BemError) as e:
        ctx.dense_solve(A, B, precision)
    assert e.value.code == bm.ERR_NUMERIC, str(e.value)
    # the context keeps working
    A[:, col] = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    X, info = ctx.dense_solve(A, B, precision)
    assert np.abs(X - np.linalg.solve(A, B)).max() < 1e-9 * np.abs(X).max()


def test_non_finite_input_rejected_on_host(ctx):
    A = np.eye(8, dtype=complex)
    B = np.ones((8, 1), dtype=complex)
    for bad in (np.nan, np.inf):
        A2 = A.copy(); A2[3, 5] = bad
        with pytest.raises(bm.BemError) as e:
            ctx.dense_solve(A2, B, 64)
        assert e.value.code == bm.ERR_ARG
        B2 = B.copy(); B2[2, 0] = bad
        with pytest.raises(bm.BemError) as e:
            ctx.dense_solve(A, B2, 32)
        assert e.value.code == bm.ERR_ARG
