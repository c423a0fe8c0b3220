"""bem_dense_solve_noalloc (the dense solve that allocates no device memory,
its buffers in one fixed device array per GPU) against bem_dense_solve, the
entry point it mirrors, on two calls in a row with the second smaller, in
both precisions; and its limit.

Tolerance: the two entry points run the same kernels on the same inputs, so
the solutions agree to 1e-13 relative (max-norm); both are also checked
against LAPACK (numpy.linalg.solve): 1e-12 (FP64), 1e-11 (FP32 LU + FP64
refinement).
"""
import numpy as np
import pytest

import bemuse_amd as bm

pytestmark = pytest.mark.gpu


def relerr(a, b):
    return np.abs(a - b).max() / np.abs(b).max()


def system(n, nrhs, seed, dominant):
    rng = np.random.default_rng(seed)
    A = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    if dominant:
        A = A / np.sqrt(n) * 0.3 + np.diag(2.0 + rng.uniform(0, 1, n))
    B = rng.standard_normal((n, nrhs)) + 1j * rng.standard_normal((n, nrhs))
    return A, B


@pytest.mark.parametrize("precision", [64, 32])
@pytest.mark.parametrize("dominant", [True, False])
def test_noalloc_matches_dense_solve_twice(ctx, precision, dominant):
    # the second call is smaller: it must not see the first call's contents
    for k, (n, nrhs) in enumerate(((bm.NOALLOC_MAX_N - 12, 11), (130, 3))):
        A, B = system(n, nrhs, 10 * k + precision + int(dominant), dominant)
        print(f"\n[gpu] noalloc n={n} nrhs={nrhs} precision={precision} dominant={dominant}", flush=True)
        Xn, info_n = ctx.dense_solve(A, B, precision, noalloc=True)
        Xd, info_d = ctx.dense_solve(A, B, precision)
        assert relerr(Xn, Xd) < 1e-13, (n, relerr(Xn, Xd))
        Xr = np.linalg.solve(A, B)
        tol = 1e-12 if precision == 64 else 1e-11
        assert relerr(Xn, Xr) < (tol if dominant else 100 * tol), (n, relerr(Xn, Xr))
        assert info_n["pivot_fallback"] == info_d["pivot_fallback"]
        assert (info_n["pivot_fallback"] > 0) == (not dominant)
        assert info_n["fell_back_fp64"] == 0 and info_n["refine_iters"] == info_d["refine_iters"]


def test_noalloc_limit_and_errors(ctx):
    n = bm.NOALLOC_MAX_N + 1
    A, B = system(n, 1, 3, True)
    with pytest.raises(bm.BemError) as e:
        ctx.dense_solve(A, B, 64, noalloc=True)
    assert e.value.code == bm.ERR_ARG
    A, B = system(200, 2, 4, False)
    A[:, 17] = 0.0
    with pytest.raises(bm.BemError) as e:
        ctx.dense_solve(A, B, 32, noalloc=True)
    assert e.value.code == bm.ERR_NUMERIC
    # and it recovers on the next call
    A[:, 17] = 1.0 + 1.0j
    X, _ = ctx.dense_solve(A, B, 32, noalloc=True)
    assert relerr(X, np.linalg.solve(A, B)) < 1e-9
