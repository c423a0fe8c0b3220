"""The Bessel evaluator of the device HavelockGF (bem_device.hpp bessel01,
constants from bemuse_amd/csrc/gen_bessel.py), run on the host through the
C-ABI: J0, J1, Y0, Y1 against 30-digit mpmath values, and against the
oracle's TR1 functions (the reference's std::tr1 calls, Kernels.cpp:31-34).
Tolerance: 6e-16 absolute against mpmath (values are at most 1 in magnitude
above x = 3 for Y, everywhere for J; about 3 ulp of 1); 2e-14 against TR1,
whose own error reaches 1e-14 near x = 80 (measured against mpmath)."""
import numpy as np
import pytest

import bemuse_amd as bm
import oracle

mp = pytest.importorskip("mpmath")


def _points():
    rng = np.random.default_rng(5)
    return np.concatenate([
        np.linspace(0.0, 3.0, 121),
        rng.uniform(3.0, 30.0, 300),
        rng.uniform(30.0, 3000.0, 100),
        [3.0, np.nextafter(3.0, 4.0), 2.404825557695773, 5.520078110286311, 3.8317059702075125,
         7.0155866698156, 1.0e5, 1.5e6],
    ])


def test_bessel01_vs_mpmath():
    mp.mp.dps = 30
    x = _points()
    got = bm.bessel01(x)
    for xi, g in zip(x, got):
        ref = [mp.besselj(0, xi), mp.besselj(1, xi)]
        ref += [mp.bessely(0, xi), mp.bessely(1, xi)] if xi > 3.0 else [0.0, 0.0]
        for k in range(4):
            assert abs(g[k] - float(ref[k])) <= 6e-16, (xi, k, g[k], float(ref[k]))


def test_bessel01_vs_oracle_tr1():
    x = np.linspace(0.0, 80.0, 4001)
    got = bm.bessel01(x)
    j0, j1, y0, y1 = oracle.bessel(x)
    assert np.abs(got[:, 0] - j0).max() < 2e-14
    assert np.abs(got[:, 1] - j1).max() < 2e-14
    big = x > 3.0
    assert np.abs(got[big, 2] - y0[big]).max() < 2e-14
    assert np.abs(got[big, 3] - y1[big]).max() < 2e-14
    assert np.all(got[~big, 2:] == 0.0)


def test_bessel01_rejects_negative_argument():
    with pytest.raises(bm.BemError):
        bm.bessel01(np.array([-1.0]))
