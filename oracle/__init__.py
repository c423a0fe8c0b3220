"""FP64 CPU oracle of the BEMUse per-frequency hot path (ctypes over liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, to check and time results against. The product
(bemuse_amd) never imports it.

Each function restates the reference arithmetic; see bem_oracle.cpp for the
file:line each one follows. Parity status of every function (which checks pin
it) is listed in tests/test_oracle.py.
"""
from __future__ import annotations

import ctypes
import os
import subprocess
from ctypes import POINTER, c_double, c_int, c_void_p

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liboracle.so")
_PD = POINTER(c_double)
_PI = POINTER(c_int)
_lib = None


def build():
    subprocess.run(["make", "-C", _HERE, "-s"], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        L.orc_body_create.restype = c_void_p
        L.orc_body_create.argtypes = [c_int, _PD, c_int, _PI]
        L.orc_body_destroy.argtypes = [c_void_p]
        L.orc_body_set_npa.argtypes = [c_void_p, c_int]
        L.orc_body_n.argtypes = [c_void_p]
        L.orc_body_geometry.argtypes = [c_void_p, _PD, _PD, _PD, _PD, _PI, _PD, _PD, _PD, _PD]
        L.orc_body_qnodes.argtypes = [c_void_p, _PI, _PD, _PD, _PD]
        L.orc_rankine_points.argtypes = [c_void_p, c_int, c_int, c_int, c_int, _PD, _PD, _PD]
        L.orc_wave_points.argtypes = [c_void_p, c_int, c_double, c_int, _PD, _PD, _PD]
        L.orc_rankine.argtypes = [c_void_p, _PD, _PD, _PD, _PD, _PD]
        L.orc_wave.argtypes = [c_void_p, c_double, _PD, _PD]
        L.orc_lu_factor.argtypes = [c_int, _PD, _PI]
        L.orc_lu_solve.argtypes = [c_int, _PD, _PI, c_int, _PD]
        L.orc_solve.argtypes = [c_void_p, c_double, c_double, c_int, c_int, c_int, _PD, _PD, _PD, _PD, _PD, _PD,
                                _PD, _PD, _PD]
        L.orc_solve.restype = c_int
        L.orc_bessel.argtypes = [c_int, _PD, _PD, _PD, _PD, _PD]
        L.orc_struve.argtypes = [c_int, _PD, _PD, _PD]
        L.orc_havelock.argtypes = [c_int, _PD, _PD, _PD, _PD]
        L.orc_poly_coeffs.argtypes = [_PD]
        L.orc_struve_coeffs.argtypes = [_PD]
        L.orc_num_threads.restype = c_int
        L.orc_set_threads.argtypes = [c_int]
        L.orc_set_diffraction_area.argtypes = [c_int]
        _lib = L
    return _lib


def _dp(a):
    return a.ctypes.data_as(_PD) if a is not None else None


def set_threads(n):
    lib().orc_set_threads(int(n))


def num_threads():
    return lib().orc_num_threads()


def bessel(x):
    x = np.ascontiguousarray(x, dtype=np.float64)
    out = [np.zeros_like(x) for _ in range(4)]
    lib().orc_bessel(len(x), _dp(x), *[_dp(o) for o in out])
    return out


def struve(x):
    x = np.ascontiguousarray(x, dtype=np.float64)
    h0, h1 = np.zeros_like(x), np.zeros_like(x)
    lib().orc_struve(len(x), _dp(x), _dp(h0), _dp(h1))
    return h0, h1


def havelock(X, Y, V):
    """HavelockGF (Kernels.cpp:342-412): returns (n, 4) complex GF, DGFX, DGFY, DGFZ."""
    X = np.ascontiguousarray(X, dtype=np.float64)
    Y = np.ascontiguousarray(Y, dtype=np.float64)
    V = np.ascontiguousarray(V, dtype=np.float64)
    out = np.zeros((len(X), 8))
    lib().orc_havelock(len(X), _dp(X), _dp(Y), _dp(V), _dp(out))
    return out[:, 0::2] + 1j * out[:, 1::2]


def poly_coeffs():
    """The GF polynomial tables cA, cB, cC, cD, cdA, cdB, cdC (64 values) then EUL."""
    out = np.zeros(65)
    lib().orc_poly_coeffs(_dp(out))
    return out


def struve_coeffs():
    """The Struve approximation literals (26): H0 P0..P5, a0..a3, b1..b3, then H1 P1..P6, a0..a3, b1..b3."""
    out = np.zeros(26)
    lib().orc_struve_coeffs(_dp(out))
    return out


def lu_factor(A):
    A = np.asfortranarray(A, dtype=np.complex128).copy(order="F")
    n = A.shape[0]
    piv = np.zeros(n, dtype=np.int32)
    lib().orc_lu_factor(n, A.ctypes.data_as(_PD), piv.ctypes.data_as(_PI))
    return A, piv


def lu_solve(LU, piv, B):
    B = np.asfortranarray(B, dtype=np.complex128).copy(order="F")
    if B.ndim == 1:
        B = B[:, None].copy(order="F")
    n = LU.shape[0]
    lib().orc_lu_solve(n, LU.ctypes.data_as(_PD), piv.ctypes.data_as(_PI), B.shape[1], B.ctypes.data_as(_PD))
    return B


class Body:
    """Oracle body built from a node/element mesh (same arrays as the GPU path)."""

    def __init__(self, nodes, elems, n_lid=0):
        """n_lid: the last n_lid elements are interior free-surface lid panels
        (irregular-frequency removal); potentials and A, B, F use the others."""
        nodes = np.ascontiguousarray(nodes, dtype=np.float64).reshape(-1, 3)
        elems = np.ascontiguousarray(elems, dtype=np.int32).reshape(-1, 4)
        self._keep = (nodes, elems)
        self._h = c_void_p(lib().orc_body_create(len(nodes), _dp(nodes), len(elems), elems.ctypes.data_as(_PI)))
        self.N_system = len(elems)
        self.N = len(elems) - n_lid
        if n_lid:
            lib().orc_body_set_npa(self._h, self.N)

    def __del__(self):
        if getattr(self, "_h", None) and _lib is not None:
            _lib.orc_body_destroy(self._h)
            self._h = None

    def geometry(self):
        """Body panel centroid, normal, area, BC point and frame; the q*
        fields (nq, qpos, qnrm, qw) are the MIRRORED panel's quadrature nodes
        (the ones the wave term integrates over)."""
        N = self.N_system
        cen = np.zeros((N, 3)); nrm = np.zeros((N, 3)); area = np.zeros(N); bc = np.zeros((N, 3))
        nq = np.zeros(N, dtype=np.int32); qpos = np.zeros((N, 4, 3)); qnrm = np.zeros((N, 4, 3)); qw = np.zeros((N, 4))
        frame = np.zeros((N, 3, 3))
        lib().orc_body_geometry(self._h, _dp(cen), _dp(nrm), _dp(area), _dp(bc), nq.ctypes.data_as(_PI), _dp(qpos),
                                _dp(qnrm), _dp(qw), _dp(frame))
        return dict(centroid=cen, normal=nrm, area=area, bc=bc, nq=nq, qpos=qpos, qnrm=qnrm, qw=qw, frame=frame)

    def body_qnodes(self):
        """Quadrature nodes of the body panels."""
        N = self.N_system
        nq = np.zeros(N, dtype=np.int32); qpos = np.zeros((N, 4, 3)); qnrm = np.zeros((N, 4, 3)); qw = np.zeros((N, 4))
        lib().orc_body_qnodes(self._h, nq.ctypes.data_as(_PI), _dp(qpos), _dp(qnrm), _dp(qw))
        return dict(nq=nq, qpos=qpos, qnrm=qnrm, qw=qw)

    def rankine_points(self, s, pts, refl=False, quadrature=False):
        pts = np.ascontiguousarray(pts, dtype=np.float64).reshape(-1, 3)
        S = np.zeros(len(pts)); D = np.zeros(len(pts))
        lib().orc_rankine_points(self._h, s, int(refl), int(quadrature), len(pts), _dp(pts), _dp(S), _dp(D))
        return S, D

    def wave_points(self, s, K, pts):
        pts = np.ascontiguousarray(pts, dtype=np.float64).reshape(-1, 3)
        S = np.zeros(2 * len(pts)); D = np.zeros(2 * len(pts))
        lib().orc_wave_points(self._h, s, K, len(pts), _dp(pts), _dp(S), _dp(D))
        return S[0::2] + 1j * S[1::2], D[0::2] + 1j * D[1::2]

    def rankine(self):
        N = self.N_system
        out = [np.zeros((N, N), order="F") for _ in range(4)]
        rn = np.zeros(N)
        lib().orc_rankine(self._h, *[_dp(o) for o in out], _dp(rn))
        return dict(Ssrc=out[0], Srefl=out[1], Dsrc=out[2], Drefl=out[3], reflnorm=rn)

    def wave(self, K):
        N = self.N_system
        S = np.zeros((N, N), dtype=np.complex128, order="F")
        D = np.zeros((N, N), dtype=np.complex128, order="F")
        lib().orc_wave(self._h, K, S.ctypes.data_as(_PD), D.ctypes.data_as(_PD))
        return S, D

    def solve(self, omega, kappa=None, betas=(), compose=0, rad_only=False, want_phi=False, diff_area=True):
        """diff_area: the reference's area factor on the diffraction right-hand
        side (Hydrodynamic_Solver.cpp:445); False solves the physical
        scattering problem D Phi_S = S (-dPhi_I/dn) instead (a check only)."""
        if kappa is None:
            kappa = omega * omega / 9.81
        betas = np.ascontiguousarray(betas, dtype=np.float64)
        nb = len(betas)
        A = np.zeros(36); B = np.zeros(36)
        F = np.zeros((6, max(nb, 1)), dtype=np.complex128)
        FK = np.zeros_like(F); SC = np.zeros_like(F)
        phiJ = np.zeros((self.N, 6), dtype=np.complex128, order="F")
        phiS = np.zeros((self.N, max(nb, 1)), dtype=np.complex128, order="F")
        t = np.zeros(3)
        lib().orc_set_diffraction_area(int(bool(diff_area)))
        try:
            lib().orc_solve(self._h, omega, kappa, compose, int(rad_only), nb, _dp(betas) if nb else None, _dp(A),
                            _dp(B), F.ctypes.data_as(_PD), FK.ctypes.data_as(_PD), SC.ctypes.data_as(_PD),
                            phiJ.ctypes.data_as(_PD), phiS.ctypes.data_as(_PD), _dp(t))
        finally:
            lib().orc_set_diffraction_area(1)
        res = dict(A=A.reshape(6, 6), B=B.reshape(6, 6), F=F[:, :nb], FK=FK[:, :nb], SC=SC[:, :nb],
                   timings=dict(assemble=t[0], factor=t[1], rest=t[2]))
        if want_phi:
            res["phiJ"] = phiJ
            res["phiS"] = phiS[:, :nb]
        return res
