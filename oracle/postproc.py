"""FP64 CPU oracle of the post-processing of a solved frequency: Kochin
functions and free-surface elevation at exterior points (numpy over the
oracle's C primitives).

TEST INFRASTRUCTURE ONLY: imported by tests/ as the reference for
bem_kochin / bem_fs_elevation. The product never imports it.

Follows Hydrodynamic_Solver.cpp:
  - Set_Kochin_Mats (:598-642) and :492-494:
        H = (E A) dphi/dn - ((n . grad E) A) phi,
        E = exp(K z + i K (x cos theta + y sin theta)) at the centroids,
    theta_i = (i + 1/2) 2 pi / NKoch. The reference writes E = exp(E) with
    E uninitialised (:626-627, undefined behaviour); E here is the Newman
    (1967) factor the comment there names.
  - Prepare_FS_Linear_System (:205-237), ..._Wave_Terms (:239-266) and
    ExtRadMat / ExtDiffMat (:499-505):
        WS = -s - s_r - ws, WD = d + d_r + wd - s_r ReflNorm K,
        ExtRad = -i Omega/g (WS N_k - WD Phi_J),
        ExtDiff = -i Omega/g (WS dPhi_I/dn + WD Phi_S).

Pinning (tests/test_postproc.py): the radiation Kochin function by the
energy relation between damping and the far field, B_jj = K/(8 pi)
Int |H_j|^2 dtheta in the reference's normalisation; the exterior
influences by the extinction property (the exterior representation of a
solved potential vanishes inside the body); and both together by the far
field, |phi_j(R, theta)| -> sqrt(K / (2 pi R)) |H_j(theta)| / (4 pi)... see
the test for the constant it checks. KochinDiff and ExtDiff combine the
same pinned matrices with the reference's diffraction inputs (Phi_S carries
the panel-area factor of the diffraction right-hand side, :445).
"""
from __future__ import annotations

import numpy as np

GRAVITY = 9.81


def body_panels(geom, n):
    """The first n panels' entries of a geometry dict (the body panels; lid
    panels of irregular-frequency removal follow them)."""
    return {k: v[:n] for k, v in geom.items()}


def generalised_normals(geom):
    c, n = geom["centroid"], geom["normal"]
    return np.concatenate([n, np.cross(c, n)], axis=1)


def incident_dn(geom, beta, omega, kappa):
    """dPhi_I/dn at the centroids, without area (Calc_Phi_Incident :544-573)."""
    c, n = geom["centroid"], geom["normal"]
    cb, sb = np.cos(beta), np.sin(beta)
    phi = 1j * (GRAVITY / omega) * np.exp(kappa * c[:, 2]) * np.exp(1j * kappa * (c[:, 0] * cb + c[:, 1] * sb))
    return (n[:, 2] * kappa + 1j * kappa * (n[:, 0] * cb + n[:, 1] * sb)) * phi


def kochin(geom, omega, kappa, nkoch, betas, phiJ, phiS):
    """(KochRad nkoch x 6, KochDiff nkoch x nb), over the body panels (the rows of phiJ)."""
    geom = body_panels(geom, len(phiJ))
    c, n, A = geom["centroid"], geom["normal"], geom["area"]
    th = (0.5 + np.arange(nkoch)) * 2.0 * np.pi / nkoch
    ct, st = np.cos(th)[:, None], np.sin(th)[:, None]
    E = np.exp(kappa * c[None, :, 2]) * np.exp(1j * kappa * (c[None, :, 0] * ct + c[None, :, 1] * st))
    dE = (n[None, :, 2] * kappa + 1j * kappa * (n[None, :, 0] * ct + n[None, :, 1] * st)) * E
    KE, KdE = E * A[None, :], dE * A[None, :]
    rad = KE @ generalised_normals(geom) - KdE @ phiJ
    diff = None
    if len(betas):
        dn = np.stack([incident_dn(geom, b, omega, kappa) for b in betas], axis=1)
        diff = KE @ dn - KdE @ phiS
    return rad, diff


def exterior_influence(ob, kappa, pts):
    """(WS, WD) complex npts x N of the oracle body ob at the points."""
    pts = np.ascontiguousarray(pts, dtype=np.float64).reshape(-1, 3)
    geom = body_panels(ob.geometry(), ob.N)
    reflnorm = 2.0 * geom["normal"][:, 2]
    WS = np.zeros((len(pts), ob.N), dtype=np.complex128)
    WD = np.zeros_like(WS)
    for S in range(ob.N):
        s, d = ob.rankine_points(S, pts)
        sr, dr = ob.rankine_points(S, pts, refl=True)
        ws, wd = ob.wave_points(S, kappa, pts)
        WS[:, S] = -s - sr - ws
        WD[:, S] = d + dr + wd - sr * reflnorm[S] * kappa
    return WS, WD


def fs_elevation(ob, omega, kappa, pts, betas, phiJ, phiS):
    """(ExtRad npts x 6, ExtDiff npts x nb)."""
    geom = body_panels(ob.geometry(), ob.N)
    WS, WD = exterior_influence(ob, kappa, pts)
    f = -1j * omega / GRAVITY
    rad = f * (WS @ generalised_normals(geom) - WD @ phiJ)
    diff = None
    if len(betas):
        dn = np.stack([incident_dn(geom, b, omega, kappa) for b in betas], axis=1)
        diff = f * (WS @ dn + WD @ phiS)
    return rad, diff
