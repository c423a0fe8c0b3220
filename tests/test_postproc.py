"""Post-processing of a solved frequency: Kochin functions and free-surface
elevation at exterior points (Hydrodynamic_Solver.cpp:205-266, 492-505,
598-642).

CPU (pinning the oracle, oracle/postproc.py), on oracle solves:
  - energy relation: B_jj = K/(4 pi) Int_0^2pi |H_j(theta)|^2 dtheta in the
    reference's normalisation (converges to 1 with the mesh: 0.976, 0.994,
    0.997 at N = 84, 336, 720);
  - far field: |eta_j(R, theta)| sqrt(R) -> (Omega/g) sqrt(K/(2 pi))
    |H_j(theta + pi)| (with E = exp(K z + i K (x cos + y sin)) as the
    reference writes it, the outgoing wave at theta carries H(theta + pi));
  - extinction: the exterior representation of a solved radiation potential
    vanishes inside the body (collocation points lie 1e-5 m inside, so the
    solved system states exactly that);
  - at the collocation points the exterior influences equal the system
    matrices S, D of the solve.
  The diffraction combinations follow the reference's formulas (Phi_S with
  the panel-area factor of the diffraction right-hand side, :445, and
  +WD Phi_S); they use the same pinned matrices but obey no physical
  identity, so beyond that they are parity unpinned.
GPU: device against the oracle, 1e-11 (Kochin) and 1e-10 (elevation)
relative to each output's max-norm.
"""
import numpy as np
import pytest

import bemuse_amd as bm
import oracle
import oracle.postproc as opp


def _cyl(na=24):
    return bm.Mesh.half_cylinder(5.0, 6.0, 6, 8, na)


@pytest.fixture(scope="module")
def solved():
    mesh = _cyl()
    nodes, elems = mesh.arrays()
    ob = oracle.Body(nodes, elems)
    om = 1.2
    K = om * om / 9.81
    r = ob.solve(om, kappa=K, betas=[0.0, 0.8], want_phi=True)
    return mesh, ob, om, K, r


def test_kochin_energy_relation(solved):
    mesh, ob, om, K, r = solved
    rad, _ = opp.kochin(ob.geometry(), om, K, 720, [], r["phiJ"], None)
    integral = (2 * np.pi / 720) * np.sum(np.abs(rad) ** 2, axis=0)
    for j in (0, 2, 4):
        ratio = r["B"][j, j] / (K / (4 * np.pi) * integral[j])
        assert abs(ratio - 1.0) < 0.015, (j, ratio)


def test_far_field_carries_kochin_function(solved):
    mesh, ob, om, K, r = solved
    nth = 8
    th = (0.5 + np.arange(nth)) * 2 * np.pi / nth
    rad, _ = opp.kochin(ob.geometry(), om, K, nth, [], r["phiJ"], None)
    R = 600.0
    pts = np.stack([R * np.cos(th), R * np.sin(th), np.zeros(nth)], axis=1)
    ext, _ = opp.fs_elevation(ob, om, K, pts, [], r["phiJ"], None)
    c = om / 9.81 * np.sqrt(K / (2 * np.pi))
    ratio = np.abs(ext[:, 2]) * np.sqrt(R) / np.abs(np.roll(rad[:, 2], nth // 2)) / c
    assert np.all(np.abs(ratio - 1.0) < 0.01), ratio


def test_far_field_direction_on_asymmetric_body():
    """Two bodies side by side: |H(theta)| is not symmetric, and the far field
    at theta follows H(theta + pi), not H(theta)."""
    m1 = bm.Mesh.half_cylinder(5.0, 6.0, 3, 4, 12)
    n1, e1 = m1.arrays()
    n2, e2 = bm.Mesh.box(8.0, 4.0, -3.0, 0.0, 4, 2, 2, True).arrays()
    n2 = n2.copy()
    n2[:, 0] += 14.0
    ob = oracle.Body(np.vstack([n1, n2]), np.vstack([e1, np.where(e2 >= 0, e2 + len(n1), -1)]))
    om = 1.2
    K = om * om / 9.81
    r = ob.solve(om, kappa=K, betas=[], want_phi=True)
    nth = 36
    th = (0.5 + np.arange(nth)) * 2 * np.pi / nth
    rad, _ = opp.kochin(ob.geometry(), om, K, nth, [], r["phiJ"], None)
    R = 600.0
    ext, _ = opp.fs_elevation(ob, om, K, np.stack([R * np.cos(th), R * np.sin(th), np.zeros(nth)], 1), [],
                              r["phiJ"], None)
    a = np.abs(ext[:, 2]) * np.sqrt(R)
    flipped = a / np.abs(np.roll(rad[:, 2], nth // 2))
    straight = a / np.abs(rad[:, 2])
    assert flipped.std() / flipped.mean() < 0.02
    assert straight.std() / straight.mean() > 0.1


def test_radiation_representation_vanishes_inside(solved):
    mesh, ob, om, K, r = solved
    inside = np.array([[0, 0, -3.0], [2.0, 1.0, -2.0], [-3.0, 0.5, -4.5], [0.5, -3.0, -1.0]])
    outside = np.array([[7.0, 0, 0.0], [0, 8.0, 0.0], [-10.0, 2.0, 0.0], [6.0, 6.0, 0.0]])
    ri, _ = opp.fs_elevation(ob, om, K, inside, [], r["phiJ"], None)
    ro, _ = opp.fs_elevation(ob, om, K, outside, [], r["phiJ"], None)
    assert np.all(np.abs(ri).max(axis=0) < 5e-3 * np.abs(ro).max(axis=0))


def test_exterior_influence_at_collocation_points_is_the_system(solved):
    mesh, ob, om, K, r = solved
    geom = ob.geometry()
    WS, WD = opp.exterior_influence(ob, K, geom["bc"][:24])
    Rk = ob.rankine()
    S, D = ob.wave(K)
    Ssys = Rk["Ssrc"][:24] + Rk["Srefl"][:24] + S[:24]
    Dsys = Rk["Dsrc"][:24] + Rk["Drefl"][:24] + D[:24]
    assert np.abs(WS - Ssys).max() <= 1e-13 * np.abs(Ssys).max()
    assert np.abs(WD - Dsys).max() <= 1e-13 * np.abs(Dsys).max()


# ------------------------------------------------------------------- GPU
def _rel(a, b):
    return np.abs(a - b).max() / np.abs(b).max()


@pytest.mark.gpu
def test_device_kochin_and_elevation_vs_oracle(ctx, solved):
    mesh, ob, om, K, r = solved
    print("\n[gpu] Kochin and free-surface elevation", flush=True)
    body = bm.Body(ctx, mesh)
    body.assemble_rankine()
    betas = [0.0, 0.8]
    g = body.solve(om, kappa=K, betas=betas, mode=bm.MODE_FP64, want_phi=True)
    kr, kd = body.kochin(om, g["phiJ"], g["phiS"], betas=betas, nkoch=72, kappa=K)
    ork, okd = opp.kochin(ob.geometry(), om, K, 72, betas, g["phiJ"], g["phiS"])
    assert _rel(kr, ork) < 1e-11 and _rel(kd, okd) < 1e-11
    rng = np.random.default_rng(9)
    ang = rng.uniform(0, 2 * np.pi, 40)
    rad = rng.uniform(6.0, 200.0, 40)
    pts = np.stack([rad * np.cos(ang), rad * np.sin(ang), np.where(np.arange(40) % 3 == 0, -1.0, 0.0)], 1)
    er, ed = body.fs_elevation(om, pts, g["phiJ"], g["phiS"], betas=betas, kappa=K)
    oer, oed = opp.fs_elevation(ob, om, K, pts, betas, g["phiJ"], g["phiS"])
    assert _rel(er, oer) < 1e-10 and _rel(ed, oed) < 1e-10
    body.close()
