"""Irregular-frequency removal (IFR): interior free-surface lid panels join
the system (Create_Panels, Hydrodynamic_Solver.cpp:45-55), while A, B, F
and the potentials use the body panels (:145-165, 450-458).

CPU (pinning the oracle's lid handling): a truncated cylinder (radius 5,
draft 8) has its first irregular frequency at k a = j01 coth(j01 d / a),
omega = 2.173 rad/s. Across it, without the lid the heave damping turns
negative (it cannot: it is the radiated power) and the heave exciting force
spikes; with the lid the damping stays positive and the force smooth.
GPU: the device path with the lid against the oracle with the lid, in both
modes (1e-8 FP64, 1e-6 mixed), at and away from the irregular frequency.
"""
import numpy as np
import pytest

import bemuse_amd as bm
import oracle

OMEGAS = np.linspace(2.05, 2.35, 9)


def _meshes():
    body = bm.Mesh.half_cylinder(5.0, 8.0, 4, 8, 16)
    lid = bm.Mesh.lid_disc(5.0, 3, 16)
    return body, lid


def _oracle_bodies():
    body, lid = _meshes()
    n1, e1 = body.arrays()
    n2, e2 = lid.arrays()
    nodes = np.vstack([n1, n2])
    elems = np.vstack([e1, np.where(e2 >= 0, e2 + len(n1), -1)])
    return oracle.Body(n1, e1), oracle.Body(nodes, elems, n_lid=len(e2))


def _roughness(f):
    f = np.asarray(f)
    return np.abs(f[2:] - 2 * f[1:-1] + f[:-2]).max() / np.abs(f).max()


def test_lid_geometry():
    lid = bm.Mesh.lid_disc(5.0, 3, 16)
    g = lid.geometry()
    assert np.allclose(g["gen"][:, 2], -1.0)                  # normals into the body
    assert np.allclose(g["recv"][:, 2], -1e-5)                # collocation points just inside
    # polygonal disc area: 16-gon of radius 5
    assert abs(g["gen"][:, 6].sum() - 0.5 * 16 * 25 * np.sin(2 * np.pi / 16)) < 1e-10
    r = bm.Mesh.lid_rect(12.0, 6.0, 6, 3).geometry()
    assert np.allclose(r["gen"][:, 2], -1.0) and abs(r["gen"][:, 6].sum() - 72.0) < 1e-10


def test_lid_removes_irregular_frequency():
    plain, ifr = _oracle_bodies()
    b0, b1, f0, f1 = [], [], [], []
    for om in OMEGAS:
        K = om * om / 9.81
        r0 = plain.solve(om, kappa=K, betas=[0.0])
        r1 = ifr.solve(om, kappa=K, betas=[0.0])
        b0.append(r0["B"][2, 2]); b1.append(r1["B"][2, 2])
        f0.append(abs(r0["F"][2, 0])); f1.append(abs(r1["F"][2, 0]))
    assert min(b0) < 0.0                  # the irregular frequency is inside the scan
    assert min(b1) > 0.0
    assert _roughness(f0) > 0.5
    assert _roughness(f1) < 0.05


def test_lid_off_irregular_frequency_changes_little():
    plain, ifr = _oracle_bodies()
    om = 1.0
    K = om * om / 9.81
    r0 = plain.solve(om, kappa=K, betas=[0.0])
    r1 = ifr.solve(om, kappa=K, betas=[0.0])
    for k in ("A", "B"):
        assert np.abs(r1[k] - r0[k]).max() < 0.05 * np.abs(r0[k]).max()


@pytest.mark.gpu
@pytest.mark.parametrize("mode,tol", [(bm.MODE_FP64, 1e-8), (bm.MODE_MIXED, 1e-6)])
def test_device_with_lid_vs_oracle(ctx, mode, tol):
    print(f"\n[gpu] IFR lid, mode {mode}", flush=True)
    body_mesh, lid = _meshes()
    body = bm.Body(ctx, body_mesh, lid=lid)
    assert body.N == 192 and body.N_system == 240
    body.assemble_rankine()
    _, ifr = _oracle_bodies()
    for om in (1.0, 2.2):
        K = om * om / 9.81
        g = body.solve(om, kappa=K, betas=[0.0, 0.5], mode=mode, want_phi=True)
        r = ifr.solve(om, kappa=K, betas=[0.0, 0.5], want_phi=True)
        tight = 1e-11 if mode == bm.MODE_FP64 else 1e-10
        for k in ("A", "B", "F"):
            den = np.abs(r[k]).max()
            assert np.abs(g[k] - r[k]).max() <= tol * den, (om, k)
            assert np.abs(g[k] - r[k]).max() <= tight * den, (om, k, "tight")
        assert g["phiJ"].shape == (192, 6) and g["phiS"].shape == (192, 2)
        assert np.abs(g["phiJ"] - r["phiJ"]).max() <= 10 * tight * np.abs(r["phiJ"]).max()
        assert np.abs(g["phiS"] - r["phiS"]).max() <= 10 * tight * np.abs(r["phiS"]).max()
    body.close()
