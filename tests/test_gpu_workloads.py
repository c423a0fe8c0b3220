"""GPU parity on the benchmark workloads themselves (the frequency loop of
BEMUser_Console.cpp:252-257, each iteration Hydrodynamic_Solver.cpp:400-477):

  - the 1,024-panel floating cylinder (r = 10 m, draft 20 m; 32 azimuthal x
    (24 side + 8 bottom rings), 992 quads + a 32-triangle centre fan), all 10
    frequencies of linspace(0.1, 2.0, 10) rad/s against the live oracle, with
    the potentials at 1.578 rad/s, next to its first irregular frequency
    (k a = j01 coth(j01 d / a), omega = 1.536 rad/s);
  - the 6018-panel three-column semi-sub of the headline measurement, against
    the committed oracle fixture (tests/golden/semisub6018_abf.npz: three of
    its 40 frequencies, 5 headings, Phi_J and Phi_S at one), and mixed
    against FP64 on the device over all 40 frequencies;
  - the batch of 64 seeded uneven bodies x 16 frequencies through
    bem_solve_batch, against per-body sweeps and, for sampled bodies, the
    oracle.

Tolerances as in test_gpu_parity.py: the bars (1e-8 FP64, 1e-6 mixed) and the
tight levels (1e-11 FP64, 1e-10 mixed) are asserted separately.
"""
import os

import numpy as np
import pytest

import bemuse_amd as bm
import oracle

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
BAR = {bm.MODE_FP64: 1e-8, bm.MODE_MIXED: 1e-6}
TIGHT = {bm.MODE_FP64: 1e-11, bm.MODE_MIXED: 1e-10}
MODES = [bm.MODE_FP64, bm.MODE_MIXED]


def relerr(a, b):
    den = np.abs(b).max()
    return np.abs(a - b).max() / (den if den > 0 else 1.0)


def assert_close(g, r, mode, keys=("A", "B", "F"), what=""):
    for key in keys:
        e = relerr(g[key], r[key])
        assert e < BAR[mode], (what, key, e, "bar")
        assert e < TIGHT[mode], (what, key, e, "tight")


def at(sw, i, keys=("A", "B", "F")):
    return {k: sw[k][i] for k in keys}


# ---------------------------------------------------------------------------
# 1,024-panel cylinder, 10 frequencies, heading 0
# ---------------------------------------------------------------------------
@pytest.fixture(scope="module")
def cyl1024(ctx):
    mesh = bm.Mesh.half_cylinder(10.0, 20.0, 8, 24, 32)
    nodes, elems = mesh.arrays()
    body = bm.Body(ctx, mesh)
    print(f"\n[gpu] cylinder r=10 d=20 N={body.N}", flush=True)
    body.assemble_rankine()
    ob = oracle.Body(nodes, elems)
    omegas = np.linspace(0.1, 2.0, 10)
    refs = [ob.solve(w, betas=[0.0], want_phi=(k == 7)) for k, w in enumerate(omegas)]
    yield body, ob, omegas, refs
    body.close()


def test_cylinder1024_topology(cyl1024):
    body, ob, omegas, refs = cyl1024
    assert body.N == 1024
    _, elems = bm.Mesh.half_cylinder(10.0, 20.0, 8, 24, 32).arrays()
    assert int((elems[:, 3] < 0).sum()) == 32            # the centre fan
    assert abs(omegas[7] - 1.5778) < 1e-4                 # next to the irregular frequency 1.536 rad/s


@pytest.mark.parametrize("mode", MODES)
def test_cylinder1024_all_frequencies(cyl1024, mode):
    body, ob, omegas, refs = cyl1024
    sw = body.solve_sweep(omegas, betas=[0.0], mode=mode, lanes=4)
    for i, w in enumerate(omegas):
        assert_close(at(sw, i, ("A", "B", "F", "FK", "SC")), refs[i], mode, ("A", "B", "F", "FK", "SC"),
                     f"omega={w:.4f}")
        if mode == bm.MODE_MIXED:
            assert sw["info"][i]["fell_back_fp64"] == 0
            assert sw["info"][i]["residual"] < 1e-13
    # potentials next to the irregular frequency
    g = body.solve(omegas[7], betas=[0.0], mode=mode, want_phi=True)
    assert relerr(g["phiJ"], refs[7]["phiJ"]) < 10 * TIGHT[mode]
    assert relerr(g["phiS"], refs[7]["phiS"]) < 10 * TIGHT[mode]


@pytest.mark.parametrize("mode", MODES)
def test_cylinder1024_lane_count_does_not_change_results(cyl1024, mode):
    """One lane factors with the latency variant of the panel kernels (two
    register stages of operand loads in FP32), several lanes with the
    throughput variant; both add the same products in the same order, so the
    results agree bit for bit."""
    body, ob, omegas, refs = cyl1024
    w = omegas[2:6]
    one = body.solve_sweep(w, betas=[0.0], mode=mode, lanes=1)
    two = body.solve_sweep(w, betas=[0.0], mode=mode, lanes=2)
    for key in ("A", "B", "F", "FK", "SC"):
        np.testing.assert_array_equal(one[key], two[key], err_msg=key)


@pytest.mark.parametrize("mode", MODES)
def test_cylinder1024_speculative_equals_checked(cyl1024, ctx, mode):
    """The solve lanes' speculative factorisation (no pivot-check launch)
    against the checked one (a pivot check after every panel): the same
    results bit for bit, alone on the device (one lane) and sharing it
    (three)."""
    body, ob, omegas, refs = cyl1024
    w = omegas[1:7]
    betas = [0.0, 0.9]
    try:
        for lanes in (1, 3):
            ctx.set_lu_speculative(True)
            spec = body.solve_sweep(w, betas=betas, mode=mode, lanes=lanes)
            ctx.set_lu_speculative(False)
            chk = body.solve_sweep(w, betas=betas, mode=mode, lanes=lanes)
            for key in ("A", "B", "F", "FK", "SC"):
                np.testing.assert_array_equal(spec[key], chk[key], err_msg=f"{key} lanes {lanes}")
            assert all(i["pivot_fallback"] == 0 for i in spec["info"])
    finally:
        ctx.set_lu_speculative(True)


def test_cylinder1024_wave_assembly_work_groups(cyl1024, ctx):
    """The wave assembly with one work-group per tile and with 1, 2 or 3
    work-groups per CU taking the 1024 tiles (64 receivers x 16 sources) in
    turn: the same system bit for bit, in FP64 and mixed mode (which also
    writes the FP32 copy of D), and the same A, B, F through a sweep. Bad
    settings are rejected."""
    body, ob, omegas, refs = cyl1024
    kappa = 1.3 ** 2 / 9.81
    try:
        for mode in MODES:
            ctx.set_wave_wg_per_cu(0)
            S0, D0 = body.get_system(kappa, mode=mode)
            for n in (1, 2, 3, -1):
                ctx.set_wave_wg_per_cu(n)
                S, D = body.get_system(kappa, mode=mode)
                np.testing.assert_array_equal(S, S0, err_msg=f"S n={n} mode {mode}")
                np.testing.assert_array_equal(D, D0, err_msg=f"D n={n} mode {mode}")
        w = omegas[2:6]
        ctx.set_wave_wg_per_cu(0)
        ref = body.solve_sweep(w, betas=[0.0], mode=bm.MODE_MIXED, lanes=2)
        ctx.set_wave_wg_per_cu(2)
        got = body.solve_sweep(w, betas=[0.0], mode=bm.MODE_MIXED, lanes=2)
        for key in ("A", "B", "F"):
            np.testing.assert_array_equal(got[key], ref[key], err_msg=key)
        for bad in (-2, 65):
            with pytest.raises(bm.BemError):
                ctx.set_wave_wg_per_cu(bad)
    finally:
        ctx.set_wave_wg_per_cu(-1)


def test_cylinder1024_refinement_steps_enqueued_later(cyl1024, ctx):
    """The refinement steps a lane leaves for after its results (all of them
    with 0 enqueued up front, one with 1) give the results of the default
    (2, enough for these systems) bit for bit, with the same step counts."""
    body, ob, omegas, refs = cyl1024
    w = omegas[[1, 5, 8]]
    ref = body.solve_sweep(w, betas=[0.0, 0.4], mode=bm.MODE_MIXED, lanes=2)
    try:
        for eager in (0, 1):
            ctx.set_refine_eager(eager)
            got = body.solve_sweep(w, betas=[0.0, 0.4], mode=bm.MODE_MIXED, lanes=2)
            for key in ("A", "B", "F", "FK", "SC"):
                np.testing.assert_array_equal(got[key], ref[key], err_msg=f"{key} eager {eager}")
            assert [i["refine_iters"] for i in got["info"]] == [i["refine_iters"] for i in ref["info"]]
            assert all(i["fell_back_fp64"] == 0 for i in got["info"])
    finally:
        ctx.set_refine_eager(2)
    with pytest.raises(bm.BemError):
        ctx.set_refine_eager(5)


@pytest.mark.parametrize("mode", MODES)
@pytest.mark.parametrize("tau", [0.0, 0.05])
def test_cylinder1024_pivot_threshold_reruns(cyl1024, ctx, mode, tau):
    """A pivot-check bound below 1 sends panels to the partial-pivoting
    fallback: a speculative factorisation that meets such a panel is run again
    with the check (the frequency's results then equal those of the checked
    path bit for bit), and the results stay at the tight level against the
    default bound (different pivots, different rounding)."""
    body, ob, omegas, refs = cyl1024
    w = omegas[[0, 4, 9]]
    betas = [0.0]
    base = body.solve_sweep(w, betas=betas, mode=mode, lanes=2)
    try:
        ctx.set_pivot_threshold(tau)
        spec = body.solve_sweep(w, betas=betas, mode=mode, lanes=2)
        ctx.set_lu_speculative(False)
        chk = body.solve_sweep(w, betas=betas, mode=mode, lanes=2)
    finally:
        ctx.set_pivot_threshold(1.0)
        ctx.set_lu_speculative(True)
    assert any(i["pivot_fallback"] > 0 for i in spec["info"]), spec["info"]
    if tau == 0.0:   # every panel of a 1024-panel system (16 panels) takes the fallback
        assert all(i["pivot_fallback"] == 16 for i in spec["info"]), spec["info"]
    for key in ("A", "B", "F"):
        np.testing.assert_array_equal(spec[key], chk[key], err_msg=key)
        for k in range(len(w)):
            e = relerr(spec[key][k], base[key][k])
            assert e < TIGHT[mode], (key, k, e)
    if mode == bm.MODE_MIXED:
        assert all(i["fell_back_fp64"] == 0 and i["residual"] < 1e-13 for i in spec["info"])
    for bad in (-0.1, 1.5, float("nan")):
        with pytest.raises(bm.BemError):
            ctx.set_pivot_threshold(bad)


# ---------------------------------------------------------------------------
# 6018-panel semi-sub, 40 frequencies, 5 headings
# ---------------------------------------------------------------------------
@pytest.fixture(scope="module")
def semisub(ctx):
    fx = np.load(os.path.join(HERE, "golden", "semisub6018_abf.npz"))
    mesh = bm.Mesh.semisub3()
    nodes, elems = mesh.arrays()
    # the library's generator still makes the fixture's mesh
    np.testing.assert_array_equal(nodes, fx["nodes"])
    np.testing.assert_array_equal(elems, fx["elems"])
    body = bm.Body(ctx, mesh)
    print(f"\n[gpu] semi-sub N={body.N}", flush=True)
    assert body.N == 6018
    body.assemble_rankine()
    yield body, fx
    body.close()


@pytest.mark.parametrize("mode", MODES)
def test_semisub6018_vs_oracle_fixture(semisub, mode):
    body, fx = semisub
    betas = fx["betas"]
    for i, w in enumerate(fx["omegas"]):
        g = body.solve(w, betas=betas, mode=mode, want_phi=bool(w == fx["phi_omega"]))
        assert_close(g, {k: fx[k][i] for k in ("A", "B", "F", "FK", "SC")}, mode, ("A", "B", "F", "FK", "SC"),
                     f"omega={w:.4f}")
        if w == fx["phi_omega"]:
            assert relerr(g["phiJ"], fx["phiJ"]) < 10 * TIGHT[mode]
            assert relerr(g["phiS"], fx["phiS"]) < 10 * TIGHT[mode]
        if mode == bm.MODE_MIXED:
            assert g["info"]["fell_back_fp64"] == 0 and g["info"]["residual"] < 1e-13
    assert any(w == fx["phi_omega"] for w in fx["omegas"])


def test_semisub6018_mixed_vs_fp64_all_frequencies(semisub):
    """All 40 frequencies x 5 headings: the mixed path (FP32-MFMA LU + FP64
    refinement) against the FP64 device path, A, B and F each."""
    body, fx = semisub
    omegas = np.linspace(0.05, 2.5, 40)
    betas = fx["betas"]
    s64 = body.solve_sweep(omegas, betas=betas, mode=bm.MODE_FP64, lanes=8)
    smx = body.solve_sweep(omegas, betas=betas, mode=bm.MODE_MIXED, lanes=8)
    worst = {k: 0.0 for k in ("A", "B", "F")}
    for i, w in enumerate(omegas):
        for k in worst:
            worst[k] = max(worst[k], relerr(smx[k][i], s64[k][i]))
        info = smx["info"][i]
        assert info["fell_back_fp64"] == 0 and info["residual"] < 1e-13 and 1 <= info["refine_iters"] <= 4, (w, info)
    print(f"\n[gpu] semi-sub mixed vs FP64 over 40 frequencies: {worst}", flush=True)
    for k, e in worst.items():
        assert e < BAR[bm.MODE_MIXED] and e < TIGHT[bm.MODE_MIXED], (k, e)
    # the sweep's fixture frequencies equal the fixture (lowest, middle, highest)
    for i, j in enumerate((0, 20, 39)):
        assert_close(at(smx, j), {k: fx[k][i] for k in ("A", "B", "F")}, bm.MODE_MIXED, what=f"sweep {j}")
        assert_close(at(s64, j), {k: fx[k][i] for k in ("A", "B", "F")}, bm.MODE_FP64, what=f"sweep {j}")
    # lanes chosen by the library: 40 frequencies run as 4 rounds of 10, with
    # the results of any other lane count above one, bit for bit
    sauto = body.solve_sweep(omegas, betas=betas, mode=bm.MODE_MIXED, lanes=0)
    assert all(i["lanes"] == 10 for i in sauto["info"]), sauto["info"][0]
    for k in ("A", "B", "F"):
        np.testing.assert_array_equal(sauto[k], smx[k], err_msg=k)
    # the checked factorisation (pivot check after every panel) on every 8th
    # frequency: the same results bit for bit
    ctx = body.ctx
    try:
        ctx.set_lu_speculative(False)
        schk = body.solve_sweep(omegas[::8], betas=betas, mode=bm.MODE_MIXED, lanes=5)
    finally:
        ctx.set_lu_speculative(True)
    for k in ("A", "B", "F"):
        np.testing.assert_array_equal(schk[k], smx[k][::8], err_msg=f"{k} checked")


# ---------------------------------------------------------------------------
# 64 seeded uneven bodies x 16 frequencies
# ---------------------------------------------------------------------------
SEED = 20261203


@pytest.fixture(scope="module")
def batch64(ctx):
    meshes = [bm.Mesh.random_body(SEED, i) for i in range(64)]
    bodies = [bm.Body(ctx, m) for m in meshes]
    ns = [b.N for b in bodies]
    print(f"\n[gpu] batch of 64 bodies, N {min(ns)}..{max(ns)}", flush=True)
    for b in bodies:
        b.assemble_rankine()
    yield meshes, bodies
    for b in bodies:
        b.close()


def test_batch64_workload_shape(batch64):
    meshes, bodies = batch64
    ns = np.array([b.N for b in bodies])
    assert ns.min() >= 400 and ns.max() <= 4200
    types = {m.params["type"] for m in meshes}
    assert types == {"cylinder", "barge"}
    for m in meshes:
        p = m.params
        assert 5.0 <= p["a"] <= 40.0 and 5.0 <= p["b"] <= 40.0 and 5.0 <= p["draft"] <= 25.0
        assert 500 <= p["target"] <= 4000 and abs(p["n"] - p["target"]) <= 0.25 * p["target"], p


@pytest.mark.parametrize("mode", MODES)
def test_batch64_vs_sweeps_and_oracle(batch64, ctx, mode):
    meshes, bodies = batch64
    omegas = np.linspace(0.1, 2.0, 16)
    betas = [0.0]
    got = ctx.solve_batch(bodies, [omegas] * len(bodies), betas=betas, mode=mode, lanes=8)
    assert len(got) == 64
    for b, g in zip(bodies, got):
        ref = b.solve_sweep(omegas, betas=betas, mode=mode, lanes=4)
        for k in range(16):
            for key in ("A", "B", "F"):
                assert relerr(g[key][k], ref[key][k]) < 1e-12, (b.N, k, key)
        if mode == bm.MODE_MIXED:
            assert all(i["fell_back_fp64"] == 0 and i["residual"] < 1e-13 for i in g["info"]), b.N
    # the oracle on three sampled bodies (the smallest cylinder, the smallest
    # barge, and the median-size body) at the lowest, a middle and the highest frequency
    ns = np.array([b.N for b in bodies])
    cyl = [i for i, m in enumerate(meshes) if m.params["type"] == "cylinder"]
    brg = [i for i, m in enumerate(meshes) if m.params["type"] == "barge"]
    picks = {min(cyl, key=lambda i: ns[i]), min(brg, key=lambda i: ns[i]), int(np.argsort(ns)[32])}
    for i in sorted(picks):
        nodes, elems = meshes[i].arrays()
        ob = oracle.Body(nodes, elems)
        for k in (0, 8, 15):
            r = ob.solve(omegas[k], betas=betas)
            assert_close({key: got[i][key][k] for key in ("A", "B", "F")}, r, mode,
                         what=f"body {i} N={ns[i]} omega={omegas[k]:.3f}")
