"""Writes the golden fixtures under tests/golden/ (run in a checkout that has
the reference sources; the fixtures are committed, the script is not needed
at test time).

gf_poly_coeffs.json
    The numeric literals of the seven polynomials of the Havelock global
    approximation (GF_FuncA..D, GF_dFuncA..C, src/Kernels.cpp:120-256), of
    the Struve approximations (StruveH0/H1, src/Kernels.cpp:35-116: the
    `Real P0 = ...;` declarations) and the Euler constant literal EUL
    (src/BEMUse_Math_Types.h:99), parsed from the reference text by a
    tokeniser (the running-power statements `S = c0` / `S += c*t;`;
    commented-out lines are skipped). The oracle's and the device's
    hand-typed tables are compared with these.

cylinder_abf.npz
    A, B, F of the oracle for the small cylinder (r = 5 m, draft 8 m,
    4 x 8 x 16 panels, N = 192) at omega = 0.4 and 1.3 rad/s, headings 0 and
    45 deg. A regression fixture of the oracle itself (the checks that pin the
    oracle live in tests/test_oracle.py); the GPU parity tests compare with
    it as well as with the live oracle.

semisub6018_abf.npz
    A, B, F, FK, SC of the oracle for the 6018-panel three-column semi-sub
    (the 6k-panel workload) at omega = 0.05, 1.3064 and 2.5 rad/s, headings
    0, 45, 90, 135, 180 deg, and Phi_J, Phi_S at 1.3064 rad/s. About a
    minute per frequency on 8 threads, which is why the GPU tests read it
    instead of running the oracle live.

Usage: python tests/golden/make_golden.py [--reference /root/reference] [--only poly|cylinder|semisub]
"""
import argparse
import json
import os
import re
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))

NUM = r"[-+]?\s*(?:\d+\.\d*|\.\d+|\d+)(?:[eE][-+]?\d+)?"


def parse_poly(text, name):
    m = re.search(r"inline\s+Real\s+" + name + r"\s*\(const Real &tt\)\s*\{(.*?)\n\}", text, re.S)
    if not m:
        raise RuntimeError(f"{name} not found")
    coeffs = []
    for line in m.group(1).splitlines():
        line = line.split("//", 1)[0].strip()
        if not line:
            continue
        a = re.match(r"Real\s+S\s*=\s*(" + NUM + r")\s*,", line)
        if a:
            coeffs.append(float(a.group(1).replace(" ", "")))
            continue
        b = re.match(r"S\s*\+=\s*(" + NUM + r")\s*\*\s*t\s*;", line)
        if b:
            coeffs.append(float(b.group(1).replace(" ", "")))
    if not coeffs:
        raise RuntimeError(f"{name}: no coefficients parsed")
    return coeffs


STRUVE_NAMES = {"StruveH0": ["P0", "P1", "P2", "P3", "P4", "P5", "a0", "a1", "a2", "a3", "b1", "b2", "b3"],
                "StruveH1": ["P1", "P2", "P3", "P4", "P5", "P6", "a0", "a1", "a2", "a3", "b1", "b2", "b3"]}


def parse_struve(text, name):
    """The `Real NAME = literal;` declarations in the body of StruveH0/H1,
    in the order listed in STRUVE_NAMES."""
    m = re.search(r"inline\s+Real\s+" + name + r"\s*\(const Real &xx\)\s*\{(.*?)\n\}", text, re.S)
    if not m:
        raise RuntimeError(f"{name} not found")
    vals = {}
    for line in m.group(1).splitlines():
        line = line.split("//", 1)[0].strip()
        d = re.match(r"Real\s+(\w+)\s*=\s*(" + NUM + r")\s*;", line)
        if d:
            vals[d.group(1)] = float(d.group(2).replace(" ", ""))
    return [vals[k] for k in STRUVE_NAMES[name]]


def poly_fixture(ref):
    text = open(os.path.join(ref, "src", "Kernels.cpp")).read()
    out = {}
    for nm in ("GF_FuncA", "GF_FuncB", "GF_FuncC", "GF_FuncD", "GF_dFuncA", "GF_dFuncB", "GF_dFuncC"):
        out[nm] = parse_poly(text, nm)
    out["StruveH0"] = parse_struve(text, "StruveH0")
    out["StruveH1"] = parse_struve(text, "StruveH1")
    mt = open(os.path.join(ref, "src", "BEMUse_Math_Types.h")).read()
    m = re.search(r"EUL\s*=\s*(" + NUM + ")", mt)
    out["EUL"] = float(m.group(1))
    out["_source"] = ("src/Kernels.cpp:120-256 (GF_Func*/GF_dFunc* literals), src/Kernels.cpp:35-116 "
                      "(StruveH0/H1: P, a, b literals), src/BEMUse_Math_Types.h:99 (EUL)")
    with open(os.path.join(HERE, "gf_poly_coeffs.json"), "w") as f:
        json.dump(out, f, indent=1)
    print("wrote gf_poly_coeffs.json:", {k: len(v) for k, v in out.items() if isinstance(v, list)})


def cylinder_fixture():
    sys.path.insert(0, ROOT)
    import oracle
    import bemuse_amd as bm
    mesh = bm.Mesh.half_cylinder(5.0, 8.0, 4, 8, 16)
    nodes, elems = mesh.arrays()
    ob = oracle.Body(nodes, elems)
    omegas = np.array([0.4, 1.3])
    betas = np.array([0.0, np.pi / 4])
    A, B, F = [], [], []
    for w in omegas:
        r = ob.solve(w, betas=betas)
        A.append(r["A"]); B.append(r["B"]); F.append(r["F"])
    np.savez(os.path.join(HERE, "cylinder_abf.npz"), nodes=nodes, elems=elems, omegas=omegas, betas=betas,
             A=np.array(A), B=np.array(B), F=np.array(F))
    print("wrote cylinder_abf.npz N =", len(elems))


def semisub_fixture():
    """The 6018-panel three-column semi-sub workload (bemuse_amd.Mesh.semisub3
    defaults): oracle A, B, F, FK, SC at three of its 40 frequencies (the
    lowest, a middle one, the highest) for its 5 headings, and the radiation
    and diffraction potentials at the middle one."""
    sys.path.insert(0, ROOT)
    import oracle
    import bemuse_amd as bm
    oracle.set_threads(os.cpu_count() or 1)
    mesh = bm.Mesh.semisub3()
    nodes, elems = mesh.arrays()
    ob = oracle.Body(nodes, elems)
    all_w = np.linspace(0.05, 2.5, 40)
    idx = np.array([0, 20, 39])
    betas = np.deg2rad([0.0, 45.0, 90.0, 135.0, 180.0])
    out = {k: [] for k in ("A", "B", "F", "FK", "SC")}
    for i in idx:
        r = ob.solve(all_w[i], betas=betas, want_phi=(i == 20))
        for k in out:
            out[k].append(r[k])
        if i == 20:
            phiJ, phiS = r["phiJ"], r["phiS"]
        print(f"semisub omega={all_w[i]:.4f} done", flush=True)
    np.savez_compressed(os.path.join(HERE, "semisub6018_abf.npz"), nodes=nodes, elems=elems, omegas=all_w[idx],
                        betas=betas, phi_omega=all_w[20], phiJ=phiJ, phiS=phiS,
                        **{k: np.array(v) for k, v in out.items()})
    print("wrote semisub6018_abf.npz N =", len(elems))


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--reference", default="/root/reference")
    ap.add_argument("--only", choices=["poly", "cylinder", "semisub"], default=None)
    args = ap.parse_args()
    if args.only in (None, "poly"):
        poly_fixture(args.reference)
    if args.only in (None, "cylinder"):
        cylinder_fixture()
    if args.only in (None, "semisub"):
        semisub_fixture()
