"""Mesh files (GDF, MAR, PNL, ASCII STL) read as the reference reads them
(src/Boundary/GDF_Geo.cpp, MAR_Geo.cpp, PNL_Geo.cpp, STL_Geo.cpp). Each file
is written here from a generated mesh, in the format's published layout,
then read back through the C-ABI."""
import numpy as np
import pytest

import bemuse_amd as bm


def panels(nodes, elems):
    """Vertex lists of every panel."""
    out = []
    for e in elems:
        k = 3 if e[3] < 0 else 4
        out.append(nodes[e[:k]])
    return out


@pytest.fixture(scope="module")
def ref_mesh():
    m = bm.Mesh.half_cylinder(5.0, 8.0, 3, 4, 12)    # quads + a triangle fan
    return m.arrays()


def write_gdf(path, nodes, elems):
    with open(path, "w") as f:
        f.write("test hull written for the reader tests\n")
        f.write("1 9.80665 ULEN GRAV\n0  0 ISX  ISY\n")
        f.write(f"{len(elems)}\n")
        for p in panels(nodes, elems):
            q = list(p) + ([p[2]] if len(p) == 3 else [])
            for v in q:
                f.write(f"  {v[0]:.17g}  {v[1]:.17g}  {v[2]:.17g}\n")


def write_mar(path, nodes, elems):
    with open(path, "w") as f:
        f.write("    2   0\n")
        for i, v in enumerate(nodes):
            f.write(f"  {i + 1}  {v[0]:.17g}  {v[1]:.17g}  {v[2]:.17g}\n")
        f.write("  0  0.00  0.00  0.00\n")
        for e in elems:
            n = [e[0] + 1, e[1] + 1, e[2] + 1, (e[3] + 1) if e[3] >= 0 else e[2] + 1]
            f.write("  " + "  ".join(str(x) for x in n) + "\n")
        f.write("  0  0  0  0\n")


def write_pnl(path, nodes, elems):
    with open(path, "w") as f:
        f.write("    --------------Hull Mesh File---------------\n\n")
        f.write("    # Number of Panels, Nodes, X-Symmetry and Y-Symmetry\n")
        f.write(f"        {len(elems)}        {len(nodes)}         0         0\n\n")
        f.write("    #Start Definition of Node Coordinates     ! node_number   x   y   z\n")
        for i, v in enumerate(nodes):
            f.write(f"     {i + 1}  {v[0]:.17g}  {v[1]:.17g}  {v[2]:.17g}\n")
        f.write("   #End Definition of Node Coordinates\n\n")
        f.write("   #Start Definition of Node Relations   ! panel_number  number_of_vertices   Vertex1_ID ...\n")
        for j, e in enumerate(elems):
            k = 3 if e[3] < 0 else 4
            f.write(f"    {j + 1}  {k}  " + "  ".join(str(x + 1) for x in e[:k]) + "\n")
        f.write("   #End Definition of Node Relations\n\n    --------------End Hull Mesh File---------------\n")


def write_stl(path, tris):
    with open(path, "w") as f:
        f.write("solid test\n")
        for t in tris:
            f.write("  facet normal 0 0 0\n    outer loop\n")
            for v in t:
                f.write(f"      vertex {v[0]:.17g} {v[1]:.17g} {v[2]:.17g}\n")
            f.write("    endloop\n  endfacet\n")
        f.write("endsolid test\n")


@pytest.mark.parametrize("fmt,writer", [("gdf", write_gdf), ("mar", write_mar), ("pnl", write_pnl)])
def test_reader_round_trip(tmp_path, ref_mesh, fmt, writer):
    nodes, elems = ref_mesh
    path = tmp_path / f"hull.{fmt}"
    writer(path, nodes, elems)
    n2, e2 = bm.Mesh.read(str(path)).arrays()
    got, want = panels(n2, e2), panels(nodes, elems)
    assert len(got) == len(want)
    for a, b in zip(got, want):
        np.testing.assert_array_equal(a, b)       # same vertices, same order, same triangle/quad split
    if fmt == "gdf":
        assert len(n2) == sum(len(p) for p in want)   # GDF nodes are per panel, not shared


def test_stl_scale_and_order(tmp_path):
    tris = [np.array([[0.0, 0.0, -0.01], [0.01, 0.0, -0.01], [0.0, 0.01, -0.01]]),
            np.array([[0.0, 0.0, -0.02], [0.0, 0.01, -0.02], [0.01, 0.0, -0.02]])]
    path = tmp_path / "part.stl"
    write_stl(path, tris)
    n2, e2 = bm.Mesh.read(str(path)).arrays()
    assert len(e2) == 2 and np.all(e2[:, 3] < 0)
    for t, e in zip(tris, e2):
        np.testing.assert_allclose(n2[e[:3]], 100.0 * t[::-1], rtol=0, atol=1e-12)   # x100, reversed


def test_reader_errors(tmp_path):
    with pytest.raises(bm.BemError):
        bm.Mesh.read(str(tmp_path / "missing.gdf"))
    bad = tmp_path / "hull.xyz"
    bad.write_text("1 2 3\n")
    with pytest.raises(bm.BemError):
        bm.Mesh.read(str(bad))
    empty = tmp_path / "empty.gdf"
    empty.write_text("title only\n")
    with pytest.raises(bm.BemError):
        bm.Mesh.read(str(empty))


@pytest.mark.gpu
def test_gdf_mesh_solves_like_the_generated_one(ctx, tmp_path):
    """Unshared GDF nodes give the same panels, hence the same A, B, F."""
    m = bm.Mesh.half_cylinder(5.0, 8.0, 4, 8, 16)
    nodes, elems = m.arrays()
    path = tmp_path / "cyl.gdf"
    write_gdf(path, nodes, elems)
    res = []
    for mesh in (m, bm.Mesh.read(str(path))):
        body = bm.Body(ctx, mesh)
        body.assemble_rankine()
        res.append(body.solve(0.9, betas=[0.0, 0.5], mode=bm.MODE_FP64))
        body.close()
    for k in ("A", "B", "F"):
        np.testing.assert_array_equal(res[0][k], res[1][k])
