"""VTK output (pmx_host_write_vtk, the C++ TOutput3D): the legacy-VTK file of
TOutput3D::WriteVtk (src/FE/Output3D.C:1773-2305) for scalar Q1/Q2
functions.

The reference itself cannot be built here for this path (Output3D.C includes
TECIO, which the image lacks), and it lists the points in the order of their
memory addresses, so no byte-for-byte fixture exists.  The file is pinned by
what fixes it:
 * the layout lines of WriteVtk (header, POINTS / CELLS / CELL_TYPES /
   POINT_DATA / SCALARS / CELL_DATA blocks, %.6e numbers);
 * every cell, in the reference's cell order, lists its vertices in the
   reference's local order: vertex j of cell i is the DOF of local index
   a + 2b + 4c (Q1) at the position (a,b,c) of hexahedron vertex j, whose
   number is GlobalNumbers[8 i + a + 2b + 4c] -- and GlobalNumbers equals
   ParMooN's own (tests/golden/reference_numbering.npz);
 * the point values are the function's nodal values (the basis of a nodal
   Q1/Q2 element is 1 at its vertex and 0 at the other vertices, so the
   reference's sum over the basis and its average over the cells of a vertex
   give the coefficient itself).
"""
import os
import subprocess

import numpy as np
import pytest

import orc
import parmoon_amd as pa

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF = np.load(os.path.join(ROOT, "tests", "golden", "reference_numbering.npz"))
VPOS = [(0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0), (0, 0, 1), (1, 0, 1), (1, 1, 1), (0, 1, 1)]


def parse(path):
    lines = open(path).read().split("\n")
    assert lines[:5] == ["# vtk DataFile Version 4.2", "file created by ParMooN", "ASCII",
                         "DATASET UNSTRUCTURED_GRID", ""]
    k = 5
    tag, npts, kind = lines[k].split()
    assert tag == "POINTS" and kind == "double"
    npts = int(npts)
    pts = np.array([[float(t) for t in lines[k + 1 + i].split()] for i in range(npts)])
    k += 1 + npts
    assert lines[k] == ""
    tag, nc, ncells_len = lines[k + 1].split()
    nc = int(nc)
    assert tag == "CELLS" and int(ncells_len) == 9 * nc
    cells = []
    for i in range(nc):
        row = lines[k + 2 + i]
        assert row.endswith(" ")  # the reference writes a blank after every number
        t = [int(v) for v in row.split()]
        assert t[0] == 8
        cells.append(t[1:])
    k += 2 + nc
    assert lines[k] == "" and lines[k + 1] == f"CELL_TYPES {nc}"
    assert lines[k + 2] == "12 " * nc and lines[k + 3] == ""
    assert lines[k + 4] == f"POINT_DATA {npts}"
    k += 5
    scal = {}
    while lines[k].startswith("SCALARS"):
        _, name, kind = lines[k].split()
        assert kind == "double" and lines[k + 1] == "LOOKUP_TABLE default"
        vals = lines[k + 2:k + 2 + npts]
        assert all(len(v.split("e")[0].lstrip("-")) == 8 for v in vals)  # %.6e: d.dddddd
        scal[name] = np.array([float(v) for v in vals])
        assert lines[k + 2 + npts] == "" and lines[k + 3 + npts] == ""
        k += 4 + npts
    assert lines[k] == "" and lines[k + 1] == f"CELL_DATA {nc}"
    assert lines[k + 2] == "SCALARS RegionID int  1" and lines[k + 3] == "LOOKUP_TABLE default  "
    assert lines[k + 4:k + 4 + nc] == ["0"] * nc and lines[k + 4 + nc:] == [""]
    return pts, np.array(cells), scal


@pytest.mark.parametrize("nref,order", [(2, 1), (1, 2), (3, 1), (2, 2)])
def test_vtk_file(tmp_path, nref, order):
    sp = orc.Space(nref, order)
    xyz = sp.dof_coords()
    u = np.sin(xyz[:, 0] + 2 * xyz[:, 1]) * np.exp(xyz[:, 2])
    w = xyz[:, 0] - 0.5 * xyz[:, 2]
    path = tmp_path / "out.vtk"
    pa.write_vtk(nref, order, {"u": u, "w": w}, path)
    pts, cells, scal = parse(path)
    n = 2 ** nref
    assert len(pts) == (n + 1) ** 3 and len(cells) == n ** 3
    assert set(scal) == {"u", "w"}
    # every mesh vertex once
    assert len(np.unique(np.rint(pts * n).astype(int), axis=0)) == len(pts)
    # cell i, vertex j is the DOF at hexahedron position VPOS[j] of cell i,
    # numbered by the reference's GlobalNumbers (pinned against ParMooN)
    _, gn, _, _ = sp.arrays()
    k = order + 1
    gn = gn.reshape(n ** 3, k ** 3)
    key = f"n{nref}_q{order}"
    if key + "_global_numbers" in REF.files:
        assert np.array_equal(gn.ravel(), REF[key + "_global_numbers"])
    for j, (a, b, c) in enumerate(VPOS):
        loc = order * (a + k * (b + k * c))
        np.testing.assert_array_equal(pts[cells[:, j]], xyz[gn[:, loc]])
    # values: the nodal values at the vertices, printed as %.6e
    vdof = {tuple(np.rint(p * order * n).astype(int)): i for i, p in enumerate(xyz)}
    at = np.array([vdof[tuple(np.rint(p * order * n).astype(int))] for p in pts])
    for name, f in (("u", u), ("w", w)):
        np.testing.assert_allclose(scal[name], f[at], rtol=5e-7, atol=1e-300)
        assert all(float(f"{v:.6e}") == s for v, s in zip(f[at], scal[name]))


def test_vtk_errors(tmp_path):
    sp = orc.Space(1, 1)
    u = np.zeros(sp.N)
    with pytest.raises(pa.PmxError):
        pa.write_vtk(1, 1, {"a b": u}, tmp_path / "x.vtk")
    with pytest.raises(pa.PmxError):
        pa.write_vtk(1, 1, {"u": u}, tmp_path / "no" / "such" / "dir.vtk")


def test_vtk_through_cpp_classes(tmp_path):
    """TFEFunction3D + TOutput3D::WriteVtk in C++ give the same bytes as the
    C-ABI call with the same values."""
    from test_cpp_shims import build
    a = tmp_path / "cpp.vtk"
    out = subprocess.run([build("shim_host", with_oracle=True), str(a)], capture_output=True, text=True, timeout=300)
    assert out.returncode == 0 and out.stdout.strip().endswith("OK"), out.stdout + out.stderr
    xyz = orc.Space(2, 2).dof_coords()
    b = tmp_path / "py.vtk"
    pa.write_vtk(2, 2, {"u": xyz[:, 0] + 2 * xyz[:, 1] + 3 * xyz[:, 2],
                        "v": xyz[:, 0] * xyz[:, 1] * xyz[:, 2]}, b)
    assert a.read_bytes() == b.read_bytes()
