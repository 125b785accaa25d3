"""Legacy VTK output of the particles and the tool (vtk_writer.cpp:20-263): the files parse,
hold the non-blanked particles with the state's values (to the %f precision of the format), and
the vertex cells index the written points (the reference indexes the particle's slot)."""
from __future__ import annotations

import numpy as np
import pytest

import pymfc

pytestmark = pytest.mark.gpu


def read_vtk(path):
    lines = open(path).read().split("\n")
    assert lines[0] == "# vtk DataFile Version 2.0" and lines[2] == "ASCII"
    out, k = {}, 0
    while k < len(lines):
        w = lines[k].split()
        if not w:
            k += 1
            continue
        if w[0] == "POINTS":
            n = int(w[1])
            out["points"] = np.array([[float(x) for x in lines[k + 1 + q].split()] for q in range(n)])
            k += 1 + n
        elif w[0] == "CELLS":
            n = int(w[1])
            out["cells"] = [[int(x) for x in lines[k + 1 + q].split()] for q in range(n)]
            assert sum(len(c) for c in out["cells"]) == int(w[2])
            k += 1 + n
        elif w[0] == "CELL_TYPES":
            n = int(w[1])
            out["cell_types"] = [int(lines[k + 1 + q]) for q in range(n)]
            k += 1 + n
        elif w[0] == "SCALARS":
            n = len(out["points"])
            out[w[1]] = np.array([float(lines[k + 2 + q]) for q in range(n)])
            k += 2 + n
        elif w[0] == "VECTORS":
            n = len(out["points"])
            out[w[1]] = np.array([[float(x) for x in lines[k + 1 + q].split()] for q in range(n)])
            k += 1 + n
        else:
            k += 1
    return out


def test_particle_vtk(tmp_path):
    s = pymfc.Sim(pymfc.cut_config(40, 12, 10.0, thermal="contact"))
    s.step(10)
    vt = s.field("vel_t")
    free = np.nonzero(s.field("flags") == 0)[0]
    vt[free[free.size // 2], 0] = np.nan  # blanks one particle in the next step
    s.set_field("vel_t", vt)
    s.step(20)
    fl = s.field("flags")
    keep = (fl & pymfc.FLAG_BLANKED) == 0
    assert (~keep).sum() >= 1
    path = str(tmp_path / "p.vtk")
    s.write_vtk(path)
    v = read_vtk(path)
    n = int(keep.sum())
    assert v["points"].shape == (n, 3)
    assert v["cells"] == [[1, k] for k in range(n)] and v["cell_types"] == [1] * n
    np.testing.assert_allclose(v["points"][:, :2], s.field("pos")[keep], atol=6e-7)
    np.testing.assert_allclose(v["density"], s.field("rho")[keep], atol=6e-7)
    np.testing.assert_allclose(v["Temperature"], s.field("T")[keep], atol=6e-7)
    assert np.array_equal(v["Fixed"], ((fl & pymfc.FLAG_FIXED) != 0)[keep].astype(float))
    assert np.array_equal(v["Tool"], ((fl & pymfc.FLAG_TOOL) != 0)[keep].astype(float))
    np.testing.assert_allclose(v["EquivAccumPlasticStrain"], s.field("eps_pl")[keep], atol=6e-7)
    np.testing.assert_allclose(v["Velocity"][:, :2], s.field("vel")[keep], atol=6e-7)
    S, p = s.field("S")[keep], s.field("p")[keep]
    c = S.copy()
    c[:, 0] -= p
    c[:, 2] -= p
    c[:, 3] -= p
    svm2 = (c[:, 0] ** 2 + c[:, 2] ** 2 + c[:, 3] ** 2 - c[:, 0] * c[:, 2] - c[:, 0] * c[:, 3] - c[:, 2] * c[:, 3]
            + 3 * c[:, 1] ** 2)
    np.testing.assert_allclose(v["SvM"], np.sqrt(np.maximum(svm2, 0.0)), atol=6e-7)


def test_tool_vtk(tmp_path):
    s = pymfc.Sim(pymfc.cut_config(40, 12, 10.0, thermal="contact"))
    s.step(5)
    path = str(tmp_path / "t.vtk")
    s.write_tool_vtk(path)
    v = read_vtk(path)
    t = s.tool()
    nseg = len(t["left"])
    ntri = (2 if nseg == 4 else 3) + (19 if t["has_fillet"] else 0)
    assert v["points"].shape == (3 * ntri, 3)
    assert v["cells"] == [[3, 3 * i, 3 * i + 1, 3 * i + 2] for i in range(ntri)]
    assert v["cell_types"] == [5] * ntri
    # the first body triangle is segment 0's end points and the next segment's right end
    right = 1 if nseg == 4 else 2
    np.testing.assert_allclose(v["points"][:3, :2], [t["left"][0], t["right"][0], t["right"][right]], atol=6e-7)
    if t["has_fillet"]:  # fan triangles start at the fillet centre
        fan = v["points"][3 * (ntri - 19):: 3, :2]
        np.testing.assert_allclose(fan, np.tile(t["fillet"][:2], (19, 1)), atol=6e-7)
