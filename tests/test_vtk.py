"""VTU output and restart input (dumux_amd/vtk.py) on the CPU: the reference's file layout,
round trips and the restart reader."""
import os

import numpy as np

import dumux_amd as dx
from dumux_amd import vtk
from tests.helpers import LENS_FIELDS, lens_problem, lens_reference


def lens_desc():
    p = dx.Problem(kind=2, dim=2, n=(48, 32))
    lens_problem(p)
    return p


def test_mesh_numbering_matches_the_reference_file():
    # test/references/test_2p_incompressible_cc-reference.vtu (Dune VTKWriter, conforming): 1536
    # quads, 1617 points numbered by first appearance, first points and connectivity as there
    pts, conn, offs, types = vtk.mesh(2, [48, 32], [0.0, 0.0], [6.0, 4.0])
    assert len(offs) == 1536 and len(pts) == 1617
    assert pts[:4].ravel().tolist() == [0, 0, 0, 0.125, 0, 0, 0, 0.125, 0, 0.125, 0.125, 0]
    assert pts[-1].tolist() == [6, 4, 0]
    assert conn[:12].tolist() == [0, 1, 3, 2, 1, 4, 5, 3, 4, 6, 7, 5]
    assert conn[-12:].tolist() == [1564, 1565, 1614, 1613, 1565, 1566, 1615, 1614, 1566, 1567, 1616, 1615]
    assert offs[:4].tolist() == [4, 8, 12, 16] and set(types.tolist()) == {vtk.VTK_QUAD}


def test_hexahedra_span_their_cells():
    n = [3, 2, 4]
    pts, conn, offs, types = vtk.mesh(3, n, [0.0, 0.0, 0.0], [1.5, 1.0, 2.0])
    assert len(pts) == 4 * 3 * 5 and set(types.tolist()) == {vtk.VTK_HEXAHEDRON}
    assert len(np.unique(pts, axis=0)) == len(pts)
    c = conn.reshape(-1, 8)
    for cell in range(24):
        i, j, k = cell % 3, (cell // 3) % 2, cell // 6
        corners = pts[c[cell]]
        lo = np.array([i * 0.5, j * 0.5, k * 0.5])
        # VTK hexahedron order: bottom face counter-clockwise, then the top face
        want = lo + 0.5 * np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0], [0, 0, 1], [1, 0, 1], [1, 1, 1], [0, 1, 1]])
        assert np.allclose(corners, want), cell


def test_round_trip_and_restart(tmp_path):
    p = lens_desc()
    rng = np.random.default_rng(1)
    u = p.u0.copy()
    u[0::2] += rng.normal(0, 300.0, p.N)
    u[1::2] = rng.uniform(0.0, 0.6, p.N)
    pp = np.empty(2 * p.N)
    pp[0::2], pp[1::2] = u[0::2], u[0::2] + rng.uniform(100, 2000, p.N)
    rm = rng.uniform(0.0, 1e6, 2 * p.N)
    fields = vtk.cell_fields(p.desc, u, pp, rm)
    assert list(fields) == ["S_aq", "p_aq", "rho_aq", "mob_aq", "S_napl", "p_napl", "rho_napl", "mob_napl", "pc",
                            "porosity", "process rank"]
    f64 = os.path.join(tmp_path, "a.vtu")
    vtk.write_vtu(f64, p.desc, fields, precision="Float64")
    back = vtk.read_vtu(f64)["cells"]
    for nm, v in fields.items():
        assert np.array_equal(back[nm], v), nm
    # restart from full precision is exact
    assert np.array_equal(vtk.load_solution(f64, p.desc), u)
    # the default Float32 ascii output prints 6 significant digits of the single-precision value,
    # as the reference's files do ("152499"): a restart from it is approximate
    f32 = os.path.join(tmp_path, "b.vtu")
    vtk.write_vtu(f32, p.desc, fields)
    back32 = vtk.read_vtu(f32)["cells"]
    for nm, v in fields.items():
        want = np.array([float(f"{float(x):.6g}") for x in v.astype(np.float32)])
        assert np.array_equal(back32[nm], want), nm
    u32 = vtk.load_solution(f32, p.desc)
    assert np.max(np.abs(u32 - u) / np.maximum(1.0, np.abs(u))) < 1e-5


def test_reference_values_survive_a_rewrite(tmp_path):
    # the fixture's Float32 prints, written and read back, are the same numbers
    ref = lens_reference()
    p = lens_desc()
    path = os.path.join(tmp_path, "r.vtu")
    vtk.write_vtu(path, p.desc, {nm: ref[nm] for nm in LENS_FIELDS})
    back = vtk.read_vtu(path)["cells"]
    for nm in LENS_FIELDS:
        assert np.array_equal(back[nm], ref[nm]), nm
    u = vtk.load_solution(path, p.desc)
    assert np.array_equal(u[0::2], ref["p_aq"]) and np.array_equal(u[1::2], ref["S_napl"])


class _StubAssembler:
    """Duck-typed assembler with a fixed state (solution, volume variables)."""

    def __init__(self, problem, u, p, rm):
        self.problem, self.m, self._u, self._p, self._rm = problem, problem.m, u, p, rm

    def solution(self):
        return self._u

    def volumeVariables(self):
        return self._p, self._rm


def test_output_module_writes_numbered_steps_and_a_collection(tmp_path):
    p = dx.Problem(kind=1, dim=3, n=(4, 3, 2))
    a = _StubAssembler(p, p.u0 + 1.0, None, None)
    out = vtk.VtkOutputModule(a, "run", directory=str(tmp_path))
    f0 = out.write(0.0)
    a._u = p.u0 + 2.0
    f1 = out.write(250.0)
    assert os.path.basename(f0) == "run-00000.vtu" and os.path.basename(f1) == "run-00001.vtu"
    assert vtk.pvd_steps(os.path.join(tmp_path, "run.pvd")) == [(0.0, "run-00000.vtu"), (250.0, "run-00001.vtu")]
    data = vtk.read_vtu(f1)
    assert list(data["cells"]) == ["p", "process rank"]
    assert np.allclose(data["cells"]["p"], p.u0 + 2.0, rtol=1e-5, atol=0)
    assert len(data["points"]) == 5 * 4 * 3
