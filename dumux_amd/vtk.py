"""VTU output and restart input of the CCTpfa path, in the reference's file layout.

Writer: VtkOutputModule (dumux/io/vtkoutputmodule.hh) with the model's IOFields
(porousmediumflow/2p/iofields.hh:30-51, porousmediumflow/1p/iofields.hh), conforming
output of a structured grid as Dune's VTKWriter writes it (dune-grid, third-party):
  * an UnstructuredGrid piece, cell data first, then points and cells;
  * points numbered in order of first appearance while walking the cells x-fastest, a cell's
    corners in Dune order (x fastest, then y, then z), connectivity in VTK order (quad
    0 1 3 2, hexahedron 0 1 3 2 4 5 7 6);
  * ascii, Float32 by default (Vtk.Precision, vtkoutputmodule.hh:72), 12 values per line;
  * a .pvd collection listing every written time step (VtkOutputModule::write(time)).
Cell fields, 2p (pw-Sn, phase names aq / napl): S_aq p_aq rho_aq mob_aq S_napl p_napl rho_napl
mob_napl pc porosity "process rank"; 1p: p "process rank".

Restart: load_solution (dumux/io/loadsolution.hh:92-353) reads the primary variables back by
their output names (2p p0s1: p_aq and S_napl, iofields.hh:53-62; 1p: p).
"""
import os
import xml.etree.ElementTree as ET

import numpy as np

VTK_QUAD, VTK_HEXAHEDRON = 9, 12
_QUAD_ORDER = (0, 1, 3, 2)
_HEX_ORDER = (0, 1, 3, 2, 4, 5, 7, 6)


def grid_of(desc):
    """(dim, n, lower, upper) of a dmx_problem_desc-like structure."""
    dim = int(desc.dim)
    n = [int(desc.n[a]) for a in range(dim)]
    return dim, n, [float(desc.lower[a]) for a in range(dim)], [float(desc.upper[a]) for a in range(dim)]


def mesh(dim, n, lower, upper):
    """Points (P x 3), connectivity, offsets and cell types of the structured grid, numbered as
    Dune's VTKWriter numbers them."""
    h = [(upper[a] - lower[a]) / n[a] for a in range(dim)]
    nv = [n[a] + 1 for a in range(dim)]
    index = -np.ones(int(np.prod(nv)), dtype=np.int64)
    points = []
    corners_per_cell = 1 << dim
    order = _QUAD_ORDER if dim == 2 else _HEX_ORDER
    conn = np.empty(int(np.prod(n)) * corners_per_cell, dtype=np.int64)
    cell = 0
    ranges = [range(n[a]) for a in range(dim)]
    for k in (ranges[2] if dim == 3 else [0]):
        for j in ranges[1]:
            for i in ranges[0]:
                ids = []
                for c in range(corners_per_cell):
                    v = [i + (c & 1), j + ((c >> 1) & 1), k + ((c >> 2) & 1)]
                    lin = v[0] + nv[0] * v[1] + (nv[0] * nv[1] * v[2] if dim == 3 else 0)
                    if index[lin] < 0:
                        index[lin] = len(points)
                        points.append([lower[a] + v[a] * h[a] for a in range(dim)] + [0.0] * (3 - dim))
                    ids.append(index[lin])
                conn[cell * corners_per_cell:(cell + 1) * corners_per_cell] = [ids[o] for o in order]
                cell += 1
    ncell = cell
    offsets = np.arange(1, ncell + 1, dtype=np.int64) * corners_per_cell
    types = np.full(ncell, VTK_QUAD if dim == 2 else VTK_HEXAHEDRON, dtype=np.int64)
    return np.array(points), conn, offsets, types


def cell_fields(desc, u, p=None, rho_mob=None):
    """Output fields of one solution: u (N*m) and the volume variables (phase pressures and
    rho*mobility, N*m each, as returned by FVAssembler.volumeVariables)."""
    N = int(np.prod([int(desc.n[a]) for a in range(int(desc.dim))]))
    m = 2 if int(desc.model) == 2 else 1
    u = np.asarray(u, dtype=np.float64)
    if m == 1:
        return {"p": u.copy(), "process rank": np.zeros(N)}
    sn = u[1::2]
    pw, pn = p[0::2], p[1::2]
    rho = (float(desc.rho[0]), float(desc.rho[1]))
    return {"S_aq": 1.0 - sn, "p_aq": pw.copy(), "rho_aq": np.full(N, rho[0]), "mob_aq": rho_mob[0::2] / rho[0],
            "S_napl": sn.copy(), "p_napl": pn.copy(), "rho_napl": np.full(N, rho[1]), "mob_napl": rho_mob[1::2] / rho[1],
            "pc": pn - pw, "porosity": np.full(N, float(desc.porosity)), "process rank": np.zeros(N)}


def _fmt(values, precision):
    if precision == "Float32":
        vals = np.asarray(values, dtype=np.float32)
        txt = [f"{float(v):.6g}" for v in vals]
    elif precision == "Float64":
        txt = [repr(float(v)) for v in np.asarray(values, dtype=np.float64)]
    else:  # integer arrays
        txt = [str(int(v)) for v in values]
    lines = [" ".join(txt[k:k + 12]) for k in range(0, len(txt), 12)]
    return "".join(f"          {ln}\n" for ln in lines)


def write_vtu(path, desc, fields, precision="Float32"):
    """Writes one VTU file of the structured grid of desc with the given cell fields."""
    if precision not in ("Float32", "Float64"):
        raise ValueError("precision must be Float32 or Float64 (Vtk.Precision)")
    dim, n, lower, upper = grid_of(desc)
    pts, conn, offs, types = mesh(dim, n, lower, upper)
    ncell = len(offs)
    names = list(fields)
    out = ['<?xml version="1.0"?>\n',
           '<VTKFile type="UnstructuredGrid" version="0.1" byte_order="LittleEndian">\n',
           '  <UnstructuredGrid>\n',
           f'    <Piece NumberOfCells="{ncell}" NumberOfPoints="{len(pts)}">\n',
           f'      <CellData Scalars="{names[0]}">\n']
    for nm in names:
        vals = np.asarray(fields[nm], dtype=np.float64)
        if vals.size != ncell:
            raise ValueError(f"field {nm}: {vals.size} values for {ncell} cells")
        out.append(f'        <DataArray type="{precision}" Name="{nm}" NumberOfComponents="1" format="ascii">\n')
        out.append(_fmt(vals, precision))
        out.append('        </DataArray>\n')
    out.append('      </CellData>\n      <Points>\n')
    out.append(f'        <DataArray type="{precision}" Name="Coordinates" NumberOfComponents="3" format="ascii">\n')
    out.append(_fmt(pts.ravel(), precision))
    out.append('        </DataArray>\n      </Points>\n      <Cells>\n')
    out.append('        <DataArray type="Int32" Name="connectivity" NumberOfComponents="1" format="ascii">\n')
    out.append(_fmt(conn, None))
    out.append('        </DataArray>\n')
    out.append('        <DataArray type="Int32" Name="offsets" NumberOfComponents="1" format="ascii">\n')
    out.append(_fmt(offs, None))
    out.append('        </DataArray>\n')
    out.append('        <DataArray type="UInt8" Name="types" NumberOfComponents="1" format="ascii">\n')
    out.append(_fmt(types, None))
    out.append('        </DataArray>\n      </Cells>\n    </Piece>\n  </UnstructuredGrid>\n</VTKFile>\n')
    with open(path, "w") as f:
        f.write("".join(out))


def read_vtu(path):
    """{"cells": {name: values}, "points": P x 3, "connectivity", "offsets", "types"} of an ascii
    VTU file."""
    root = ET.parse(path).getroot()
    piece = root.find("UnstructuredGrid/Piece")
    if piece is None:
        raise ValueError(f"{path}: no UnstructuredGrid piece")

    def arr(node):
        kind = node.get("type", "Float32")
        dtype = np.int64 if kind.startswith(("Int", "UInt")) else np.float64
        if node.get("format", "ascii") != "ascii":
            raise ValueError(f"{path}: only ascii data arrays are read")
        return np.array((node.text or "").split(), dtype=dtype)

    cells = {}
    cd = piece.find("CellData")
    if cd is not None:
        for da in cd.findall("DataArray"):
            cells[da.get("Name")] = arr(da)
    pts = arr(piece.find("Points/DataArray")).reshape(-1, 3)
    out = {"cells": cells, "points": pts}
    for da in piece.findall("Cells/DataArray"):
        out[da.get("Name")] = arr(da)
    return out


def primary_variable_names(desc):
    """Output names of the primary variables (IOFields::primaryVariableName)."""
    return ["p_aq", "S_napl"] if int(desc.model) == 2 else ["p"]


def load_solution(path, desc):
    """The solution vector (N*m, cells x-fastest) stored in a VTU file (restart,
    dumux/io/loadsolution.hh:92-353): the primary variables by their output names."""
    data = read_vtu(path)["cells"]
    names = primary_variable_names(desc)
    missing = [nm for nm in names if nm not in data]
    if missing:
        raise KeyError(f"{path}: missing primary variable fields {missing}")
    cols = [data[nm] for nm in names]
    u = np.empty(cols[0].size * len(cols))
    for k, c in enumerate(cols):
        u[k::len(cols)] = c
    return u


class VtkOutputModule:
    """VtkOutputModule (dumux/io/vtkoutputmodule.hh) of one assembler: write(time) writes
    <name>-<k:05d>.vtu of the current solution and rewrites <name>.pvd with every step so far."""

    def __init__(self, assembler, name, directory=".", precision="Float32"):
        self.assembler = assembler
        self.name = name
        self.directory = directory
        self.precision = precision
        self.steps = []

    def write(self, time):
        a = self.assembler
        u = a.solution()
        if a.m == 2:
            p, rm = a.volumeVariables()
            fields = cell_fields(a.problem.desc, u, p, rm)
        else:
            fields = cell_fields(a.problem.desc, u)
        fname = f"{self.name}-{len(self.steps):05d}.vtu"
        write_vtu(os.path.join(self.directory, fname), a.problem.desc, fields, self.precision)
        self.steps.append((float(time), fname))
        lines = ['<?xml version="1.0"?>\n', '<VTKFile type="Collection" version="0.1" byte_order="LittleEndian">\n',
                 '<Collection>\n']
        for t, fn in self.steps:
            lines.append(f'<DataSet timestep="{t!r}" group="" part="0" name="" file="{fn}"/>\n')
        lines.append('</Collection>\n</VTKFile>\n')
        with open(os.path.join(self.directory, f"{self.name}.pvd"), "w") as f:
            f.write("".join(lines))
        return os.path.join(self.directory, fname)


def pvd_steps(path):
    """[(time, file)] of a .pvd collection."""
    root = ET.parse(path).getroot()
    return [(float(d.get("timestep")), d.get("file")) for d in root.iter("DataSet")]

