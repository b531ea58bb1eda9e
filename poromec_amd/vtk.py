"""The results file of the ``pres`` macro: mesh, displacements, pressures,
stresses, Darcy flux and porosity of the current state at the nodes, as
``write_mesh_res_pm`` (``Write_vtk.f:912-1393``) writes it with the
``Vtk.f`` writers.

Three encodings, chosen by the deck's config keys as the reference chooses
them:

* ``legacy_vtk true``: legacy VTK ``<prefix>_istep_<n>.vtk``, ASCII, or
  BINARY with ``bvtk true`` (big-endian stream: 4-byte integers, 8-byte
  reals, the text lines the reference writes around them);
* default: XML VTU ``<prefix>_istep_<n>.vtu``, ASCII.  The reference has no
  binary VTU writer (``head_vtu`` stops), so that pair is rejected.

Which fields are written is ``setprint``'s choice (``set_print_vtk_pm``,
``Rdata.f:3383-3520``): ``desloc`` and ``pressure`` by default.  With
``quadratic`` the mid-edge nodes are printed too, as hexa20 cells (VTK type
25, the corner and edge order of ``elm_vtk``); their coordinates, pressures
and recovered fields are the averages of the edge's two vertices
(``mk_coor_quad``, ``split_u_p``, ``mk_field_quad``).  Otherwise the vertices
are printed, as hexa8 cells (type 12).  Stress tensors are written in full
(``make_full_tensor``: xx xy xz yx yy yz zx zy zz).

Numbers are formatted as the reference's edit descriptors write them
(``poromec_amd.ftext``)."""
from __future__ import annotations

import numpy as np

from .ftext import fortran_a, fortran_e, fortran_f, fortran_i, list_record

# setprint words (set_print_vtk_pm), 1-based as print_flag
PRINT_WORDS = ("quadratic", "desloc", "pressure", "dpressure", "totalstress", "biotstress",
               "terzaghistress", "darcyflux", "porosity", "pconsolidation", "elplastic",
               "temperature", "heatFlux")
DEFAULT_FLAGS = (False, True, True) + (False,) * 10  # Main.f:220-232

# hexa20 corners and edge nodes in VTK's quadratic-hexahedron order (elm_vtk)
HEX20_VTK = np.array([5, 6, 7, 8, 1, 2, 3, 4, 13, 14, 15, 16, 9, 10, 11, 12, 17, 18, 19, 20]) - 1
# (vertex, vertex, edge node) of the 12 edges (hexa20edgeNod, Adjacency.f:974)
HEX20_EDGES = np.array([(1, 2, 9), (2, 3, 10), (3, 4, 11), (4, 1, 12), (5, 6, 13), (6, 7, 14),
                        (7, 8, 15), (8, 5, 16), (1, 5, 17), (2, 6, 18), (3, 7, 19),
                        (4, 8, 20)]) - 1
HEAD = "Malha poro mec".ljust(30)


class VtkError(ValueError):
    pass


def results_name(prefix: str, istep: int, legacy: bool) -> str:
    """``name(prename, istep, 0, 2|3)`` (Filenames.f)."""
    return f"{prefix}_istep_{istep}.{'vtk' if legacy else 'vtu'}"


def quad_field(field, ix, nprint, quad):
    """The field at the printed nodes: the first nprint rows, with each
    edge node set to the average of its two vertices when quad."""
    f = np.array(field[:nprint], dtype=float, copy=True)
    if quad:
        for a, b, c in HEX20_EDGES:
            f[ix[:, c] - 1] = 0.5 * (field[ix[:, a] - 1] + field[ix[:, b] - 1])
    return f


def full_tensor(t6):
    """(n, 6) xx yy zz xy yz xz -> (n, 9) xx xy xz yx yy yz zx zy zz."""
    return t6[:, [0, 3, 5, 3, 1, 4, 5, 4, 2]]


def _fields(ix, x, u, dp, porosity, tx, txb, txe, flux, flags, quad, nnodev):
    nnode = u.shape[0]
    nprint = nnode if quad else nnodev
    if quad:
        xf = np.zeros((nnode, 3))
        xf[:nnodev] = x
        xq = quad_field(xf, ix, nprint, True)
    else:
        xq = np.array(x[:nprint], dtype=float)
    pts = []  # (kind, name, values)
    pts.append(("iscalar", "noIdG", np.arange(1, nprint + 1)))
    if flags[1]:
        pts.append(("vector", "desloc", np.asarray(u[:nprint, :3], dtype=float)))
    if flags[2]:
        pts.append(("scalar", "pressao", quad_field(u[:, 3], ix, nprint, quad)))
    if flags[3]:
        pts.append(("scalar", "deltaPressao", quad_field(dp, ix, nprint, quad)))
    if flags[8]:
        pts.append(("scalar", "porosity", quad_field(porosity, ix, nprint, quad)))
    if flags[7]:
        pts.append(("vector", "darcyFlux", quad_field(flux, ix, nprint, quad)))
    for flag, name, t6 in ((4, "totalStress", tx), (6, "terzaghiStress", txe),
                           (5, "BiotStress", txb)):
        if flags[flag]:
            pts.append(("tensor", name, quad_field(full_tensor(t6), ix, nprint, quad)))
    if quad:
        conn, ctype = ix[:, HEX20_VTK] - 1, 25
    else:
        conn, ctype = ix[:, :8] - 1, 12
    return xq, conn, ctype, pts


def write_results(path, ix, x, u, dp, porosity, tx, txb, txe, flux, istep, t, flags=DEFAULT_FLAGS,
                  quad=False, legacy=False, binary=False, nnodev=None):
    """Writes one results file.  ix: (numel, 21) hexa20 connectivity (1-based)
    with the material last; x: (nnodev, 3) vertex coordinates; u: (nnode, 4);
    dp, porosity: (nnode); tx, txb, txe: (nnode, 6); flux: (nnode, 3).
    Fields whose flag is off may be None."""
    if binary and not legacy:
        raise VtkError("binary VTU is not written by the reference (head_vtu: nao implementado)")
    if flags[9] or flags[10]:
        raise VtkError("pconsolidation / elplastic belong to the plastic path")
    ix = np.asarray(ix)
    numel = ix.shape[0]
    nnodev = x.shape[0] if nnodev is None else nnodev
    xq, conn, ctype, pts = _fields(ix, x, u, dp, porosity, tx, txb, txe, flux, flags, quad,
                                   nnodev)
    mats = ix[:, 20]
    if legacy and binary:
        data = _legacy_binary(xq, conn, ctype, mats, pts, istep, t)
        with open(path, "wb") as f:
            f.write(data)
        return path
    lines = (_legacy_ascii if legacy else _vtu_ascii)(xq, conn, ctype, mats, pts, istep, t)
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")
    return path


def _e155(v):
    return fortran_e(v, 15, 5)


def _e155e3(v):
    return fortran_e(v, 15, 5, 3)


def _i10(v):
    return fortran_i(v, 10)


def _scalar_head(name15, kind):
    """'(a,1x,a15,1x,a8,1x,i3)' of 'SCALARS', the name, the type, 1."""
    return ("SCALARS " + fortran_a(name15, 15) + " " + fortran_a(kind, 8) + " "
            + fortran_i(1, 3))


def _legacy_ascii(xq, conn, ctype, mats, pts, istep, t):
    L = ["# vtk DataFile Version 3.0", HEAD, "ASCII", "DATASET UNSTRUCTURED_GRID",
         "FIELD FieldData 4", "TIME_S 1 1 double", fortran_f(t, 16, 6),
         "TIME_H 1 1 double", fortran_f(t / 3600.0, 16, 6),
         "TIME_D 1 1 double", fortran_f(t / (24.0 * 3600.0), 16, 6),
         "CYCLE 1 1 int", fortran_i(istep, 9)]
    n, numel, nen = xq.shape[0], conn.shape[0], conn.shape[1]
    L.append("POINTS " + _i10(n) + " double")
    L.extend("".join(_e155(v) for v in r) for r in xq)
    L.append("CELLS " + _i10(numel) + _i10(numel * (nen + 1)))
    L.extend(_i10(nen) + "".join(_i10(v) for v in r) for r in conn)
    L.append("CELL_TYPES " + _i10(numel))
    L.extend([fortran_i(ctype, 3)] * numel)
    L.append("CELL_DATA " + fortran_i(numel, 9))
    for name, vals in (("mat", mats), ("elIdG", np.arange(1, numel + 1))):
        L.append(list_record("SCALARS ", name.ljust(15), " int ", 1))
        L.append("LOOKUP_TABLE default")
        L.extend(_i10(v) for v in vals)
    L.append("POINT_DATA " + fortran_i(n, 9))
    for kind, name, vals in pts:
        nm = name.ljust(15)
        if kind == "iscalar":
            L.append(_scalar_head(nm, "int    "))
            L.append("LOOKUP_TABLE default")
            L.extend(_i10(v) for v in vals)
        elif kind == "scalar":
            L.append(_scalar_head(nm, "double "))
            L.append("LOOKUP_TABLE default")
            L.extend(_e155e3(v) for v in vals)
        elif kind == "vector":
            L.append("VECTORS " + nm + " double")
            L.extend("".join(_e155e3(v) for v in r) for r in vals)
        else:
            L.append("TENSORS " + nm + " double")
            L.extend("".join(_e155e3(v) for v in r) for r in vals)
    return L


def _legacy_binary(xq, conn, ctype, mats, pts, istep, t):
    out = []
    a = out.append

    def txt(s):
        a(s.encode("ascii"))

    def ints(v):
        a(np.asarray(v, dtype=">i4").tobytes())

    def reals(v):
        a(np.asarray(v, dtype=">f8").tobytes())

    txt("# vtk DataFile Version 3.0\n")
    txt(HEAD)
    txt("\nBINARY\n")
    txt("DATASET UNSTRUCTURED_GRID\n")
    txt("FIELD FieldData 4\n")
    txt("TIME_S 1 1 double\n")
    reals([t])
    txt("TIME_H 1 1 double\n")
    reals([t / 3600.0])
    txt("TIME_D 1 1 double\n")
    reals([t / (24.0 * 3600.0)])
    txt("CYCLE 1 1 int\n")
    ints([istep])
    n, numel, nen = xq.shape[0], conn.shape[0], conn.shape[1]
    txt("POINTS " + fortran_i(n, 15) + " double\n")
    reals(xq)
    txt("\nCELLS " + fortran_i(numel, 15) + fortran_i(numel * (nen + 1), 15) + "\n")
    ints(np.column_stack([np.full(numel, nen), conn]))
    txt("\nCELL_TYPES " + fortran_i(numel, 15) + "\n")
    ints(np.full(numel, ctype))
    txt("\nCELL_DATA " + fortran_i(numel, 8) + "\n")
    for name, vals in (("mat", mats), ("elIdG", np.arange(1, numel + 1))):
        txt("SCALARS " + name.ljust(15) + " int" + fortran_i(1, 8) + "\n")
        txt("LOOKUP_TABLE default\n")
        ints(vals)
    txt("\n POINT_DATA " + fortran_i(n, 8) + "\n")
    for kind, name, vals in pts:
        nm = name.ljust(15)
        if kind == "iscalar":
            txt(" SCALARS " + nm + " int " + fortran_i(1, 6) + "\n")
            txt(" LOOKUP_TABLE default \n")
            ints(vals)
        elif kind == "scalar":
            txt("\nSCALARS " + nm + " double" + fortran_i(1, 6) + "\n")
            txt("\nLOOKUP_TABLE default \n")
            reals(vals)
        elif kind == "vector":
            txt(" VECTORS " + nm + " double\n")
            reals(vals)
        else:
            txt(" TENSORS " + nm + " double\n")
            reals(vals)
    return b"".join(out)


def _vtu_ascii(xq, conn, ctype, mats, pts, istep, t):
    n, numel, nen = xq.shape[0], conn.shape[0], conn.shape[1]
    L = ['<?xml version="1.0"?>',
         '<VTKFile type="UnstructuredGrid" version="0.1" byte_order="BigEndian">',
         "<UnstructuredGrid>", "<FieldData>",
         '<DataArray type="Int32" Name="CYCLE" NumberOfTuples="1" format="ascii">',
         list_record(int(istep)), "</DataArray>",
         '<DataArray type="Float64" Name="TIME_S" NumberOfTuples="1" format="ascii">',
         list_record(float(t)), "</DataArray>",
         '<DataArray type="Float64" Name="TIME_H" NumberOfTuples="1" format="ascii">',
         list_record(t / 3600.0), "</DataArray>",
         '<DataArray type="Float64" Name="TIME_D" NumberOfTuples="1" format="ascii">',
         list_record(t / (24.0 * 3600.0)), "</DataArray>",
         "</FieldData>",
         f'<Piece NumberOfPoints="{fortran_i(n, 15)}" NumberOfCells="{fortran_i(numel, 15)}">',
         "<Points>", '<DataArray type="Float64" NumberOfComponents ="3" format="ascii">']
    L.extend("".join(_e155(v) for v in r) for r in xq)
    L += ["</DataArray>", "</Points>", "<Cells> ",
          '<DataArray type="Int32" Name="connectivity" format="ascii">']
    L.extend("".join(_i10(v) for v in r) for r in conn)
    L += ["</DataArray>", '<DataArray type="Int64" Name="offsets" format="ascii">']
    L.extend(fortran_i(nen * (k + 1), 20) for k in range(numel))
    L += ["</DataArray>", '<DataArray type="UInt8" Name="types" format="ascii">']
    L.extend([fortran_i(ctype, 3)] * numel)
    L += ["</DataArray>", "</Cells>", '<CellData Scalars="scalars">']
    for name, vals in (("mat", mats), ("elIdG", np.arange(1, numel + 1))):
        L.append(f'<DataArray type="Int32" NumberOfComponents="1" Name="{name}" format="ascii">')
        L.extend(fortran_i(v, 9) for v in vals)
        L.append("</DataArray>")
    L += ["</CellData>", "<PointData>"]
    for kind, name, vals in pts:
        if kind == "iscalar":
            L.append(f'<DataArray type="Int32" NumberOfComponents=" 1" Name="{name}" '
                     'format="ascii">')
            L.extend(_i10(v) for v in vals)
        elif kind == "scalar":
            L.append(f'<DataArray type="Float64" NumberOfComponents=" 1" Name="{name}" '
                     'format="ascii">')
            L.extend(_e155e3(v) for v in vals)
        else:
            nc = " 3" if kind == "vector" else "9"
            L.append(f'<DataArray type="Float64" NumberOfComponents="{nc}" Name="{name}" '
                     'format="ascii">')
            L.extend("".join(_e155e3(v) for v in r) for r in vals)
        L.append("</DataArray>")
    L += ["</PointData>", "</Piece>", "</UnstructuredGrid>", "</VTKFile>"]
    return L
