"""Legacy-VTK output of a model's current state: the library's counterpart of nla3d's
VtkProcessor (VtkProcessor.cpp:110-297) with every result the u/p elements answer
(writeAllResults / revealAllResults, VtkProcessor.cpp:300-369).

The file has the reference's layout and labels: POINTS (undeformed positions), CELLS
(0-based node numbers), CELL_TYPES (12 hexahedron, 9 quadrilateral, 3 line), POINT_DATA
with the "disp" and "reaction" vectors when a displacement DoF is in use and one scalar per
other node DoF type in use, CELL_DATA with one scalar per element DoF type in use and then the
element queries (scalars by their scalarQuery label, vectors by their vectorQuery label,
tensors as six scalars <label>_XX ... <label>_ZZ), numbers as std::ostream prints them by
default (six significant digits).  The element queries are evaluated on the device
(nla3d_gpu_element_query, GP_MEAN); elements that do not answer a query get 0, as the
reference's zero-initialised arrays do.

Differences from the reference: VtkProcessor::process (VtkProcessor.cpp:58-65) truncates the
per-load-step file without writing to it, so only the file before the solution has content;
here every call writes the whole state.  PLANE41's IC query indexes past its argument
(PLANE41.cpp:205-207) and is left out."""
import numpy as np

from . import PLANE41, QUADTH, SOLID81, SURFACE_LINETH

DOF_LABELS = ["UX", "UY", "UZ", "ROTX", "ROTY", "ROTZ", "HYDRO_PRESSURE", "TEMP"]  # Dof.cpp:6-7
SCALAR_LABELS = {1: "S_P", 2: "W", 3: "WU", 4: "WP", 5: "VOL"}  # query.h:19
VECTOR_LABELS = {1: "IC", 2: "FLUX", 3: "GRADT"}  # query.h:34
TENSOR_LABELS = {1: "COUCHY", 2: "PK2", 3: "E", 4: "C", 5: "TSTRAIN"}  # query.h:55
COMPONENTS = ["XX", "XY", "XZ", "YY", "YZ", "ZZ"]  # solidmech.h:40
CELL_TYPES = {SOLID81: 12, PLANE41: 9, QUADTH: 9, SURFACE_LINETH: 3}  # VtkProcessor.cpp:145-166
# the queries each u/p element answers: (scalar, vector, tensor) query numbers
ANSWERS = {SOLID81: ((1, 3, 4, 5), (1,), (1, 2, 3, 4)), PLANE41: ((1,), (), (1, 2, 3, 4))}


def _num(x):
    """a double as std::ostream << prints it by default (%g, six significant digits)"""
    return f"{float(x):g}"


def write_vtk(path, model, results=True):
    """Write `model` (an initialised nla3d_amd.GpuModel) at its current solution to `path`.
    results=False leaves out the element queries (VtkProcessor without writeAllResults)."""
    xyz = np.asarray(model.xyz, dtype=np.float64)
    cells = [(kind, row) for kind, conn in model.cells for row in conn]
    U = model.get_solution()
    R = model.get_R()
    ne, ee = model.equation_numbers()
    nn, en = len(xyz), len(cells)
    out = ["# vtk DataFile Version 2.0", "This file is generated by nla3D program", "ASCII",
           "DATASET UNSTRUCTURED_GRID", f"POINTS {nn} float"]
    out += [" ".join(_num(v) for v in p) for p in xyz]
    out.append(f"CELLS {en} {sum(len(c) + 1 for _, c in cells)}")
    out += [" ".join([str(len(c))] + [str(int(n) - 1) for n in c]) for _, c in cells]
    out.append(f"CELL_TYPES {en}")
    out += [str(CELL_TYPES[k]) for k, _ in cells]

    def node_values(t):
        q = ne[:, t].astype(np.int64)
        return np.where(q > 0, U[np.maximum(q, 1) - 1], 0.0), np.where(q > 0, R[np.maximum(q, 1) - 1], 0.0)

    def scalars(name, vals):
        out.extend([f"SCALARS {name} float 1", "LOOKUP_TABLE default"])
        out.extend(_num(v) for v in vals)
        out.append("")

    def vectors(name, vals):
        out.append(f"VECTORS {name} float")
        out.extend(" ".join(_num(x) for x in v) for v in vals)
        out.append("")

    out.append(f"POINT_DATA {nn}")
    used = [t for t in range(8) if (ne[:, t] > 0).any()]
    if any(t in used for t in (0, 1, 2)):
        disp = np.zeros((nn, 3))
        react = np.zeros((nn, 3))
        for t in (0, 1, 2):
            disp[:, t], react[:, t] = node_values(t)
        vectors("disp", disp)
        vectors("reaction", react)
    for t in used:
        if t > 2:
            scalars(DOF_LABELS[t], node_values(t)[0])

    eused = [t for t in range(8) if (ee[:, t] > 0).any()]
    up = [i for i, (k, _) in enumerate(cells) if k in ANSWERS]
    answers = ANSWERS[cells[up[0]][0]] if (results and up) else ((), (), ())
    if eused or any(answers):
        out.append(f"CELL_DATA {en}")
        for t in eused:
            q = ee[:, t].astype(np.int64)
            scalars(DOF_LABELS[t], np.where(q > 0, U[np.maximum(q, 1) - 1], 0.0))
        first, n_up = (up[0], len(up)) if up else (0, 0)
        for kind, queries in enumerate(answers):
            for query in queries:
                width = (1, 3, 6)[kind]
                vals = np.zeros((en, width))
                vals[first:first + n_up] = model.element_query(n_up, kind, query, -1)
                if kind == 0:
                    scalars(SCALAR_LABELS[query], vals[:, 0])
                elif kind == 1:
                    vectors(VECTOR_LABELS[query], vals)
                else:
                    for j in range(6):
                        scalars(f"{TENSOR_LABELS[query]}_{COMPONENTS[j]}", vals[:, j])
    with open(path, "w") as f:
        f.write("\n".join(out) + "\n")
