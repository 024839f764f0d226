"""Element result queries on the device (nla3d_gpu_element_query: Element::getScalar /
getVector / getTensor of SOLID81 and PLANE41) against the oracle's restatement, every query
each element answers, at GP_MEAN and at single Gauss points, on jittered meshes at a random
solution.  Tolerance (FP64): max |gpu - ref| <= 1e-12 max |ref| per query."""
import numpy as np
import pytest

import nla3d_amd
from nla3d_amd import meshgen
from oracle import pyoracle as po
from test_gpu_plane41 import models as plane_models
from test_gpu_plane41 import random_state as plane_state

pytestmark = pytest.mark.gpu

SOLID_QUERIES = [(0, 1), (0, 3), (0, 4), (0, 5), (1, 1), (2, 1), (2, 2), (2, 3), (2, 4)]
PLANE_QUERIES = [(0, 1), (2, 1), (2, 2), (2, 3), (2, 4)]
MAT = (2, [0.4, -0.05, 0.02, 0.1, 50.0])


def close(g, r):
    return np.abs(g - r).max() <= 1e-12 * max(np.abs(r).max(), 1e-300)


@pytest.mark.parametrize("gp", [-1, 0, 6])
def test_solid81_queries(gpu_lib, gp):
    print(f"[gpu] solid81 element queries gp {gp}", flush=True)
    xyz, conn, (I, J, K) = meshgen.hex_box(3, 2, 2, jitter=0.1, seed=9)
    m = nla3d_amd.GpuModel(0)
    m.set_nodes(xyz)
    m.add_elements(nla3d_amd.SOLID81, conn)
    m.set_material(*MAT)
    fixes = [(int(n), d, 0.0) for n in np.nonzero(K == 0)[0] + 1 for d in range(3)]
    for n, d, v in fixes:
        m.add_fix(n, d, v)
    m.init_solution()
    o = po.OracleModel(xyz, [(1, c) for c in conn], material=MAT, fixes=fixes).setup()
    rng = np.random.default_rng(13)
    U = np.zeros(o.nAll)
    for n in range(len(xyz)):
        for d in range(3):
            U[o.nodeEq[n][d] - 1] = 0.04 * rng.standard_normal()
    for e in range(len(conn)):
        U[o.elemEq[e][po.HYDRO_PRESSURE] - 1] = 0.3 * rng.standard_normal()
    m.set_solution(U)
    for kind, query in SOLID_QUERIES:
        g = m.element_query(len(conn), kind, query, gp)
        for e, c in enumerate(conn):
            Ue = [U[o.nodeEq[n - 1][d] - 1] for n in c for d in range(3)]
            r = po.element_query(po.SOLID81, xyz[np.array(c) - 1], Ue, U[o.elemEq[e][po.HYDRO_PRESSURE] - 1], *MAT,
                                 kind, query, gp)
            assert close(g[e], r), (kind, query, e, g[e], r)


@pytest.mark.parametrize("gp", [-1, 2])
def test_plane41_queries(gpu_lib, gp):
    print(f"[gpu] plane41 element queries gp {gp}", flush=True)
    m, o, xyz, conn = plane_models(MAT)
    U = plane_state(o, xyz, 0.03, seed=17)
    m.set_solution(U)
    for kind, query in PLANE_QUERIES:
        g = m.element_query(len(conn), kind, query, gp)
        for e, c in enumerate(conn):
            Ue = [U[o.nodeEq[n - 1][d] - 1] for n in c for d in range(2)]
            r = po.element_query(po.PLANE41, xyz[np.array(c) - 1], Ue, U[o.elemEq[e][po.HYDRO_PRESSURE] - 1], *MAT,
                                 kind, query, gp)
            assert close(g[e], r), (kind, query, e, g[e], r)
    for kind, query in ((0, 3), (0, 5), (1, 1)):
        with pytest.raises(nla3d_amd.Nla3dGpuError):
            m.element_query(len(conn), kind, query, gp)
    with pytest.raises(nla3d_amd.Nla3dGpuError):
        m.element_query(len(conn), 2, 1, 4)  # Gauss point out of range


def test_vtk_file_of_a_solved_plane_model(gpu_lib, tmp_path):
    """write_vtk on a device model after a Newton load step: displacements, reactions and the
    device queries appear under their labels (nla3d_amd.vtk, VtkProcessor.cpp:110-297)"""
    from nla3d_amd.vtk import _num, write_vtk
    print("[gpu] vtk output of a PLANE41 model", flush=True)
    m, o, xyz, conn = plane_models((1, [1.0, 500.0]), top_uy=-0.1)
    m.solve_load_step(1.0, m.newton_defaults(criterion=1, convergence=1e-10, krylov_tol=1e-13,
                                             krylov_max_iter=20000, max_iterations=20))
    path = tmp_path / "plate.vtk"
    write_vtk(str(path), m)
    lines = path.read_text().split("\n")
    U, R = m.get_solution(), m.get_R()
    ne, ee = m.equation_numbers()
    d = lines.index("VECTORS disp float")
    r = lines.index("VECTORS reaction float")
    for n in range(len(xyz)):
        assert lines[d + 1 + n] == " ".join(_num(U[ne[n, k] - 1]) if ne[n, k] else "0" for k in range(3))
        assert lines[r + 1 + n] == " ".join(_num(R[ne[n, k] - 1]) if ne[n, k] else "0" for k in range(3))
    c = lines.index("SCALARS COUCHY_XY float 1")
    want = m.element_query(len(conn), 2, 1, -1)[:, 1]
    assert lines[c + 2:c + 2 + len(conn)] == [_num(v) for v in want]
    assert "SCALARS IC float 1" not in lines and "VECTORS IC float" not in lines
