"""PLANE41 (plane-strain u/p quadrilateral, elements/PLANE41.cpp) in the oracle, pinned
without reference fixtures: the element tangent equals minus the derivative of the element
forces by central differences for all three materials; rigid motion leaves no residual;
the nodal forces sum to zero; a Newton solve of a small mesh converges quadratically."""
import numpy as np
import pytest

from nla3d_amd import meshgen
from oracle import pyoracle as po

X = np.array([[0.0, 0.0, 0.0], [1.1, 0.05, 0.0], [1.0, 0.9, 0.0], [-0.05, 1.0, 0.0]])
MATS = [(1, [1.0, 500.0]), (2, [0.4, 0.02, 0.001, 0.1, 200.0]), (3, [0.5, 0.1, 100.0])]


def full(out):
    K = np.zeros((9, 9))
    for s, (a, b) in enumerate(po.plane41_slot_pairs()):
        K[a, b] = K[b, a] = out[s]
    return K, out[45:]


@pytest.mark.parametrize("mat", MATS)
def test_tangent_is_minus_force_derivative(mat):
    rng = np.random.default_rng(3)
    U, p = 0.05 * rng.standard_normal(8), 0.3
    K, F = full(po.plane41_element(X, U, p, *mat))
    h = 1e-6
    Jfd = np.zeros((9, 9))
    for j in range(9):
        Up, Um, pp, pm = U.copy(), U.copy(), p, p
        if j < 8:
            Up[j] += h
            Um[j] -= h
        else:
            pp += h
            pm -= h
        Jfd[:, j] = -(full(po.plane41_element(X, Up, pp, *mat))[1] - full(po.plane41_element(X, Um, pm, *mat))[1]) / (2 * h)
    assert np.abs(Jfd - K).max() <= 1e-7 * np.abs(K).max()
    assert abs(F[0:8:2].sum()) <= 1e-12 * np.abs(F).max() and abs(F[1:8:2].sum()) <= 1e-12 * np.abs(F).max()


def test_rigid_motion_has_no_residual():
    th = 0.4
    R = np.array([[np.cos(th), -np.sin(th)], [np.sin(th), np.cos(th)]])
    U = ((R - np.eye(2)) @ X[:, :2].T).T.ravel() + np.tile([0.3, -0.2], 4)
    assert np.abs(full(po.plane41_element(X, U, 0.0, 3, [0.5, 0.1, 100.0]))[1]).max() <= 1e-14


def test_newton_converges_quadratically():
    xyz, conn, (I, J) = meshgen.quad_grid(4, 3, 1.0, 0.75)
    nodes = np.arange(1, len(xyz) + 1)
    fixes = [(int(n), d, 0.0) for n in nodes[J == 0] for d in (0, 1)]
    fixes += [(int(n), 1, -0.1) for n in nodes[J == 3]] + [(int(n), 0, 0.0) for n in nodes[J == 3]]
    om = po.OracleModel(xyz, [(po.PLANE41, list(map(int, c))) for c in conn], material=(1, [1.0, 500.0]),
                        fixes=fixes).setup()
    om.newton_init()
    for t in (0.5, 1.0):
        assert om.solve_load_step(t, criterion=1, convergence=1e-10, max_iterations=20) <= 5


def test_a2000_damper_reaction_curve_matches_ansys():
    """the reference's a2000_damper benchmark (test/CMakeLists.txt:142-146) on the oracle:
    3,101 PLANE41 elements, Neo-Hookean G = 1, K = 500, 20 load steps; the DISP_BC UY
    reaction curve against ANSYS's (readRefCurveData / compareCurves, main.cpp:326-361:
    mean absolute difference, the reference's threshold 0.2)"""
    import os

    from nla3d_amd.cdb import read_cdb

    gold = os.path.join(os.path.dirname(__file__), "golden", "a2000_damper")
    d = read_cdb(os.path.join(gold, "model.cdb.gz"))
    om = po.OracleModel(np.array(d["nodes"]), [(po.PLANE41, list(n)) for t, n in d["cells"] if t == 1],
                        material=(1, [1.0, 500.0]), fixes=d["fixes"]).setup()
    om.newton_init()
    eqs = [om.nodeEq[n - 1][po.UY] - 1 for n in d["comps"]["DISP_BC"]]
    curve = [0.0]
    for t in po.load_steps(20):
        om.solve_load_step(t, criterion=0, convergence=1e-3, max_iterations=15)
        curve.append(float(om.R[eqs].sum()))
    ref = a2000_curve(gold)
    assert len(curve) == len(ref)
    assert np.abs(np.array(curve) - ref).mean() < 0.2


def a2000_curve(gold):
    """readRefCurveData (main.cpp:342-361): third column, consecutive repeats dropped"""
    out = []
    with open(gold + "/loading_curve_ansys.txt") as f:
        next(f)
        for line in f:
            if line.split():
                v = float(line.split()[2])
                if not out or v != out[-1]:
                    out.append(v)
    return np.array(out)


def test_a2000_cdb_records():
    """readCdbFile on the a2000_damper model (FEReaders.cpp:321-729): the mesh, the three node
    components and the D records the benchmark runs on"""
    import os

    from nla3d_amd.cdb import read_cdb

    d = read_cdb(os.path.join(os.path.dirname(__file__), "golden", "a2000_damper", "model.cdb.gz"))
    assert len(d["nodes"]) == 3225 and len(d["cells"]) == 3101
    assert all(t == 1 and len(n) == 4 for t, n in d["cells"])
    assert np.abs(np.array(d["nodes"])[:, 2]).max() == 0.0
    comps = {k: set(v) for k, v in d["comps"].items()}
    assert {k: len(v) for k, v in comps.items()} == {"DISP_BC": 79, "FIX_BC": 73, "SYM_BC": 7}
    fixes = {(n, dof): v for n, dof, v in d["fixes"]}
    assert len(fixes) == len(d["fixes"]) == 308
    assert all(fixes.get((n, 1)) == -9.0 for n in comps["DISP_BC"])
    assert all(fixes.get((n, 0)) == 0.0 and fixes.get((n, 1)) == 0.0 for n in comps["FIX_BC"])
    assert all(fixes.get((n, 0)) == 0.0 for n in comps["SYM_BC"])
    assert not d["mpcs"]
