"""GPU parity of the PLANE41 path (plane-strain u/p quadrilateral, elements/PLANE41.cpp)
against the CPU oracle.

Tolerances (FP64): per element, max|K_gpu - K_ref| / max|K_ref| <= 1e-12 over the tangent
(Kuu, Kup, Kpp) and separately over the residual (Fe); the assembled K blocks and F to 1e-12
of each block's max |value|; equation numbers exactly; a Newton solve with the reference's
criterion replaced by a 1e-10 residual ratio equals the oracle's (direct solve) to 1e-8 of
max |U| at every load step.  And the reference's a2000_damper benchmark
(test/CMakeLists.txt:142-146) end to end: the DISP_BC UY reaction curve against ANSYS's,
mean absolute difference < 0.2 (the reference's threshold)."""
import os

import numpy as np
import pytest

import nla3d_amd
from nla3d_amd import meshgen
from nla3d_amd.cdb import read_cdb
from oracle import pyoracle as po
from oracle.pyoracle import load_steps
from test_oracle_plane41 import a2000_curve

pytestmark = pytest.mark.gpu

MATS = [(1, [1.0, 500.0]), (3, [0.5, 0.1, 100.0]), (2, [0.4, -0.05, 0.02, 0.1, 50.0])]
TOL = 1e-12
GOLD = os.path.join(os.path.dirname(__file__), "golden", "a2000_damper")


def plate(nx=5, ny=4, jitter=0.15, seed=5):
    """jittered quad grid of [0,1]x[0,0.8]: interior nodes moved by up to jitter * h"""
    xyz, conn, (I, J) = meshgen.quad_grid(nx, ny, 1.0, 0.8)
    rng = np.random.default_rng(seed)
    inner = (I > 0) & (I < nx) & (J > 0) & (J < ny)
    xyz[inner, 0] += jitter / nx * rng.uniform(-1, 1, inner.sum())
    xyz[inner, 1] += jitter * 0.8 / ny * rng.uniform(-1, 1, inner.sum())
    return xyz, conn, I, J


def models(material, top_uy=0.0, nx=5, ny=4):
    xyz, conn, I, J = plate(nx, ny)
    nodes = np.arange(1, len(xyz) + 1)
    fixes = [(int(n), d, 0.0) for n in nodes[J == 0] for d in (0, 1)]
    fixes += [(int(n), 0, 0.0) for n in nodes[J == ny]] + [(int(n), 1, top_uy) for n in nodes[J == ny]]
    m = nla3d_amd.GpuModel(0)
    m.set_nodes(xyz)
    m.add_elements(nla3d_amd.PLANE41, conn)
    m.set_material(*material)
    for n, d, v in fixes:
        m.add_fix(n, d, v)
    m.init_solution()
    o = po.OracleModel(xyz, [(po.PLANE41, list(map(int, c))) for c in conn], material=material, fixes=fixes).setup()
    return m, o, xyz, conn


def random_state(o, xyz, amp, seed):
    rng = np.random.default_rng(seed)
    U = np.zeros(o.nAll)
    for node in range(len(xyz)):
        for d in range(2):
            U[o.nodeEq[node][d] - 1] = amp * rng.standard_normal()
    for e in range(len(o.elemEq)):
        U[o.elemEq[e][po.HYDRO_PRESSURE] - 1] = 0.5 * rng.standard_normal()
    return U


def rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


@pytest.mark.parametrize("mi", range(len(MATS)))
@pytest.mark.parametrize("amp", [0.0, 0.03])
def test_element_matrices(gpu_lib, mi, amp):
    print(f"[gpu] plane41 element matrices material {MATS[mi][0]} amp {amp}", flush=True)
    m, o, xyz, conn = models(MATS[mi])
    U = random_state(o, xyz, amp, seed=7 + mi)
    m.set_solution(U)
    ke = m.plane41_element_matrices(len(conn))
    for e, c in enumerate(conn):
        eqn = [o.nodeEq[n - 1][d] for n in c for d in (0, 1)]
        ref = po.plane41_element(xyz[np.array(c) - 1], U[np.array(eqn) - 1], U[o.elemEq[e][po.HYDRO_PRESSURE] - 1],
                                 *MATS[mi])
        assert rel(ke[e, :45], ref[:45]) <= TOL, (e, rel(ke[e, :45], ref[:45]))
        if np.abs(ref[45:]).max() > 0:
            assert np.abs(ke[e, 45:] - ref[45:]).max() <= TOL * np.abs(ref).max(), e


@pytest.mark.parametrize("amp", [0.0, 0.03])
def test_assembly_matches_oracle(gpu_lib, amp):
    print(f"[gpu] plane41 assembly amp {amp}", flush=True)
    m, o, xyz, conn = models(MATS[0])
    ne, ee = m.equation_numbers()
    for n in range(len(xyz)):
        for d in (0, 1, 2):
            assert int(ne[n, d]) == (o.nodeEq[n][d] if d < 2 else 0)
    for e in range(len(conn)):
        assert int(ee[e, po.HYDRO_PRESSURE]) == o.elemEq[e][po.HYDRO_PRESSURE]
    U = random_state(o, xyz, amp, seed=31)
    m.set_solution(U)
    m.assemble()
    Kref, Fref = o.assemble(U)
    for b in (1, 12, 2):
        g = m.get_K(b)
        assert g.shape == Kref[b].shape
        assert rel(g, Kref[b]) <= TOL, (b, rel(g, Kref[b]))
    F = m.get_F()
    assert rel(F, Fref) <= TOL or np.abs(Fref).max() == 0.0


@pytest.mark.parametrize("pc", [0, 2])
def test_newton_matches_oracle(gpu_lib, pc):
    print(f"[gpu] plane41 newton, preconditioner {pc}", flush=True)
    m, o, xyz, conn = models(MATS[0], top_uy=-0.15, nx=8, ny=6)
    o.newton_init()
    opts = m.newton_defaults(criterion=1, convergence=1e-10, krylov_tol=1e-13, krylov_max_iter=20000,
                             max_iterations=20, preconditioner=pc)
    for t in (0.5, 1.0):
        st = m.solve_load_step(t, opts)
        o.solve_load_step(t, criterion=1, convergence=1e-10, max_iterations=20)
        U, Uo = m.get_solution(), o.U
        assert np.abs(U - Uo).max() <= 1e-8 * np.abs(Uo).max(), t
        R, Ro = m.get_R()[:o.nC], o.R[:o.nC]
        assert np.abs(R - Ro).max() <= 1e-8 * np.abs(Ro).max(), t
    assert st.precond_fallbacks == 0


def test_plane41_entry_points_reject_other_kinds(gpu_lib):
    m, o, xyz, conn = models(MATS[0])
    with pytest.raises(nla3d_amd.Nla3dGpuError):
        m.solid81_element_matrices(len(conn))
    with pytest.raises(nla3d_amd.Nla3dGpuError):
        m.solid81_gp_state(len(conn))


def test_a2000_damper_reaction_curve(gpu_lib):
    print("[gpu] a2000_damper benchmark", flush=True)
    d = read_cdb(os.path.join(GOLD, "model.cdb.gz"))
    m = nla3d_amd.GpuModel(0)
    m.set_nodes(d["nodes"])
    m.add_elements(nla3d_amd.PLANE41, [n for t, n in d["cells"] if t == 1])
    m.set_material(nla3d_amd.NEO_HOOKEAN, [1.0, 500.0])
    for n, dof, v in d["fixes"]:
        m.add_fix(n, dof, v)
    m.init_solution()
    opts = m.newton_defaults(criterion=0, convergence=1e-3, max_iterations=15, krylov_tol=1e-10,
                             krylov_max_iter=20000, preconditioner=2)
    ne, _ = m.equation_numbers()
    eqs = np.array([int(ne[n - 1, 1]) for n in d["comps"]["DISP_BC"]]) - 1
    curve = [0.0]
    st = nla3d_amd.NewtonStats()
    for k, t in enumerate(load_steps(20)):
        m.solve_load_step(t, opts, st)
        curve.append(float(m.get_R()[eqs].sum()))
        print(f"load step {k + 1}: t = {t:.3f}, Newton {st.last_step_iterations}, "
              f"Krylov {st.last_krylov_iterations}, fallbacks {st.precond_fallbacks}, reaction {curve[-1]:.3f}",
              flush=True)
    curve, ref = np.array(curve), a2000_curve(GOLD)
    err = np.abs(curve - ref).mean()
    print("reaction curve", np.round(curve, 3), "ANSYS", ref, "mean abs", err, flush=True)
    assert len(curve) == len(ref)
    assert err < 0.2
    # every solve ran the multigrid (a broken cycle would be hidden by the block-Jacobi fallback)
    assert st.precond_fallbacks == 0
