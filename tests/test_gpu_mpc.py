"""Constant linear MPCs on the device: a Newton solve with CE-style constraints equals the
oracle's load step by load step (1e-8), with block Jacobi and with the Schur multigrid; and
the reference's 3d_damper benchmark end to end (test/CMakeLists.txt:149-153): 7,016
SOLID81 elements, 497 CE constraints, Neo-Hookean G = 10, K = 5000, TOP_SIDE pushed
120 units over 20 load steps; the TOP_SIDE UY reaction curve against ANSYS's
(mean absolute difference < 3.0, the reference's threshold)."""
import os

import numpy as np
import pytest

import nla3d_amd
from nla3d_amd.cdb import read_cdb
from oracle.pyoracle import load_steps
from test_oracle_mpc import ce_case, library_model, oracle_model

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(__file__), "golden", "3d_damper")


@pytest.mark.parametrize("pc", [0, 2])
def test_newton_with_mpcs_matches_oracle(gpu_lib, pc):
    print(f"[gpu] newton with constant MPCs, preconditioner {pc}", flush=True)
    xyz, conn, fixes, mpcs = ce_case(4)
    m = library_model(xyz, conn, fixes, mpcs, device=0)
    om = oracle_model(xyz, conn, fixes, mpcs)
    om.newton_init()
    opts = m.newton_defaults(criterion=1, convergence=1e-10, krylov_tol=1e-13, krylov_max_iter=20000,
                             max_iterations=20, preconditioner=pc)
    for t in (0.5, 1.0):
        st = m.solve_load_step(t, opts)
        om.solve_load_step(t, criterion=1, convergence=1e-10, max_iterations=20)
        U, Uo = m.get_solution(), om.U
        assert np.abs(U - Uo).max() <= 1e-8 * np.abs(Uo).max(), t
        R, Ro = m.get_R()[:om.nC], om.R[:om.nC]
        assert np.abs(R - Ro).max() <= 1e-8 * np.abs(Ro).max(), t
    assert st.precond_fallbacks == 0


def damper_curve():
    """readRefCurveData (main.cpp:342-361): third column, consecutive repeats dropped"""
    out = []
    with open(os.path.join(GOLD, "loading_curve_ansys.txt")) as f:
        next(f)
        for line in f:
            if line.split():
                v = float(line.split()[2])
                if not out or v != out[-1]:
                    out.append(v)
    return np.array(out)


def test_3d_damper_reaction_curve(gpu_lib):
    print("[gpu] 3d_damper benchmark", flush=True)
    d = read_cdb(os.path.join(GOLD, "model.cdb.gz"))
    m = nla3d_amd.GpuModel(0)
    m.set_nodes(d["nodes"])
    m.add_elements(nla3d_amd.SOLID81, [n for t, n in d["cells"]])
    m.set_material(nla3d_amd.NEO_HOOKEAN, [10.0, 5000.0])
    for n, dof, v in d["fixes"]:
        m.add_fix(n, dof, v)
    for terms, b in d["mpcs"]:
        m.add_mpc(terms, b)
    m.init_solution()
    top = d["comps"]["TOP_SIDE"]
    opts = m.newton_defaults(criterion=0, convergence=1e-3, max_iterations=15, krylov_tol=1e-10,
                             krylov_max_iter=20000, preconditioner=2)
    ne, _ = m.equation_numbers()
    eqs = np.array([int(ne[n - 1, 1]) for n in top]) - 1
    curve = [0.0]
    st = nla3d_amd.NewtonStats()
    for k, t in enumerate(load_steps(20)):
        m.solve_load_step(t, opts, st)
        R = m.get_R()
        curve.append(float(R[eqs].sum()))
        print(f"load step {k + 1}: t = {t:.3f}, Newton {st.last_step_iterations}, "
              f"Krylov {st.last_krylov_iterations}, fallbacks {st.precond_fallbacks}, reaction {curve[-1]:.2f}",
              flush=True)
    curve, ref = np.array(curve), damper_curve()
    err = np.abs(curve - ref).mean()
    print("reaction curve", np.round(curve, 2), "ANSYS", ref, "mean abs", err, flush=True)
    assert len(curve) == len(ref)
    assert err < 3.0
    # every solve ran the multigrid (a broken cycle would be hidden by the block-Jacobi fallback)
    assert st.precond_fallbacks == 0
