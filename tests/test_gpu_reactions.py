"""Reactions on the device (FESolver.cpp:464-467) through the reference's rigid-body
test: block_ROTX.cdb, neo-Hookean G = 1, K = 500, TOP_SIDE tied to MASTER_NODE,
master ROTX driven to 0.7853 over 20 load steps; the master's ROTX reaction after
each load step against ANSYS's curve (mean abs < 1e-4, test/CMakeLists.txt:52-55)
and against the oracle's curve (1e-8 of the largest value)."""
import os

import numpy as np
import pytest

import nla3d_amd
from nla3d_amd.cdb import read_cdb
from oracle.pyoracle import load_steps

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(__file__), "golden", "rigid_body_mpc")


def test_block_rotx_reaction_curve(gpu_lib):
    print("[gpu] block_ROTX reaction curve", flush=True)
    from test_oracle_reactions import ROTX, block_rotx_oracle, reference_curve
    d = read_cdb(os.path.join(GOLD, "block_ROTX.cdb"))
    master = d["comps"]["MASTER_NODE"][0]
    m = nla3d_amd.GpuModel(0)
    m.set_nodes(d["nodes"])
    m.add_elements(nla3d_amd.SOLID81, [n for t, n in d["cells"] if t == 1])
    m.set_material(nla3d_amd.NEO_HOOKEAN, [1.0, 500.0])
    for n, dof, v in d["fixes"]:
        m.add_fix(n, dof, v)
    m.add_rigid_body_mpc(master, d["comps"]["TOP_SIDE"], (0, 1, 2))
    m.init_solution()
    opts = m.newton_defaults(criterion=0, convergence=1e-3, max_iterations=15, krylov_tol=1e-13,
                             krylov_max_iter=5000)
    o, _ = block_rotx_oracle()
    o.newton_init()
    curve, ocurve = [0.0], [0.0]
    for t in load_steps(20):
        st = m.solve_load_step(t, opts)
        curve.append(m.reaction(master, ROTX))
        o.solve_load_step(t, criterion=0, convergence=1e-3, max_iterations=15)
        ocurve.append(o.reaction(master, ROTX))
    curve, ocurve, ref = np.array(curve), np.array(ocurve), reference_curve()
    print("GPU curve", curve[-3:], "mean abs vs ANSYS", np.abs(curve - ref).mean(), flush=True)
    assert np.abs(curve - ocurve).max() <= 1e-8 * np.abs(ocurve).max()
    assert np.abs(curve - ref).mean() < 1e-4
