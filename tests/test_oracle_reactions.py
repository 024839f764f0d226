"""Reaction recovery (FESolver.cpp:464-467, ReactionProcessor.cpp:91-117) in the
oracle, pinned by the reference's rigid-body test: one SOLID81 element,
neo-Hookean G = 1, K = 500, the TOP_SIDE nodes tied by a RigidBodyMpc to
MASTER_NODE (node 9), whose ROTX is driven to 0.7853 over 20 load steps; the
ROTX reaction of the master after every load step against ANSYS's curve
(test/rigid_body_mpc/reference_MOMZ_reaction.txt), mean abs difference < 1e-4
as test/CMakeLists.txt:52-55 (compareCurves, main.cpp:328-339)."""
import os

import numpy as np

from nla3d_amd.cdb import read_cdb
from oracle.pyoracle import OracleModel, load_steps

GOLD = os.path.join(os.path.dirname(__file__), "golden", "rigid_body_mpc")
ROTX = 3


def reference_curve():
    """readRefCurveData (main.cpp:342-361): third column, consecutive repeats dropped"""
    out = []
    with open(os.path.join(GOLD, "reference_MOMZ_reaction.txt")) as f:
        next(f)
        for line in f:
            if line.split():
                v = float(line.split()[2])
                if not out or v != out[-1]:
                    out.append(v)
    return np.array(out)


def block_rotx_oracle():
    d = read_cdb(os.path.join(GOLD, "block_ROTX.cdb"))
    master = d["comps"]["MASTER_NODE"][0]
    o = OracleModel(d["nodes"], [(1, n) for t, n in d["cells"] if t == 1], material=(1, [1.0, 500.0]),
                    fixes=d["fixes"], rigid_bodies=[(master, d["comps"]["TOP_SIDE"], [0, 1, 2])]).setup()
    return o, master


def test_rigid_body_rotx_reaction_curve_matches_ansys(orc):
    o, master = block_rotx_oracle()
    o.newton_init()
    curve = [0.0]
    for t in load_steps(20):
        o.solve_load_step(t, criterion=0, convergence=1e-3, max_iterations=15)
        curve.append(o.reaction(master, ROTX))
    ref = reference_curve()
    assert len(curve) == len(ref) == 21
    err = np.abs(np.array(curve) - ref).mean()
    assert err < 1e-4, err
