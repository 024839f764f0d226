"""Constant linear MPCs (FEStorage::addMpc, FEStorage.cpp:175-179; ANSYS CE records,
FEReaders.cpp:549-573): the library's numbering and sparsity with such rows equal the
oracle's bit for bit; the oracle's Newton solution satisfies every constraint; the CE
records of the reference's 3d_damper fixture are read as the reference reads them."""
import os

import numpy as np

import nla3d_amd
from nla3d_amd import meshgen
from nla3d_amd.cdb import read_cdb
from oracle.pyoracle import OracleModel

GOLD = os.path.join(os.path.dirname(__file__), "golden", "3d_damper")
MAT = (3, [0.5, 0.1, 100.0])  # Mooney-Rivlin


def ce_case(n=3):
    """n^3 cube, bottom fixed, top UZ = 0.15, the x = 0 face (inside in z) tied by
    UX + UZ = 0 as the damper's CEs, and one three-term constraint"""
    xyz, conn, (I, J, K) = meshgen.hex_box(n, n, n, jitter=0.1, seed=11)
    nodes = np.arange(1, len(xyz) + 1)
    fixes = [(int(v), d, 0.0) for v in nodes[K == 0] for d in range(3)] + [(int(v), 2, 0.15) for v in nodes[K == n]]
    side = nodes[(I == 0) & (K > 0) & (K < n)]
    mpcs = [([(int(s), 0, 1.0), (int(s), 2, 1.0)], 0.0) for s in side]
    a, b, c = (int(v) for v in nodes[(I == n) & (J == 1) & (K > 0) & (K < n)][:1].tolist() +
               nodes[(I == n) & (J == 2) & (K > 0) & (K < n)][:2].tolist())
    mpcs.append(([(a, 0, 1.0), (b, 0, -1.0), (c, 1, 0.5)], 0.0))
    return xyz, conn, fixes, mpcs


def library_model(xyz, conn, fixes, mpcs, device=-1):
    m = nla3d_amd.GpuModel(device)
    m.set_nodes(xyz)
    m.add_elements(nla3d_amd.SOLID81, conn)
    m.set_material(*MAT)
    for node, d, v in fixes:
        m.add_fix(node, d, v)
    for terms, b in mpcs:
        m.add_mpc(terms, b)
    m.init_solution()
    return m


def oracle_model(xyz, conn, fixes, mpcs):
    return OracleModel(xyz, [(1, list(map(int, c))) for c in conn], material=MAT, fixes=fixes, mpcs=mpcs).setup()


def test_numbering_and_sparsity_with_mpcs():
    xyz, conn, fixes, mpcs = ce_case()
    m = library_model(xyz, conn, fixes, mpcs)
    om = oracle_model(xyz, conn, fixes, mpcs)
    assert m.sizes["nMpc"] == om.nMpc == len(mpcs)
    for blk in (1, 12, 2):
        io, cols = m.pattern(blk)
        assert np.array_equal(io, np.asarray(om.blocks[blk].iofeir)), blk
        assert np.array_equal(cols, np.asarray(om.blocks[blk].columns)), blk


def test_oracle_newton_satisfies_the_constraints():
    xyz, conn, fixes, mpcs = ce_case()
    om = oracle_model(xyz, conn, fixes, mpcs)
    om.newton_init()
    for t in (0.5, 1.0):
        om.solve_load_step(t, criterion=1, convergence=1e-10, max_iterations=20)
    U = om.U
    scale = np.abs(U[:om.nDofs]).max()
    for terms, b in mpcs:
        s = sum(c * U[om.nodeEq[n - 1][d] - 1] for n, d, c in terms)
        assert abs(s - b) <= 1e-12 * scale
    # the constraints change the answer (a solution that ignores them violates them)
    free = oracle_model(xyz, conn, fixes, [])
    free.newton_init()
    for t in (0.5, 1.0):
        free.solve_load_step(t, criterion=1, convergence=1e-10, max_iterations=20)
    terms = mpcs[0][0]
    s = sum(c * free.U[free.nodeEq[n - 1][d] - 1] for n, d, c in terms)
    assert abs(s) > 1e-6 * scale


def test_damper_ce_records():
    d = read_cdb(os.path.join(GOLD, "model.cdb.gz"))
    assert len(d["nodes"]) == 9355 and len(d["cells"]) == 7016
    assert len(d["mpcs"]) == 497
    for terms, b in d["mpcs"]:
        assert b == 0.0 and len(terms) == 2
        (n0, d0, c0), (n1, d1, c1) = terms
        assert n0 == n1 and (d0, d1) == (0, 2) and c0 == c1 == 1.0
    assert len(d["comps"]["TOP_SIDE"]) == 338
    assert sorted(set(v for n, dd, v in d["fixes"] if dd == 1 and n in set(d["comps"]["TOP_SIDE"]))) == [-120.0]
