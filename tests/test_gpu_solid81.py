"""GPU parity of the SOLID81 path against the CPU oracle.

Tolerance (FP64): per element, max|K_gpu - K_ref| / max|K_ref| <= 1e-12 over the
tangent (Kuu, Kup, Kpp) and, separately, over the residual (Fu, Fp); the
assembled K blocks and F agree to 1e-12 of each block's max |value|.  The
GPU restructures the contractions (see solid81.hip), so the comparison is not
bitwise; CSR patterns and equation numbers are compared exactly.
"""
import numpy as np
import pytest

import nla3d_amd
from nla3d_amd import meshgen
from oracle.pyoracle import OracleModel

pytestmark = pytest.mark.gpu

MATS = [(3, [0.5, 0.1, 100.0]), (1, [1.0, 500.0]), (2, [0.4, -0.05, 0.02, 0.1, 50.0])]
TOL = 1e-12


def build(n=(3, 3, 2), material=MATS[0], seed=1):
    m = nla3d_amd.GpuModel(0)
    xyz, conn, (I, J, K) = meshgen.hex_box(*n, jitter=0.1, seed=seed)
    m.set_nodes(xyz)
    m.add_elements(1, conn)
    m.set_material(*material)
    for node in np.nonzero(K == 0)[0] + 1:
        for d in range(3):
            m.add_fix(node, d, 0.0)
    m.init_solution()
    fixes = [(int(nd), d, 0.0) for nd in np.nonzero(K == 0)[0] + 1 for d in range(3)]
    o = OracleModel(xyz, [(1, c) for c in conn], material=material, fixes=fixes).setup()
    return m, o, xyz, conn


def random_state(o, xyz, amp, seed):
    rng = np.random.default_rng(seed)
    U = np.zeros(o.nAll)
    for node in range(len(xyz)):
        for d in range(3):
            U[o.nodeEq[node][d] - 1] = amp * rng.standard_normal()
    for e in range(len(o.elemEq)):
        U[o.elemEq[e][6] - 1] = 0.5 * rng.standard_normal()
    return U


def rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


@pytest.mark.parametrize("mi", range(len(MATS)))
@pytest.mark.parametrize("amp", [0.0, 0.02, 0.08])
def test_element_matrices(gpu_lib, mi, amp):
    print(f"[gpu] element matrices material={MATS[mi][0]} amp={amp}", flush=True)
    m, o, xyz, conn = build(material=MATS[mi], seed=mi + 3)
    U = random_state(o, xyz, amp, seed=17 + mi)
    m.set_solution(U)
    g = m.solid81_element_matrices(len(conn))
    r = o.solid81_elements_at(U)
    for e in range(len(conn)):
        k_idx = np.r_[0:325]
        f_idx = np.r_[325:350]
        assert rel(g[e, k_idx], r[e, k_idx]) <= TOL, (e, rel(g[e, k_idx], r[e, k_idx]))
        if np.abs(r[e, f_idx]).max() > 0:
            assert rel(g[e, f_idx], r[e, f_idx]) <= TOL, (e, rel(g[e, f_idx], r[e, f_idx]))
        else:
            assert np.abs(g[e, f_idx]).max() < 1e-14


def test_gauss_point_state(gpu_lib, orc):
    print("[gpu] gauss point state", flush=True)
    m, o, xyz, conn = build(seed=5)
    U = random_state(o, xyz, 0.05, seed=9)
    m.set_solution(U)
    st = m.solid81_gp_state(len(conn))
    for e, c in enumerate(conn):
        N, det = orc.hex8_jacob(xyz[c - 1])
        Ue = np.array([U[o.nodeEq[n - 1][d] - 1] for n in c for d in range(3)])
        O, C, S = orc.solid81_update(N, Ue, U[o.elemEq[e][6] - 1], *o.material)
        assert rel(st[e, :, 0:9], O) <= TOL
        assert rel(st[e, :, 9:15], C) <= TOL
        assert rel(st[e, :, 15:21], S) <= 1e-11


@pytest.mark.parametrize("amp", [0.0, 0.05])
def test_assembly_matches_reference_loop(gpu_lib, amp):
    print(f"[gpu] assembly amp={amp}", flush=True)
    m, o, xyz, conn = build(n=(3, 2, 3), seed=11)
    U = random_state(o, xyz, amp, seed=23)
    m.set_solution(U)
    m.assemble()
    Kref, Fref = o.assemble(U)
    for b in (1, 12, 2):
        g = m.get_K(b)
        assert g.shape == Kref[b].shape
        assert rel(g, Kref[b]) <= TOL, (b, rel(g, Kref[b]))
    F = m.get_F()
    assert rel(F, Fref) <= TOL or np.abs(Fref).max() == 0.0


def test_rigid_body_mpc_rows(gpu_lib):
    print("[gpu] rigid body MPC assembly", flush=True)
    m = nla3d_amd.GpuModel(0)
    xyz, conn = meshgen.solid81_rigid_top_case(m, 3, 2, 2, 2.0, 2.0, 1.0, 0.15, 0.349)
    m.init_solution()
    master = len(xyz)
    z = xyz[:-1, 2]
    fixes = [(int(nd), d, 0.0) for nd in np.nonzero(z == 0.0)[0] + 1 for d in range(3)]
    fixes += [(master, d, 0.0) for d in range(6)]
    top = list(np.nonzero(z == 1.0)[0] + 1)
    o = OracleModel(xyz, [(1, c) for c in conn], material=(3, [0.5, 0.1, 1000.0]), fixes=fixes,
                    rigid_bodies=[(master, top, [0, 1, 2])]).setup()
    rng = np.random.default_rng(4)
    U = 0.02 * rng.standard_normal(o.nAll)
    for t, v in zip(range(6), [0.01, -0.02, 0.15, 0.05, -0.03, 0.349]):
        U[o.nodeEq[master - 1][t] - 1] = v  # master rotation beyond the small-angle branch
    m.set_solution(U)
    m.assemble()
    Kref, Fref = o.assemble(U)
    for b in (1, 12, 2):
        assert rel(m.get_K(b), Kref[b]) <= TOL, b
    assert rel(m.get_F(), Fref) <= TOL
