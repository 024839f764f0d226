"""The solver-level entry points (EquationSolver on the caller's 1-based upper CSR,
EquationSolver.h:19-78) serving an nla3d-style caller whose FEStorage assembles on the host:
the oracle's Newton loop (host assembly, the reference's numbering and CSR) with its direct
solve replaced by the device solver gives the direct-solve Newton's displacements after every
load step within 1e-8; a changed pattern is analysed again; the error codes replace the
reference's CHECKs."""
import numpy as np
import pytest

import nla3d_amd
from nla3d_amd import meshgen
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu


def rigid_oracle(nx, ny, nz, seed=5):
    xyz, conn, (I, J, K) = meshgen.hex_box(nx, ny, nz, 2.0, 2.0, 1.0, jitter=0.1, seed=seed)
    xyz = np.vstack([xyz, [[1.0, 1.0, 1.0]]])
    master = len(xyz)
    fixes = [(int(n), d, 0.0) for n in np.nonzero(K == 0)[0] + 1 for d in range(3)]
    fixes += [(master, d, v) for d, v in enumerate([0.0, 0.0, 0.15, 0.0, 0.0, 0.349])]
    return po.OracleModel(xyz, [(po.SOLID81, c) for c in conn], material=(3, [0.5, 0.1, 1000.0]), fixes=fixes,
                          rigid_bodies=[(master, list(np.nonzero(K == nz)[0] + 1), [0, 1, 2])], max_in_row=0).setup()


def dof_map(o):
    n2 = o.nU + o.nMpc
    rn, rd = np.zeros(n2, dtype=np.int64), np.full(n2, -1, dtype=np.int64)
    for n, d in enumerate(o.nodeEq):
        for t, q in d.items():
            if q > o.nC:
                rn[q - o.nC - 1], rd[q - o.nC - 1] = n + 1, t
    for d in o.elemEq:
        for t, q in d.items():
            rd[q - o.nC - 1] = t
    return rn, rd


@pytest.mark.parametrize("use_map", [True, False])
def test_host_assembled_newton_with_the_device_solver(gpu_lib, use_map):
    print(f"[gpu] oracle Newton with the device equation solver (DoF map {use_map})", flush=True)
    a, b = rigid_oracle(6, 6, 3), rigid_oracle(6, 6, 3)
    s = nla3d_amd.EqSolver(0, tol=1e-13, max_iter=50000, preconditioner=2)
    if use_map:
        s.set_dof_map(*dof_map(a))
    si = a.blocks[2]
    iters = []

    def gpu_solve(vals, rhs):
        x = s.solve(vals, si.iofeir, si.columns, rhs)
        st = s.stats()
        assert st.precond_fallbacks == 0 and st.breakdown == 0
        iters.append(st.iterations)
        return x

    a.newton_init()
    b.newton_init()
    for t in po.load_steps(3):
        ia = a.solve_load_step(t, criterion=1, convergence=1e-10, max_iterations=15, linear_solve=gpu_solve)
        ib = b.solve_load_step(t, criterion=1, convergence=1e-10, max_iterations=15)
        err = np.abs(a.U - b.U).max() / np.abs(b.U).max()
        print(f"t={t:.3f}: {ia} / {ib} solves, MINRES {iters[-ia:]}, error {err:.2e}", flush=True)
        assert ia == ib and err <= 1e-8, (t, err)


def test_pattern_change_and_error_codes(gpu_lib):
    print("[gpu] equation solver: re-analysis and error codes", flush=True)
    s = nla3d_amd.EqSolver(0, tol=1e-13, max_iter=50000)
    with pytest.raises(nla3d_amd.Nla3dGpuError) as e:
        s.substitute(np.ones(4))
    assert e.value.code == -3  # NLA3D_GPU_ESTATE
    # two different models through one solver: the second pattern is analysed afresh
    for nx in (3, 4):
        o = rigid_oracle(nx, 3, 2, seed=nx)
        U = 0.01 * np.random.default_rng(nx).standard_normal(o.nAll)
        Kv, _ = o.assemble(U)
        si = o.blocks[2]
        rhs = np.random.default_rng(nx + 10).standard_normal(len(si.iofeir) - 1)
        s.factorize(Kv[2], si.iofeir, si.columns)
        x = s.substitute(rhs)
        xr = np.linalg.solve(po.dense_sym(si, Kv[2]), rhs)
        assert np.abs(x - xr).max() <= 1e-8 * np.abs(xr).max(), nx
    # SparsityInfo's invariants: 1-based, sorted, upper triangle (NLA3D_GPU_EPATTERN = -4)
    io = np.array([1, 3, 4, 5], dtype=np.uint32)
    for cols in ([2, 1, 2, 3], [1, 2, 1, 3], [1, 4, 2, 3]):  # unsorted, lower entry, column > n
        with pytest.raises(nla3d_amd.Nla3dGpuError) as e:
            s.factorize(np.ones(4), io, np.array(cols, dtype=np.uint32))
        assert e.value.code == -4, cols
    with pytest.raises(nla3d_amd.Nla3dGpuError) as e:
        s.factorize(np.ones(4), np.array([0, 2, 3, 4], dtype=np.uint32), np.array([1, 2, 2, 3], dtype=np.uint32))
    assert e.value.code == -4
    with pytest.raises(nla3d_amd.Nla3dGpuError) as e:
        s.set_options(tol=1e-12, max_iter=100, preconditioner=1)
    assert e.value.code == -1
    # a 3x3 SPD system with a tolerance MINRES cannot reach in one iteration: ENOTCONV (-5)
    s.set_options(tol=1e-30, max_iter=1)
    with pytest.raises(nla3d_amd.Nla3dGpuError) as e:
        s.solve(np.array([4.0, 1.0, 3.0, 1.0]), io, np.array([1, 2, 2, 3], dtype=np.uint32), np.ones(3))
    assert e.value.code == -5


@pytest.mark.parametrize("use_map,node_dia", [(True, True), (False, True), (False, False)])
def test_structured_interior_through_the_drop_in(gpu_lib, use_map, node_dia):
    """a 24 x 24 x 12 rigid-top system assembled on the host (the oracle's C element loop; the
    first tangent, then a state with 1e-3 noise): the drop-in builds the SpMV's node-diagonal
    copy from the caller's CSR (the 23 x 23 x 9 interior nodes, with or without a DoF map; none
    with NLA3D_GPU_NODE_DIA=0) and solves to the direct solve's answer within 1e-8, for two
    right-hand sides and after new values"""
    import os
    import scipy.sparse as sp
    import scipy.sparse.linalg as spla
    print(f"[gpu] equation solver on a 24x24x12 system (DoF map {use_map}, node diagonals {node_dia})", flush=True)
    o = rigid_oracle(24, 24, 12)
    old = os.environ.get("NLA3D_GPU_NODE_DIA")
    os.environ["NLA3D_GPU_NODE_DIA"] = "1" if node_dia else "0"
    try:
        s = nla3d_amd.EqSolver(0, tol=1e-12, max_iter=50000, preconditioner=2)
        if use_map:
            s.set_dof_map(*dof_map(o))
        si = o.blocks[2]
        n2 = len(si.iofeir) - 1
        rows = np.repeat(np.arange(n2), np.diff(si.iofeir.astype(np.int64)))
        its = []
        for k, scale in enumerate((0.0, 1e-3)):
            U = scale * np.random.default_rng(40 + k).standard_normal(o.nAll)
            Kv, _ = o.assemble(U)
            A = sp.csr_matrix((Kv[2], (rows, si.columns.astype(np.int64) - 1)), shape=(n2, n2))
            A = (A + sp.triu(A, 1).T).tocsc()
            lu = spla.splu(A)
            s.factorize(Kv[2], si.iofeir, si.columns)  # the pattern is analysed here, on the first call
            for j in range(2):
                rhs = np.random.default_rng(50 + 2 * k + j).standard_normal(n2)
                x = s.substitute(rhs)
                xr = lu.solve(rhs)
                st = s.stats()
                err = np.abs(x - xr).max() / np.abs(xr).max()
                its.append(st.iterations)
                print(f"values {k}, rhs {j}: MINRES {st.iterations}, node-diagonal nodes {st.spmv_dia_nodes}, "
                      f"error {err:.2e}", flush=True)
                assert st.spmv_dia_nodes == (23 * 23 * 9 if node_dia else 0) and st.precond_fallbacks == 0
                assert err <= 1e-8, (k, j, err)
    finally:
        if old is None:
            del os.environ["NLA3D_GPU_NODE_DIA"]
        else:
            os.environ["NLA3D_GPU_NODE_DIA"] = old
