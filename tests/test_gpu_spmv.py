"""The Krylov solve's block(2) operator on the device: y = K.block(2) x through the SpMV
kernel (node-diagonal copy for the interior nodes of a structured block, grouped CSR for
the rest) against the product with the assembled upper CSR (scipy, FP64), with the copy on
and off, on models where it is built and where it is not; and MINRES through both layouts
reaching the same solution."""
import ctypes
import os

import numpy as np
import pytest
import scipy.sparse as sp

import nla3d_amd
from nla3d_amd import meshgen

pytestmark = pytest.mark.gpu


def full_k2(m):
    sz = m.sizes
    n2 = sz["nUnknown"] + sz["nMpc"]
    io, cols = m.pattern(2)
    rows = np.repeat(np.arange(n2), np.diff(io.astype(np.int64)))
    A = sp.csr_matrix((m.get_K(2), (rows, cols.astype(np.int64) - 1)), shape=(n2, n2))
    return (A + sp.triu(A, 1).T).tocsr()


def newton_state(m, t):
    opts = m.newton_defaults(criterion=1, convergence=1e-8, krylov_tol=1e-10, krylov_max_iter=20000)
    m.newton_iteration(t, 1, ctypes.c_double(0.0), opts, nla3d_amd.NewtonStats())
    m.assemble()


def rigid_box(nx, ny, nz, node_dia=True):
    old = os.environ.get("NLA3D_GPU_NODE_DIA")
    os.environ["NLA3D_GPU_NODE_DIA"] = "1" if node_dia else "0"
    try:
        m = nla3d_amd.GpuModel(0)
        meshgen.solid81_rigid_top_case(m, nx, ny, nz, 2.0, 2.0, 1.0, 0.15, 0.349)
        m.set_max_in_row(0)
        m.init_solution()
    finally:
        if old is None:
            del os.environ["NLA3D_GPU_NODE_DIA"]
        else:
            os.environ["NLA3D_GPU_NODE_DIA"] = old
    newton_state(m, 0.5)
    return m


def check_product(m, seed):
    """|y - K x| <= 1e-14 (|K| |x|) row by row: FP64 sums of the same products in another
    order"""
    A = full_k2(m)
    rng = np.random.default_rng(seed)
    x = rng.standard_normal(A.shape[0])
    y = m.apply_block2(x)
    ref = A @ x
    bound = abs(A) @ np.abs(x)
    err = np.abs(y - ref)
    assert np.all(err <= 1e-14 * bound + 1e-300), (err / np.maximum(bound, 1e-300)).max()
    return y, x


def test_node_diagonal_spmv_matches_assembled_matrix(gpu_lib):
    """a rigid-top box (element pressures, multipliers of the rigid-body MPC, the bottom
    fixed): the nodes with all 26 neighbours in the run of 3-row node groups go through the
    14-offset node-diagonal copy: i, j in 1..23 and k in 2..10 (the layer k = 1 couples to
    the fixed bottom; the top-face rows each carry their own multiplier column, so they are
    single-row groups outside the run and the layer k = 11 below them does not fit)"""
    print("[gpu] node-diagonal SpMV, rigid-top box 24x24x12", flush=True)
    m = rigid_box(24, 24, 12)
    nd_nodes, nd_off, n_csr = m.spmv_info()
    free_nodes = 25 * 25 * 12
    print("node-diagonal nodes", nd_nodes, "offsets", nd_off, "CSR groups", n_csr, flush=True)
    assert nd_off == 14
    assert nd_nodes == 23 * 23 * 9 and nd_nodes < free_nodes
    for seed in (1, 2):
        check_product(m, seed)


def test_node_diagonal_copy_off_gives_the_same_operator(gpu_lib):
    print("[gpu] SpMV with and without the node-diagonal copy", flush=True)
    a = rigid_box(24, 24, 12, node_dia=True)
    b = rigid_box(24, 24, 12, node_dia=False)
    assert a.spmv_info()[0] > 0 and b.spmv_info()[0] == 0
    b.set_solution(a.get_solution())  # the same state, so the same assembled K
    b.assemble()
    assert np.array_equal(a.get_K(2), b.get_K(2))
    ya, x = check_product(a, 3)
    yb, _ = check_product(b, 3)
    scale = (abs(full_k2(a)) @ np.abs(x)).max()
    assert np.abs(ya - yb).max() <= 1e-14 * scale
    rhs = a.get_rhs()
    xa, ia, _ = a.solve_block2(rhs, tol=1e-11)
    xb, ib, _ = b.solve_block2(rhs, tol=1e-11)
    print("MINRES iterations", ia, ib, flush=True)
    assert np.abs(xa - xb).max() <= 1e-8 * np.abs(xb).max()


def test_node_diagonal_copy_follows_reassembly(gpu_lib):
    """the copy is refreshed after every assembly: a new state's operator is the new K"""
    print("[gpu] node-diagonal copy after reassembly", flush=True)
    m = rigid_box(24, 24, 12)
    check_product(m, 4)
    newton_state(m, 1.0)
    check_product(m, 5)


def test_spmv_on_batch_cube_and_small_models(gpu_lib):
    """a compression cube of the batched workload (copy built) and a small box (no copy:
    under the size threshold) through the same kernel"""
    print("[gpu] SpMV on a batch cube and a small box", flush=True)
    (m, p), = meshgen.batch_models(1, seed=11, n_range=(20, 20))
    newton_state(m, 0.5)
    assert m.spmv_info()[0] > 0
    check_product(m, 6)
    s = rigid_box(6, 6, 3)
    assert s.spmv_info()[:2] == (0, 0)
    check_product(s, 7)
