"""Pins of the oracle's integer work, of its Schur-matrix definition and of its node
features, by known answers that do not come from another implementation of this build:

* the reference's own SparsityInfo / SparseMatrix / SparseSymMatrix known answers
  (test/sparse_math.cpp: CSR arrays after compress() and exact mat-vec products) replayed
  through oracle/pyoracle.py's SparsityInfo and its dense and sparse CSR products (the
  ones its Newton loop uses);
* the DoF numbering rule (FEStorage.cpp:382-432) on a model small enough to number by
  hand: the expected equation numbers are typed into the test, and the library's host-only
  numbering must give the same;
* the sparsity rule (FEStorage.cpp:434-550) restated as a set: every pair of DoFs of one
  element plus every (MPC equation, term DoF) pair, nothing else (no automatic diagonal);
* pyschur.schur_matrix against the exact Schur complement of the matrix whose displacement
  block is replaced by its block diagonal, formed by dense elimination;
* pylearned.node_features by the invariances its definition claims (scaling K, renumbering
  nodes) and by closed-form values on a single undeformed unit cube.
"""
import math

import numpy as np
import pytest
import scipy.sparse as sp

from oracle import pylearned as pl
from oracle import pyoracle as po
from oracle import pyschur as ps


def _values(si, entries):
    """values array of a compressed SparsityInfo with value v put at (i, j) through getIndex"""
    v = np.zeros(len(si.columns))
    for i, j, x in entries:
        k = si.getIndex(i, j)
        assert k >= 0, (i, j)
        v[k] = x
    return v


# ---------------------------------------------------------------------------
# test/sparse_math.cpp known answers
# ---------------------------------------------------------------------------
def test_sparse_matrix_3x4_arrays():
    """sparse_math.cpp:82-124: SparseMatrix(3, 4) [0 1 0 4; 0 0 0 0; 2 8 5 9]"""
    si = po.SparsityInfo(3, 4)
    for i, j in ((1, 2), (1, 4), (3, 1), (3, 2), (3, 3), (3, 4)):
        si.addEntry(i, j)
    si.compress()
    v = _values(si, [(1, 2, 1.0), (1, 4, 4.0), (3, 1, 2.0), (3, 2, 8.0), (3, 3, 5.0), (3, 4, 9.0)])
    assert list(v) == [1.0, 4.0, 2.0, 8.0, 5.0, 9.0]
    assert list(si.columns) == [2, 4, 1, 2, 3, 4]
    assert list(si.iofeir) == [1, 3, 3, 7]


def test_sparse_sym_matrix_4x4_arrays_with_constructor_diagonal():
    """sparse_math.cpp:140-186: SparseSymMatrix(4) adds every diagonal entry in its
    constructor, so (2, 2) is stored (value 0) although the test never adds it; (3, 2) is
    stored as (2, 3)"""
    si = po.sparse_sym_pattern(4)
    for i, j in ((1, 1), (1, 2), (1, 4), (3, 2), (3, 3), (4, 4)):
        si.addEntrySym(i, j)
    si.compress()
    v = _values(si, [(1, 1, 2.0), (1, 2, 3.0), (1, 4, 4.0), (2, 3, 8.0), (3, 3, 1.0), (4, 4, 9.0)])
    assert list(v) == [2.0, 3.0, 4.0, 0.0, 8.0, 1.0, 9.0]
    assert list(si.columns) == [1, 2, 4, 2, 3, 3, 4]
    assert list(si.iofeir) == [1, 4, 6, 7, 8]


def test_shared_sparsity_info():
    """sparse_math.cpp:189-300: A and B share one SparsityInfo, so both hold the union of the
    entries either added before compress()"""
    si = po.SparsityInfo(4, 4)
    for i, j in ((1, 3), (2, 1), (2, 3), (4, 3)):  # A's entries
        si.addEntry(i, j)
    for i, j in ((1, 1), (1, 4), (2, 3), (3, 2)):  # B's entries (2, 3 already present)
        si.addEntry(i, j)
    si.compress()
    vb = _values(si, [(1, 1, 2.0), (1, 4, 2.0), (2, 3, 2.0), (3, 2, 2.0)])
    va = _values(si, [(1, 3, 1.0), (2, 1, 1.0), (2, 3, 1.0), (4, 3, 1.0)])
    assert list(vb) == [2.0, 0.0, 2.0, 0.0, 2.0, 2.0, 0.0]
    assert list(va) == [0.0, 1.0, 0.0, 1.0, 1.0, 0.0, 1.0]
    assert list(si.columns) == [1, 3, 4, 1, 3, 2, 3]
    assert list(si.iofeir) == [1, 4, 6, 7, 8]


def test_sparse_matrix_products():
    """sparse_math.cpp:302-364: matBVprod(A, [1 2 3 4]) = [18 0 69] and
    matBTVprod(A, [3 2 1]) = [2 11 5 21], through the oracle's dense and sparse CSR forms"""
    si = po.SparsityInfo(3, 4)
    for i, j in ((1, 2), (1, 4), (3, 1), (3, 2), (3, 3), (3, 4)):
        si.addEntry(i, j)
    si.compress()
    v = _values(si, [(1, 2, 1.0), (1, 4, 4.0), (3, 1, 2.0), (3, 2, 8.0), (3, 3, 5.0), (3, 4, 9.0)])
    for A in (po.dense_rect(si, v, 4), po.sparse_rect(si, v, 4)):
        assert list(A @ np.array([1.0, 2.0, 3.0, 4.0])) == [18.0, 0.0, 69.0]
        assert list(A.T @ np.array([3.0, 2.0, 1.0])) == [2.0, 11.0, 5.0, 21.0]


def test_sparse_sym_matrix_product():
    """sparse_math.cpp:367-453: the 8x8 symmetric matrix given by its upper entries (some
    added lower-first), matBVprod(A, b) = [66 34 36 105 67 19 163 50]"""
    ent = [(1, 1, 7.0), (1, 3, 1.0), (1, 6, 2.0), (1, 7, 7.0), (2, 5, 2.0), (2, 3, 8.0), (2, 2, -4.0),
           (3, 3, 1.0), (3, 8, 5.0), (4, 7, 9.0), (4, 4, 7.0), (5, 5, 5.0), (5, 6, 1.0), (5, 7, 5.0),
           (6, 8, 5.0), (6, 6, -1.0), (7, 7, 11.0), (8, 8, 5.0)]
    si = po.sparse_sym_pattern(8)
    for i, j, _ in ent:
        si.addEntrySym(i, j)
    si.compress()
    v = _values(si, ent)
    b = np.array([1.0, 2.0, 4.0, 6.0, 5.0, 3.0, 7.0, 3.0])
    for A in (po.dense_sym(si, v), po.sparse_sym(si, v)):
        assert list(A @ b) == [66.0, 34.0, 36.0, 105.0, 67.0, 19.0, 163.0, 50.0]


def test_max_in_row_overflow_and_column_clamp():
    """SparsityInfo::reinit clamps maxInRow to nColumns (SparseMatrix.cpp:50); addEntry
    beyond it is "maxInRow overflow!" (SparseMatrix.cpp:84); a repeated entry is a no-op"""
    si = po.SparsityInfo(2, 3, max_in_row=100)
    assert si.maxInRow == 3
    for j in (3, 1, 2, 1):
        si.addEntry(1, j)
    si2 = po.SparsityInfo(2, 8, max_in_row=2)
    si2.addEntry(2, 5)
    si2.addEntry(2, 5)
    si2.addEntry(2, 1)
    with pytest.raises(RuntimeError, match="maxInRow overflow"):
        si2.addEntry(2, 7)
    si.compress()
    assert list(si.columns) == [1, 2, 3] and list(si.iofeir) == [1, 4, 4]


# ---------------------------------------------------------------------------
# numbering and sparsity on a model numbered by hand
# ---------------------------------------------------------------------------
UX, UY, UZ, ROTX, ROTY, ROTZ, P = 0, 1, 2, 3, 4, 5, 6
# two unit hexahedra stacked in z (nodes 1-4 at z=0, 5-8 at z=1, 9-12 at z=2) and an
# element-less rigid-body master, node 13, whose only slave is node 9's UZ
HAND_XYZ = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0], [0, 0, 1], [1, 0, 1], [1, 1, 1], [0, 1, 1],
                     [0, 0, 2], [1, 0, 2], [1, 1, 2], [0, 1, 2], [0.5, 0.5, 2]], dtype=float)
HAND_CONN = [[1, 2, 3, 4, 5, 6, 7, 8], [5, 6, 7, 8, 9, 10, 11, 12]]
HAND_FIXES = [(1, UX, 0.0), (1, UY, 0.0), (1, UZ, 0.0), (2, UZ, 0.0),
              (13, UX, 0.0), (13, UY, 0.0), (13, UZ, 0.1), (13, ROTY, 0.0), (13, ROTZ, 0.0)]
HAND_RB = (13, [9], [UZ])
# FEStorage::assignEquationNumbers by hand: 9 constrained DoFs take 1..9 in the order met
# (node 1 UX UY UZ, node 2 UZ, node 13 UX UY UZ ROTY ROTZ); the free ones count from 10:
# the element pressures first (10, 11), then the nodes in order, DoF types in enum order;
# the master's free ROTX is 44, the one MPC row 45.
HAND_NODE_EQ = {
    1: {UX: 1, UY: 2, UZ: 3}, 2: {UX: 12, UY: 13, UZ: 4}, 3: {UX: 14, UY: 15, UZ: 16},
    4: {UX: 17, UY: 18, UZ: 19}, 5: {UX: 20, UY: 21, UZ: 22}, 6: {UX: 23, UY: 24, UZ: 25},
    7: {UX: 26, UY: 27, UZ: 28}, 8: {UX: 29, UY: 30, UZ: 31}, 9: {UX: 32, UY: 33, UZ: 34},
    10: {UX: 35, UY: 36, UZ: 37}, 11: {UX: 38, UY: 39, UZ: 40}, 12: {UX: 41, UY: 42, UZ: 43},
    13: {UX: 5, UY: 6, UZ: 7, ROTX: 44, ROTY: 8, ROTZ: 9},
}
HAND_ELEM_EQ = {1: {P: 10}, 2: {P: 11}}
HAND_MPC_EQ = 45


def hand_oracle():
    return po.OracleModel(HAND_XYZ, [(po.SOLID81, c) for c in HAND_CONN], material=(3, [0.5, 0.1, 100.0]),
                          fixes=HAND_FIXES, rigid_bodies=[HAND_RB]).setup()


def hand_library():
    import nla3d_amd
    m = nla3d_amd.GpuModel(device=-1)
    m.set_nodes(HAND_XYZ)
    m.add_elements(nla3d_amd.SOLID81, HAND_CONN)
    m.set_material(3, [0.5, 0.1, 100.0])
    for n, d, v in HAND_FIXES:
        m.add_fix(n, d, v)
    m.add_rigid_body_mpc(HAND_RB[0], HAND_RB[1], HAND_RB[2])
    m.init_solution()
    return m


def test_numbering_by_hand_oracle():
    o = hand_oracle()
    assert (o.nDofs, o.nC, o.nU, o.nMpc, o.nAll) == (44, 9, 35, 1, 45)
    for n, eqs in HAND_NODE_EQ.items():
        assert o.nodeEq[n - 1] == eqs, n
    for e, eqs in HAND_ELEM_EQ.items():
        assert o.elemEq[e - 1] == eqs, e
    assert [m[2] for m in o.mpcs] == [HAND_MPC_EQ]


def test_numbering_by_hand_library(gpu_lib):
    m = hand_library()
    s = m.sizes
    assert (s["nDofs"], s["nConstrained"], s["nUnknown"], s["nMpc"]) == (44, 9, 35, 1)
    ne, ee = m.equation_numbers()
    for n in range(1, 14):
        assert {t: int(ne[n - 1, t]) for t in range(8) if ne[n - 1, t]} == HAND_NODE_EQ[n], n
    for e in (1, 2):
        assert {t: int(ee[e - 1, t]) for t in range(8) if ee[e - 1, t]} == HAND_ELEM_EQ[e], e


def rule_pattern(node_eq, elem_eq, conn, mpcs, nC, nAll):
    """the sparsity rule as a set: all pairs of DoFs of one element (its nodes' registered
    DoFs and its own), all (MPC equation, term DoF) pairs; split into the three upper blocks
    and returned as {block: (iofeir, columns)} 1-based"""
    pairs = set()
    for e, nodes in enumerate(conn):
        dofs = [q for n in nodes for q in node_eq[n].values()] + list(elem_eq[e + 1].values())
        pairs |= {(min(a, b), max(a, b)) for a in dofs for b in dofs}
    for eq, terms in mpcs:
        pairs |= {(min(eq, node_eq[n][d]), max(eq, node_eq[n][d])) for n, d in terms}
    out = {}
    for blk, rows, off_r, off_c in ((1, nC, 0, 0), (12, nC, 0, nC), (2, nAll - nC, nC, nC)):
        io, cols = [1], []
        for r in range(1, rows + 1):
            i = r + off_r
            row = sorted(j - off_c for a, j in pairs if a == i and (j <= nC) == (blk == 1))
            cols += row
            io.append(len(cols) + 1)
        out[blk] = (np.array(io, dtype=np.uint32), np.array(cols, dtype=np.uint32))
    return out


def test_sparsity_rule_by_hand(gpu_lib):
    node_eq = {n: dict(v) for n, v in HAND_NODE_EQ.items()}
    elem_eq = {e: dict(v) for e, v in HAND_ELEM_EQ.items()}
    mpcs = [(HAND_MPC_EQ, [(9, UZ), (13, UZ), (13, ROTX), (13, ROTY), (13, ROTZ)])]
    want = rule_pattern(node_eq, elem_eq, HAND_CONN, mpcs, 9, 45)
    o = hand_oracle()
    m = hand_library()
    for b in (1, 12, 2):
        assert np.array_equal(o.blocks[b].iofeir, want[b][0]), b
        assert np.array_equal(o.blocks[b].columns, want[b][1]), b
        io, cols = m.pattern(b)
        assert np.array_equal(io, want[b][0]) and np.array_equal(cols, want[b][1]), b
    # the element-less master's free ROTX (44 -> block(2) row 35) and the MPC row (45 -> 36)
    # have no diagonal entry (BlockSparseMatrix.h:73-85): row 35 holds only (35, 36)
    io, cols = want[2]
    assert list(cols[io[34] - 1:io[35] - 1]) == [36] and io[35] == io[36]
    # node 5 UX (eq 20 -> row 11) belongs to both elements: its upper row holds the free DoFs
    # numbered from it on, those of nodes 5..12 (eqs 20..43 -> rows 11..34); the pressures
    # (rows 1, 2) couple to it from their own rows: element 1's to nodes 2..8 (node 2 UX UY,
    # eqs 12..31 -> rows 3..22), element 2's to nodes 5..12
    row = lambda r: list(cols[io[r - 1] - 1:io[r] - 1])
    assert row(11) == list(range(11, 35))
    assert row(1) == [1] + list(range(3, 23)) and row(2) == [2] + list(range(11, 35))
    # the MPC row's couplings sit in the slave's and the master's rows: node 9 UZ (eq 34 -> row 25)
    assert row(25)[-1] == 36 and row(36) == []
    # block(1,2): node 2 UZ (eq 4) couples to element 1's pressure and its free DoFs
    io12, cols12 = want[12]
    assert list(cols12[io12[3] - 1:io12[4] - 1]) == [1] + list(range(3, 23))
    # the master's constrained UZ, ROTY, ROTZ (eqs 7, 8, 9) couple only to the MPC row
    for r in (7, 8, 9):
        assert list(cols12[io12[r - 1] - 1:io12[r] - 1]) == [36]


# ---------------------------------------------------------------------------
# pyschur.schur_matrix: exact elimination
# ---------------------------------------------------------------------------
def oracle_box(nx, ny, nz, rigid=True, bulk=1000.0, jitter=0.1, seed=7, U_scale=0.0):
    """oracle model of a jittered box (bottom fixed; rigid-body top or prescribed top UZ),
    its block(2) at a seeded state, and the learned-path block structure"""
    from nla3d_amd import meshgen
    xyz, conn, (I, J, K) = meshgen.hex_box(nx, ny, nz, 2.0, 2.0, 1.0, jitter=jitter, seed=seed)
    fixes = [(int(n), d, 0.0) for n in np.nonzero(K == 0)[0] + 1 for d in range(3)]
    rbs = []
    if rigid:
        xyz = np.vstack([xyz, [[1.0, 1.0, 1.0]]])
        master = len(xyz)
        fixes += [(master, d, v) for d, v in enumerate([0.0, 0.0, 0.15, 0.0, 0.0, 0.349])]
        rbs = [(master, list(np.nonzero(K == nz)[0] + 1), [0, 1, 2])]
    else:
        fixes += [(int(n), 2, 0.1) for n in np.nonzero(K == nz)[0] + 1]
    o = po.OracleModel(xyz, [(po.SOLID81, c) for c in conn], material=(3, [0.5, 0.1, bulk]), fixes=fixes,
                       rigid_bodies=rbs).setup()
    U = U_scale * np.random.default_rng(seed).standard_normal(o.nAll)
    K, _ = o.assemble(U)
    return o, U, o.sparse_sym(o.blocks[2], K[2], o.nU + o.nMpc), K


def eq_arrays(o):
    ne = np.zeros((len(o.xyz), 8), dtype=np.int64)
    ee = np.zeros((len(o.elements), 8), dtype=np.int64)
    for n, d in enumerate(o.nodeEq):
        for t, q in d.items():
            ne[n, t] = q
    for e, d in enumerate(o.elemEq):
        for t, q in d.items():
            ee[e, t] = q
    return ne, ee


def structure(o, K2):
    ne, ee = eq_arrays(o)
    n2 = o.nU + o.nMpc
    mpc = [m[2] for m in o.mpcs]
    groups, lrows, kind, rnode = pl.block_structure(ne, ee, o.nC, n2, mpc, K2)
    return groups, lrows, kind, rnode


@pytest.mark.parametrize("alpha", [1.0, ps.SCHUR_ALPHA])
def test_schur_matrix_is_the_exact_schur_complement_of_the_block_diagonal_system(alpha):
    """With A_u the displacement block of block(2), D its node-block diagonal, B the couplings
    of the Schur rows q (pressures, multipliers) to the node rows and C their own block,
    M = [[D / alpha, B], [B^T, C]] has (M^-1)_qq = (C - alpha B^T D^-1 B)^-1 (block
    elimination).  schur_matrix must equal minus that complement: a wrong sign, a transposed
    B, a missing alpha or a missing C fails."""
    o, U, K2, _ = oracle_box(2, 2, 1, U_scale=0.01)
    groups, _, kind, _ = structure(o, K2)
    bj = pl.block_jacobi(K2, groups)
    q, S = ps.schur_matrix(K2, groups, bj, alpha=alpha)
    n2 = K2.shape[0]
    Kd = K2.toarray()
    M = Kd.copy()
    urows = [r for r0, sz, kd in groups if kd == 0 for r in range(r0, r0 + sz)]
    M[np.ix_(urows, urows)] = 0.0
    for r0, sz, kd in groups:
        if kd == 0:
            blk = Kd[r0:r0 + sz, r0:r0 + sz]
            M[r0:r0 + sz, r0:r0 + sz] = 0.5 * (blk + blk.T) / alpha
    Minv = np.linalg.solve(M, np.eye(n2))
    S_exact = np.linalg.inv(Minv[np.ix_(q, q)])
    assert len(q) == len(o.elements) + o.nMpc and np.any(np.diag(Kd)[q] == 0.0)
    assert np.abs(S + S_exact).max() <= 1e-9 * np.abs(S_exact).max()


# ---------------------------------------------------------------------------
# pylearned.node_features: invariances and closed forms
# ---------------------------------------------------------------------------
def features(o, U, K2, bulk, t):
    groups, lrows, kind, rnode = structure(o, K2)
    bj = pl.block_jacobi(K2, groups)
    k12 = np.bincount(o.blocks[12].columns.astype(np.int64) - 1, minlength=K2.shape[0])
    X = pl.node_features(K2, lrows, kind, rnode, k12, bj, U, o.nC, bulk, t)
    return X, rnode[lrows]


def test_node_features_are_scale_free():
    """K -> s K (the block-Jacobi data recomputed from s K) leaves every feature unchanged:
    each one is a ratio of matrix entries, a count, a state or the load time"""
    o, U, K2, _ = oracle_box(3, 2, 2, U_scale=0.02)
    X, nodes = features(o, U, K2, 1000.0, 0.4)
    Xs, nodes_s = features(o, U, 37.5 * K2, 1000.0, 0.4)
    assert np.array_equal(nodes, nodes_s) and len(nodes) > 0
    assert np.abs(Xs - X).max() <= 1e-12 * max(1.0, np.abs(X).max())


def test_node_features_follow_a_renumbering_of_the_nodes():
    """the same mesh with its nodes numbered in another order (elements unchanged) gives each
    node the same features; the learned nodes come in the new order"""
    from nla3d_amd import meshgen
    xyz, conn, (I, J, K) = meshgen.hex_box(2, 2, 2, jitter=0.1, seed=3)
    nN = len(xyz)
    perm = np.random.default_rng(4).permutation(nN)  # new id of old node n is perm[n] + 1
    inv = np.argsort(perm)

    def model(ids, pos):
        bottom = [int(ids[n]) for n in np.nonzero(K == 0)[0]]
        top = [int(ids[n]) for n in np.nonzero(K == 2)[0]]
        fixes = [(b, d, 0.0) for b in bottom for d in range(3)] + [(t, 2, 0.1) for t in top]
        cn = [[int(ids[n - 1]) for n in c] for c in conn]
        o = po.OracleModel(pos, [(po.SOLID81, c) for c in cn], material=(3, [0.5, 0.1, 200.0]), fixes=fixes).setup()
        return o

    oa = model(np.arange(nN) + 1, xyz)
    ob = model(perm + 1, xyz[inv])
    # the same physical state: U by (node, dof) and pressures by element
    rng = np.random.default_rng(8)
    disp = 0.02 * rng.standard_normal((nN, 3))
    press = 0.5 * rng.standard_normal(len(conn))

    def state(o, ids):
        U = np.zeros(o.nAll)
        for n in range(nN):
            for d in range(3):
                U[o.nodeEq[int(ids[n]) - 1][d] - 1] = disp[n, d]
        for e in range(len(conn)):
            U[o.elemEq[e][6] - 1] = press[e]
        return U

    Ua, Ub = state(oa, np.arange(nN) + 1), state(ob, perm + 1)
    Ka, _ = oa.assemble(Ua)
    Kb, _ = ob.assemble(Ub)
    n2 = oa.nU + oa.nMpc
    Xa, na = features(oa, Ua, oa.sparse_sym(oa.blocks[2], Ka[2], n2), 200.0, 0.7)
    Xb, nb = features(ob, Ub, ob.sparse_sym(ob.blocks[2], Kb[2], n2), 200.0, 0.7)
    assert len(na) == len(nb) > 0
    new_of_old = {int(n): int(perm[n]) for n in na}
    pos_b = {int(n): i for i, n in enumerate(nb)}
    for i, n in enumerate(na):
        assert np.abs(Xa[i] - Xb[pos_b[new_of_old[int(n)]]]).max() <= 1e-12 * max(1.0, np.abs(Xa[i]).max()), n


def test_node_features_closed_form_on_a_unit_cube():
    """one undeformed unit cube, bottom face fixed, Mooney-Rivlin (mu = 2 (C10 + C01)), at
    U = 0: a top node's block is the small-strain deviatoric stiffness 10 mu / 27 on its
    diagonal and +-mu / 36 off it (cubic symmetry), its pressure coupling is
    int dN/dx_d dV = +-1/4 per direction, K_pp = -V / k, three other free nodes, 12
    constrained couplings per row, one element, no multiplier"""
    C10, C01, k, t = 0.5, 0.1, 100.0, 0.25
    mu = 2.0 * (C10 + C01)
    X8 = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0], [0, 0, 1], [1, 0, 1], [1, 1, 1], [0, 1, 1]], float)
    fixes = [(n, d, 0.0) for n in (1, 2, 3, 4) for d in range(3)]
    o = po.OracleModel(X8, [(po.SOLID81, list(range(1, 9)))], material=(3, [C10, C01, k]), fixes=fixes).setup()
    U = np.zeros(o.nAll)
    K, _ = o.assemble(U)
    K2 = o.sparse_sym(o.blocks[2], K[2], o.nU + o.nMpc)
    X, nodes = features(o, U, K2, k, t)
    assert sorted(int(n) for n in nodes) == [4, 5, 6, 7]  # 0-based top nodes
    s = 10.0 * mu / 27.0
    for f, n in zip(X, nodes):
        sx, sy = (-1.0 if X8[n, 0] == 0 else 1.0), (-1.0 if X8[n, 1] == 0 else 1.0)
        kp = np.array([sx, sy, 1.0]) / 4.0  # int dN/dx_d over the cube
        assert np.abs(f[[0, 3, 5]] - 1.0).max() <= 1e-14
        # off-diagonal block entries: (mu/3) int N,a N,b = +-mu/36, sign of kp_a kp_b
        off = (mu / 36.0) / s * np.array([sx * sy, sx, sy])
        assert np.abs(f[[1, 2, 4]] - off).max() <= 1e-14
        assert f[8] == 3.0 / 26.0 and f[17] == 1.0 / 8.0 and f[18] == 0.0 and f[19] == 36.0 / 81.0
        assert f[20] == 0.0 and f[21] == 0.0 and f[22] == t
        G = np.outer(kp, kp) / (kp @ kp)
        assert np.abs(f[10:16] - G[[0, 0, 0, 1, 1, 2], [0, 1, 2, 1, 2, 2]]).max() <= 1e-14
        trG = (kp @ kp) * k  # |kp|^2 / |K_pp|, V = 1
        assert abs(f[16] - math.log10(trG / (3.0 * s))) <= 1e-14
        # Schur estimate of the pressure row: |K_pp| + sum over its 12 free columns (1/16) / s
        d = 1.0 / k + 12.0 * (1.0 / 16.0) / s
        assert abs(f[23] - (1.0 / k) / d) <= 1e-14


def test_oracle_assembly_c_loop_is_the_python_loop():
    """OracleModel.assemble runs SOLID81-only models through the C restatement of the element
    loop; the Python loop (per-entry binary-search addValue in assemble3 order) must give the
    same K blocks and F bit for bit, with rigid-body rows"""
    o, _, _, _ = oracle_box(3, 2, 2, U_scale=0.0)
    U = 0.02 * np.random.default_rng(1).standard_normal(o.nAll)
    Kc, Fc = o.assemble(U)
    o.fast_solid81 = False
    Kp, Fp = o.assemble(U)
    assert all(np.array_equal(Kc[b], Kp[b]) for b in (1, 12, 2)) and np.array_equal(Fc, Fp)
