"""Nested-dissection multifrontal LU of K.block(2) on structured hexahedral boxes (TEST
INFRASTRUCTURE ONLY: the CPU baseline's stand-in for the reference's PARDISO).

Only bench.py's cpu_baseline leg and tests/ use this module; the product never does.

The reference factorises K.block(2) with MKL PARDISO (EquationSolver.cpp:141-251: phase 11
with a METIS nested-dissection ordering, iparm[1] = 2, then phases 22 and 33 on every
solveEquations call, FESolver.cpp:457), multithreaded.  MKL is absent, so this restates that
work with the same ingredients: a nested-dissection ordering (here geometric: recursive
bisection of the box by planes of nodes, the textbook ordering of a structured grid), a
multifrontal factorisation in the elimination tree's postorder (dense frontal matrices,
partial LU with row pivoting inside each front's fully summed block, extend-add of the
children's update matrices), and the forward / backward substitution through the tree.  The
dense kernels are LAPACK getrf/getrs and BLAS gemm through scipy/numpy, which run on all host
cores (OpenBLAS threads).

Variables: the block(2) rows of a model built by nla3d_amd.meshgen.solid81_rigid_top_case-like
boxes: node DoFs (by node grid index), element pressures (by element grid index) and MPC
multipliers (with the node they constrain).  Pinned by tests/test_oracle_ndsolve.py: the
solution of a C3-like saddle-point system agrees with SuperLU's to 1e-10 and its residual is
below 1e-12; the ordering is a valid elimination tree (every variable in one front; the
separators separate).
"""
import numpy as np
import scipy.linalg as sla
import scipy.sparse as sp


class _Tree:
    __slots__ = ("V", "children", "B", "lu", "X", "F21", "post")

    def __init__(self, V, children):
        self.V, self.children = V, children
        self.B = self.lu = self.X = self.F21 = None
        self.post = -1


def _dissect(node_rows, elem_rows, nbox, ebox, leaf_nodes):
    """elimination tree of the variables in node box nbox = ((i0, i1), (j0, j1), (k0, k1))
    (inclusive node index ranges) and element box ebox (inclusive element index ranges):
    bisection by the middle plane of nodes along the longest axis"""
    ext = [b - a + 1 for a, b in nbox]
    if min(ext) <= 0:
        return None
    if ext[0] * ext[1] * ext[2] <= leaf_nodes or max(ext) < 3:
        V = _gather(node_rows, nbox) + _gather(elem_rows, ebox)
        return _Tree(np.array(sorted(set(V)), dtype=np.int64), [])
    ax = int(np.argmax(ext))
    s = (nbox[ax][0] + nbox[ax][1]) // 2
    sep = list(nbox)
    sep[ax] = (s, s)
    V = _gather(node_rows, sep)
    kids = []
    for nrange, erange in (((nbox[ax][0], s - 1), (ebox[ax][0], s - 1)), ((s + 1, nbox[ax][1]), (s, ebox[ax][1]))):
        nb, eb = list(nbox), list(ebox)
        nb[ax], eb[ax] = nrange, erange
        c = _dissect(node_rows, elem_rows, tuple(nb), tuple(eb), leaf_nodes)
        if c is not None:
            kids.append(c)
    return _Tree(np.array(sorted(set(V)), dtype=np.int64), kids)


def _gather(rows, box):
    (i0, i1), (j0, j1), (k0, k1) = box
    out = []
    for k in range(max(k0, 0), min(k1, rows.shape[2] - 1) + 1):
        for j in range(max(j0, 0), min(j1, rows.shape[1] - 1) + 1):
            for i in range(max(i0, 0), min(i1, rows.shape[0] - 1) + 1):
                out.extend(rows[i, j, k])
    return out


class NestedDissectionLU:
    """factorize(A) then solve(b) for a full symmetric (n x n, scipy) block(2) whose rows are
    described by node_rows[i][j][k] (block(2) rows of node (i, j, k): its free DoFs and the
    multipliers of its constraints) and elem_rows[i][j][k] (the pressure row of element
    (i, j, k)), nested arrays of lists"""

    def __init__(self, node_rows, elem_rows, leaf_nodes=64):
        self.node_rows, self.elem_rows = node_rows, elem_rows
        nn = node_rows.shape
        ne = elem_rows.shape
        self.root = _dissect(node_rows, elem_rows, tuple((0, d - 1) for d in nn), tuple((0, d - 1) for d in ne),
                             leaf_nodes)
        self.order = []
        self._postorder(self.root)

    def _postorder(self, t):
        for c in t.children:
            self._postorder(c)
        t.post = len(self.order)
        self.order.append(t)

    def check_tree(self, n):
        """every variable in exactly one front"""
        seen = np.zeros(n, dtype=np.int64)
        for t in self.order:
            seen[t.V] += 1
        return bool(np.all(seen == 1))

    def _front(self, t, A, owner, pos, updates):
        """assemble, partially factorise and pass on the update matrix of front t"""
        V = t.V
        rows = A[V]
        cand = [rows.indices]
        for c in t.children:
            cand.append(c.B)
        cand = np.unique(np.concatenate(cand))
        B = cand[owner[cand] > t.post]
        t.B = B
        nv, nb = len(V), len(B)
        F = np.concatenate([V, B])
        pos[F] = np.arange(nv + nb)
        M = np.zeros((nv + nb, nv + nb))
        r = np.repeat(np.arange(nv), np.diff(rows.indptr))
        cp = pos[rows.indices]
        keep = cp >= 0
        M[r[keep], cp[keep]] = rows.data[keep]
        M[nv:, :nv] = M[:nv, nv:].T
        for c in t.children:
            cb, S = updates.pop(c.post)
            pc = pos[cb]
            M[np.ix_(pc, pc)] += S
        pos[F] = -1
        t.lu = sla.lu_factor(M[:nv, :nv], check_finite=False)
        if nb:
            t.X = sla.lu_solve(t.lu, M[:nv, nv:], check_finite=False)
            t.F21 = M[nv:, :nv].copy()
            updates[t.post] = (B, M[nv:, nv:] - t.F21 @ t.X)
        return 2.0 / 3.0 * nv ** 3 + 2.0 * nv * nv * nb + 2.0 * nv * nb * nb

    def _subtree(self, t, A, owner, updates):
        pos = np.full(A.shape[0], -1, dtype=np.int64)
        stack, out, flops = [t], [], 0.0
        while stack:  # postorder of the subtree
            u = stack.pop()
            out.append(u)
            stack.extend(u.children)
        for u in sorted(out, key=lambda u: u.post):
            flops += self._front(u, A, owner, pos, updates)
        return flops

    def factorize(self, A, threads=None):
        """the multifrontal factorisation: the subtrees below the first tree level holding at
        least 2 * threads fronts run concurrently with one BLAS thread each, the fronts above
        them in postorder with all BLAS threads"""
        import os
        from concurrent.futures import ThreadPoolExecutor
        from threadpoolctl import threadpool_limits
        A = sp.csr_matrix(A)
        n = A.shape[0]
        threads = threads or os.cpu_count() or 1
        owner = np.full(n, -1, dtype=np.int64)
        for t in self.order:
            owner[t.V] = t.post
        if np.any(owner < 0):
            raise ValueError("a variable belongs to no front")
        frontier = [self.root]
        while len(frontier) < 2 * threads and any(t.children for t in frontier):
            nxt = []
            for t in frontier:
                nxt.extend(t.children if t.children else [t])
            frontier = nxt
        below = set()
        for f in frontier:
            stack = [f]
            while stack:
                u = stack.pop()
                below.add(u.post)
                stack.extend(u.children)
        updates = {}
        self.threads = threads
        with threadpool_limits(limits=1, user_api="blas"):
            with ThreadPoolExecutor(max_workers=threads) as ex:
                self.flops = sum(ex.map(lambda f: self._subtree(f, A, owner, updates), frontier))
        pos = np.full(n, -1, dtype=np.int64)
        for t in self.order:
            if t.post not in below:
                self.flops += self._front(t, A, owner, pos, updates)
        return self

    def solve(self, b):
        y = np.array(b, dtype=np.float64, copy=True)
        w = {}
        for t in self.order:  # forward: eliminate each front's variables from its boundary rows
            wv = sla.lu_solve(t.lu, y[t.V], check_finite=False)
            w[t.post] = wv
            if len(t.B):
                y[t.B] -= t.F21 @ wv
        x = np.zeros_like(y)
        for t in reversed(self.order):  # backward
            xv = w[t.post]
            if len(t.B):
                xv = xv - t.X @ x[t.B]
            x[t.V] = xv
        return x


def box_rows(ne, ee, nC, nDofs, nMpc, nx, ny, nz):
    """node_rows / elem_rows of a rigid-top box of meshgen.hex_box numbering (node (i, j, k) is
    1 + i + (nx+1)(j + (ny+1)k), element (i, j, k) likewise with nx, ny) from its 1-based
    equation numbers by node and DoF type (ne) and by element (ee); the multipliers of a
    rigid-body top (nMpc rows after the nDofs DoFs, top-face slaves in node order times
    UX UY UZ) follow their slave node"""
    node_rows = np.empty((nx + 1, ny + 1, nz + 1), dtype=object)
    elem_rows = np.empty((nx, ny, nz), dtype=object)
    for k in range(nz + 1):
        for j in range(ny + 1):
            for i in range(nx + 1):
                nid = i + (nx + 1) * (j + (ny + 1) * k)
                node_rows[i, j, k] = [int(q) - nC - 1 for q in ne[nid, :3] if q > nC]
    for k in range(nz):
        for j in range(ny):
            for i in range(nx):
                e = i + nx * (j + ny * k)
                elem_rows[i, j, k] = [int(ee[e, 6]) - nC - 1]
    # multiplier rows: the MPC rows after the DoFs, in the rigid body's slave order (top face,
    # x fastest, then y) times UX UY UZ
    n_mpc = nMpc
    if n_mpc:
        first = nDofs - nC
        q = 0
        for j in range(ny + 1):
            for i in range(nx + 1):
                for d in range(3):
                    node_rows[i, j, nz].append(first + q)
                    q += 1
        assert q == n_mpc
    return node_rows, elem_rows
