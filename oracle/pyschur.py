"""CPU oracle of the Schur-row multigrid preconditioner (test infrastructure only).

No counterpart in the reference (its solver is PARDISO).  Restates the definitions
the device path implements:

* schur_matrix: S = B^T D^-1 B - C on the rows that block Jacobi treats by a scalar
  Schur estimate (element pressures, multipliers), B their couplings to the node
  rows, D^-1 the block-Jacobi inverse node blocks, C their own block;
* vcycle: one V-cycle over given aggregation maps with Jacobi-Chebyshev smoothing
  on [lmax/div, 1.05 lmax], the same polynomial before and after (finest level:
  16 terms, div 400; coarser levels: 3 terms, div 10), exact solve on the
  coarsest level.  The smoother's and the residual's products use each level's
  matrix with its entries rounded to FP32 (products and sums in FP64), the diagonal
  scaling and the coarse matrices the FP64 values.

Pinned by: schur_matrix equals minus the exact Schur complement (dense elimination) of
block(2) with its displacement block replaced by its block diagonal over alpha; a
one-level hierarchy gives S^-1 exactly; the cycle is a symmetric
positive-definite operator; the smoother alone reduces the energy-norm error; S is
symmetric positive definite on the models of the tests.
"""
import math

import numpy as np
import scipy.sparse as sp


def schur_matrix(K2, groups, bj, alpha=None):
    """(q rows, S as a dense array) from the full block(2) matrix, the block-Jacobi groups
    [(r0, size, kind)] and their inverse blocks [(r0, size, B)]; alpha defaults to the
    library's SCHUR_ALPHA"""
    alpha = SCHUR_ALPHA if alpha is None else alpha
    q = np.array([r0 for r0, sz, kind in groups if kind == 1], dtype=np.int64)
    urows, blocks = [], []
    for (r0, sz, kind), (_, _, B) in zip(groups, bj):
        if kind == 0:
            urows.extend(range(r0, r0 + sz))
            blocks.append(B)
    Dinv = sp.block_diag(blocks).tocsr()
    K = K2.tocsr()
    Bq = K[urows][:, q]
    C = K[q][:, q]
    return q, (alpha * (Bq.T @ Dinv @ Bq) - C).toarray()


SCHUR_ALPHA = 0.5  # S = alpha B^T D^-1 B - C: the weight of the block-Jacobi Schur term
STEPS0, DIV0 = 16, 400.0   # finest level of a 10^6-row matrix (the longest smoother)
STEPSC, DIVC = 3, 10.0     # coarser levels


def finest_steps(n):
    """finest-level smoother length for n rows: 8 terms up to 10^4 rows, + 4 per decade, at
    most STEPS0 (the library's finestSteps)"""
    s = 8 + 4.0 * math.log10(max(n, 1) / 1.0e4)
    return max(8, min(STEPS0, int(math.floor(s + 0.5))))


def finest_div(steps):
    """the interval [lmax / div, 1.05 lmax] widens with the square of the length"""
    return 100.0 * (steps / 8.0) ** 2


def cheb_coef(lmax, steps, div=10.0):
    hi, lo = 1.05 * lmax, lmax / div
    theta, delta = 0.5 * (hi + lo), 0.5 * (hi - lo)
    sigma = theta / delta
    c0 = 1.0 / theta
    rho = 1.0 / sigma
    c1, c2 = [0.0], [0.0]
    for _ in range(1, steps):
        rn = 1.0 / (2.0 * sigma - rho)
        c1.append(rn * rho)
        c2.append(2.0 * rn / delta)
        rho = rn
    return c0, c1, c2


def smooth(A, dinv, r, x, lmax, steps, div=10.0):
    """Jacobi-Chebyshev: x + p(D^-1 A) D^-1 (r - A x)"""
    c0, c1, c2 = cheb_coef(lmax, steps, div)
    res = r - A @ x
    d = c0 * dinv * res
    for i in range(1, steps):
        x = x + d
        res = res - A @ d
        d = c1[i] * d + c2[i] * dinv * res
    return x + d


def hierarchy(S, aggs):
    """levels [A_0, A_1, ...] from the aggregate maps of every level but the coarsest"""
    mats = [sp.csr_matrix(S)]
    for agg in aggs:
        n, nc = len(agg), int(agg.max()) + 1
        P = sp.csr_matrix((np.ones(n), (np.arange(n), agg)), shape=(n, nc))
        mats.append((P.T @ mats[-1] @ P).tocsr())
    return mats


def rounded_f32(A):
    """A with its stored entries rounded to FP32 (kept as FP64)"""
    A = sp.csr_matrix(A, dtype=np.float64, copy=True)
    A.data = A.data.astype(np.float32).astype(np.float64)
    return A


def vcycle(mats, aggs, lmaxs, r, level=0):
    A = mats[level]
    if level == len(mats) - 1:
        return np.linalg.solve(A.toarray(), r)
    steps = finest_steps(A.shape[0]) if level == 0 else STEPSC
    div = finest_div(steps) if level == 0 else DIVC
    dinv = 1.0 / A.diagonal()
    As = rounded_f32(A)
    x = smooth(As, dinv, r, np.zeros_like(r), lmaxs[level], steps, div)
    agg = aggs[level]
    rc = np.bincount(agg, weights=r - As @ x, minlength=mats[level + 1].shape[0])
    x = x + vcycle(mats, aggs, lmaxs, rc, level + 1)[agg]
    return smooth(As, dinv, r, x, lmaxs[level], steps, div)
