"""CPU oracle of the Schur-row multigrid preconditioner (test infrastructure only).

No counterpart in the reference (its solver is PARDISO).  Restates the definitions
the device path implements:

* schur_matrix: S = B^T D^-1 B - C on the rows that block Jacobi treats by a scalar
  Schur estimate (element pressures, multipliers), B their couplings to the node
  rows, D^-1 the block-Jacobi inverse node blocks, C their own block;
* polynomial: the default stage, one Jacobi-Chebyshev sweep on S from zero (no coarse
  levels) weighted by poly_beta, 8 terms up to 10^4 rows and 8 more per decade to at most 24,
  on [lmax/div, 1.05 lmax] with div = 25 terms (at most 400), on S's FP32-rounded entries
  or, for large matrices, on its half-precision operator (below);
* vcycle: one V-cycle over given aggregation maps with Jacobi-Chebyshev smoothing
  on [lmax/div, 1.05 lmax], the same polynomial before and after (finest level:
  16 terms, div 400; coarser levels: 3 terms, div 10), exact solve on the
  coarsest level.  The coarser levels' smoother and residual products use the
  level's matrix with its entries rounded to FP32 (products and sums in FP64), the
  diagonal scaling and the coarse matrices the FP64 values.  The finest level of a
  hierarchy with more than one level smooths either the same way or (finest_half,
  the library's choice for large levels) with its half-precision operator
  (half_operator: the symmetrically scaled matrix with FP16 off-diagonal entries and
  a diagonal raised by their rounding errors) in the scaled variables, by the
  three-term form of the same polynomial (smooth_scaled).

Pinned by: schur_matrix equals minus the exact Schur complement (dense elimination) of
block(2) with its displacement block replaced by its block diagonal over alpha; a
one-level hierarchy gives S^-1 exactly; the cycle is a symmetric
positive-definite operator; the smoother alone reduces the energy-norm error; S is
symmetric positive definite on the models of the tests; the two-term smoother is the
scaled minimax Chebyshev polynomial in closed form; half_operator is exact on a matrix
whose scaled entries FP16 holds, dominates the scaled matrix, and stays positive definite
where plain FP16 rounding does not; the scaled three-term sweep on it equals the two-term
Jacobi-Chebyshev sweep on the unscaled operator; the polynomial stage is a symmetric
positive-definite operator whose unweighted sweep reduces the energy-norm error, and on a
diagonal matrix it is the closed-form minimax polynomial (the sweep above) times the weight.
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


SCHUR_ALPHA = 0.5  # S = alpha B^T D^-1 B - C: the weight of the block-Jacobi Schur term (V-cycle)
SCHUR_ALPHA_SWEEP = 0.3  # the same under the default one-sweep stage
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


POLY_MIN, POLY_MAX = 8, 24  # the polynomial stage's shortest and longest sweep



def poly_steps(n):
    """length of the polynomial stage's one sweep for n rows: 8 terms up to 10^4 rows, + 8 per
    decade, at most POLY_MAX (the library's polySteps)"""
    s = POLY_MIN + 8.0 * math.log10(max(n, 1) / 1.0e4)
    return max(POLY_MIN, min(POLY_MAX, int(math.floor(s + 0.5))))


def poly_beta(steps):
    """the stage's output weight against the block-Jacobi rows: 1 at 8 terms rising to 2.5 at 24
    (the library's polyBeta)"""
    return 1.0 + 1.5 * (steps - POLY_MIN) / (POLY_MAX - POLY_MIN)


def poly_div(steps):
    """its interval [lmax / div, 1.05 lmax]: div = 25 steps, at most 400"""
    return min(400.0, 25.0 * steps)


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


def round_half(v):
    """FP64 values rounded to FP16 in one step, to nearest even (kept as FP64)"""
    return np.asarray(v, dtype=np.float64).astype(np.float16).astype(np.float64)


def round_up_f32(v):
    """the smallest FP32 values >= v (kept as FP64)"""
    v = np.asarray(v, dtype=np.float64)
    f = v.astype(np.float32)
    low = f.astype(np.float64) < v
    f[low] = np.nextafter(f[low], np.float32(np.inf))
    return f.astype(np.float64)


def half_operator(A):
    """(s, H) for the finest level's half-precision smoother: s = D^-1/2 (D = diag A, s = 0 where
    D <= 0), the scaled matrix A^ = diag(s) A diag(s) (entry (i, c) computed as A_ic (s_i s_c)),
    H = its off-diagonal entries rounded to FP16 plus the diagonal 1 + Delta_i rounded up to
    FP32, Delta_i = |1 - A^_ii| + sum over the row's other entries of |round_half(A^_ic) - A^_ic|"""
    A = sp.coo_matrix(A)
    n = A.shape[0]
    d = sp.csr_matrix(A).diagonal()
    s = np.sqrt(np.where(d > 0, 1.0 / np.where(d > 0, d, 1.0), 0.0))
    rows, cols = A.row.astype(np.int64), A.col.astype(np.int64)
    v = A.data * (s[rows] * s[cols])
    off = rows != cols
    h = round_half(v[off])
    err = np.zeros(len(v))
    err[off] = np.abs(h - v[off])
    err[~off] = np.abs(1.0 - v[~off])
    delta = np.bincount(rows, weights=err, minlength=n)
    H = sp.csr_matrix((h, (rows[off], cols[off])), shape=(n, n)) + sp.diags(round_up_f32(1.0 + delta))
    return s, H.tocsr()


def smooth_scaled(H, rhat, x0, lmax, steps, div):
    """the Chebyshev sweep of `smooth` in the scaled variables, three-term form:
    x_1 = x_0 + c0 (rhat - H x_0), x_(k+1) = x_k + c1_k (x_k - x_(k-1)) + c2_k (rhat - H x_k)"""
    c0, c1, c2 = cheb_coef(lmax, steps, div)
    prev, x = x0, x0 + c0 * (rhat - H @ x0)
    for i in range(1, steps):
        prev, x = x, x + c1[i] * (x - prev) + c2[i] * (rhat - H @ x)
    return x


def polynomial(S, lmax, r, half=None, use_half=False, beta=None):
    """the default Schur stage: one Jacobi-Chebyshev sweep from zero on S weighted by beta
    (default poly_beta(terms)), z = beta p(D^-1 S) D^-1 r, poly_steps(rows) terms on
    [lmax / poly_div, 1.05 lmax]; use_half: on the half-precision operator in the scaled
    variables (half = (s, H) replaces half_operator(S)), else with S's entries rounded to FP32
    (products and sums in FP64, D the FP64 diagonal)"""
    S = sp.csr_matrix(S)
    steps = poly_steps(S.shape[0])
    div = poly_div(steps)
    beta = poly_beta(steps) if beta is None else beta
    if use_half:
        s, H = half_operator(S) if half is None else half
        return s * (beta * smooth_scaled(H, s * r, np.zeros_like(r), lmax, steps, div))
    return beta * smooth(rounded_f32(S), 1.0 / S.diagonal(), r, np.zeros_like(r), lmax, steps, div)


def vcycle(mats, aggs, lmaxs, r, level=0, half=None, finest_half=True):
    """one V-cycle from `level`; finest_half: the finest level smooths with its half-precision
    operator (else with its FP32-rounded matrix, as the coarser levels); half = (s, H) replaces
    half_operator(mats[0]) (to run the cycle on a given stored operator)"""
    A = mats[level]
    if level == len(mats) - 1:
        return np.linalg.solve(A.toarray(), r)
    steps = finest_steps(A.shape[0]) if level == 0 else STEPSC
    div = finest_div(steps) if level == 0 else DIVC
    agg = aggs[level]
    if level == 0 and finest_half:
        s, H = half_operator(A) if half is None else half
        rhat = s * r
        x = smooth_scaled(H, rhat, np.zeros_like(r), lmaxs[0], steps, div)
        pos = s > 0
        t = np.zeros_like(r)
        t[pos] = (rhat - H @ x)[pos] / s[pos]
        rc = np.bincount(agg, weights=t, minlength=mats[1].shape[0])
        xc = vcycle(mats, aggs, lmaxs, rc, 1, finest_half=finest_half)[agg]
        x[pos] += xc[pos] / s[pos]
        return s * smooth_scaled(H, rhat, x, lmaxs[0], steps, div)
    dinv = 1.0 / A.diagonal()
    As = rounded_f32(A)
    x = smooth(As, dinv, r, np.zeros_like(r), lmaxs[level], steps, div)
    rc = np.bincount(agg, weights=r - As @ x, minlength=mats[level + 1].shape[0])
    x = x + vcycle(mats, aggs, lmaxs, rc, level + 1, finest_half=finest_half)[agg]
    return smooth(As, dinv, r, x, lmaxs[level], steps, div)
