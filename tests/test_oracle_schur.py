"""Pins of the Schur-multigrid oracle (oracle/pyschur.py) on the CPU."""
import numpy as np
import scipy.sparse as sp

from oracle import pyschur as ps


def laplacian(n):
    T = sp.diags([-np.ones(n - 1), 2 * np.ones(n), -np.ones(n - 1)], [-1, 0, 1])
    I = sp.identity(n)
    return (sp.kron(T, I) + sp.kron(I, T) + 0.01 * sp.identity(n * n)).tocsr()


def test_one_level_is_exact_inverse():
    A = laplacian(6)
    r = np.random.default_rng(0).standard_normal(36)
    z = ps.vcycle([A], [], [], r)
    assert np.abs(A @ z - r).max() <= 1e-12 * np.abs(r).max()


def test_vcycle_is_symmetric_positive_definite():
    A = laplacian(8)
    agg = np.arange(64) // 4
    mats = ps.hierarchy(A, [agg])
    lmax = float(np.max(np.linalg.eigvals((sp.diags(1 / A.diagonal()) @ A).toarray()).real))
    B = np.stack([ps.vcycle(mats, [agg], [lmax], e) for e in np.eye(64)], axis=1)
    assert np.abs(B - B.T).max() <= 1e-12 * np.abs(B).max()
    assert np.linalg.eigvalsh(0.5 * (B + B.T)).min() > 0


def test_smoother_reduces_energy_error():
    A = laplacian(8).toarray()
    rng = np.random.default_rng(1)
    x_true = rng.standard_normal(64)
    r = A @ x_true
    dinv = 1 / np.diag(A)
    lmax = float(np.max(np.linalg.eigvals(np.diag(dinv) @ A).real))
    x = ps.smooth(A, dinv, r, np.zeros(64), lmax, 3)
    e0, e1 = x_true, x_true - x
    assert e1 @ A @ e1 < e0 @ A @ e0


def test_chebyshev_error_polynomial_is_the_minimax_one():
    """On A = diag(mu) (D = I) the smoother maps an error e to (1 - mu p(mu)) e; the
    three-term recurrence must reproduce the scaled Chebyshev polynomial
    T_k((theta - mu) / delta) / T_k(theta / delta) on [lo, hi] = [lmax/div, 1.05 lmax]:
    its closed form, |.| <= 1 / T_k(theta/delta) inside the interval and < 1 below it
    (for both the finest-level and the coarse-level settings)"""
    lmax = 2.0
    mu = np.linspace(1e-6, 1.05 * lmax, 4001)
    for steps, div in ((ps.STEPS0, ps.DIV0), (8, ps.finest_div(8)), (ps.STEPSC, ps.DIVC)):
        e = ps.smooth(np.diag(mu), np.ones_like(mu), np.zeros_like(mu), np.ones_like(mu), lmax, steps, div)
        hi, lo = 1.05 * lmax, lmax / div
        theta, delta = 0.5 * (hi + lo), 0.5 * (hi - lo)
        T = lambda x: np.cosh(steps * np.arccosh(x)) if np.ndim(x) == 0 else np.where(
            np.abs(x) <= 1, np.cos(steps * np.arccos(np.clip(x, -1, 1))),
            np.sign(x) ** steps * np.cosh(steps * np.arccosh(np.abs(x))))
        closed = T((theta - mu) / delta) / T(theta / delta)
        assert np.abs(e - closed).max() <= 1e-10
        inside = mu >= lo
        assert np.abs(e[inside]).max() <= 1.0 / T(theta / delta) + 1e-12
        assert np.abs(e[~inside]).max() < 1.0


def test_finest_smoother_length_follows_the_size():
    """8 terms on [lmax/100, .] up to 10^4 rows, 4 more per decade, 16 on [lmax/400, .] from
    10^6 rows (the C3 Schur matrix has 907,923 rows)"""
    assert [ps.finest_steps(n) for n in (1, 3101, 7513, 10000, 32768, 100000, 907923, 10 ** 7)] == \
        [8, 8, 8, 8, 10, 12, 16, 16]
    assert ps.finest_div(8) == 100.0 and ps.finest_div(16) == ps.DIV0 == 400.0
