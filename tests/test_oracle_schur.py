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


def jittered_spd(n, lam_min, seed):
    """a dense SPD matrix with a unit-free spread of entries and smallest eigenvalue lam_min
    relative to its diagonal scale (entries not representable in FP16)"""
    rng = np.random.default_rng(seed)
    Q, _ = np.linalg.qr(rng.standard_normal((n, n)))
    lam = np.concatenate([[lam_min], rng.uniform(0.05, 2.0, n - 1)])
    A = (Q * lam) @ Q.T
    d = 1.0 + rng.uniform(0.0, 3.0, n)  # a non-trivial diagonal scaling
    return sp.csr_matrix(np.sqrt(d)[:, None] * A * np.sqrt(d)[None, :])


def test_half_operator_is_exact_when_fp16_holds_the_scaled_entries():
    """the 2-D Laplacian without shift: every diagonal 4, s = 1/2, scaled entries -1/4 and 1,
    all exact in FP16 and FP64, so Delta = 0 and H is the scaled matrix itself"""
    n = 8
    T = sp.diags([-np.ones(n - 1), 2 * np.ones(n), -np.ones(n - 1)], [-1, 0, 1])
    A = (sp.kron(T, sp.identity(n)) + sp.kron(sp.identity(n), T)).tocsr()
    s, H = ps.half_operator(A)
    assert np.all(s == 0.5)
    assert np.array_equal(H.toarray(), 0.25 * A.toarray())


def test_half_operator_dominates_the_scaled_matrix_and_stays_definite():
    """H - A^ is positive semi-definite (diagonally dominant by construction), so H is SPD
    with A; on a matrix whose smallest scaled eigenvalue lies below the FP16 rounding, plain
    rounding of the off-diagonal entries (unit diagonal) gives an indefinite matrix and H
    does not.  The seed is fixed where plain rounding fails (asserted)."""
    A = jittered_spd(40, 2e-6, seed=3)
    s, H = ps.half_operator(A)
    Ad = A.toarray()
    Ah = s[:, None] * Ad * s[None, :]
    Hd = H.toarray()
    assert np.array_equal(Hd, Hd.T)
    off = ~np.eye(40, dtype=bool)
    # within half an FP16 spacing (relative 2^-11, absolute 2^-25 among the subnormals)
    assert np.all(np.abs(Hd - Ah)[off] <= 2.0 ** -11 * np.abs(Ah)[off] + 2.0 ** -25)
    assert np.linalg.eigvalsh(Hd - Ah).min() >= -1e-14
    plain = np.where(off, ps.round_half(Ah), 1.0)
    assert np.linalg.eigvalsh(0.5 * (plain + plain.T)).min() < 0
    assert np.linalg.eigvalsh(Hd).min() > 0
    # dropping the compensation (a unit diagonal) is what fails: the diagonal carries all of it
    assert np.all(np.diag(Hd) > 1.0)


def test_round_half_is_one_correctly_rounded_step():
    """FP64 -> FP16 rounds once: a value just above an FP16 midpoint goes up (two roundings
    through FP32 would land on the midpoint and go to the even neighbour below), ties go to the
    even neighbour, and the subnormal spacing is 2^-24"""
    lo, hi = 1.0, 1.0 + 2.0 ** -10          # neighbouring FP16 values, lo even
    mid = 0.5 * (lo + hi)
    assert ps.round_half(mid + 2.0 ** -40) == hi
    assert np.float64(np.float32(mid + 2.0 ** -40)) == mid  # (FP32 alone would make it a tie)
    assert ps.round_half(mid) == lo and ps.round_half(mid - 2.0 ** -40) == lo
    assert ps.round_half(0.5 * (hi + hi + 2.0 ** -10)) == hi + 2.0 ** -10  # tie up to the even one
    assert ps.round_half(3 * 2.0 ** -24) == 3 * 2.0 ** -24 and ps.round_half(2.0 ** -26) == 0.0


def test_scaled_three_term_sweep_is_the_two_term_sweep():
    """smooth_scaled on H in the scaled variables equals smooth (Jacobi-Chebyshev, two-term)
    on the unscaled operator diag(1/s) H diag(1/s) with D^-1 = s^2, from zero and from a
    non-zero start"""
    A = jittered_spd(30, 1e-2, seed=5)
    s, H = ps.half_operator(A)
    Sp = sp.diags(1 / s) @ H @ sp.diags(1 / s)
    rng = np.random.default_rng(7)
    r, x0 = rng.standard_normal(30), rng.standard_normal(30)
    lmax = float(np.max(np.linalg.eigvalsh((sp.diags(s) @ Sp @ sp.diags(s)).toarray())))
    for steps, div in ((16, 400.0), (9, ps.finest_div(9))):
        for start in (np.zeros(30), x0):
            two = ps.smooth(Sp, s * s, r, start, lmax, steps, div)
            three = s * ps.smooth_scaled(H, s * r, start / s, lmax, steps, div)
            assert np.abs(two - three).max() <= 1e-10 * np.abs(two).max()


def test_polynomial_length_follows_the_size():
    """8 terms on [lmax/200, .] up to 10^4 rows, 8 more per decade, 24 on [lmax/400, .] from
    about 9 10^5 rows (the 120x120x60 box's Schur matrix has 907,923 rows)"""
    assert [ps.poly_steps(n) for n in (1, 3101, 10000, 31623, 100000, 907923, 10 ** 7)] == \
        [8, 8, 8, 12, 16, 24, 24]
    assert [ps.poly_div(k) for k in (8, 12, 16, 24)] == [200.0, 300.0, 400.0, 400.0]
    assert [ps.poly_beta(k) for k in (8, 12, 16, 24)] == [1.0, 1.375, 1.75, 2.5]


def test_polynomial_stage_is_spd_and_contracts():
    """z = B r of the one-sweep stage, both forms: B symmetric positive definite, and without
    its weight (beta = 1) I - B A contracts in the energy norm (the minimax polynomial on
    [lmax/div, 1.05 lmax] with lmax the largest eigenvalue of D^-1 A); the weighted stage is
    poly_beta(terms) times the unweighted one"""
    A = laplacian(8)
    n = A.shape[0]
    lmax = float(np.max(np.linalg.eigvals((sp.diags(1 / A.diagonal()) @ A).toarray()).real))
    B = np.stack([ps.polynomial(A, lmax, e, beta=1.0) for e in np.eye(n)], axis=1)
    assert np.abs(B - B.T).max() <= 1e-12 * np.abs(B).max()
    assert np.linalg.eigvalsh(0.5 * (B + B.T)).min() > 0
    Bw = np.stack([ps.polynomial(A, lmax, e) for e in np.eye(n)], axis=1)
    assert np.abs(Bw - ps.poly_beta(ps.poly_steps(n)) * B).max() <= 1e-14 * np.abs(Bw).max()
    Ad = ps.rounded_f32(A).toarray()
    L = np.linalg.cholesky(Ad)
    E = L.T @ (np.eye(n) - B @ Ad) @ np.linalg.inv(L.T)
    assert np.linalg.norm(E, 2) < 1.0
    s, H = ps.half_operator(A)
    lh = float(np.max(np.linalg.eigvalsh(H.toarray())))
    Bh = np.stack([ps.polynomial(A, lh, e, half=(s, H), use_half=True) for e in np.eye(n)], axis=1)
    assert np.abs(Bh - Bh.T).max() <= 1e-12 * np.abs(Bh).max()
    assert np.linalg.eigvalsh(0.5 * (Bh + Bh.T)).min() > 0


def test_polynomial_stage_is_the_sweep_from_zero():
    """the one-sweep stage equals `smooth` from x = 0 with the stage's length and interval on the
    FP32-rounded matrix, and the scaled three-term sweep on the half-precision operator"""
    A = jittered_spd(30, 1e-2, seed=11)
    r = np.random.default_rng(2).standard_normal(30)
    lmax = 1.7
    k = ps.poly_steps(30)
    ref = ps.smooth(ps.rounded_f32(A), 1 / A.diagonal(), r, np.zeros(30), lmax, k, ps.poly_div(k))
    assert np.array_equal(ps.polynomial(A, lmax, r, beta=1.0), ref)
    assert np.array_equal(ps.polynomial(A, lmax, r), ps.poly_beta(k) * ref)
    s, H = ps.half_operator(A)
    refh = s * ps.smooth_scaled(H, s * r, np.zeros(30), lmax, k, ps.poly_div(k))
    assert np.array_equal(ps.polynomial(A, lmax, r, use_half=True, beta=1.0), refh)
