"""Pins of the learned-correction oracle (oracle/pylearned.py) on the CPU:
the MLP backward against central differences of the forward, the MINRES
log-rate against the true preconditioned residual of the iterate after k
steps, the recurrence against a dense solve, and the corrected block in
closed form (C = I gives block Jacobi; C = e^a I scales it by e^2a)."""
import math

import numpy as np
import scipy.sparse as sp

from oracle import pylearned as pl


def random_params(rng, out_scale=0.3):
    w = np.zeros(pl.N_PARAMS)
    for l, (o, i) in enumerate(pl.SHAPES):
        w[pl.OFF_W[l]:pl.OFF_W[l] + o * i] = rng.uniform(-1, 1, o * i) * math.sqrt(6.0 / (o + i))
        w[pl.OFF_B[l]:pl.OFF_B[l] + o] = 0.1 * rng.standard_normal(o)
    w[pl.OFF_W[4]:pl.OFF_B[4]] *= out_scale
    return w


def test_mlp_backward_matches_central_differences():
    rng = np.random.default_rng(1)
    w = random_params(rng)
    X = rng.standard_normal((5, pl.N_IN))
    dY = rng.standard_normal((5, pl.N_OUT))
    Y, cache = pl.mlp_forward(w, X)
    g = pl.mlp_backward(w, cache, dY)
    idx = np.concatenate([rng.choice(np.arange(pl.OFF_W[l], pl.OFF_B[l] + pl.SHAPES[l][0]), 4, replace=False)
                          for l in range(5)])
    for j in idx:
        h = 1e-6
        wp, wm = w.copy(), w.copy()
        wp[j] += h
        wm[j] -= h
        fd = ((pl.mlp_forward(wp, X)[0] - pl.mlp_forward(wm, X)[0]) * dY).sum() / (2 * h)
        assert abs(fd - g[j]) <= 1e-6 * max(1.0, abs(fd)), (j, fd, g[j])


def saddle(rng, n=40, m=8):
    A = rng.standard_normal((n, n))
    A = A @ A.T / n + np.eye(n)
    B = rng.standard_normal((n, m))
    C = 1e-3 * np.eye(m)
    K = np.block([[A, B], [B.T, -C]])
    M = np.diag(np.r_[1.0 / np.diag(A), np.ones(m) / (np.sum(B * B / np.diag(A)[:, None], 0) + 1e-3)])
    return sp.csr_matrix(K), sp.csr_matrix(M)


def test_minres_recurrence_solves_and_rate_is_true_residual():
    rng = np.random.default_rng(2)
    K, M = saddle(rng)
    b = rng.standard_normal(K.shape[0])
    x, it = pl.minres_solve(K, M, b, 1e-13, 500)
    assert np.linalg.norm(K @ x - b) <= 1e-9 * np.linalg.norm(b)
    assert np.abs(x - np.linalg.solve(K.toarray(), b)).max() <= 1e-8 * np.abs(x).max()
    for k in (1, 5, 17):
        xk, _ = pl.minres_solve(K, M, b, 0.0, k)
        r = b - K @ xk
        true = math.log(math.sqrt(r @ (M @ r)) / math.sqrt(b @ (M @ b))) / k
        assert abs(pl.minres_log_rate(K, M, b, k) - true) <= 1e-9


def test_corrected_block_closed_forms():
    rng = np.random.default_rng(3)
    A = rng.standard_normal((9, 9))
    A = A @ A.T + 9 * np.eye(9)
    K = sp.csr_matrix(A)
    rows = np.array([0, 3, 6])
    P0 = pl.corrected_blocks(K, rows, np.zeros((3, 6)))
    for r, P in zip(rows, P0):
        assert np.abs(P - np.linalg.inv(A[r:r + 3, r:r + 3])).max() <= 1e-12 * np.abs(P).max()
    a = 0.37
    Pa = pl.corrected_blocks(K, rows, np.tile([a, a, a, 0, 0, 0], (3, 1)))
    assert np.abs(Pa - math.exp(2 * a) * P0).max() <= 1e-12 * np.abs(Pa).max()
    Pr = pl.corrected_blocks(K, rows, rng.standard_normal((3, 6)))
    for P in Pr:
        assert np.abs(P - P.T).max() <= 1e-12 * np.abs(P).max()
        assert np.linalg.eigvalsh(P).min() > 0
