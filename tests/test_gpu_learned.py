"""GPU parity of the learned block-Jacobi correction against oracle/pylearned.py:
the FP32 forward/backward (1e-5 / 1e-4 of the largest value), the BF16 MFMA
forward (3e-2 of the largest output), the node features (1e-5), the corrected
blocks (1e-12), the unrolled-MINRES loss (1e-6) and its gradient against
central differences of the oracle loss; then training and the block-Jacobi
judge: the converged solution does not depend on the preconditioner (1e-8)."""
import ctypes
import math

import numpy as np
import pytest
import scipy.sparse as sp

import nla3d_amd
from nla3d_amd import learned, meshgen
from oracle import pylearned as pl

pytestmark = pytest.mark.gpu


def random_params(rng, out_scale=0.3):
    w = np.zeros(pl.N_PARAMS)
    for l, (o, i) in enumerate(pl.SHAPES):
        w[pl.OFF_W[l]:pl.OFF_W[l] + o * i] = rng.uniform(-1, 1, o * i) * math.sqrt(6.0 / (o + i))
        w[pl.OFF_B[l]:pl.OFF_B[l] + o] = 0.1 * rng.standard_normal(o)
    w[pl.OFF_W[4]:pl.OFF_B[4]] *= out_scale
    return w.astype(np.float32)


def rigid_model(nx, ny, nz, t=0.5, seed=meshgen.DEFAULT_SEED):
    m = nla3d_amd.GpuModel(0)
    meshgen.solid81_rigid_top_case(m, nx, ny, nz, 2.0, 2.0, 1.0, 0.15, 0.349, material=(3, (0.5, 0.1, 1000.0)),
                                   seed=seed)
    m.set_max_in_row(0)
    m.init_solution()
    opts = m.newton_defaults(criterion=1, convergence=1e-8, krylov_tol=1e-12, krylov_max_iter=20000)
    m.newton_iteration(t, 1, ctypes.c_double(0.0), opts, nla3d_amd.NewtonStats())
    return m


def oracle_inputs(m):
    sz = m.sizes
    nC, n2 = sz["nConstrained"], sz["nUnknown"] + sz["nMpc"]
    io, cols = m.pattern(2)
    rows = np.repeat(np.arange(n2), np.diff(io.astype(np.int64)))
    A = sp.csr_matrix((m.get_K(2), (rows, cols.astype(np.int64) - 1)), shape=(n2, n2))
    K2 = (A + sp.triu(A, 1).T).tocsr()
    ne, ee = m.equation_numbers()
    mpc = list(range(sz["nDofs"] + 1, sz["nDofs"] + 1 + sz["nMpc"]))
    _, c12 = m.pattern(12)
    k12 = np.bincount(c12.astype(np.int64) - 1, minlength=n2)
    groups, lrows, kind, rnode = pl.block_structure(ne, ee, nC, n2, mpc, K2)
    bj = pl.block_jacobi(K2, groups)
    return dict(K2=K2, nC=nC, n2=n2, k12=k12, groups=groups, lrows=lrows, kind=kind, rnode=rnode, bj=bj,
                U=m.get_solution())


def test_mlp_forward_fp32_and_bf16_mfma(gpu_lib):
    print("[gpu] mlp forward", flush=True)
    rng = np.random.default_rng(11)
    w = random_params(rng, out_scale=1.0)
    net = nla3d_amd.Mlp(0, seed=1)
    net.set_params(w)
    mean = rng.standard_normal(24).astype(np.float32) * 0.1
    istd = rng.uniform(0.5, 2.0, 24).astype(np.float32)
    net.set_normalization(mean, istd)
    X = rng.standard_normal((1000, 24)).astype(np.float32)
    Yo, _ = pl.mlp_forward(w.astype(np.float64), X.astype(np.float64), mean, istd)
    Y32 = net.forward(X, bf16=False)
    Ybf = net.forward(X, bf16=True)
    scale = np.abs(Yo).max()
    assert np.abs(Y32 - Yo).max() <= 1e-5 * scale
    assert np.abs(Ybf - Yo).max() <= 3e-2 * scale
    assert np.sqrt(np.mean((Ybf - Yo) ** 2)) <= 1e-2 * np.sqrt(np.mean(Yo ** 2))


def test_mlp_backward(gpu_lib):
    print("[gpu] mlp backward", flush=True)
    rng = np.random.default_rng(12)
    w = random_params(rng)
    net = nla3d_amd.Mlp(0, seed=2)
    net.set_params(w)
    X = rng.standard_normal((2500, 24)).astype(np.float32)  # more than one reduction chunk
    dY = rng.standard_normal((2500, 6)).astype(np.float32)
    g = net.backward(X, dY)
    _, cache = pl.mlp_forward(w.astype(np.float64), X.astype(np.float64))
    go = pl.mlp_backward(w.astype(np.float64), cache, dY.astype(np.float64))
    assert np.abs(g - go).max() <= 1e-4 * np.abs(go).max()


def test_features_and_corrected_blocks(gpu_lib):
    print("[gpu] node features and corrected blocks", flush=True)
    m = rigid_model(4, 4, 2)
    o = oracle_inputs(m)
    X, rows = m.node_features(0.5)
    assert np.array_equal(rows, o["lrows"])
    Xo = pl.node_features(o["K2"], o["lrows"], o["kind"], o["rnode"], o["k12"], o["bj"], o["U"], o["nC"], 1000.0, 0.5)
    assert np.abs(X - Xo).max() <= 1e-5 * max(1.0, np.abs(Xo).max())
    Y = (0.3 * np.random.default_rng(5).standard_normal((len(rows), 6))).astype(np.float32)
    B = m.learned_blocks(Y)
    Bo = pl.corrected_blocks(o["K2"], o["lrows"], Y.astype(np.float64))
    assert np.abs(B - Bo).max() <= 1e-12 * np.abs(Bo).max()


def test_unrolled_minres_loss_and_gradient(gpu_lib):
    print("[gpu] unrolled MINRES loss gradient", flush=True)
    m = rigid_model(3, 3, 2)
    o = oracle_inputs(m)
    rng = np.random.default_rng(13)
    net = nla3d_amd.Mlp(0, seed=3)
    w = random_params(rng, out_scale=0.2)
    net.set_params(w)
    X, rows = m.node_features(0.5)
    mean, std = X.mean(0), X.std(0)
    istd = np.where(std > 1e-6, 1.0 / np.maximum(std, 1e-6), 1.0).astype(np.float32)
    net.set_normalization(mean.astype(np.float32), istd)
    k = 12
    rhs = rng.standard_normal(o["n2"])
    loss, g = net.loss_grad(m, rhs, k=k, t=0.5)

    def oracle_loss(wv):
        Y, _ = pl.mlp_forward(wv, X.astype(np.float64), mean.astype(np.float32), istd)
        P = pl.corrected_blocks(o["K2"], o["lrows"], Y)
        M = pl.precond_matrix(o["n2"], o["bj"], o["lrows"], P)
        return pl.minres_log_rate(o["K2"], M, rhs, k)

    w64 = w.astype(np.float64)
    Lo = oracle_loss(w64)
    assert abs(loss - Lo) <= 1e-6 * max(1.0, abs(Lo))
    idx = np.concatenate([rng.choice(np.arange(pl.OFF_W[l], pl.OFF_B[l] + pl.SHAPES[l][0]), 3, replace=False)
                          for l in range(5)])
    fd = []
    for j in idx:
        h = 1e-4
        wp, wm = w64.copy(), w64.copy()
        wp[j] += h
        wm[j] -= h
        fd.append((oracle_loss(wp) - oracle_loss(wm)) / (2 * h))
    fd = np.array(fd)
    print("grad", g[idx], "fd", fd, flush=True)
    assert np.all(np.abs(g[idx] - fd) <= 1e-3 * np.abs(fd).max() + 1e-2 * np.abs(fd))


def test_unrolled_minres_gradient_under_schur_multigrid(gpu_lib):
    """preconditioner 3's operator: corrected node blocks, the V-cycle (taken from the device
    column by column, a fixed symmetric matrix) on the Schur rows, block Jacobi elsewhere;
    loss against the oracle's recurrence on that matrix, gradient against its finite differences"""
    print("[gpu] unrolled MINRES loss gradient, multigrid", flush=True)
    m = rigid_model(4, 4, 2)
    o = oracle_inputs(m)
    rng = np.random.default_rng(17)
    net = nla3d_amd.Mlp(0, seed=4)
    w = random_params(rng, out_scale=0.2)
    net.set_params(w)
    X, rows = m.node_features(0.5)
    mean, std = X.mean(0), X.std(0)
    istd = np.where(std > 1e-6, 1.0 / np.maximum(std, 1e-6), 1.0).astype(np.float32)
    net.set_normalization(mean.astype(np.float32), istd)
    k = 12
    rhs = rng.standard_normal(o["n2"])
    loss, g = net.loss_grad(m, rhs, k=k, t=0.5, multigrid=True)
    qrow, _, _, _ = m.schur_matrix()
    nq = len(qrow)
    V = np.column_stack([m.schur_apply(np.eye(nq)[:, j]) for j in range(nq)])
    assert np.abs(V - V.T).max() <= 1e-10 * np.abs(V).max()
    q = qrow.astype(np.int64)

    def oracle_loss(wv):
        Y, _ = pl.mlp_forward(wv, X.astype(np.float64), mean.astype(np.float32), istd)
        P = pl.corrected_blocks(o["K2"], o["lrows"], Y)
        M = pl.precond_matrix(o["n2"], o["bj"], o["lrows"], P).toarray()
        M[np.ix_(q, q)] = V
        return pl.minres_log_rate(o["K2"], M, rhs, k)

    w64 = w.astype(np.float64)
    Lo = oracle_loss(w64)
    assert abs(loss - Lo) <= 1e-6 * max(1.0, abs(Lo))
    idx = np.concatenate([rng.choice(np.arange(pl.OFF_W[l], pl.OFF_B[l] + pl.SHAPES[l][0]), 3, replace=False)
                          for l in range(5)])
    fd = []
    for j in idx:
        h = 1e-4
        wp, wm = w64.copy(), w64.copy()
        wp[j] += h
        wm[j] -= h
        fd.append((oracle_loss(wp) - oracle_loss(wm)) / (2 * h))
    fd = np.array(fd)
    print("grad", g[idx], "fd", fd, flush=True)
    assert np.all(np.abs(g[idx] - fd) <= 1e-3 * np.abs(fd).max() + 1e-2 * np.abs(fd))


def test_training_gain_and_block_jacobi_judge(gpu_lib):
    """train on seeded small models, checkpoint by the block-Jacobi judge on one
    validation system, and measure on a held-out model of another size and seed"""
    print("[gpu] train and judge", flush=True)
    net = nla3d_amd.Mlp(0, seed=7)
    val = rigid_model(8, 8, 4, seed=99)
    test = rigid_model(10, 10, 5, seed=123)
    n = lambda mm: mm.sizes["nUnknown"] + mm.sizes["nMpc"]
    rv = np.random.default_rng(21).standard_normal(n(val))
    rt = np.random.default_rng(22).standard_normal(n(test))
    it0, it_untrained, _ = learned.judge(test, net, rt, tol=1e-9, t=0.5)
    # zero output layer: block Jacobi up to rounding (L^-T L^-1 is formed another way than the
    # block-Jacobi inverse), so the counts agree to within a couple of iterations
    assert abs(it_untrained - it0) <= max(2, it0 // 100)
    loss0 = net.loss_grad(test, rt, k=100, t=0.5)[0]
    info = learned.train(net, n_pairs=1000, k=100, lr=1e-4, seed=5, sizes=(4, 5, 6, 7, 8), random_rhs_per_tangent=3,
                         validation=[(val, rv, 0.5)], log=print)
    loss1 = net.loss_grad(test, rt, k=100, t=0.5)[0]
    it_bj, it_l, diff = learned.judge(test, net, rt, tol=1e-9, t=0.5)
    print(f"training {info['pairs']} pairs in {info['seconds']:.1f}s; validation {info['validation']}", flush=True)
    print(f"held-out 10x10x5 solve to 1e-9: block Jacobi {it_bj} iterations, learned {it_l} "
          f"({100.0 * (it_bj - it_l) / it_bj:.1f}% fewer); 100-step log-rate {loss0:.5f} -> {loss1:.5f}", flush=True)
    assert loss1 < loss0
    assert it_l < it_bj
    assert diff <= 1e-6  # both preconditioners converge to the same solution


def test_newton_with_learned_preconditioner_matches_block_jacobi(gpu_lib):
    """the learned corrections change iteration counts, never the converged state"""
    print("[gpu] newton learned vs block Jacobi", flush=True)
    net = nla3d_amd.Mlp(0, seed=8)
    learned.train(net, n_pairs=200, k=100, lr=1e-4, seed=6, sizes=(4, 6))
    out = []
    for pc in (0, 1):
        m = nla3d_amd.GpuModel(0)
        meshgen.solid81_rigid_top_case(m, 6, 6, 3, 2.0, 2.0, 1.0, 0.15, 0.349, material=(3, (0.5, 0.1, 1000.0)))
        m.set_max_in_row(0)
        m.init_solution()
        m.attach_mlp(net)
        opts = m.newton_defaults(criterion=1, convergence=1e-10, krylov_tol=1e-12, krylov_max_iter=50000,
                                 preconditioner=pc, n_load_steps=2, max_iterations=15)
        st = m.solve_nonlinear(opts)
        out.append((m.get_solution(), st.krylov_iterations, st.precond_fallbacks))
    (u0, k0, _), (u1, k1, fb) = out
    print(f"Krylov iterations: block Jacobi {k0}, learned {k1}, fallbacks {fb}", flush=True)
    assert np.abs(u1 - u0).max() <= 1e-8 * np.abs(u0).max()
