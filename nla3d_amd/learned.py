"""Training and judging of the learned block-Jacobi correction.

The network (nla3d_amd.Mlp) is trained on tangent matrices and right-hand sides
that the solver itself produces: SOLID81 rigid-top boxes of 8..16 elements
across with seeded material and load variations (Mooney-Rivlin C10 ~ U[0.2, 1],
C01 ~ U[0, 0.3], K/C10 ~ U[50, 2000], top UZ ~ U[5%, 30%], twist ~ U[0, 20 deg])
are driven through Newton iterations with plain block Jacobi, and after each
iteration's assembly the current (tangent, rhs) pair (optionally with seeded
random right-hand sides on the same tangent) feeds an Adam step of the
unrolled-MINRES loss.
Block Jacobi is the judge: `judge` solves one system both ways.
"""
import ctypes
import math
import time

import numpy as np

import nla3d_amd
from nla3d_amd import meshgen


def training_model(rng, n, device=0):
    """one seeded rigid-top box of n x n x max(2, n//2) hexahedra; returns (model, sizes)"""
    c10 = rng.uniform(0.2, 1.0)
    c01 = rng.uniform(0.0, 0.3)
    bulk = c10 * rng.uniform(50.0, 2000.0)
    uz = rng.uniform(0.05, 0.30)
    rotz = rng.uniform(0.0, math.radians(20.0))
    m = nla3d_amd.GpuModel(device)
    meshgen.solid81_rigid_top_case(m, n, n, max(2, n // 2), 2.0, 2.0, 1.0, uz, rotz, jitter=0.1,
                                   seed=int(rng.integers(1, 2**31)), material=(3, (c10, c01, bulk)))
    m.set_max_in_row(0)
    m.init_solution()
    return m


def validation_system(n=12, seed=777, t=0.5, device=0):
    """(model, rhs, t): an n x n x n/2 rigid-top box (C10 0.5, C01 0.1, K 1000, +15% UZ,
    20 deg twist) after one Newton iteration at time t; rhs is that iteration's right-hand
    side and the assembled tangent stays on the device"""
    m = nla3d_amd.GpuModel(device)
    meshgen.solid81_rigid_top_case(m, n, n, max(2, n // 2), 2.0, 2.0, 1.0, 0.15, math.radians(20.0), seed=seed,
                                   material=(3, (0.5, 0.1, 1000.0)))
    m.set_max_in_row(0)
    m.init_solution()
    opts = m.newton_defaults(criterion=1, convergence=1e-8, krylov_tol=1e-10, krylov_max_iter=50000)
    m.newton_iteration(t, 1, ctypes.c_double(0.0), opts, nla3d_amd.NewtonStats())
    return m, m.get_rhs(), t


def set_normalization_from(net, model, t=0.5):
    X, _ = model.node_features(t)
    mean = X.mean(0)
    std = X.std(0)
    net.set_normalization(mean, np.where(std > 1e-6, 1.0 / np.maximum(std, 1e-6), 1.0))


def train(net, n_pairs=2000, k=100, lr=1e-4, seed=20261203, sizes=(8, 10, 12, 16), load_steps=4,
          random_rhs_per_tangent=0, device=0, validation=(), check_every=250, log=None, multigrid=False):
    """Adam steps over n_pairs (tangent, rhs) pairs harvested from Newton iterations of
    seeded training models.

    validation: (model, rhs, t) systems judged every check_every pairs by full MINRES
    solves to 1e-9; the parameters with the fewest summed iterations are kept, the
    untrained network (exactly block Jacobi) being the first candidate.  multigrid: train
    and judge under the Schur-row multigrid (preconditioner 3 against 2).  Returns a dict
    with the loss history and the validation record."""
    rng = np.random.default_rng(seed)
    opts = nla3d_amd.GpuModel.newton_defaults(criterion=1, convergence=1e-6, krylov_tol=1e-8,
                                              krylov_max_iter=20000)
    losses, record, t0 = [], [], time.perf_counter()
    best = None

    def evaluate():
        nonlocal best
        its = [judge(vm, net, vr, t=vt, base=2 if multigrid else 0)[1] for vm, vr, vt in validation]
        record.append((pairs, its))
        if best is None or sum(its) < best[0]:
            best = (sum(its), net.get_params())
        if log:
            log(f"[train] {pairs} pairs: validation iterations {its}")

    pairs = 0
    normalized = False
    while pairs < n_pairs:
        m = training_model(rng, int(rng.choice(sizes)), device)
        if not normalized:
            set_normalization_from(net, m)
            normalized = True
            if validation:
                evaluate()
        st = nla3d_amd.NewtonStats()
        rhs1 = ctypes.c_double(0.0)
        n2 = m.sizes["nUnknown"] + m.sizes["nMpc"]
        conv = False
        for s in range(1, load_steps + 1):
            t = s / load_steps
            for it in range(1, 6):
                solves = st.linear_solves
                try:
                    conv = m.newton_iteration(t, it, rhs1, opts, st)
                except nla3d_amd.Nla3dGpuError:
                    conv = None  # a diverged training model is dropped
                    break
                if st.linear_solves > solves:
                    # the tangent and rhs of this iteration are still on the device
                    for q in range(1 + random_rhs_per_tangent):
                        rhs = None if q == 0 else rng.standard_normal(n2)
                        try:
                            losses.append(net.train_step(m, rhs, k=k, t=t, lr=lr, multigrid=multigrid))
                        except nla3d_amd.Nla3dGpuError:
                            continue  # Krylov space exhausted on a tiny model: skip the sample
                        pairs += 1
                        if validation and pairs % check_every == 0:
                            evaluate()
                        if pairs >= n_pairs:
                            break
                if conv or pairs >= n_pairs:
                    break
            if pairs >= n_pairs or conv is None:
                break
        m.close()
    if validation:
        if pairs % check_every:
            evaluate()
        net.set_params(best[1])
    return {"losses": losses, "pairs": pairs, "seconds": time.perf_counter() - t0, "validation": record,
            "best_validation_iterations": best[0] if best else None}


def judge(model, net, rhs, tol=1e-9, max_iter=100000, t=1.0, base=0):
    """solve K.block(2) x = rhs without (preconditioner base: 0 block Jacobi, 2 with the
    Schur-row multigrid) and with the learned corrections (base + 1); returns
    (iterations without, iterations with, max |x_without - x_with| / max |x_without|)"""
    model.attach_mlp(net)
    x0, it0, _ = model.solve_block2_pc(rhs, tol, max_iter, preconditioner=base, t=t)
    x1, it1, _ = model.solve_block2_pc(rhs, tol, max_iter, preconditioner=base + 1, t=t)
    return it0, it1, float(np.abs(x0 - x1).max() / max(np.abs(x0).max(), 1e-300))
