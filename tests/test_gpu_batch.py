"""Batched path: the batched assembly is bit-identical to the single-model
assembly; the batched Newton solve (persistent team MINRES, large systems on
the whole device, per-model convergence) reaches the single-model solution
within 1e-8; the autotuner returns a valid configuration."""
import numpy as np
import pytest

import nla3d_amd
from nla3d_amd import meshgen

pytestmark = pytest.mark.gpu


def pair(count, seed, n_range):
    return meshgen.batch_models(count, seed, n_range), meshgen.batch_models(count, seed, n_range)


def test_batch_assembly_is_bitwise_single_assembly(gpu_lib):
    print("[gpu] batch assembly", flush=True)
    A, B = pair(5, 3, (2, 6))
    batch = nla3d_amd.Batch([m for m, _ in A])
    batch.assemble(0.0)
    for (ma, _), (mb, _) in zip(A, B):
        mb.assemble()
        for blk in (1, 12, 2):
            assert np.array_equal(ma.get_K(blk), mb.get_K(blk))
        assert np.array_equal(ma.get_F(), mb.get_F())


def solve_pair(A, B, opts, **cfg):
    batch = nla3d_amd.Batch([m for m, _ in A])
    if cfg:
        batch.set_config(**cfg)
    stats, conv = batch.solve_nonlinear(opts)
    assert np.all(conv == 1), conv
    for k, ((ma, p), (mb, _)) in enumerate(zip(A, B)):
        st = mb.solve_nonlinear(opts)
        ua, ub = ma.get_solution(), mb.get_solution()
        assert np.abs(ua - ub).max() <= 1e-8 * np.abs(ub).max(), (k, p)
        assert stats[k].load_steps == st.load_steps
    return stats


def test_batch_newton_matches_single_models(gpu_lib):
    print("[gpu] batch newton (team solves)", flush=True)
    A, B = pair(6, 4, (2, 7))
    opts = nla3d_amd.GpuModel.newton_defaults(criterion=1, convergence=1e-9, krylov_tol=1e-12, krylov_max_iter=20000,
                                              n_load_steps=3, max_iterations=12)
    stats = solve_pair(A, B, opts)
    print("Krylov iterations per model", [s.krylov_iterations for s in stats], flush=True)


def test_batch_split_and_autotune(gpu_lib):
    print("[gpu] batch autotune and large-system split", flush=True)
    A, B = pair(6, 5, (2, 7))
    batch = nla3d_amd.Batch([m for m, _ in A])
    cfg, rep = batch.autotune(t=0.5, probe_iterations=8)
    print("autotuned", cfg.as_dict(), flush=True)
    assert cfg.team_threads in (256, 512, 1024) and cfg.team_lanes in (4, 8, 16)
    batch.close()
    sizes = sorted(m.sizes["nUnknown"] for m, _ in A)
    opts = nla3d_amd.GpuModel.newton_defaults(criterion=0, convergence=1e-3, krylov_tol=1e-12, krylov_max_iter=20000,
                                              n_load_steps=2, max_iterations=15)
    # the two largest systems on the whole device, the rest in the team launch
    solve_pair(A, B, opts, big_rows=sizes[-3])


def test_batch_whole_device_solves_with_schur_multigrid(gpu_lib):
    """preconditioner 2 in a batch: the three largest systems on the whole device with the
    Schur-row multigrid, the rest in the team launch; every model reaches the solution of
    its own single-model Newton solve with block Jacobi (1e-8), with no fallback"""
    print("[gpu] batch newton, whole-device solves with the Schur multigrid", flush=True)
    A, B = pair(6, 6, (3, 8))
    sizes = sorted(m.sizes["nUnknown"] for m, _ in A)
    opts = nla3d_amd.GpuModel.newton_defaults(criterion=1, convergence=1e-9, krylov_tol=1e-12, krylov_max_iter=20000,
                                              n_load_steps=2, max_iterations=12)
    opts.preconditioner = 2
    batch = nla3d_amd.Batch([m for m, _ in A])
    batch.set_config(big_rows=sizes[-4])
    stats, conv = batch.solve_nonlinear(opts)
    assert np.all(conv == 1), conv
    opts.preconditioner = 0
    for k, ((ma, p), (mb, _)) in enumerate(zip(A, B)):
        st = mb.solve_nonlinear(opts)
        ua, ub = ma.get_solution(), mb.get_solution()
        assert np.abs(ua - ub).max() <= 1e-8 * np.abs(ub).max(), (k, p)
        assert stats[k].precond_fallbacks == 0, (k, p)
    big = [k for k, (m, _) in enumerate(A) if m.sizes["nUnknown"] > sizes[-4]]
    print("whole-device models", big, "Krylov iterations (batch, multigrid on those):",
          [s.krylov_iterations for s in stats], flush=True)


def test_autotune_leaves_models_unchanged(gpu_lib):
    """the tuning probes assemble and apply the prescribed values at t; every model's
    solution vector must be bit-identical afterwards, and a Newton solve after tuning
    must reach the untuned models' solution (1e-8)"""
    print("[gpu] autotune leaves the models' state", flush=True)
    A, B = pair(5, 8, (3, 9))
    before = [m.get_solution().copy() for m, _ in A]
    batch = nla3d_amd.Batch([m for m, _ in A])
    batch.autotune(t=0.5, probe_iterations=8, preconditioner=2)  # multigrid probes on the whole-device side
    for (m, _), u0 in zip(A, before):
        assert np.array_equal(m.get_solution(), u0)
    opts = nla3d_amd.GpuModel.newton_defaults(criterion=1, convergence=1e-9, krylov_tol=1e-12, krylov_max_iter=20000,
                                              n_load_steps=2, max_iterations=12)
    stats, conv = batch.solve_nonlinear(opts)
    assert np.all(conv == 1), conv
    for (ma, p), (mb, _) in zip(A, B):
        mb.solve_nonlinear(opts)
        ua, ub = ma.get_solution(), mb.get_solution()
        assert np.abs(ua - ub).max() <= 1e-8 * np.abs(ub).max(), p
