"""The Schur-row stage on the device against oracle/pyschur.py: S (1e-12 of its
largest entry), the default one-sweep Chebyshev stage (1e-12 of the oracle's sweep on the
device's stored operator and eigenvalue estimate), one V-cycle over the device's aggregation
maps and eigenvalue estimates (1e-9), both stages' symmetry, and MINRES / Newton with each: the
same solution as a dense solve and as block Jacobi (1e-8), in fewer iterations."""
import ctypes

import numpy as np
import pytest
import scipy.sparse as sp

import nla3d_amd
from nla3d_amd import meshgen
from oracle import pylearned as pl
from oracle import pyschur as ps

pytestmark = pytest.mark.gpu


def rigid_model(nx, ny, nz, t=0.5):
    m = nla3d_amd.GpuModel(0)
    meshgen.solid81_rigid_top_case(m, nx, ny, nz, 2.0, 2.0, 1.0, 0.15, 0.349, material=(3, (0.5, 0.1, 1000.0)))
    m.set_max_in_row(0)
    m.init_solution()
    opts = m.newton_defaults(criterion=1, convergence=1e-8, krylov_tol=1e-12, krylov_max_iter=20000)
    m.newton_iteration(t, 1, ctypes.c_double(0.0), opts, nla3d_amd.NewtonStats())
    return m


def full_k2(m):
    sz = m.sizes
    n2 = sz["nUnknown"] + sz["nMpc"]
    io, cols = m.pattern(2)
    rows = np.repeat(np.arange(n2), np.diff(io.astype(np.int64)))
    A = sp.csr_matrix((m.get_K(2), (rows, cols.astype(np.int64) - 1)), shape=(n2, n2))
    return (A + sp.triu(A, 1).T).tocsr(), n2


def oracle_parts(m):
    K2, n2 = full_k2(m)
    sz = m.sizes
    ne, ee = m.equation_numbers()
    mpc = list(range(sz["nDofs"] + 1, sz["nDofs"] + 1 + sz["nMpc"]))
    groups, _, _, _ = pl.block_structure(ne, ee, sz["nConstrained"], n2, mpc, K2)
    return K2, n2, groups, pl.block_jacobi(K2, groups)


def gpu_s_dense(m):
    qrow, rp, col, val = m.schur_matrix()
    n = len(qrow)
    return qrow, sp.csr_matrix((val, col, rp), shape=(n, n)).toarray()


def check_half_operator(m):
    """the finest level's stored smoother operator against the oracle's half_operator(S): s
    bitwise, the FP16 entries and the compensated diagonal bitwise except (at most two) entries
    whose FP64 inputs differ because the device reads a diagonal-form row's lower entries from
    its neighbour's upper ones (S is symmetric to rounding, not bitwise); returns (s, H)"""
    qrow, rp, col, val = m.schur_matrix()
    n = len(qrow)
    S = sp.csr_matrix((val, col, rp), shape=(n, n))
    sd, hv = m.schur_smoother_operator()
    so, Ho = ps.half_operator(S)
    assert np.array_equal(sd, so)
    Hd = sp.csr_matrix((hv, S.indices, S.indptr), shape=(n, n))
    D = (Hd - Ho).tocoo()
    bad = D.data != 0
    assert bad.sum() <= 2, bad.sum()
    ref = Ho[D.row, D.col]
    ref = ref.toarray().ravel() if sp.issparse(ref) else np.asarray(ref).ravel()
    assert np.all(np.abs(D.data) <= 2.0 ** -10 * np.abs(ref) + 2.0 ** -24)
    assert abs(Hd - Hd.T).max() == 0.0
    return sd, Hd


def test_schur_matrix_matches_oracle(gpu_lib):
    print("[gpu] schur matrix", flush=True)
    m = rigid_model(4, 4, 2)
    K2, n2, groups, bj = oracle_parts(m)
    q, So = ps.schur_matrix(K2, groups, bj)
    qrow, S = gpu_s_dense(m)
    assert np.array_equal(qrow, q)
    assert np.abs(S - So).max() <= 1e-12 * np.abs(So).max()
    assert np.linalg.eigvalsh(So).min() > 0


def test_schur_weight_follows_the_stage(gpu_lib, monkeypatch):
    """S = alpha B^T D^-1 B - C with alpha = 0.3 under the default one-sweep stage and 1/2
    under the V-cycle (the switch rebuilds S), against the oracle's dense S at 1e-12"""
    print("[gpu] schur weight per stage", flush=True)
    monkeypatch.delenv("NLA3D_GPU_SCHUR_CYCLE", raising=False)
    m = rigid_model(8, 8, 4)
    assert m.schur_info()["levels"] >= 2 and not m.schur_vcycle()
    K2, n2, groups, bj = oracle_parts(m)
    for vcycle, alpha in ((False, ps.SCHUR_ALPHA_SWEEP), (True, ps.SCHUR_ALPHA)):
        m.schur_set_cycle(vcycle)
        q, So = ps.schur_matrix(K2, groups, bj, alpha)
        qrow, S = gpu_s_dense(m)
        assert np.array_equal(qrow, q)
        assert np.abs(S - So).max() <= 1e-12 * np.abs(So).max(), alpha


@pytest.mark.parametrize("smoother", ["auto", "half"])
def test_vcycle_matches_oracle_and_is_symmetric(gpu_lib, smoother, monkeypatch):
    """auto: this small finest level smooths with its FP32 copy; half: forced to the
    half-precision operator"""
    print(f"[gpu] schur v-cycle ({smoother})", flush=True)
    if smoother == "half":
        monkeypatch.setenv("NLA3D_GPU_SCHUR_SMOOTHER", "half")
    else:
        monkeypatch.delenv("NLA3D_GPU_SCHUR_SMOOTHER", raising=False)
    m = rigid_model(8, 8, 4)
    m.schur_set_cycle(True)
    assert m.schur_vcycle()
    fh = m.schur_finest_precision() == 16
    assert fh == (smoother == "half")
    info = m.schur_info()
    print("hierarchy", info, flush=True)
    assert info["levels"] >= 2
    qrow, S = gpu_s_dense(m)
    aggs, lmaxs = [], []
    for l in range(info["levels"] - 1):
        agg, lm = m.schur_level(l, info["sizes"][l])
        aggs.append(agg.astype(np.int64))
        lmaxs.append(lm)
    mats = ps.hierarchy(S, aggs)
    rng = np.random.default_rng(3)
    y, z = rng.standard_normal(len(qrow)), rng.standard_normal(len(qrow))
    by, bz = m.schur_apply(y), m.schur_apply(z)
    if fh:
        half = check_half_operator(m)
        assert np.abs(by - ps.vcycle(mats, aggs, lmaxs, y, half=half)).max() <= 1e-12 * np.abs(by).max()
    bo = ps.vcycle(mats, aggs, lmaxs, y, finest_half=fh)
    # the coarse levels' FP32 copies are rounded from values that the device sums by a gather
    # and the oracle by P^T A P: an entry an ulp apart can round to the neighbouring FP32
    # value, which moves the cycle by far less than 1e-9
    assert np.abs(by - bo).max() <= 1e-9 * np.abs(bo).max()
    assert abs(z @ by - y @ bz) <= 1e-10 * abs(z @ by)
    assert y @ by > 0 and z @ bz > 0


@pytest.mark.parametrize("cycle", ["polynomial", "polynomial-half", "vcycle"])
def test_minres_and_newton_with_schur_multigrid(gpu_lib, cycle, monkeypatch):
    """the sweep (on the FP32 copy, or polynomial-half: on the half-precision operator) runs
    fused with MINRES's preconditioner pass (the start takes the Schur rows' update from the
    block-Jacobi pass, the exit writes their share of the r . z partials); the V-cycle runs
    after the pass"""
    print(f"[gpu] minres / newton with the schur stage ({cycle})", flush=True)
    if cycle == "vcycle":
        monkeypatch.setenv("NLA3D_GPU_SCHUR_CYCLE", "vcycle")
    else:
        monkeypatch.delenv("NLA3D_GPU_SCHUR_CYCLE", raising=False)
    if cycle == "polynomial-half":
        monkeypatch.setenv("NLA3D_GPU_SCHUR_SMOOTHER", "half")
    else:
        monkeypatch.delenv("NLA3D_GPU_SCHUR_SMOOTHER", raising=False)
    m = rigid_model(8, 8, 4)
    assert m.schur_vcycle() == (cycle == "vcycle")
    assert m.schur_finest_precision() == (16 if cycle == "polynomial-half" else 32)
    K2, n2 = full_k2(m)
    rhs = np.random.default_rng(4).standard_normal(n2)
    x0, it0, _ = m.solve_block2_pc(rhs, 1e-10, 50000, preconditioner=0)
    x2, it2, _ = m.solve_block2_pc(rhs, 1e-10, 50000, preconditioner=2)
    xd = np.linalg.solve(K2.toarray(), rhs)
    print(f"MINRES to 1e-10: block Jacobi {it0}, Schur multigrid {it2} iterations", flush=True)
    assert np.abs(x2 - xd).max() <= 1e-8 * np.abs(xd).max()
    assert it2 < it0
    out = []
    for pc in (0, 2):
        mm = nla3d_amd.GpuModel(0)
        meshgen.solid81_rigid_top_case(mm, 6, 6, 3, 2.0, 2.0, 1.0, 0.15, 0.349, material=(3, (0.5, 0.1, 1000.0)))
        mm.set_max_in_row(0)
        mm.init_solution()
        opts = mm.newton_defaults(criterion=1, convergence=1e-10, krylov_tol=1e-12, krylov_max_iter=50000,
                                  preconditioner=pc, n_load_steps=2, max_iterations=15)
        st = mm.solve_nonlinear(opts)
        out.append((mm.get_solution(), st.krylov_iterations, st.precond_fallbacks))
    (u0, k0, _), (u2, k2, fb) = out
    print(f"Newton Krylov iterations: block Jacobi {k0}, Schur multigrid {k2}, fallbacks {fb}", flush=True)
    assert fb == 0
    assert np.abs(u2 - u0).max() <= 1e-8 * np.abs(u0).max()


@pytest.mark.parametrize("smoother", ["auto", "half"])
def test_fused_sweep_matches_the_sweep_after_the_pass(gpu_lib, smoother, monkeypatch):
    """one system (12 x 12 x 6 rigid top, 864 Schur rows: two partials of the fused exit) solved
    with the sweep fused into MINRES's block-Jacobi pass and with the same sweep run after the
    pass (NLA3D_GPU_SCHUR_FUSED=0): the r . z partials are summed in another order, nothing else
    differs, so the iteration counts agree within one and the solutions within 1e-9; the drop-in
    equation solver given the model's DoF map builds the same stage (fused too) and the same
    iteration count within one; all three within 1e-8 of a direct solve"""
    import scipy.sparse.linalg as spla
    print(f"[gpu] fused against unfused Schur sweep ({smoother})", flush=True)
    monkeypatch.delenv("NLA3D_GPU_SCHUR_CYCLE", raising=False)
    monkeypatch.delenv("NLA3D_GPU_SCHUR_FUSED", raising=False)
    if smoother == "half":
        monkeypatch.setenv("NLA3D_GPU_SCHUR_SMOOTHER", "half")
    else:
        monkeypatch.delenv("NLA3D_GPU_SCHUR_SMOOTHER", raising=False)
    m = rigid_model(12, 12, 6)
    assert not m.schur_vcycle()
    assert m.schur_finest_precision() == (16 if smoother == "half" else 32)
    K2, n2 = full_k2(m)
    rhs = np.random.default_rng(6).standard_normal(n2)
    xd = spla.spsolve(K2.tocsc(), rhs)
    scale = np.abs(xd).max()
    xf, itf, _ = m.solve_block2_pc(rhs, 1e-10, 50000, preconditioner=2)
    monkeypatch.setenv("NLA3D_GPU_SCHUR_FUSED", "0")
    xu, itu, _ = m.solve_block2_pc(rhs, 1e-10, 50000, preconditioner=2)
    monkeypatch.delenv("NLA3D_GPU_SCHUR_FUSED")
    sz = m.sizes
    nC = sz["nConstrained"]
    ne, ee = m.equation_numbers()
    rn, rd = np.zeros(n2, dtype=np.uint32), np.full(n2, -1, dtype=np.int32)
    nodes, types = np.nonzero(ne > nC)
    q = ne[nodes, types].astype(np.int64) - nC - 1
    rn[q], rd[q] = nodes + 1, types
    els, types = np.nonzero(ee > nC)
    rd[ee[els, types].astype(np.int64) - nC - 1] = types
    s = nla3d_amd.EqSolver(0, tol=1e-10, max_iter=50000, preconditioner=2)
    s.set_dof_map(rn, rd)
    io, cols = m.pattern(2)
    xe = s.solve(m.get_K(2), io, cols, rhs)
    st = s.stats()
    print(f"MINRES iterations: fused {itf}, after the pass {itu}, equation solver {st.iterations}; "
          f"fused - unfused {np.abs(xf - xu).max() / scale:.2e}", flush=True)
    assert st.precond_fallbacks == 0 and st.breakdown == 0
    assert abs(itf - itu) <= 1 and abs(st.iterations - itf) <= 1
    assert np.abs(xf - xu).max() <= 1e-9 * scale
    for x in (xf, xu, xe):
        assert np.abs(x - xd).max() <= 1e-8 * scale


def test_smoother_timing_entry(gpu_lib):
    print("[gpu] smoother timing entry", flush=True)
    m = rigid_model(8, 8, 4)  # large enough for a smoothed level below the dense coarsest one
    ms, nbytes = m.time_smoother(4)
    info = m.schur_info()
    assert ms > 0.0
    # at least one byte per entry (the half-precision operator stores each off-diagonal entry of
    # the diagonal-form rows once, in two bytes) and the step's four vectors
    assert nbytes >= info["nnz0"] + 4 * 8 * info["nq"]


def a2000_damper_model():
    import os
    from nla3d_amd.cdb import read_cdb
    d = read_cdb(os.path.join(os.path.dirname(__file__), "golden", "a2000_damper", "model.cdb.gz"))
    m = nla3d_amd.GpuModel(0)
    m.set_nodes(d["nodes"])
    m.add_elements(nla3d_amd.PLANE41, [n for t, n in d["cells"] if t == 1])
    m.set_material(nla3d_amd.NEO_HOOKEAN, [1.0, 500.0])
    for n, dof, v in d["fixes"]:
        m.add_fix(n, dof, v)
    m.init_solution()
    opts = m.newton_defaults(criterion=1, convergence=1e-8, krylov_tol=1e-12, krylov_max_iter=20000)
    m.newton_iteration(0.3, 1, ctypes.c_double(0.0), opts, nla3d_amd.NewtonStats())
    return m


@pytest.mark.parametrize("case,smoother", [("box_30x30x15", "auto"), ("box_30x30x15", "half"),
                                           ("a2000_damper", "auto"), ("a2000_damper", "half")])
def test_vcycle_matches_oracle_on_deep_hierarchies(gpu_lib, case, smoother, monkeypatch):
    """The V-cycle against the oracle where the short tests do not reach: a 30x30x15
    rigid-top box (16,383 Schur rows: an odd, 9-term finest sweep, which ends on a step that
    leaves x alone) and the reference's a2000_damper model (PLANE41, 3,101 Schur rows: the
    coarse levels below a smoothed one run as one workgroup launch, fuse_from >= 1); each with
    the finest level's FP32 copy (the choice at these sizes) and forced to the half-precision
    operator"""
    print(f"[gpu] schur v-cycle, {case} ({smoother})", flush=True)
    if smoother == "half":
        monkeypatch.setenv("NLA3D_GPU_SCHUR_SMOOTHER", "half")
    else:
        monkeypatch.delenv("NLA3D_GPU_SCHUR_SMOOTHER", raising=False)
    m = rigid_model(30, 30, 15) if case == "box_30x30x15" else a2000_damper_model()
    m.schur_set_cycle(True)
    fh = m.schur_finest_precision() == 16
    assert fh == (smoother == "half")
    info = m.schur_info()
    steps, fuse = m.schur_cycle_info()
    print("hierarchy", info, "finest steps", steps, "fused from", fuse, flush=True)
    assert info["levels"] >= 3
    if case == "box_30x30x15":
        assert steps == ps.finest_steps(info["nq"]) == 9
    else:
        assert steps == ps.finest_steps(info["nq"]) and 1 <= fuse < info["levels"] - 1
    qrow, rp, col, val = m.schur_matrix()
    S = sp.csr_matrix((val, col, rp), shape=(len(qrow), len(qrow)))
    aggs, lmaxs = [], []
    for l in range(info["levels"] - 1):
        agg, lm = m.schur_level(l, info["sizes"][l])
        aggs.append(agg.astype(np.int64))
        lmaxs.append(lm)
    mats = ps.hierarchy(S, aggs)
    assert [A.shape[0] for A in mats] == info["sizes"]
    rng = np.random.default_rng(9)
    y, z = rng.standard_normal(len(qrow)), rng.standard_normal(len(qrow))
    by, bz = m.schur_apply(y), m.schur_apply(z)
    if fh:
        half = check_half_operator(m)
        assert np.abs(by - ps.vcycle(mats, aggs, lmaxs, y, half=half)).max() <= 1e-12 * np.abs(by).max()
    bo = ps.vcycle(mats, aggs, lmaxs, y, finest_half=fh)
    assert np.abs(by - bo).max() <= 1e-9 * np.abs(bo).max()
    assert abs(z @ by - y @ bz) <= 1e-10 * abs(z @ by)
    assert y @ by > 0 and z @ bz > 0


@pytest.mark.parametrize("case,smoother", [("box_8x8x4", "auto"), ("box_8x8x4", "half"), ("box_30x30x15", "auto"),
                                           ("box_30x30x15", "half"), ("box_28x28x14", "auto"),
                                           ("box_28x28x14", "half"), ("a2000_damper", "auto")])
def test_polynomial_stage_matches_oracle(gpu_lib, case, smoother, monkeypatch):
    """the default Schur stage, one Chebyshev sweep on S (no coarse levels), against the
    oracle's sweep on the device's S and eigenvalue estimate: its length follows the rows
    (8 terms at 499 and 3,101 rows, an odd 9 at 13,499, 10 at 16,383), on the FP32 copy (the
    choice at these sizes) and forced to the half-precision operator (then on the stored
    operator, 1e-12); symmetric and positive; the eigenvalue estimate bounds the operator the
    sweep applies (its power iteration runs on that operator)"""
    print(f"[gpu] schur polynomial stage, {case} ({smoother})", flush=True)
    if smoother == "half":
        monkeypatch.setenv("NLA3D_GPU_SCHUR_SMOOTHER", "half")
    else:
        monkeypatch.delenv("NLA3D_GPU_SCHUR_SMOOTHER", raising=False)
    monkeypatch.delenv("NLA3D_GPU_SCHUR_CYCLE", raising=False)
    m = a2000_damper_model() if case == "a2000_damper" else \
        rigid_model(*[int(v) for v in case.split("_")[1].split("x")])
    assert not m.schur_vcycle()
    fh = m.schur_finest_precision() == 16
    assert fh == (smoother == "half")
    info = m.schur_info()
    steps, _ = m.schur_cycle_info()
    print("rows", info["nq"], "terms", steps, flush=True)
    assert info["levels"] >= 2 and steps == ps.poly_steps(info["nq"])
    qrow, rp, col, val = m.schur_matrix()
    n = len(qrow)
    S = sp.csr_matrix((val, col, rp), shape=(n, n))
    _, lmax = m.schur_level(0, n)
    rng = np.random.default_rng(21)
    y, z = rng.standard_normal(n), rng.standard_normal(n)
    by, bz = m.schur_apply(y), m.schur_apply(z)
    if fh:
        bo = ps.polynomial(S, lmax, y, half=check_half_operator(m), use_half=True)
    else:
        bo = ps.polynomial(S, lmax, y)
    assert np.abs(by - bo).max() <= 1e-12 * np.abs(bo).max()
    assert abs(z @ by - y @ bz) <= 1e-10 * abs(z @ by)
    assert y @ by > 0 and z @ bz > 0
    if case == "box_8x8x4":  # small enough for a dense eigenvalue check
        if fh:
            sh, Hh = check_half_operator(m)
            lam = np.linalg.eigvalsh(Hh.toarray()).max()
        else:
            d = S.diagonal()
            lam = np.linalg.eigvalsh((S / np.sqrt(np.outer(d, d))).toarray()).max()
        # 40 power steps approach lambda_max from below; the stored estimate is 1.05 times theirs
        assert lam * 0.999 <= lmax <= 1.05 * lam * (1 + 1e-12), (lam, lmax)
