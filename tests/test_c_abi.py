"""The C ABI from plain C: tests/c_abi/plane_model.c includes only include/nla3d_gpu.h, links
libnla3d_gpu.so and builds a PLANE41 plate.  On the CPU (host-only handle) its sizes and
equation numbers equal the oracle's; on the GPU its two Newton load steps equal the oracle's
direct-solve Newton to 1e-8 of max |U|."""
import os
import shutil
import subprocess

import numpy as np
import pytest

import nla3d_amd
from nla3d_amd import meshgen
from oracle import pyoracle as po

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def compile_caller(tmp_path):
    """the caller built by __graft_entry__.build() beside its source (it travels with the
    built library); rebuilt into tmp_path here when absent and gcc is present"""
    built = os.path.join(ROOT, "tests", "c_abi", "plane_model")
    nla3d_amd.lib()  # the library must be built
    if os.path.exists(built):
        return built
    if shutil.which("gcc") is None:
        pytest.skip("the C caller is not built and gcc is absent")
    libdir = os.path.dirname(nla3d_amd.LIB_PATH)
    exe = str(tmp_path / "plane_model")
    subprocess.run(["gcc", "-std=c99", "-O1", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"),
                    os.path.join(ROOT, "tests", "c_abi", "plane_model.c"), "-L", libdir, "-lnla3d_gpu",
                    f"-Wl,-rpath,{libdir}", "-o", exe], check=True)
    return exe


def oracle(nx, ny):
    xyz, conn, (I, J) = meshgen.quad_grid(nx, ny, 1.0, 1.0)
    nodes = np.arange(1, len(xyz) + 1)
    fixes = []
    for i in range(nx + 1):
        fixes += [(int(nodes[(J == 0) & (I == i)][0]), 0, 0.0), (int(nodes[(J == 0) & (I == i)][0]), 1, 0.0)]
        top = int(nodes[(J == ny) & (I == i)][0])
        fixes += [(top, 0, 0.0), (top, 1, -0.1)]
    return po.OracleModel(xyz, [(po.PLANE41, list(map(int, c))) for c in conn], material=(1, [1.0, 500.0]),
                          fixes=fixes).setup()


def parse(out):
    sizes, eqs, U = None, {}, {}
    for line in out.splitlines():
        f = line.split()
        if f[0] == "sizes":
            sizes = [int(v) for v in f[1:]]
        elif f[0] == "eq":
            eqs[int(f[1])] = (int(f[2]), int(f[3]))
        elif f[0] == "U":
            U[int(f[1])] = float(f[2])
    return sizes, eqs, np.array([U[k] for k in sorted(U)]) if U else None


def test_host_only_numbering_from_c(tmp_path):
    exe = compile_caller(tmp_path)
    r = subprocess.run([exe, "-1", "5", "3"], capture_output=True, text=True, check=True)
    sizes, eqs, _ = parse(r.stdout)
    o = oracle(5, 3)
    assert sizes[:6] == [24, 15, o.nDofs, o.nC, o.nU, o.nMpc]
    assert sizes[6:] == [len(o.blocks[b].columns) for b in (1, 12, 2)]
    for n in range(1, 25):
        assert eqs[n] == (o.nodeEq[n - 1][po.UX], o.nodeEq[n - 1][po.UY])


@pytest.mark.gpu
def test_newton_from_c_matches_oracle(gpu_lib, tmp_path):
    print("[gpu] C caller: PLANE41 plate, two load steps", flush=True)
    exe = compile_caller(tmp_path)
    r = subprocess.run([exe, "0", "6", "4"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    _, _, U = parse(r.stdout)
    o = oracle(6, 4)
    o.newton_init()
    for t in (0.5, 1.0):
        o.solve_load_step(t, criterion=1, convergence=1e-10, max_iterations=20)
    assert U.shape == o.U.shape
    assert np.abs(U - o.U).max() <= 1e-8 * np.abs(o.U).max()


# ---------------------------------------------------------------------------
# the solver-level entry points from plain C (tests/c_abi/csr_solve.c)
# ---------------------------------------------------------------------------
def csr_caller():
    exe = os.path.join(ROOT, "tests", "c_abi", "csr_solve")
    if not os.path.exists(exe):
        pytest.skip("tests/c_abi/csr_solve is not built (run __graft_entry__.build())")
    return exe


def saddle_case(seed=3):
    """K.block(2) of a jittered 3x3x2 SOLID81 box with a rigid-body top (pressure rows, node
    rows, multiplier rows without a diagonal), assembled by the oracle at a seeded state, its
    DoF map read off the oracle's numbering, and two right-hand sides"""
    xyz, conn, (I, J, K) = meshgen.hex_box(3, 3, 2, 2.0, 2.0, 1.0, jitter=0.1, seed=seed)
    xyz = np.vstack([xyz, [[1.0, 1.0, 1.0]]])
    master = len(xyz)
    fixes = [(int(n), d, 0.0) for n in np.nonzero(K == 0)[0] + 1 for d in range(3)]
    fixes += [(master, d, v) for d, v in enumerate([0.0, 0.0, 0.15, 0.0, 0.0, 0.349])]
    o = po.OracleModel(xyz, [(po.SOLID81, c) for c in conn], material=(3, [0.5, 0.1, 1000.0]), fixes=fixes,
                       rigid_bodies=[(master, list(np.nonzero(K == 2)[0] + 1), [0, 1, 2])]).setup()
    U = 0.01 * np.random.default_rng(seed).standard_normal(o.nAll)
    Kv, _ = o.assemble(U)
    n2 = o.nU + o.nMpc
    row_node, row_dof = np.zeros(n2, dtype=np.int64), np.full(n2, -1, dtype=np.int64)
    for n, d in enumerate(o.nodeEq):
        for t, q in d.items():
            if q > o.nC:
                row_node[q - o.nC - 1], row_dof[q - o.nC - 1] = n + 1, t
    for d in o.elemEq:
        for t, q in d.items():
            row_dof[q - o.nC - 1] = t
    rng = np.random.default_rng(seed + 1)
    return o, Kv[2], row_node, row_dof, rng.standard_normal(n2), rng.standard_normal(n2)


def write_case(path, o, vals, rhs1, rhs2, row_node, row_dof):
    si = o.blocks[2]
    n = len(si.iofeir) - 1
    with open(path, "w") as f:
        f.write(f"{n} {len(si.columns)}\n")
        f.write(" ".join(str(int(v)) for v in si.iofeir) + "\n")
        f.write(" ".join(str(int(v)) for v in si.columns) + "\n")
        f.write(" ".join(repr(float(v)) for v in vals) + "\n")
        f.write(" ".join(repr(float(v)) for v in np.concatenate([rhs1, rhs2])) + "\n")
        f.write(" ".join(f"{int(a)} {int(b)}" for a, b in zip(row_node, row_dof)) + "\n")


def test_csr_caller_reports_a_missing_device(tmp_path):
    """without a usable device the caller stops with the library's status (no crash)"""
    o, vals, rn, rd, b1, b2 = saddle_case()
    path = tmp_path / "case.txt"
    write_case(path, o, vals, b1, b2, rn, rd)
    r = subprocess.run([csr_caller(), "-1", str(path), "0", "1"], capture_output=True, text=True)
    assert r.returncode == 3 and "nla3d_gpu_eqsolver_create failed" in r.stderr


@pytest.mark.gpu
@pytest.mark.parametrize("pc,use_map", [(0, 0), (0, 1), (2, 0), (2, 1)])
def test_csr_caller_solves_match_dense_solve(tmp_path, pc, use_map):
    """factorize once + substitute two right-hand sides, then solve with every value doubled, on
    the caller's 1-based upper CSR: each solution within 1e-8 (relative, max-abs) of the
    oracle's dense FP64 solve of the same matrix"""
    print(f"[gpu] C caller csr_solve, preconditioner {pc}, DoF map {use_map}", flush=True)
    o, vals, rn, rd, b1, b2 = saddle_case()
    path = tmp_path / "case.txt"
    write_case(path, o, vals, b1, b2, rn, rd)
    r = subprocess.run([csr_caller(), "0", str(path), str(pc), str(use_map)], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stderr
    xs = {0: {}, 1: {}, 2: {}}
    for line in r.stdout.splitlines():
        f = line.split()
        if f[0].startswith("x"):
            xs[int(f[0][1:])][int(f[1])] = float(f[2])
        elif f[0] == "iterations":
            print(line, flush=True)
            assert int(f[3]) == 0  # no fallback: the multigrid (pc 2) ran every solve
    A = po.dense_sym(o.blocks[2], vals)
    for k, (Ak, b) in enumerate(((A, b1), (A, b2), (2.0 * A, b1))):
        x = np.array([xs[k][i] for i in range(len(b))])
        xr = np.linalg.solve(Ak, b)
        assert np.abs(x - xr).max() <= 1e-8 * np.abs(xr).max(), k
