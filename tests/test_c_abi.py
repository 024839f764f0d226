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
