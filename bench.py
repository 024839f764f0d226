#!/usr/bin/env python3
"""SOLID81 Newton-iteration hot path on one MI355X: elements/s of the fused
update + buildK + deterministic scatter at the 3.5M-DoF shape.

Workload (synthetic, seeded): [0,2]x[0,2]x[0,1] box, 120x120x60 hexahedra
(864,000 SOLID81 u/p elements, 893,101 nodes + 1 rigid-body master), interior
nodes jittered by +-0.1 h (splitmix64, seed 20261203), Mooney-Rivlin
C10=0.5 C01=0.1 K=1000, bottom face fixed, top face tied by a RigidBodyMpc to
a master carrying +15% UZ and a 20 degree ROTZ.  The solution vector is a
smooth state at half of that load (twist growing with z), so every Gauss
point is in the nonlinear range.

A step is one FEStorage::assembleGlobalEqMatrices at the current solution
(FEStorage.cpp:26-78): update() + buildK() of every element, scatter into the
three CSR blocks and F, and the rigid-body MPC rows -- all on the device with
the solution already resident in HBM.  value = elements / s over the timed
steps.  The roofline entry is for the element kernel, from HIP events on the
library's stream.  cpu_baseline times the reference assembly loop
(oracle/nla3d_oracle.c: buildK + binary-search addValue, one core) on a
bounded sample of the same model's elements.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

# Algorithmic FP64 work of one SOLID81 element per Newton iteration (update +
# buildK, 8 Gauss points), counted on the reference formulation: per Gauss
# point 8,420 flop (FMA = 2) -- see the breakdown in DESIGN.md.
FLOP_PER_ELEMENT = 8 * 8420
PEAK_FP64_TFLOPS = 78.6
PEAK_HBM_GBS = 8000.0


def synthetic_state(m, xyz, frac=0.5, uz=0.15, rotz=np.deg2rad(20.0)):
    ne, _ = m.equation_numbers()
    U = np.zeros(m.nAll)
    X = xyz
    z = X[:, 2]
    th = frac * rotz * z
    c, s = np.cos(th), np.sin(th)
    dx, dy = X[:, 0] - 1.0, X[:, 1] - 1.0
    disp = np.stack([c * dx - s * dy - dx, s * dx + c * dy - dy, frac * uz * z], axis=1)
    for d in range(3):
        eq = ne[:, d].astype(np.int64)
        ok = eq > 0
        U[eq[ok] - 1] = disp[ok, d]
    # master node: UZ and ROTZ at the same load fraction
    master = len(xyz) - 1
    U[int(ne[master, 2]) - 1] = frac * uz
    U[int(ne[master, 5]) - 1] = frac * rotz
    return U


def cpu_baseline(m, xyz, conn, U, n_sample):
    from oracle import pyoracle
    ne, ee = m.equation_numbers()
    sel = np.arange(min(n_sample, len(conn)))
    c = conn[sel]
    eq = np.zeros((len(sel), 25), dtype=np.uint32)
    eq[:, :24] = ne[c - 1][:, :, :3].reshape(len(sel), 24)
    eq[:, 24] = ee[sel, 6]
    Xe = xyz[c - 1].reshape(len(sel), 24)
    Ue = U[eq[:, :24].astype(np.int64) - 1]
    pe = U[eq[:, 24].astype(np.int64) - 1]
    pats = [m.pattern(b) for b in (1, 12, 2)]
    t0 = time.perf_counter()
    pyoracle.solid81_assemble_ref(Xe, Ue, pe, 3, [0.5, 0.1, 1000.0], eq, m.sizes["nConstrained"], *pats, m.nAll)
    dt = time.perf_counter() - t0
    return {"value": len(sel) / dt, "unit": "elements/s", "cores": 1, "kind": "port",
            "sample": f"reference assembly loop (update+buildK+binary-search addValue) over the first {len(sel)} "
                      f"elements of the same model, {dt:.1f} s, 1 thread"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--shape", type=str, default="120,120,60")
    ap.add_argument("--cpu-sample", type=int, default=250000)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; multi-GPU sharding is a design "
                                   "sketch (DESIGN.md), no figure at several GPUs is produced"}))
        sys.exit(2)

    import nla3d_amd
    from nla3d_amd import meshgen

    nx, ny, nz = (int(v) for v in args.shape.split(","))
    t_setup = time.perf_counter()
    m = nla3d_amd.GpuModel(0)
    xyz, conn = meshgen.solid81_rigid_top_case(m, nx, ny, nz, 2.0, 2.0, 1.0, 0.15, np.deg2rad(20.0))
    # the rigid-body master rows couple to every top-face MPC row: lift the reference's
    # 100-entry row capacity, which would abort this model (SparseMatrix.cpp:84)
    m.set_max_in_row(0)
    sizes = m.init_solution()
    U = synthetic_state(m, xyz)
    m.set_solution(U)
    t_setup = time.perf_counter() - t_setup
    print(f"[bench] setup {t_setup:.1f}s sizes {sizes}", file=sys.stderr, flush=True)
    nE = sizes["nElements"]

    for _ in range(args.warmup):
        m.assemble()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        m.assemble()
    t1 = time.perf_counter()
    ms_step = (t1 - t0) * 1e3 / args.steps
    value = nE * args.steps / (t1 - t0)
    print(f"[bench] {ms_step:.3f} ms/step", file=sys.stderr, flush=True)

    ms_elem, ms_scatter = m.time_assembly(max(5, args.steps))
    achieved = FLOP_PER_ELEMENT * nE / (ms_elem * 1e-3) / 1e12
    roofline = {"bound": "mfma", "achieved": achieved, "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
                "frac": achieved / PEAK_FP64_TFLOPS, "traffic": None,
                "kernel": "solid81_kernel", "kernel_ms": ms_elem, "scatter_ms": ms_scatter}

    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(m, xyz, conn, U, args.cpu_sample)

    n_dof = sizes["nUnknown"] + sizes["nMpc"]
    out = {
        "metric": "SOLID81 Newton-iter wall @3.5M DoF, learned vs block-Jacobi; elements/s FP64",
        "value": value, "unit": "elements/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": ms_step, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"SOLID81 u/p assembly (update+buildK+scatter+MPC), {nx}x{ny}x{nz} hex box, "
                               f"{nE} elements, {n_dof} equations in block(2), Mooney-Rivlin K/C10=2000",
                   "elements": nE, "equations": n_dof, "nnz_block2": sizes["nnz2"]},
        "roofline": roofline, "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
