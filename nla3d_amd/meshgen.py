"""Synthetic structured meshes and load cases for the SOLID81 / QUADTH workloads.

Node and element numbering follow the trilinear hexahedron convention of nla3d
(node order of ElementIsoParamHEXAHEDRON::formFunc, isoparametric.cpp:197-208):
x fastest, then y, then z; all numbers 1-based.

Jitter is counter-based splitmix64: the value for node n (0-based) and
coordinate c is mix64(seed + GOLDEN * (3n + c + 1)), mapped to u in [0, 1)
from its top 53 bits and scaled to (2u - 1) * amp * h_c.  Only interior nodes
(on no boundary face) move, so boundary faces stay planar.
"""
import numpy as np

GOLDEN = np.uint64(0x9E3779B97F4A7C15)
M1 = np.uint64(0xBF58476D1CE4E5B9)
M2 = np.uint64(0x94D049BB133111EB)
DEFAULT_SEED = 20261203


def splitmix64(x):
    """splitmix64 finaliser on a uint64 array (wrapping arithmetic)."""
    x = np.asarray(x, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = x.copy()
        z = (z ^ (z >> np.uint64(30))) * M1
        z = (z ^ (z >> np.uint64(27))) * M2
        z = z ^ (z >> np.uint64(31))
    return z


def uniform01(seed, idx):
    """u in [0,1) for counter idx (uint64 array) under seed."""
    with np.errstate(over="ignore"):
        z = splitmix64(np.uint64(seed) + GOLDEN * (np.asarray(idx, dtype=np.uint64) + np.uint64(1)))
    return (z >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def hex_box(nx, ny, nz, lx=1.0, ly=1.0, lz=1.0, jitter=0.1, seed=DEFAULT_SEED):
    """Structured hexahedral mesh of [0,lx]x[0,ly]x[0,lz].

    Returns (xyz[nN,3], conn[nE,8] 1-based, grid index arrays (i,j,k) per node)."""
    hx, hy, hz = lx / nx, ly / ny, lz / nz
    i = np.arange(nx + 1)
    j = np.arange(ny + 1)
    k = np.arange(nz + 1)
    K, J, I = np.meshgrid(k, j, i, indexing="ij")
    I = I.ravel(); J = J.ravel(); K = K.ravel()
    xyz = np.stack([I * hx, J * hy, K * hz], axis=1).astype(np.float64)
    if jitter:
        interior = (I > 0) & (I < nx) & (J > 0) & (J < ny) & (K > 0) & (K < nz)
        n = np.nonzero(interior)[0].astype(np.uint64)
        h = np.array([hx, hy, hz])
        for c in range(3):
            u = uniform01(seed, np.uint64(3) * n + np.uint64(c))
            xyz[n.astype(np.int64), c] += (2.0 * u - 1.0) * jitter * h[c]
    ei = np.arange(nx)
    ej = np.arange(ny)
    ek = np.arange(nz)
    EK, EJ, EI = np.meshgrid(ek, ej, ei, indexing="ij")
    EI = EI.ravel(); EJ = EJ.ravel(); EK = EK.ravel()

    def nid(a, b, c):
        return 1 + a + (nx + 1) * (b + (ny + 1) * c)

    conn = np.stack([
        nid(EI, EJ, EK), nid(EI + 1, EJ, EK), nid(EI + 1, EJ + 1, EK), nid(EI, EJ + 1, EK),
        nid(EI, EJ, EK + 1), nid(EI + 1, EJ, EK + 1), nid(EI + 1, EJ + 1, EK + 1), nid(EI, EJ + 1, EK + 1),
    ], axis=1).astype(np.uint32)
    return xyz, conn, (I, J, K)


def solid81_compression_case(model, n, jitter=0.1, seed=DEFAULT_SEED, material=(3, (0.5, 0.1, 100.0)),
                             top_uz=0.2):
    """Unit cube n^3 SOLID81, bottom fixed, top face UZ prescribed (its full value; the
    solver ramps it with the normalised time, FESolver.cpp:229-247)."""
    xyz, conn, (I, J, K) = hex_box(n, n, n, jitter=jitter, seed=seed)
    model.set_nodes(xyz)
    model.add_elements(1, conn)
    model.set_material(material[0], material[1])
    bottom = np.nonzero(K == 0)[0] + 1
    top = np.nonzero(K == n)[0] + 1
    for node in bottom:
        for d in range(3):
            model.add_fix(node, d, 0.0)
    for node in top:
        model.add_fix(node, 2, top_uz)
    return xyz, conn


def solid81_rigid_top_case(model, nx, ny, nz, lx, ly, lz, uz, rotz, jitter=0.1, seed=DEFAULT_SEED,
                           material=(3, (0.5, 0.1, 1000.0))):
    """Box with the bottom fixed and the top face tied by a RigidBodyMpc to a master
    node at the top-face centre, which carries the prescribed UZ and ROTZ."""
    xyz, conn, (I, J, K) = hex_box(nx, ny, nz, lx, ly, lz, jitter=jitter, seed=seed)
    master_pos = np.array([[lx / 2, ly / 2, lz]])
    xyz_all = np.vstack([xyz, master_pos])
    master = len(xyz_all)
    model.set_nodes(xyz_all)
    model.add_elements(1, conn)
    model.set_material(material[0], material[1])
    bottom = np.nonzero(K == 0)[0] + 1
    top = np.nonzero(K == nz)[0] + 1
    for node in bottom:
        for d in range(3):
            model.add_fix(node, d, 0.0)
    model.add_rigid_body_mpc(master, top, (0, 1, 2))
    vals = {0: 0.0, 1: 0.0, 2: uz, 3: 0.0, 4: 0.0, 5: rotz}
    for d in range(6):
        model.add_fix(master, d, vals[d])
    return xyz_all, conn


def batch_models(count, seed=DEFAULT_SEED, n_range=(4, 32), device=0, make=None):
    """Seeded independent compression cubes for the batched path: n^3 hexahedra with n
    log-uniform in n_range, Mooney-Rivlin C10 ~ U[0.2, 1], C01 ~ U[0, 0.3],
    K/C10 ~ U[50, 2000], top UZ strain ~ U[5%, 30%]; interior jitter 0.1 h.
    Returns a list of (model, params) with the models initialised."""
    import nla3d_amd
    rng = np.random.default_rng(seed)
    out = []
    lo, hi = np.log(n_range[0]), np.log(n_range[1])
    for i in range(count):
        n = int(round(np.exp(rng.uniform(lo, hi))))
        c10 = rng.uniform(0.2, 1.0)
        c01 = rng.uniform(0.0, 0.3)
        bulk = c10 * rng.uniform(50.0, 2000.0)
        strain = rng.uniform(0.05, 0.30)
        m = (make or nla3d_amd.GpuModel)(device)
        solid81_compression_case(m, n, jitter=0.1, seed=seed + 7919 * (i + 1), material=(3, (c10, c01, bulk)),
                                 top_uz=strain)
        m.init_solution()
        out.append((m, {"n": n, "c10": c10, "c01": c01, "bulk": bulk, "strain": strain}))
    return out


def quad_grid(nx, ny, lx=1.0, ly=1.0):
    """Structured 4-node quadrilaterals of [0,lx]x[0,ly] (counter-clockwise).
    Returns (xyz[nN,3] with z = 0, conn[nE,4] 1-based, node grid indices (i, j))."""
    i = np.arange(nx + 1)
    j = np.arange(ny + 1)
    J, I = np.meshgrid(j, i, indexing="ij")
    I = I.ravel()
    J = J.ravel()
    xyz = np.stack([I * (lx / nx), J * (ly / ny), np.zeros(I.size)], axis=1)
    EJ, EI = np.meshgrid(np.arange(ny), np.arange(nx), indexing="ij")
    EI = EI.ravel()
    EJ = EJ.ravel()

    def nid(a, b):
        return 1 + a + (nx + 1) * b

    conn = np.stack([nid(EI, EJ), nid(EI + 1, EJ), nid(EI + 1, EJ + 1), nid(EI, EJ + 1)], axis=1).astype(np.uint32)
    return xyz, conn, (I, J)


def quadth_transient_case(model, n, seed=DEFAULT_SEED, h=10.0, t_left=1.0):
    """Unit square of n x n QUADTH elements, conductivity exp(N(0, 0.5^2)) per element
    (Box-Muller on splitmix64 uniforms), rho c = 1, T = t_left fixed on the left edge,
    SurfaceLINETH convection (h, T_inf = 0) on the right edge."""
    xyz, conn, (I, J) = quad_grid(n, n)
    idx = np.arange(len(conn), dtype=np.uint64)
    u1 = uniform01(seed, np.uint64(2) * idx)
    u2 = uniform01(seed, np.uint64(2) * idx + np.uint64(1))
    z = np.sqrt(-2.0 * np.log(np.maximum(u1, 1e-300))) * np.cos(2.0 * np.pi * u2)
    k = np.exp(0.5 * z)
    qp = np.stack([k, np.ones_like(k), np.ones_like(k), np.zeros_like(k)], axis=1)
    right = np.nonzero(I == n)[0] + 1
    right = right[np.argsort(J[right - 1])]
    lines = np.stack([right[:-1], right[1:]], axis=1).astype(np.uint32)
    lp = np.tile([0.0, h, 0.0, 0.0], (len(lines), 1))
    model.set_nodes(xyz)
    model.add_elements(2, conn, qp)
    model.add_elements(3, lines, lp)
    for node in np.nonzero(I == 0)[0] + 1:
        model.add_fix(int(node), 7, t_left)
    return xyz, conn, k
