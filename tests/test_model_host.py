"""Host-side model work of the library (host-only handle, no GPU): reference DoF
numbering (FEStorage.cpp:382-432) and the 1-based upper CSR of the three K blocks
(FEStorage.cpp:434-550) must be bit-identical to the oracle restatement."""
import numpy as np
import pytest

import nla3d_amd
from nla3d_amd import meshgen
from oracle.pyoracle import OracleModel


def oracle_for(xyz, conn, fixes, rbs=(), material=(3, [0.5, 0.1, 100.0])):
    return OracleModel(xyz, [(1, c) for c in conn], material=material, fixes=fixes, rigid_bodies=rbs).setup()


def compare(m, o):
    s = m.sizes
    assert s["nDofs"] == o.nDofs and s["nConstrained"] == o.nC and s["nMpc"] == o.nMpc
    ne, ee = m.equation_numbers()
    for n in range(len(o.nodeEq)):
        for t in range(8):
            assert ne[n, t] == o.nodeEq[n].get(t, 0)
    for e in range(len(o.elemEq)):
        for t in range(8):
            assert ee[e, t] == o.elemEq[e].get(t, 0)
    for b in (1, 12, 2):
        io, cols = m.pattern(b)
        assert np.array_equal(io, o.blocks[b].iofeir), b
        assert np.array_equal(cols, o.blocks[b].columns), b


def collect_fixes(xyz, conn, top_uz=0.2, n=None):
    fixes = []
    z = xyz[:, 2]
    for node in np.nonzero(z == 0.0)[0] + 1:
        for d in range(3):
            fixes.append((int(node), d, 0.0))
    for node in np.nonzero(z == z.max())[0] + 1:
        fixes.append((int(node), 2, top_uz))
    return fixes


@pytest.mark.parametrize("n", [1, 2, 3])
def test_compression_case_numbering_and_pattern(gpu_lib, n):
    m = nla3d_amd.GpuModel(device=-1)
    xyz, conn = meshgen.solid81_compression_case(m, n)
    m.init_solution()
    compare(m, oracle_for(xyz, conn, collect_fixes(xyz, conn)))


def test_rigid_body_case_numbering_and_pattern(gpu_lib):
    m = nla3d_amd.GpuModel(device=-1)
    xyz, conn = meshgen.solid81_rigid_top_case(m, 3, 2, 2, 2.0, 2.0, 1.0, 0.15, 0.349)
    m.init_solution()
    master = len(xyz)
    z = xyz[:-1, 2]
    fixes = [(int(nd), d, 0.0) for nd in np.nonzero(z == 0.0)[0] + 1 for d in range(3)]
    fixes += [(master, d, 0.0) for d in range(6)]
    top = list(np.nonzero(z == 1.0)[0] + 1)
    o = oracle_for(xyz, conn, fixes, rbs=[(master, top, [0, 1, 2])], material=(3, [0.5, 0.1, 1000.0]))
    compare(m, o)
    assert m.sizes["nMpc"] == 3 * len(top)


def test_unconstrained_no_fix_reports_sizes(gpu_lib):
    m = nla3d_amd.GpuModel(device=-1)
    xyz, conn, _ = meshgen.hex_box(2, 1, 1)
    m.set_nodes(xyz)
    m.add_elements(1, conn)
    s = m.init_solution()
    assert s["nConstrained"] == 0 and s["nnz1"] == 0 and s["nDofs"] == 12 * 3 + 2


def test_empty_model_is_an_error_like_the_reference(gpu_lib):
    """FEStorage::initDofs stops on a model without nodes and elements (FEStorage.cpp:340-341)"""
    m = nla3d_amd.GpuModel(device=-1)
    m.set_nodes(np.zeros((0, 3)))
    with pytest.raises(nla3d_amd.Nla3dGpuError, match="No any nodes or elements"):
        m.init_solution()


def test_all_node_dofs_fixed_leaves_pressure_rows(gpu_lib):
    """every displacement constrained: block(2) holds only the elements' HYDRO_PRESSURE rows,
    numbered and patterned as the oracle does"""
    m = nla3d_amd.GpuModel(device=-1)
    xyz, conn, _ = meshgen.hex_box(2, 2, 1, jitter=0.1, seed=3)
    m.set_nodes(xyz)
    m.add_elements(nla3d_amd.SOLID81, conn)
    m.set_material(1, [1.0, 100.0])
    fixes = [(n, d, 0.01 * d) for n in range(1, len(xyz) + 1) for d in range(3)]
    for n, d, v in fixes:
        m.add_fix(n, d, v)
    s = m.init_solution()
    o = oracle_for(xyz, conn, fixes, material=(1, [1.0, 100.0]))
    compare(m, o)
    assert s["nConstrained"] == 3 * len(xyz) and s["nDofs"] == 3 * len(xyz) + len(conn)


def test_bad_node_number_is_an_error(gpu_lib):
    m = nla3d_amd.GpuModel(device=-1)
    xyz, conn, _ = meshgen.hex_box(1, 1, 1)
    m.set_nodes(xyz)
    with pytest.raises(nla3d_amd.Nla3dGpuError):
        m.add_elements(1, conn + 100)


def test_device_call_on_host_handle_is_an_error(gpu_lib):
    m = nla3d_amd.GpuModel(device=-1)
    meshgen.solid81_compression_case(m, 1)
    m.init_solution()
    with pytest.raises(nla3d_amd.Nla3dGpuError):
        m.assemble()


def test_reference_assembly_loop_in_c_matches_python_restatement(gpu_lib, orc):
    """oracle C assembly loop (cpu baseline) == python restatement, bitwise (same summation order)"""
    m = nla3d_amd.GpuModel(device=-1)
    xyz, conn = meshgen.solid81_compression_case(m, 2)
    m.init_solution()
    o = oracle_for(xyz, conn, collect_fixes(xyz, conn))
    rng = np.random.default_rng(1)
    U = 0.03 * rng.standard_normal(o.nAll)
    Kp, Fp = o.assemble(U)
    ne, ee = m.equation_numbers()
    eq = np.zeros((len(conn), 25), dtype=np.uint32)
    eq[:, :24] = ne[conn - 1][:, :, :3].reshape(len(conn), 24)
    eq[:, 24] = ee[:, 6]
    Xe = xyz[conn - 1].reshape(len(conn), 24)
    Kc, Fc = orc.solid81_assemble_ref(Xe, U[eq[:, :24].astype(int) - 1], U[eq[:, 24].astype(int) - 1], 3,
                                      [0.5, 0.1, 100.0], eq, m.sizes["nConstrained"], m.pattern(1), m.pattern(12),
                                      m.pattern(2), m.nAll)
    for b in (1, 12, 2):
        assert np.array_equal(Kc[b], Kp[b])
    assert np.array_equal(Fc, Fp)


def test_max_in_row_overflow_is_reported_like_the_reference(gpu_lib):
    """a rigid body with more than 100 slave rows per master DoF overflows maxInRow=100
    (SparseMatrix.cpp:84); set_max_in_row(0) lifts the limit without changing anything else"""
    m = nla3d_amd.GpuModel(device=-1)
    meshgen.solid81_rigid_top_case(m, 6, 6, 2, 2.0, 2.0, 1.0, 0.15, 0.3)  # 49 top nodes -> 147 MPC rows
    with pytest.raises(nla3d_amd.Nla3dGpuError, match="maxInRow"):
        m.init_solution()
    m2 = nla3d_amd.GpuModel(device=-1)
    meshgen.solid81_rigid_top_case(m2, 6, 6, 2, 2.0, 2.0, 1.0, 0.15, 0.3)
    m2.set_max_in_row(0)
    s = m2.init_solution()
    io, cols = m2.pattern(12)
    assert np.diff(io.astype(np.int64)).max() >= 147


def test_plane41_numbering_and_pattern(gpu_lib):
    """PLANE41 registers UX UY per node and HYDRO_PRESSURE per element (PLANE41.cpp:24-28)"""
    xyz, conn, (I, J) = meshgen.quad_grid(5, 3, 1.0, 0.6)
    fixes = [(int(n), d, 0.0) for n in np.nonzero(J == 0)[0] + 1 for d in (0, 1)]
    fixes += [(int(n), 1, -0.1) for n in np.nonzero(J == 3)[0] + 1]
    m = nla3d_amd.GpuModel(device=-1)
    m.set_nodes(xyz)
    m.add_elements(nla3d_amd.PLANE41, conn)
    for n, d, v in fixes:
        m.add_fix(n, d, v)
    m.init_solution()
    compare(m, OracleModel(xyz, [(4, list(map(int, c))) for c in conn], fixes=fixes).setup())


def test_solid81_and_plane41_cannot_be_mixed(gpu_lib):
    xyz, conn, _ = meshgen.hex_box(1, 1, 1)
    m = nla3d_amd.GpuModel(device=-1)
    m.set_nodes(xyz)
    m.add_elements(nla3d_amd.SOLID81, conn)
    with pytest.raises(nla3d_amd.Nla3dGpuError):
        m.add_elements(nla3d_amd.PLANE41, [[1, 2, 4, 3]])
