"""nla3d_gpu_solve_block2_noalloc, the variant of nla3d_gpu_solve_block2 that allocates nothing
(its MINRES work space is a static per-device array): through the C ABI it gives the same
iterations and the same solution (1e-12 relative, max-abs) as nla3d_gpu_solve_block2 on two
calls in a row, the second on a smaller model, with right-hand side and unknowns in device
memory the test allocates itself; its argument checks return the ABI's codes."""
import ctypes

import numpy as np
import pytest

import nla3d_amd
from nla3d_amd import meshgen

pytestmark = pytest.mark.gpu


def hip_runtime():
    """the HIP runtime the library is linked to (the test's own device buffers)"""
    nla3d_amd.lib()
    with open("/proc/self/maps") as f:
        paths = [l.split()[-1] for l in f if "libamdhip64.so" in l]
    assert paths, "libamdhip64 not loaded"
    hip = ctypes.CDLL(paths[0])
    hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    hip.hipFree.argtypes = [ctypes.c_void_p]
    hip.hipSetDevice.argtypes = [ctypes.c_int]
    return hip


class DeviceArray:
    def __init__(self, hip, n):
        self.hip, self.n, self.ptr = hip, n, ctypes.c_void_p()
        assert hip.hipMalloc(ctypes.byref(self.ptr), n * 8) == 0

    def put(self, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        assert self.hip.hipMemcpy(self.ptr, a.ctypes.data, self.n * 8, 1) == 0  # host to device

    def get(self):
        a = np.empty(self.n)
        assert self.hip.hipMemcpy(a.ctypes.data, self.ptr, self.n * 8, 2) == 0  # device to host
        return a

    def free(self):
        self.hip.hipFree(self.ptr)


def assembled_model(nx, ny, nz, seed):
    m = nla3d_amd.GpuModel(0)
    meshgen.solid81_rigid_top_case(m, nx, ny, nz, 2.0, 2.0, 1.0, 0.15, 0.349, material=(3, (0.5, 0.1, 1000.0)))
    m.set_max_in_row(0)
    m.init_solution()
    m.set_solution(0.01 * np.random.default_rng(seed).standard_normal(m.nAll))
    m.assemble()
    return m


def test_noalloc_solve_matches_solve_block2(gpu_lib):
    hip = hip_runtime()
    assert hip.hipSetDevice(0) == 0
    for nx, ny, nz in ((8, 8, 4), (3, 3, 2)):  # the second call on a smaller system
        print(f"[gpu] solve_block2_noalloc, {nx}x{ny}x{nz} rigid-top box", flush=True)
        m = assembled_model(nx, ny, nz, seed=nx)
        n2 = m.sizes["nUnknown"] + m.sizes["nMpc"]
        assert n2 <= nla3d_amd.NOALLOC_MAX_ROWS
        rhs = np.random.default_rng(nx + 1).standard_normal(n2)
        x_ref, it_ref, _ = m.solve_block2(rhs, tol=1e-12, max_iter=20000)
        r_d, x_d = DeviceArray(hip, n2), DeviceArray(hip, n2)
        r_d.put(rhs)
        x_d.put(np.full(n2, np.nan))  # the entry point must overwrite every unknown
        it, rr = m.solve_block2_noalloc(r_d.ptr.value, x_d.ptr.value, tol=1e-12, max_iter=20000)
        x = x_d.get()
        r_d.free()
        x_d.free()
        print(f"iterations {it} (solve_block2 {it_ref}), relres {rr:.2e}, "
              f"max |x - x_ref| / max |x_ref| = {np.abs(x - x_ref).max() / np.abs(x_ref).max():.2e}", flush=True)
        assert it == it_ref and rr <= 1e-12
        assert np.abs(x - x_ref).max() <= 1e-12 * np.abs(x_ref).max()
        m.close()


def test_noalloc_argument_checks(gpu_lib):
    print("[gpu] solve_block2_noalloc argument checks", flush=True)
    m = assembled_model(2, 2, 1, seed=1)
    L = nla3d_amd.lib()
    it, rr = ctypes.c_int(), ctypes.c_double()
    assert L.nla3d_gpu_solve_block2_noalloc(m._h, None, None, 1e-12, 100, ctypes.byref(it), ctypes.byref(rr)) == -1
    host_only = nla3d_amd.GpuModel(-1)
    assert L.nla3d_gpu_solve_block2_noalloc(host_only._h, ctypes.c_void_p(8), ctypes.c_void_p(8), 1e-12, 100,
                                            ctypes.byref(it), ctypes.byref(rr)) == -3
