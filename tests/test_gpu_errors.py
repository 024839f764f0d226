"""Error paths of the Newton loop on the device, the reference's (FESolver.cpp:474-490):
criterion 0 above 1e6 stops with "The solution is diverged!", and running out of
equilibrium iterations with "The solution is not converged ...", both as ABI error codes
(NLA3D_GPU_EDIVERGED = -6, NLA3D_GPU_ENOTCONV = -5), the oracle raising the same."""
import pytest

import nla3d_amd
from test_gpu_plane41 import MATS, models

pytestmark = pytest.mark.gpu


def test_diverged_is_reported_like_the_reference(gpu_lib):
    print("[gpu] newton divergence error path", flush=True)
    m, o, _, _ = models(MATS[0], top_uy=-0.1)
    # criterion 0 divides mean |dU| by the convergence value twice (FESolver.cpp:511-521):
    # with 1e-12 the first increment already measures above 1e6
    opts = m.newton_defaults(criterion=0, convergence=1e-12, max_iterations=5, krylov_tol=1e-12)
    with pytest.raises(nla3d_amd.Nla3dGpuError) as e:
        m.solve_load_step(1.0, opts)
    assert e.value.code == -6 and "diverged" in str(e.value)
    o.newton_init()
    with pytest.raises(RuntimeError, match="diverged"):
        o.solve_load_step(1.0, criterion=0, convergence=1e-12, max_iterations=5)


def test_not_converged_is_reported_like_the_reference(gpu_lib):
    print("[gpu] newton iteration-limit error path", flush=True)
    m, o, _, _ = models(MATS[0], top_uy=-0.1)
    opts = m.newton_defaults(criterion=1, convergence=1e-14, max_iterations=1, krylov_tol=1e-12)
    with pytest.raises(nla3d_amd.Nla3dGpuError) as e:
        m.solve_load_step(1.0, opts)
    assert e.value.code == -5 and "not converged" in str(e.value)
    o.newton_init()
    with pytest.raises(RuntimeError, match="not converged"):
        o.solve_load_step(1.0, criterion=1, convergence=1e-14, max_iterations=1)
