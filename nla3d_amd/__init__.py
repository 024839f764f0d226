"""nla3d_amd: MI355X (gfx950) implementation of nla3d's Newton-Raphson equilibrium iteration.

The compute lives in the in-tree shared library ``libnla3d_gpu.so`` (HIP kernels
plus the host model store), reached through the C ABI declared in
``include/nla3d_gpu.h``.  This module is a thin ctypes binding of that ABI.
There is no CPU fallback: if the library is not built, loading fails loudly.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libnla3d_gpu.so")

# Dof::dofType (Dof.h:15-25)
UX, UY, UZ, ROTX, ROTY, ROTZ, HYDRO_PRESSURE, TEMP = range(8)
# element kinds
SOLID81, QUADTH, SURFACE_LINETH, PLANE41 = 1, 2, 3, 4
# MaterialFactory ids (MaterialFactory.cpp:9-13)
NEO_HOOKEAN, BIDERMAN, MOONEY_RIVLIN = 1, 2, 3
# element result queries (query.h): kinds, and scalarQuery / vectorQuery / tensorQuery numbers
QUERY_SCALAR, QUERY_VECTOR, QUERY_TENSOR = 0, 1, 2
SP, WU, WP, VOL = 1, 3, 4, 5
IC = 1
COUCHY, PK2, E, C = 1, 2, 3, 4
GP_MEAN = -1

_lib = None


class NewtonOpts(ctypes.Structure):
    """nla3d_gpu_newton_opts (include/nla3d_gpu.h)"""
    _fields_ = [("n_load_steps", ctypes.c_int), ("max_iterations", ctypes.c_int), ("criterion", ctypes.c_int),
                ("convergence", ctypes.c_double), ("krylov_tol", ctypes.c_double),
                ("krylov_max_iter", ctypes.c_int), ("preconditioner", ctypes.c_int),
                ("krylov_tol_loose", ctypes.c_double), ("krylov_forcing", ctypes.c_double),
                ("first_solve_predictor", ctypes.c_int)]


class TransientOpts(ctypes.Structure):
    """nla3d_gpu_transient_opts (include/nla3d_gpu.h)"""
    _fields_ = [("alpha", ctypes.c_double), ("delta", ctypes.c_double), ("time0", ctypes.c_double),
                ("time1", ctypes.c_double), ("n_timesteps", ctypes.c_int), ("init_value", ctypes.c_double),
                ("dirichlet_coupling", ctypes.c_int), ("krylov_tol", ctypes.c_double),
                ("krylov_max_iter", ctypes.c_int), ("predictor", ctypes.c_int)]


class NewtonStats(ctypes.Structure):
    """nla3d_gpu_newton_stats (include/nla3d_gpu.h)"""
    _fields_ = [("load_steps", ctypes.c_int), ("newton_iterations", ctypes.c_int),
                ("last_step_iterations", ctypes.c_int), ("linear_solves", ctypes.c_int),
                ("krylov_iterations", ctypes.c_longlong), ("last_krylov_iterations", ctypes.c_int),
                ("krylov_breakdown", ctypes.c_int),
                ("last_criterion", ctypes.c_double), ("last_time", ctypes.c_double),
                ("ms_assemble", ctypes.c_double), ("ms_solve", ctypes.c_double), ("ms_total", ctypes.c_double),
                ("precond_fallbacks", ctypes.c_int)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


class AdamOpts(ctypes.Structure):
    """nla3d_gpu_adam_opts (include/nla3d_gpu.h)"""
    _fields_ = [("lr", ctypes.c_float), ("beta1", ctypes.c_float), ("beta2", ctypes.c_float), ("eps", ctypes.c_float)]


class BatchConfig(ctypes.Structure):
    """nla3d_gpu_batch_config (include/nla3d_gpu.h)"""
    _fields_ = [("elem_grid", ctypes.c_int), ("gather_threads", ctypes.c_int), ("team_threads", ctypes.c_int),
                ("team_lanes", ctypes.c_int), ("team_grid", ctypes.c_int), ("big_rows", ctypes.c_longlong)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


# rows of the largest block(2) nla3d_gpu_solve_block2_noalloc takes (include/nla3d_gpu.h)
NOALLOC_MAX_ROWS = 1 << 20


class EqSolverOpts(ctypes.Structure):
    """nla3d_gpu_eqsolver_opts (include/nla3d_gpu.h)"""
    _fields_ = [("tol", ctypes.c_double), ("max_iter", ctypes.c_int), ("preconditioner", ctypes.c_int)]


class EqSolverStats(ctypes.Structure):
    """nla3d_gpu_eqsolver_stats (include/nla3d_gpu.h)"""
    _fields_ = [("iterations", ctypes.c_int), ("relres", ctypes.c_double), ("precond_fallbacks", ctypes.c_int),
                ("breakdown", ctypes.c_int), ("ms_factorize", ctypes.c_double), ("ms_substitute", ctypes.c_double),
                ("spmv_dia_nodes", ctypes.c_int)]


_P = ctypes.c_void_p
_U32P = ctypes.POINTER(ctypes.c_uint32)
_U64P = ctypes.POINTER(ctypes.c_uint64)
_DP = ctypes.POINTER(ctypes.c_double)
_IP = ctypes.POINTER(ctypes.c_int)

# name -> argtypes; every function returns int (0 = OK)
ABI = {
    "nla3d_gpu_create": [ctypes.POINTER(_P), ctypes.c_int],
    "nla3d_gpu_destroy": [_P],
    "nla3d_gpu_set_nodes": [_P, ctypes.c_uint32, _DP],
    "nla3d_gpu_add_elements": [_P, ctypes.c_int, ctypes.c_uint32, _U32P, _DP],
    "nla3d_gpu_set_material": [_P, ctypes.c_int, _DP, ctypes.c_int],
    "nla3d_gpu_add_fix": [_P, ctypes.c_uint32, ctypes.c_int, ctypes.c_double],
    "nla3d_gpu_add_load": [_P, ctypes.c_uint32, ctypes.c_int, ctypes.c_double],
    "nla3d_gpu_add_mpc": [_P, ctypes.c_int, _U32P, _IP, _DP, ctypes.c_double],
    "nla3d_gpu_add_rigid_body_mpc": [_P, ctypes.c_uint32, ctypes.c_uint32, _U32P, ctypes.c_int, _IP],
    "nla3d_gpu_set_max_in_row": [_P, ctypes.c_uint32],
    "nla3d_gpu_init_solution": [_P],
    "nla3d_gpu_get_sizes": [_P, _U64P],
    "nla3d_gpu_get_equation_numbers": [_P, _U32P, _U32P],
    "nla3d_gpu_get_pattern": [_P, ctypes.c_int, _U32P, _U32P],
    "nla3d_gpu_set_solution": [_P, _DP],
    "nla3d_gpu_get_solution": [_P, _DP],
    "nla3d_gpu_assemble": [_P],
    "nla3d_gpu_get_K": [_P, ctypes.c_int, _DP],
    "nla3d_gpu_get_F": [_P, _DP],
    "nla3d_gpu_solid81_element_matrices": [_P, _DP],
    "nla3d_gpu_plane41_element_matrices": [_P, _DP],
    "nla3d_gpu_element_query": [_P, ctypes.c_int, ctypes.c_int, ctypes.c_int, _DP],
    "nla3d_gpu_solid81_gp_state": [_P, _DP],
    "nla3d_gpu_time_assembly": [_P, ctypes.c_int, _DP, _DP],
    "nla3d_gpu_time_spmv": [_P, ctypes.c_int, _DP],
    "nla3d_gpu_apply_block2": [_P, _DP, _DP],
    "nla3d_gpu_spmv_info": [_P, _IP, _IP, _IP],
    "nla3d_gpu_copy_calibration": [_P, ctypes.c_uint64, _DP],
    "nla3d_gpu_solve_load_step": [_P, ctypes.c_double, ctypes.POINTER(NewtonOpts), ctypes.POINTER(NewtonStats)],
    "nla3d_gpu_solve_nonlinear": [_P, ctypes.POINTER(NewtonOpts), ctypes.POINTER(NewtonStats)],
    "nla3d_gpu_newton_iteration": [_P, ctypes.c_double, ctypes.c_int, _DP, ctypes.POINTER(NewtonOpts),
                                   ctypes.POINTER(NewtonStats), _IP],
    "nla3d_gpu_solve_block2": [_P, _DP, _DP, ctypes.c_double, ctypes.c_int, _IP, _DP],
    # device pointers (rhs, x) as plain addresses
    "nla3d_gpu_solve_block2_noalloc": [_P, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_double, ctypes.c_int, _IP,
                                       _DP],
}
ABI.update({
    "nla3d_gpu_assemble_thermal": [_P],
    "nla3d_gpu_get_C": [_P, ctypes.c_int, _DP],
    "nla3d_gpu_solve_linear": [_P, ctypes.c_double, ctypes.c_int, ctypes.POINTER(NewtonStats)],
    "nla3d_gpu_transient_begin": [_P, ctypes.POINTER(TransientOpts)],
    "nla3d_gpu_transient_step": [_P, ctypes.POINTER(NewtonStats), _IP],
    "nla3d_gpu_solve_transient": [_P, ctypes.POINTER(TransientOpts), ctypes.POINTER(NewtonStats)],
})
_FP = ctypes.POINTER(ctypes.c_float)
# learned block-Jacobi correction (first argument: network or model handle)
ABI.update({
    "nla3d_gpu_mlp_destroy": [_P],
    "nla3d_gpu_mlp_get_params": [_P, _FP, ctypes.c_int],
    "nla3d_gpu_mlp_set_params": [_P, _FP, ctypes.c_int],
    "nla3d_gpu_mlp_set_normalization": [_P, _FP, _FP],
    "nla3d_gpu_mlp_get_normalization": [_P, _FP, _FP],
    "nla3d_gpu_mlp_forward": [_P, _FP, ctypes.c_int, _FP, ctypes.c_int],
    "nla3d_gpu_mlp_backward": [_P, _FP, _FP, ctypes.c_int, _FP],
    "nla3d_gpu_attach_mlp": [_P, _P],
    "nla3d_gpu_node_features": [_P, ctypes.c_double, _FP, _IP, ctypes.c_int, _IP],
    "nla3d_gpu_learned_blocks": [_P, _FP, ctypes.c_int, _DP],
    "nla3d_gpu_mlp_loss_grad": [_P, _P, _DP, ctypes.c_int, ctypes.c_double, ctypes.c_int, _DP, _FP],
    "nla3d_gpu_mlp_train_step": [_P, _P, _DP, ctypes.c_int, ctypes.c_double, ctypes.c_int, ctypes.POINTER(AdamOpts),
                                 _DP],
    "nla3d_gpu_get_rhs": [_P, _DP],
    "nla3d_gpu_schur_info": [_P, _IP, _IP, _IP, _IP, _IP, ctypes.c_int],
    "nla3d_gpu_schur_matrix": [_P, _IP, _IP, _IP, _DP],
    "nla3d_gpu_schur_level": [_P, ctypes.c_int, _IP, _DP],
    "nla3d_gpu_schur_cycle_info": [_P, _IP, _IP],
    "nla3d_gpu_schur_set_cycle": [_P, ctypes.c_int],
    "nla3d_gpu_schur_get_cycle": [_P, _IP],
    "nla3d_gpu_schur_apply": [_P, _DP, _DP, ctypes.c_int],
    "nla3d_gpu_schur_smoother_operator": [_P, _DP, _DP],
    "nla3d_gpu_schur_finest_precision": [_P, _IP],
    "nla3d_gpu_time_smoother": [_P, ctypes.c_int, _DP, ctypes.POINTER(ctypes.c_longlong)],
    "nla3d_gpu_get_R": [_P, _DP],
    "nla3d_gpu_solve_block2_pc": [_P, _DP, _DP, ctypes.c_double, ctypes.c_int, ctypes.c_int, ctypes.c_double, _IP,
                                  _DP],
})
# batched models (first argument: batch handle)
ABI.update({
    "nla3d_gpu_batch_destroy": [_P],
    "nla3d_gpu_batch_get_config": [_P, ctypes.POINTER(BatchConfig)],
    "nla3d_gpu_batch_set_config": [_P, ctypes.POINTER(BatchConfig)],
    "nla3d_gpu_batch_autotune": [_P, ctypes.c_double, ctypes.c_int, ctypes.c_int, ctypes.POINTER(BatchConfig), _DP,
                                 ctypes.c_int],
    "nla3d_gpu_batch_assemble": [_P, ctypes.c_double],
    "nla3d_gpu_batch_newton_iteration": [_P, ctypes.c_double, ctypes.c_int, ctypes.POINTER(NewtonOpts),
                                         ctypes.POINTER(NewtonStats), _IP],
    "nla3d_gpu_batch_solve_load_step": [_P, ctypes.c_double, ctypes.POINTER(NewtonOpts), ctypes.POINTER(NewtonStats),
                                        _IP],
    "nla3d_gpu_batch_solve_nonlinear": [_P, ctypes.POINTER(NewtonOpts), ctypes.POINTER(NewtonStats), _IP],
})
# equation solver on a caller's matrix (first argument: solver handle)
ABI.update({
    "nla3d_gpu_eqsolver_destroy": [_P],
    "nla3d_gpu_eqsolver_set_options": [_P, ctypes.POINTER(EqSolverOpts)],
    "nla3d_gpu_eqsolver_set_dof_map": [_P, ctypes.c_uint32, _U32P, _IP],
    "nla3d_gpu_eqsolver_factorize": [_P, ctypes.c_uint32, _DP, _U32P, _U32P],
    "nla3d_gpu_eqsolver_substitute": [_P, _DP, _DP],
    "nla3d_gpu_eqsolver_solve": [_P, ctypes.c_uint32, _DP, _U32P, _U32P, _DP, _DP],
    "nla3d_gpu_eqsolver_get_stats": [_P, ctypes.POINTER(EqSolverStats)],
})
# functions that take no handle
ABI_FREE = {"nla3d_gpu_eqsolver_create": [ctypes.POINTER(_P), ctypes.c_int],
            "nla3d_gpu_eqsolver_last_error": [_P],
            "nla3d_gpu_newton_defaults": [ctypes.POINTER(NewtonOpts)],
            "nla3d_gpu_transient_defaults": [ctypes.POINTER(TransientOpts)],
            "nla3d_gpu_mlp_num_params": [],
            "nla3d_gpu_mlp_create": [ctypes.POINTER(_P), ctypes.c_int, ctypes.c_ulonglong],
            "nla3d_gpu_adam_defaults": [ctypes.POINTER(AdamOpts)],
            "nla3d_gpu_mlp_last_error": [_P],
            "nla3d_gpu_batch_create": [ctypes.POINTER(_P), ctypes.POINTER(_P), ctypes.c_int],
            "nla3d_gpu_batch_default_config": [ctypes.POINTER(BatchConfig)],
            "nla3d_gpu_batch_last_error": [_P]}


def lib():
    """Load the in-tree libnla3d_gpu.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built; run __graft_entry__.build() first")
        L = ctypes.CDLL(LIB_PATH)
        for name, args in ABI.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = ctypes.c_int
        for name, args in ABI_FREE.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = ctypes.c_int
        L.nla3d_gpu_last_error.argtypes = [_P]
        L.nla3d_gpu_last_error.restype = ctypes.c_char_p
        L.nla3d_gpu_mlp_last_error.restype = ctypes.c_char_p
        L.nla3d_gpu_batch_last_error.restype = ctypes.c_char_p
        L.nla3d_gpu_eqsolver_last_error.restype = ctypes.c_char_p
        L.nla3d_gpu_eqsolver_defaults.argtypes = [ctypes.POINTER(EqSolverOpts)]
        L.nla3d_gpu_eqsolver_defaults.restype = None
        _lib = L
    return _lib


class Nla3dGpuError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"nla3d_gpu error {code}: {msg}")
        self.code = code


def _ptr(a, ct):
    return a.ctypes.data_as(ctypes.POINTER(ct))


class GpuModel:
    """One nla3d_gpu handle: an FEStorage-like model plus its device state.

    device=-1 gives a host-only handle (DoF numbering and sparsity, no GPU)."""

    def __init__(self, device=0):
        self._L = lib()
        self._h = _P()
        rc = self._L.nla3d_gpu_create(ctypes.byref(self._h), device)
        if rc != 0:
            raise Nla3dGpuError(rc, f"cannot create a handle on device {device}")
        self.device = device

    def _call(self, name, *args):
        rc = getattr(self._L, name)(self._h, *args)
        if rc != 0:
            raise Nla3dGpuError(rc, self._L.nla3d_gpu_last_error(self._h).decode())

    def close(self):
        if self._h:
            self._L.nla3d_gpu_destroy(self._h)
            self._h = _P()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- model set-up (FEStorage / FESolver) ----
    def set_nodes(self, xyz):
        xyz = np.ascontiguousarray(xyz, dtype=np.float64).reshape(-1, 3)
        self._call("nla3d_gpu_set_nodes", len(xyz), _ptr(xyz, ctypes.c_double))
        self.xyz = xyz.copy()

    def add_elements(self, kind, conn, params=None):
        conn = np.ascontiguousarray(conn, dtype=np.uint32)
        n = conn.shape[0]
        pp = None
        if params is not None:
            params = np.ascontiguousarray(params, dtype=np.float64).reshape(n, 4)
            pp = _ptr(params, ctypes.c_double)
        self._call("nla3d_gpu_add_elements", kind, n, _ptr(conn, ctypes.c_uint32), pp)
        if not hasattr(self, "cells"):
            self.cells = []  # (kind, 1-based connectivity [n, nodes]) in the order added
        self.cells.append((int(kind), conn.reshape(n, -1).copy()))

    def set_material(self, mtype, consts):
        mc = np.ascontiguousarray(consts, dtype=np.float64)
        self._call("nla3d_gpu_set_material", mtype, _ptr(mc, ctypes.c_double), len(mc))

    def add_fix(self, node, dof, value=0.0):
        self._call("nla3d_gpu_add_fix", int(node), int(dof), float(value))

    def add_load(self, node, dof, value):
        self._call("nla3d_gpu_add_load", int(node), int(dof), float(value))

    def add_mpc(self, terms, b=0.0):
        """FEStorage::addMpc: sum coef * U(node, dof) = b over terms [(node, dof, coef), ...]"""
        n = np.ascontiguousarray([t[0] for t in terms], dtype=np.uint32)
        d = np.ascontiguousarray([t[1] for t in terms], dtype=np.int32)
        c = np.ascontiguousarray([t[2] for t in terms], dtype=np.float64)
        self._call("nla3d_gpu_add_mpc", len(n), _ptr(n, ctypes.c_uint32), _ptr(d, ctypes.c_int),
                   _ptr(c, ctypes.c_double), float(b))

    def add_rigid_body_mpc(self, master, slaves, dofs=(UX, UY, UZ)):
        s = np.ascontiguousarray(slaves, dtype=np.uint32)
        d = np.ascontiguousarray(dofs, dtype=np.int32)
        self._call("nla3d_gpu_add_rigid_body_mpc", int(master), len(s), _ptr(s, ctypes.c_uint32), len(d),
                   _ptr(d, ctypes.c_int))

    def set_max_in_row(self, n):
        """per-row sparsity capacity (reference default 100; 0 = unlimited)"""
        self._call("nla3d_gpu_set_max_in_row", int(n))

    def init_solution(self):
        self._call("nla3d_gpu_init_solution")
        s = np.zeros(9, dtype=np.uint64)
        self._call("nla3d_gpu_get_sizes", _ptr(s, ctypes.c_uint64))
        keys = ["nNodes", "nElements", "nDofs", "nConstrained", "nUnknown", "nMpc", "nnz1", "nnz12", "nnz2"]
        self.sizes = {k: int(v) for k, v in zip(keys, s)}
        self.nAll = self.sizes["nDofs"] + self.sizes["nMpc"]
        return self.sizes

    def equation_numbers(self):
        ne = np.zeros((self.sizes["nNodes"], 8), dtype=np.uint32)
        ee = np.zeros((self.sizes["nElements"], 8), dtype=np.uint32)
        self._call("nla3d_gpu_get_equation_numbers", _ptr(ne, ctypes.c_uint32), _ptr(ee, ctypes.c_uint32))
        return ne, ee

    def pattern(self, block):
        nC = self.sizes["nConstrained"]
        nrows = nC if block in (1, 12) else self.nAll - nC
        nnz = self.sizes[{1: "nnz1", 12: "nnz12", 2: "nnz2"}[block]]
        io = np.zeros(nrows + 1, dtype=np.uint32)
        cols = np.zeros(max(nnz, 1), dtype=np.uint32)
        self._call("nla3d_gpu_get_pattern", block, _ptr(io, ctypes.c_uint32), _ptr(cols, ctypes.c_uint32))
        return io, cols[:nnz]

    # ---- solution and assembly ----
    def set_solution(self, U):
        U = np.ascontiguousarray(U, dtype=np.float64)
        assert U.size == self.nAll
        self._call("nla3d_gpu_set_solution", _ptr(U, ctypes.c_double))

    def get_solution(self):
        U = np.zeros(self.nAll)
        self._call("nla3d_gpu_get_solution", _ptr(U, ctypes.c_double))
        return U

    def assemble(self):
        self._call("nla3d_gpu_assemble")

    def get_K(self, block):
        nnz = self.sizes[{1: "nnz1", 12: "nnz12", 2: "nnz2"}[block]]
        v = np.zeros(max(nnz, 1))
        self._call("nla3d_gpu_get_K", block, _ptr(v, ctypes.c_double))
        return v[:nnz]

    def get_F(self):
        F = np.zeros(self.nAll)
        self._call("nla3d_gpu_get_F", _ptr(F, ctypes.c_double))
        return F

    def solid81_element_matrices(self, n_solid):
        out = np.zeros((n_solid, 350))
        self._call("nla3d_gpu_solid81_element_matrices", _ptr(out, ctypes.c_double))
        return out

    def plane41_element_matrices(self, n_plane):
        """per-element PLANE41 outputs [n, 54] in assembleK order (PLANE41.h:55-89)"""
        out = np.zeros((n_plane, 54))
        self._call("nla3d_gpu_plane41_element_matrices", _ptr(out, ctypes.c_double))
        return out

    def element_query(self, n_elements, kind, query, gp=-1):
        """Element::getScalar / getVector / getTensor for every u/p element (kind 0 / 1 / 2 with
        the reference's query numbers, gp = Gauss point or -1 for GP_MEAN) -> [n, 1 | 3 | 6]"""
        width = (1, 3, 6)[kind] if kind in (0, 1, 2) else 1
        out = np.zeros((n_elements, width))
        self._call("nla3d_gpu_element_query", int(kind), int(query), int(gp), _ptr(out, ctypes.c_double))
        return out

    def solid81_gp_state(self, n_solid):
        out = np.zeros((n_solid, 8, 21))
        self._call("nla3d_gpu_solid81_gp_state", _ptr(out, ctypes.c_double))
        return out

    # ---- solution procedures ----
    @staticmethod
    def newton_defaults(**kw):
        o = NewtonOpts()
        lib().nla3d_gpu_newton_defaults(ctypes.byref(o))
        for k, v in kw.items():
            setattr(o, k, v)
        return o

    def solve_load_step(self, t, opts, stats=None):
        stats = stats or NewtonStats()
        self._call("nla3d_gpu_solve_load_step", float(t), ctypes.byref(opts), ctypes.byref(stats))
        return stats

    def newton_iteration(self, t, iteration, rhs1, opts, stats):
        """one equilibrium iteration; rhs1 is a ctypes.c_double carried across the load step"""
        conv = ctypes.c_int()
        self._call("nla3d_gpu_newton_iteration", float(t), int(iteration), ctypes.byref(rhs1), ctypes.byref(opts),
                   ctypes.byref(stats), ctypes.byref(conv))
        return bool(conv.value)

    def solve_nonlinear(self, opts, stats=None):
        stats = stats or NewtonStats()
        self._call("nla3d_gpu_solve_nonlinear", ctypes.byref(opts), ctypes.byref(stats))
        return stats

    # ---- learned block-Jacobi correction ----
    def attach_mlp(self, mlp):
        self._call("nla3d_gpu_attach_mlp", mlp._h if mlp is not None else None)
        self._mlp = mlp  # keep the network alive while attached

    def node_features(self, t=1.0):
        n = ctypes.c_int()
        self._call("nla3d_gpu_node_features", float(t), None, None, 0, ctypes.byref(n))
        X = np.zeros((max(n.value, 1), 24), dtype=np.float32)
        rows = np.zeros(max(n.value, 1), dtype=np.int32)
        self._call("nla3d_gpu_node_features", float(t), _ptr(X, ctypes.c_float), _ptr(rows, ctypes.c_int), n.value,
                   ctypes.byref(n))
        return X[:n.value], rows[:n.value]

    def learned_blocks(self, Y):
        Y = np.ascontiguousarray(Y, dtype=np.float32).reshape(-1, 6)
        B = np.zeros((len(Y), 9))
        self._call("nla3d_gpu_learned_blocks", _ptr(Y, ctypes.c_float), len(Y), _ptr(B, ctypes.c_double))
        return B.reshape(-1, 3, 3)

    # ---- Schur-row multigrid ----
    def schur_info(self):
        av, nq, nnz, lv = (ctypes.c_int() for _ in range(4))
        sizes = np.zeros(16, dtype=np.int32)
        self._call("nla3d_gpu_schur_info", ctypes.byref(av), ctypes.byref(nq), ctypes.byref(nnz), ctypes.byref(lv),
                   _ptr(sizes, ctypes.c_int), 16)
        return {"available": bool(av.value), "nq": nq.value, "nnz0": nnz.value, "levels": lv.value,
                "sizes": sizes[:lv.value].tolist()}

    def schur_matrix(self):
        """(qrow, scipy-style CSR (rowPtr, col, val)) of S on the Schur rows for the current tangent"""
        inf = self.schur_info()
        qrow = np.zeros(inf["nq"], dtype=np.int32)
        rp = np.zeros(inf["nq"] + 1, dtype=np.int32)
        col = np.zeros(inf["nnz0"], dtype=np.int32)
        val = np.zeros(inf["nnz0"])
        self._call("nla3d_gpu_schur_matrix", _ptr(qrow, ctypes.c_int), _ptr(rp, ctypes.c_int), _ptr(col, ctypes.c_int),
                   _ptr(val, ctypes.c_double))
        return qrow, rp, col, val

    def schur_cycle_info(self):
        """(finest-level Chebyshev terms, first level of the one-workgroup coarse range or -1)"""
        st, ff = ctypes.c_int(), ctypes.c_int()
        self._call("nla3d_gpu_schur_cycle_info", ctypes.byref(st), ctypes.byref(ff))
        return st.value, ff.value

    def schur_set_cycle(self, vcycle):
        """the Schur stage: False (default) one Chebyshev sweep on S, True the V-cycle"""
        self._call("nla3d_gpu_schur_set_cycle", int(bool(vcycle)))

    def schur_vcycle(self):
        """True when the Schur stage is the V-cycle"""
        v = ctypes.c_int()
        self._call("nla3d_gpu_schur_get_cycle", ctypes.byref(v))
        return bool(v.value)

    def schur_level(self, level, n):
        agg = np.zeros(max(n, 1), dtype=np.int32)
        lm = ctypes.c_double()
        self._call("nla3d_gpu_schur_level", int(level), _ptr(agg, ctypes.c_int), ctypes.byref(lm))
        return agg[:n], lm.value

    def schur_finest_precision(self):
        """16 (half-precision finest smoother operator), 32 (FP32 copy) or 64 (one dense level)"""
        b = ctypes.c_int()
        self._call("nla3d_gpu_schur_finest_precision", ctypes.byref(b))
        return b.value

    def schur_smoother_operator(self):
        """(s, CSR values in schur_matrix()'s pattern) of the finest level's stored half-precision
        smoother operator"""
        inf = self.schur_info()
        sc = np.zeros(inf["nq"])
        val = np.zeros(inf["nnz0"])
        self._call("nla3d_gpu_schur_smoother_operator", _ptr(sc, ctypes.c_double), _ptr(val, ctypes.c_double))
        return sc, val

    def schur_apply(self, r, refresh=False):
        r = np.ascontiguousarray(r, dtype=np.float64)
        z = np.zeros_like(r)
        self._call("nla3d_gpu_schur_apply", _ptr(r, ctypes.c_double), _ptr(z, ctypes.c_double), int(refresh))
        return z

    def get_R(self):
        """vecR: reactions at the constrained equations, applied loads elsewhere"""
        r = np.zeros(self.nAll)
        self._call("nla3d_gpu_get_R", _ptr(r, ctypes.c_double))
        return r

    def reaction(self, node, dof, R=None):
        """FEStorage::getReaction(node, dof) (FEStorage.cpp:105-119)"""
        ne, _ = self.equation_numbers()
        eq = int(ne[node - 1, dof])
        if eq == 0:
            return 0.0
        return float((self.get_R() if R is None else R)[eq - 1])

    def get_rhs(self):
        r = np.zeros(self.sizes["nUnknown"] + self.sizes["nMpc"])
        self._call("nla3d_gpu_get_rhs", _ptr(r, ctypes.c_double))
        return r

    def solve_block2_pc(self, rhs, tol=1e-12, max_iter=20000, preconditioner=0, t=1.0):
        rhs = np.ascontiguousarray(rhs, dtype=np.float64)
        x = np.zeros_like(rhs)
        it = ctypes.c_int()
        rr = ctypes.c_double()
        self._call("nla3d_gpu_solve_block2_pc", _ptr(rhs, ctypes.c_double), _ptr(x, ctypes.c_double), float(tol),
                   int(max_iter), int(preconditioner), float(t), ctypes.byref(it), ctypes.byref(rr))
        return x, it.value, rr.value

    def solve_block2_noalloc(self, rhs_dev, x_dev, tol=1e-12, max_iter=20000):
        """nla3d_gpu_solve_block2_noalloc on device addresses (ints) of rhs and x"""
        it, rr = ctypes.c_int(), ctypes.c_double()
        self._call("nla3d_gpu_solve_block2_noalloc", ctypes.c_void_p(rhs_dev), ctypes.c_void_p(x_dev), float(tol),
                   int(max_iter), ctypes.byref(it), ctypes.byref(rr))
        return it.value, rr.value

    def solve_block2(self, rhs, tol=1e-12, max_iter=20000):
        rhs = np.ascontiguousarray(rhs, dtype=np.float64)
        x = np.zeros_like(rhs)
        it = ctypes.c_int()
        rr = ctypes.c_double()
        self._call("nla3d_gpu_solve_block2", _ptr(rhs, ctypes.c_double), _ptr(x, ctypes.c_double), float(tol),
                   int(max_iter), ctypes.byref(it), ctypes.byref(rr))
        return x, it.value, rr.value

    def copy_calibration(self, nbytes):
        a = ctypes.c_double()
        self._call("nla3d_gpu_copy_calibration", int(nbytes), ctypes.byref(a))
        return a.value

    # ---- thermal path ----
    def assemble_thermal(self):
        self._call("nla3d_gpu_assemble_thermal")

    def get_C(self, block):
        nnz = self.sizes[{1: "nnz1", 12: "nnz12", 2: "nnz2"}[block]]
        v = np.zeros(max(nnz, 1))
        self._call("nla3d_gpu_get_C", block, _ptr(v, ctypes.c_double))
        return v[:nnz]

    def solve_linear(self, tol=1e-12, max_iter=100000):
        st = NewtonStats()
        self._call("nla3d_gpu_solve_linear", float(tol), int(max_iter), ctypes.byref(st))
        return st

    @staticmethod
    def transient_defaults(**kw):
        o = TransientOpts()
        lib().nla3d_gpu_transient_defaults(ctypes.byref(o))
        for k, v in kw.items():
            setattr(o, k, v)
        return o

    def transient_begin(self, opts):
        self._call("nla3d_gpu_transient_begin", ctypes.byref(opts))

    def transient_step(self, stats):
        done = ctypes.c_int()
        self._call("nla3d_gpu_transient_step", ctypes.byref(stats), ctypes.byref(done))
        return not done.value

    def solve_transient(self, opts, stats=None):
        stats = stats or NewtonStats()
        self._call("nla3d_gpu_solve_transient", ctypes.byref(opts), ctypes.byref(stats))
        return stats

    def time_smoother(self, iters):
        """(ms per finest-level smoother step, bytes it moves)"""
        a = ctypes.c_double()
        nb = ctypes.c_longlong()
        self._call("nla3d_gpu_time_smoother", int(iters), ctypes.byref(a), ctypes.byref(nb))
        return a.value, nb.value

    def apply_block2(self, x):
        """K.block(2) x through the Krylov solve's SpMV (host arrays)"""
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.zeros_like(x)
        self._call("nla3d_gpu_apply_block2", _ptr(x, ctypes.c_double), _ptr(y, ctypes.c_double))
        return y

    def spmv_info(self):
        """(nodes read through the node-diagonal copy, its upper node offsets, row groups read
        through the grouped CSR)"""
        a, b, c = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        self._call("nla3d_gpu_spmv_info", ctypes.byref(a), ctypes.byref(b), ctypes.byref(c))
        return a.value, b.value, c.value

    def time_spmv(self, iters):
        a = ctypes.c_double()
        self._call("nla3d_gpu_time_spmv", int(iters), ctypes.byref(a))
        return a.value

    def time_assembly(self, iters):
        a = ctypes.c_double()
        b = ctypes.c_double()
        self._call("nla3d_gpu_time_assembly", int(iters), ctypes.byref(a), ctypes.byref(b))
        return a.value, b.value


class Mlp:
    """nla3d_gpu_mlp_t: the per-node network of the learned block-Jacobi correction."""
    N_PARAMS = 53510

    def __init__(self, device=0, seed=0):
        self._L = lib()
        self._h = _P()
        rc = self._L.nla3d_gpu_mlp_create(ctypes.byref(self._h), device, int(seed))
        if rc != 0:
            raise Nla3dGpuError(rc, f"cannot create a network on device {device}")

    def _call(self, name, *args):
        rc = getattr(self._L, name)(self._h, *args)
        if rc != 0:
            raise Nla3dGpuError(rc, self._L.nla3d_gpu_mlp_last_error(self._h).decode())

    def close(self):
        if self._h:
            self._L.nla3d_gpu_mlp_destroy(self._h)
            self._h = _P()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def get_params(self):
        w = np.zeros(self.N_PARAMS, dtype=np.float32)
        self._call("nla3d_gpu_mlp_get_params", _ptr(w, ctypes.c_float), self.N_PARAMS)
        return w

    def set_params(self, w):
        w = np.ascontiguousarray(w, dtype=np.float32)
        self._call("nla3d_gpu_mlp_set_params", _ptr(w, ctypes.c_float), len(w))

    def set_normalization(self, mean, inv_std):
        m = np.ascontiguousarray(mean, dtype=np.float32)
        s = np.ascontiguousarray(inv_std, dtype=np.float32)
        self._call("nla3d_gpu_mlp_set_normalization", _ptr(m, ctypes.c_float), _ptr(s, ctypes.c_float))

    def get_normalization(self):
        m = np.zeros(24, dtype=np.float32)
        s = np.zeros(24, dtype=np.float32)
        self._call("nla3d_gpu_mlp_get_normalization", _ptr(m, ctypes.c_float), _ptr(s, ctypes.c_float))
        return m, s

    def forward(self, X, bf16=True):
        X = np.ascontiguousarray(X, dtype=np.float32).reshape(-1, 24)
        Y = np.zeros((len(X), 6), dtype=np.float32)
        self._call("nla3d_gpu_mlp_forward", _ptr(X, ctypes.c_float), len(X), _ptr(Y, ctypes.c_float), int(bf16))
        return Y

    def backward(self, X, dY):
        X = np.ascontiguousarray(X, dtype=np.float32).reshape(-1, 24)
        dY = np.ascontiguousarray(dY, dtype=np.float32).reshape(-1, 6)
        g = np.zeros(self.N_PARAMS, dtype=np.float32)
        self._call("nla3d_gpu_mlp_backward", _ptr(X, ctypes.c_float), _ptr(dY, ctypes.c_float), len(X),
                   _ptr(g, ctypes.c_float))
        return g

    def _rhs(self, rhs):
        if rhs is None:
            return None, None
        r = np.ascontiguousarray(rhs, dtype=np.float64)
        return r, _ptr(r, ctypes.c_double)

    def loss_grad(self, model, rhs=None, k=40, t=1.0, multigrid=False):
        keep, rp = self._rhs(rhs)
        loss = ctypes.c_double()
        g = np.zeros(self.N_PARAMS, dtype=np.float32)
        self._call("nla3d_gpu_mlp_loss_grad", model._h, rp, int(k), float(t), int(multigrid), ctypes.byref(loss),
                   _ptr(g, ctypes.c_float))
        return loss.value, g

    def train_step(self, model, rhs=None, k=40, t=1.0, lr=1e-3, multigrid=False):
        keep, rp = self._rhs(rhs)
        o = AdamOpts()
        self._L.nla3d_gpu_adam_defaults(ctypes.byref(o))
        o.lr = lr
        loss = ctypes.c_double()
        self._call("nla3d_gpu_mlp_train_step", model._h, rp, int(k), float(t), int(multigrid), ctypes.byref(o),
                   ctypes.byref(loss))
        return loss.value


class Batch:
    """nla3d_gpu_batch_t: independent SOLID81 models (initialised GpuModel handles on one
    device) assembled and solved together; the models stay owned by the caller."""

    def __init__(self, models):
        self._L = lib()
        self._h = _P()
        self.models = list(models)
        arr = (_P * len(self.models))(*[m._h.value for m in self.models])
        rc = self._L.nla3d_gpu_batch_create(ctypes.byref(self._h), arr, len(self.models))
        if rc != 0:
            raise Nla3dGpuError(rc, "cannot create the batch (models must be initialised SOLID81 device handles "
                                    "without MPCs, on one device)")
        self.n = len(self.models)

    def _call(self, name, *args):
        rc = getattr(self._L, name)(self._h, *args)
        if rc != 0:
            raise Nla3dGpuError(rc, self._L.nla3d_gpu_batch_last_error(self._h).decode())

    def close(self):
        if self._h:
            self._L.nla3d_gpu_batch_destroy(self._h)
            self._h = _P()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def config(self):
        c = BatchConfig()
        self._call("nla3d_gpu_batch_get_config", ctypes.byref(c))
        return c

    def set_config(self, **kw):
        c = self.config()
        for k, v in kw.items():
            setattr(c, k, v)
        self._call("nla3d_gpu_batch_set_config", ctypes.byref(c))

    def autotune(self, t=0.1, probe_iterations=24, preconditioner=0):
        c = BatchConfig()
        rep = np.zeros(64)
        self._call("nla3d_gpu_batch_autotune", float(t), int(probe_iterations), int(preconditioner), ctypes.byref(c),
                   _ptr(rep, ctypes.c_double), len(rep))
        return c, rep

    def assemble(self, t):
        self._call("nla3d_gpu_batch_assemble", float(t))

    def new_stats(self):
        return (NewtonStats * self.n)()

    def newton_iteration(self, t, iteration, opts, stats, conv):
        self._call("nla3d_gpu_batch_newton_iteration", float(t), int(iteration), ctypes.byref(opts), stats,
                   _ptr(conv, ctypes.c_int))

    def solve_load_step(self, t, opts, stats=None):
        stats = stats if stats is not None else self.new_stats()
        conv = np.zeros(self.n, dtype=np.int32)
        self._call("nla3d_gpu_batch_solve_load_step", float(t), ctypes.byref(opts), stats, _ptr(conv, ctypes.c_int))
        return stats, conv

    def solve_nonlinear(self, opts, stats=None):
        stats = stats if stats is not None else self.new_stats()
        conv = np.zeros(self.n, dtype=np.int32)
        self._call("nla3d_gpu_batch_solve_nonlinear", ctypes.byref(opts), stats, _ptr(conv, ctypes.c_int))
        return stats, conv


class EqSolver:
    """nla3d_gpu_eqsolver_t: EquationSolver (math/EquationSolver.h:19-78) on a caller's
    K.block(2) given as the reference's 1-based upper CSR (values, iofeir, columns)."""

    def __init__(self, device=0, tol=1e-12, max_iter=20000, preconditioner=0):
        self._L = lib()
        self._h = _P()
        rc = self._L.nla3d_gpu_eqsolver_create(ctypes.byref(self._h), int(device))
        if rc != 0:
            raise Nla3dGpuError(rc, f"cannot create an equation solver on device {device}")
        self.set_options(tol=tol, max_iter=max_iter, preconditioner=preconditioner)

    def _call(self, name, *args):
        rc = getattr(self._L, name)(self._h, *args)
        if rc != 0:
            raise Nla3dGpuError(rc, self._L.nla3d_gpu_eqsolver_last_error(self._h).decode())

    def close(self):
        if self._h:
            self._L.nla3d_gpu_eqsolver_destroy(self._h)
            self._h = _P()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_options(self, tol=1e-12, max_iter=20000, preconditioner=0):
        o = EqSolverOpts(float(tol), int(max_iter), int(preconditioner))
        self._call("nla3d_gpu_eqsolver_set_options", ctypes.byref(o))

    def set_dof_map(self, row_node=None, row_dof=None):
        """row_node (1-based node, 0 for element DoFs / multipliers) and row_dof (Dof type, 6 for an
        element pressure, -1 for a multiplier) per block(2) row; None clears the map"""
        if row_node is None:
            self._call("nla3d_gpu_eqsolver_set_dof_map", 0, None, None)
            return
        rn = np.ascontiguousarray(row_node, dtype=np.uint32)
        rd = np.ascontiguousarray(row_dof, dtype=np.int32)
        self._call("nla3d_gpu_eqsolver_set_dof_map", len(rn), _ptr(rn, ctypes.c_uint32), _ptr(rd, ctypes.c_int))

    @staticmethod
    def _csr(values, iofeir, columns):
        v = np.ascontiguousarray(values, dtype=np.float64)
        io = np.ascontiguousarray(iofeir, dtype=np.uint32)
        c = np.ascontiguousarray(columns, dtype=np.uint32)
        return v, io, c

    def factorize(self, values, iofeir, columns):
        v, io, c = self._csr(values, iofeir, columns)
        self._keep = (v, io, c)
        self._n = len(io) - 1
        self._call("nla3d_gpu_eqsolver_factorize", self._n, _ptr(v, ctypes.c_double), _ptr(io, ctypes.c_uint32),
                   _ptr(c, ctypes.c_uint32))

    def substitute(self, rhs):
        b = np.ascontiguousarray(rhs, dtype=np.float64)
        x = np.zeros_like(b)
        self._call("nla3d_gpu_eqsolver_substitute", _ptr(b, ctypes.c_double), _ptr(x, ctypes.c_double))
        return x

    def solve(self, values, iofeir, columns, rhs):
        v, io, c = self._csr(values, iofeir, columns)
        b = np.ascontiguousarray(rhs, dtype=np.float64)
        x = np.zeros_like(b)
        self._call("nla3d_gpu_eqsolver_solve", len(io) - 1, _ptr(v, ctypes.c_double), _ptr(io, ctypes.c_uint32),
                   _ptr(c, ctypes.c_uint32), _ptr(b, ctypes.c_double), _ptr(x, ctypes.c_double))
        return x

    def stats(self):
        st = EqSolverStats()
        self._call("nla3d_gpu_eqsolver_get_stats", ctypes.byref(st))
        return st
