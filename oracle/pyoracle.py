"""Python half of the CPU oracle (TEST INFRASTRUCTURE ONLY: the checker).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this module.  The product never does.

* ctypes bindings of liboracle.so (nla3d_oracle.c: element, material and dense
  kernels restated in the reference's arithmetic order);
* a restatement of the reference's host-side integer work for small models:
  DoF registration and numbering (FEStorage.cpp:338-432), the SparsityInfo
  build with maxInRow slots and compress (SparseMatrix.cpp:64-112;
  FEStorage.cpp:434-550), BlockSparseSymMatrix<2> addValue with the binary
  search getIndex (BlockSparseMatrix.h:115-168; SparseMatrix.cpp:117-154,
  467-479), the element-order assembly loop (FEStorage.cpp:26-78;
  SOLID81.h:54-94) and RigidBodyMpc::update (Mpc.cpp:70-164);
* TimeControl (FESolver.cpp:34-56) and the Newton loop of
  NonlinearFESolver::solve (FESolver.cpp:407-508) with a dense FP64 solve in
  place of PARDISO (PARDISO itself is unavailable: parity unpinned there).

Pure-Python loops: intended for models of at most a few thousand DoFs.
"""
import ctypes
import math
import os

import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
_lib = None

UX, UY, UZ, ROTX, ROTY, ROTZ, HYDRO_PRESSURE, TEMP = range(8)
NUM_DOF_TYPES = 8
SOLID81, QUADTH, SURFACE_LINETH, PLANE41 = 1, 2, 3, 4
# block(2) size above which the Newton step uses a sparse direct solve (SuperLU) instead of a
# dense one; both are FP64 direct solves standing in for PARDISO
SPARSE_ABOVE = 4000

_D = ctypes.POINTER(ctypes.c_double)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} missing: run `make -C oracle`")
        L = ctypes.CDLL(LIB_PATH)
        for name, args in {
            "orc_matBTDBprod": [ctypes.c_int, ctypes.c_int, _D, _D, ctypes.c_double, _D],
            "orc_matBTVprod": [ctypes.c_int, ctypes.c_int, _D, _D, ctypes.c_double, _D],
            "orc_matBVprod": [ctypes.c_int, ctypes.c_int, _D, _D, ctypes.c_double, _D],
            "orc_matABprod": [ctypes.c_int, ctypes.c_int, ctypes.c_int, _D, _D, ctypes.c_double, _D],
            "orc_matATBprod": [ctypes.c_int, ctypes.c_int, ctypes.c_int, _D, _D, ctypes.c_double, _D],
            "orc_hex8_jacob": [_D, _D, _D],
            "orc_quad4_jacob": [_D, _D, _D],
            "orc_getS_UP": [ctypes.c_int, _D, _D, ctypes.c_double, _D],
            "orc_getDdDp_UP": [ctypes.c_int, _D, _D, ctypes.c_double, _D, _D],
            "orc_solid81_update": [_D, _D, ctypes.c_double, ctypes.c_int, _D, _D, _D, _D],
            "orc_solid81_buildK": [_D, _D, _D, _D, _D, ctypes.c_double, ctypes.c_int, _D, _D],
            "orc_solid81_elements": [ctypes.c_int, _D, _D, _D, ctypes.c_int, _D, _D],
            "orc_quadth_element": [_D, ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_double, _D, _D, _D],
            "orc_lineth_element": [_D, ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_double, _D, _D],
            "orc_timecontrol": [ctypes.c_double, ctypes.c_double, ctypes.c_double, _D, ctypes.c_int],
            "orc_plane41_element": [_D, _D, ctypes.c_double, ctypes.c_int, _D, _D],
        }.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = ctypes.c_int if name in ("orc_hex8_jacob", "orc_quad4_jacob", "orc_timecontrol") else None
        _U = ctypes.POINTER(ctypes.c_uint32)
        L.orc_solid81_assemble_ref.argtypes = [ctypes.c_int, _D, _D, _D, ctypes.c_int, _D, _U, ctypes.c_uint32,
                                               _U, _U, _D, _U, _U, _D, _U, _U, _D, _D]
        L.orc_solid81_assemble_ref.restype = ctypes.c_int
        L.orc_mat_K.argtypes = [ctypes.c_int, _D]
        L.orc_mat_K.restype = ctypes.c_double
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(_D)


def _f64(a, shape=None):
    a = np.ascontiguousarray(a, dtype=np.float64)
    return a if shape is None else a.reshape(shape)


# ---------------- element-level wrappers ----------------
def matBTDBprod(B, Dsym, coef, Rsym):
    B = _f64(B); Dsym = _f64(Dsym); R = _f64(Rsym).copy()
    lib().orc_matBTDBprod(B.shape[0], B.shape[1], _p(B), _p(Dsym), coef, _p(R))
    return R


def matBTVprod(B, V, coef, R):
    B = _f64(B); V = _f64(V); R = _f64(R).copy()
    lib().orc_matBTVprod(B.shape[0], B.shape[1], _p(B), _p(V), coef, _p(R))
    return R


def matBVprod(B, V, coef, R):
    B = _f64(B); V = _f64(V); R = _f64(R).copy()
    lib().orc_matBVprod(B.shape[0], B.shape[1], _p(B), _p(V), coef, _p(R))
    return R


def matABprod(A, B, coef, R):
    A = _f64(A); B = _f64(B); R = _f64(R).copy()
    lib().orc_matABprod(A.shape[0], A.shape[1], B.shape[1], _p(A), _p(B), coef, _p(R))
    return R


def matATBprod(A, B, coef, R):
    A = _f64(A); B = _f64(B); R = _f64(R).copy()
    lib().orc_matATBprod(A.shape[0], A.shape[1], B.shape[1], _p(A), _p(B), coef, _p(R))
    return R


def hex8_jacob(X):
    X = _f64(X, (8, 3)); N = np.zeros((8, 8, 3)); det = np.zeros(8)
    lib().orc_hex8_jacob(_p(X), _p(N), _p(det))
    return N, det


def quad4_jacob(X):
    X = _f64(X, (4, 3)); N = np.zeros((4, 4, 2)); det = np.zeros(4)
    lib().orc_quad4_jacob(_p(X), _p(N), _p(det))
    return N, det


def mat_K(mtype, MC):
    MC = _f64(MC)
    return lib().orc_mat_K(mtype, _p(MC))


def getS_UP(mtype, MC, C, p):
    MC = _f64(MC); C = _f64(C); S = np.zeros(6)
    lib().orc_getS_UP(mtype, _p(MC), _p(C), float(p), _p(S))
    return S


def getDdDp_UP(mtype, MC, C, p):
    MC = _f64(MC); C = _f64(C); Dd = np.zeros(21); Dp = np.zeros(6)
    lib().orc_getDdDp_UP(mtype, _p(MC), _p(C), float(p), _p(Dd), _p(Dp))
    return Dd, Dp


def solid81_update(NiXj, U, p, mtype, MC):
    NiXj = _f64(NiXj); U = _f64(U); MC = _f64(MC)
    O = np.zeros((8, 9)); C = np.zeros((8, 6)); S = np.zeros((8, 6))
    lib().orc_solid81_update(_p(NiXj), _p(U), float(p), mtype, _p(MC), _p(O), _p(C), _p(S))
    return O, C, S


def solid81_elements(Xe, Ue, pe, mtype, MC):
    """Per-element makeJacob + update + buildK -> (ne, 350) in assemble3 order."""
    Xe = _f64(Xe); Ue = _f64(Ue); pe = _f64(pe); MC = _f64(MC)
    ne = pe.size
    out = np.zeros((ne, 350))
    lib().orc_solid81_elements(ne, _p(Xe), _p(Ue), _p(pe), mtype, _p(MC), _p(out))
    return out


def solid81_assemble_ref(Xe, Ue, pe, mtype, MC, eq, nC, pat1, pat12, pat2, nAll):
    """reference assembly loop in C; pat* = (iofeir, columns) uint32 1-based.
    Returns ({1: v1, 12: v12, 2: v2}, F)."""
    U32 = ctypes.POINTER(ctypes.c_uint32)
    Xe = _f64(Xe); Ue = _f64(Ue); pe = _f64(pe); MC = _f64(MC)
    eq = np.ascontiguousarray(eq, dtype=np.uint32)
    pats = [tuple(np.ascontiguousarray(a, dtype=np.uint32) for a in p) for p in (pat1, pat12, pat2)]
    vals = [np.zeros(max(len(p[1]), 1)) for p in pats]
    F = np.zeros(nAll)
    args = []
    for (io, c), v in zip(pats, vals):
        args += [io.ctypes.data_as(U32), c.ctypes.data_as(U32), _p(v)]
    rc = lib().orc_solid81_assemble_ref(pe.size, _p(Xe), _p(Ue), _p(pe), mtype, _p(MC), eq.ctypes.data_as(U32),
                                        nC, *args, _p(F))
    if rc != 0:
        raise RuntimeError("position absent from the pattern")
    return {1: vals[0][:len(pats[0][1])], 12: vals[1][:len(pats[1][1])], 2: vals[2][:len(pats[2][1])]}, F


def quadth_element(X, k, rho, c, volFlux=0.0):
    X = _f64(X, (4, 3)); Ke = np.zeros(10); Fe = np.zeros(4); Ce = np.zeros(10)
    lib().orc_quadth_element(_p(X), k, rho, c, volFlux, _p(Ke), _p(Fe), _p(Ce))
    return Ke, Fe, Ce


def plane41_element(X, U, p, mtype, MC):
    """update() at (U, p) then buildK() of one PLANE41 element: 54 values (assembleK order:
    36 node-node, 8 node-pressure, 1 pressure-pressure, then Fe: 8 node, 1 pressure)"""
    X = _f64(X, (4, 3)); U = _f64(U, (8,)); out = np.zeros(54)
    lib().orc_plane41_element(_p(X), _p(U), float(p), int(mtype), _p(_f64(MC)), _p(out))
    return out


# local DoF pairs of the 45 matrix slots of plane41_element (0..7 node DoFs 2i+d, 8 pressure)
def plane41_slot_pairs():
    pairs = []
    for i in range(4):
        for j in range(i, 4):
            for di in range(2):
                for dj in range(2):
                    if i == j and dj > di:
                        continue
                    pairs.append((2 * i + di, 2 * j + dj))
    pairs += [(a, 8) for a in range(8)] + [(8, 8)]
    return pairs


# ---------------- element result queries (Element::getScalar/getVector/getTensor) ----------------
# query numbers as query.h: scalar SP 1, WU 3, WP 4, VOL 5; vector IC 1; tensor COUCHY 1, PK2 2, E 3, C 4
def _ic_c(C):
    """solidmech::IC_C (solidmech.cpp:33-38): I1, I2 and I3 = J_C = sqrt(det C)"""
    I1 = C[0] + C[3] + C[5]
    I2 = C[0] * C[3] + C[3] * C[5] + C[0] * C[5] - C[1] * C[1] - C[4] * C[4] - C[2] * C[2]
    J = math.sqrt(C[0] * (C[3] * C[5] - C[4] * C[4]) - C[1] * (C[1] * C[5] - C[4] * C[2])
                  + C[2] * (C[1] * C[4] - C[3] * C[2]))
    return I1, I2, J


def _inv_c(C, J):
    """solidmech::invC_C (solidmech.cpp:21-30)"""
    oo = 1.0 / (J * J)
    return [oo * (C[3] * C[5] - C[4] * C[4]), oo * (C[2] * C[4] - C[1] * C[5]), oo * (C[1] * C[4] - C[2] * C[3]),
            oo * (C[0] * C[5] - C[2] * C[2]), oo * (C[1] * C[2] - C[0] * C[4]), oo * (C[0] * C[3] - C[1] * C[1])]


def _W(mtype, MC, I1, I2):
    """Mat_Comp_*::W (materials_hyperelastic.cpp:150-200)"""
    if mtype == 1:
        return 0.5 * MC[0] * (I1 - 3.0)
    if mtype == 2:
        return (MC[0] * (I1 - 3.0) + MC[1] * (I1 - 3.0) * (I1 - 3.0) + MC[2] * (I1 - 3.0) * (I1 - 3.0) * (I1 - 3.0)
                + MC[3] * (I2 - 3.0))
    return MC[0] * (I1 - 3.0) + MC[1] * (I2 - 3.0)


def _btdb(B, D, coef, R):
    """matBTDBprod (math/Mat.h:687-741): R (packed upper 3x3) += coef B^T D B, D packed upper"""
    A = [[0.0] * 3 for _ in range(3)]
    Dp = 0
    for k in range(3):  # B^T times the upper part of D, diagonal included
        for j in range(k, 3):
            for i in range(3):
                A[i][j] += B[k][i] * D[Dp]
            Dp += 1
    Dp = 0
    for j in range(3):  # B^T times the strict upper part of D transposed
        Dp += 1
        for k in range(j + 1, 3):
            for i in range(3):
                A[i][j] += B[k][i] * D[Dp]
            Dp += 1
    r = 0
    for i in range(3):
        for j in range(i, 3):
            for k in range(3):
                R[r] += A[i][k] * B[k][j] * coef
            r += 1


def element_query(etype, X, U, p, mtype, MC, kind, query, gp=-1):
    """getScalar / getVector / getTensor of one SOLID81 (SOLID81.cpp:191-345) or PLANE41
    (PLANE41.cpp:153-300) element at update() state (U node-major, p the element pressure);
    gp = Gauss point or -1 (GP_MEAN: each point weighted by intWeight / volume).  None where the
    reference element returns false (and for PLANE41 IC, which indexes past its argument)."""
    hexa = etype == SOLID81
    if hexa:
        NiXj, det = hex8_jacob(X)
        O, Cs, Ss = solid81_update(NiXj, U, p, mtype, MC)
    else:
        NiXj, det = quad4_jacob(X)
        U = np.asarray(U, dtype=np.float64)
        O = np.zeros((4, 4))
        Cs = np.zeros((4, 6))
        for np_ in range(4):  # update (PLANE41.cpp:136-149): O = Bomega U, C from O
            N = NiXj[np_]
            o = [sum(N[a][0] * U[2 * a] for a in range(4)), sum(N[a][1] * U[2 * a] for a in range(4)),
                 sum(N[a][0] * U[2 * a + 1] for a in range(4)), sum(N[a][1] * U[2 * a + 1] for a in range(4))]
            O[np_] = o
            cxx = 1.0 + 2 * o[0] + 1.0 * (o[0] * o[0] + o[2] * o[2])
            cyy = 1.0 + 2 * o[3] + 1.0 * (o[3] * o[3] + o[1] * o[1])
            cxy = o[1] + o[2] + o[0] * o[1] + o[2] * o[3]
            Cs[np_] = [cxx, cxy, 0.0, cyy, 0.0, 1.0]  # CVec
    ngp = 8 if hexa else 4
    vol = float(sum(det))  # volume(): Gauss weights 1
    width = (1, 3, 6)[kind]
    if hexa:
        ok = {0: (1, 3, 4, 5), 1: (1,), 2: (1, 2, 3, 4)}[kind]
    else:
        ok = {0: (1,), 1: (), 2: (1, 2, 3, 4)}[kind]
    if query not in ok:
        return None
    R = [0.0] * width
    pts = range(ngp) if gp < 0 else [gp]
    for g in pts:
        scale = det[g] / vol if gp < 0 else 1.0
        C = list(Cs[g])
        if kind == 0:
            if query == 1:
                R[0] += p * scale
            elif query == 3:
                I1, I2, J = _ic_c(C)
                R[0] += _W(mtype, MC, I1, I2) * scale
            elif query == 4:
                J = _ic_c(C)[2]
                R[0] += 0.5 * mat_K(mtype, MC) * (J - 1.0) * (J - 1.0) * scale
            else:
                R[0] += vol * scale
        elif kind == 1:
            I1, I2, J = _ic_c(C)
            R[0] += I1 * scale
            R[1] += I2 * scale
            R[2] += J * scale
        elif query == 4:
            for v in range(6):
                R[v] += C[v] * scale
        elif query == 3:
            for v in range(6):
                R[v] += (C[v] - (1.0 if v in (0, 3, 5) else 0.0)) * 0.5 * scale
        elif hexa:
            J = _ic_c(C)[2]
            ci = _inv_c(C, J)
            S = [Ss[g][v] + p * J * ci[v] for v in range(6)]
            if query == 2:
                for v in range(6):
                    R[v] += S[v] * scale
            else:  # matF holds F^T (SOLID81.cpp:284-293): R += scale/J F S F^T
                o = O[g]
                Ft = [[1 + o[0], o[3], o[6]], [o[1], 1 + o[4], o[7]], [o[2], o[5], 1 + o[8]]]
                _btdb(Ft, S, 1.0 / J * scale, R)
        else:
            S = list(getS_UP(mtype, MC, C, p))
            if query == 2:
                for v in range(6):
                    R[v] += S[v] * scale
            else:  # matF = F (PLANE41.cpp:247-252), no weight
                o = O[g]
                F = [[1 + o[0], o[1], 0.0], [o[2], 1 + o[3], 0.0], [0.0, 0.0, 1.0]]
                J = (F[0][0] * (F[1][1] * F[2][2] - F[1][2] * F[2][1]) - F[0][1] * (F[1][0] * F[2][2] - F[1][2] * F[2][0])
                     + F[0][2] * (F[1][0] * F[2][1] - F[1][1] * F[2][0]))
                _btdb(F, S, 1.0 / J, R)
    return np.array(R)


def lineth_element(X, flux, htc, e0, e1):
    X = _f64(X, (2, 3)); Ke = np.zeros(3); Fe = np.zeros(2)
    lib().orc_lineth_element(_p(X), flux, htc, e0, e1, _p(Ke), _p(Fe))
    return Ke, Fe


def timecontrol(start, end, delta):
    """step end times of TimeControl::nextStep(delta) from currentTime = 0"""
    t = np.zeros(int((end - start) / delta) * 2 + 8)
    n = lib().orc_timecontrol(start, end, delta, _p(t), len(t))
    return t[:n]


def load_steps(nsteps, start=0.0, end=1.0):
    """NonlinearFESolver schedule: delta = (end - start) / N (FESolver.cpp:439)"""
    return timecontrol(start, end, (end - start) / nsteps)


# ---------------- host-side model restatement ----------------
class SparsityInfo:
    """SparsityInfo (SparseMatrix.cpp:44-154): per-row slots of maxInRow, compress sorts."""

    def __init__(self, nrows, ncols, max_in_row=100):
        self.nRows, self.nColumns = nrows, ncols
        self.maxInRow = min(max_in_row, ncols)
        self.rows = [[] for _ in range(nrows)]
        self.compressed = False

    def addEntry(self, i, j):  # 1-based
        assert 0 < i <= self.nRows and 0 < j <= self.nColumns and not self.compressed
        r = self.rows[i - 1]
        if j in r:
            return
        if len(r) >= self.maxInRow:
            raise RuntimeError("maxInRow overflow!")
        r.append(j)

    def addEntrySym(self, i, j):
        """SparseSymMatrix::addEntry (SparseMatrix.h:259-265): the upper triangle only"""
        self.addEntry(min(i, j), max(i, j))

    def compress(self):
        self.iofeir = np.zeros(self.nRows + 1, dtype=np.uint32)
        cols = []
        nxt = 0
        for i in range(self.nRows):
            self.iofeir[i] = nxt + 1
            rr = sorted(self.rows[i])
            cols.extend(rr)
            nxt += len(rr)
        self.iofeir[self.nRows] = nxt + 1
        self.columns = np.array(cols, dtype=np.uint32)
        self.values = np.zeros(len(cols))
        self.compressed = True

    def getIndex(self, i, j):
        """binary search exactly as SparsityInfo::getIndex (SparseMatrix.cpp:117-154); -1 = invalid"""
        st = int(self.iofeir[i - 1]) - 1
        en = int(self.iofeir[i]) - 1
        if st == en:
            return -1
        en -= 1
        cols = self.columns
        while True:
            if en - st == 1:
                if cols[st] == j:
                    return st
                if cols[en] == j:
                    return en
                return -1
            ind = int((en + st) * 0.5)
            if cols[ind] == j:
                return ind
            if en == st:
                return -1
            if cols[ind] > j:
                en = ind
            else:
                st = ind


def dense_rect(si, vals, ncols):
    """SparseMatrix (SparseMatrix.h:131-159) as a dense array: values[k] at (i, columns[k]) for
    k in [iofeir[i] - 1, iofeir[i + 1] - 1), 1-based arrays; A @ v is matBVprod and A.T @ v
    matBTVprod (SparseMatrix.cpp:522-561)"""
    n = si.nRows
    A = np.zeros((n, ncols))
    for i in range(n):
        for k in range(int(si.iofeir[i]) - 1, int(si.iofeir[i + 1]) - 1):
            A[i, int(si.columns[k]) - 1] = vals[k]
    return A


def dense_sym(si, vals):
    """SparseSymMatrix (SparseMatrix.h:162-186) as a dense symmetric array: the upper entries
    mirrored (value(i, j) swaps to i <= j, SparseMatrix.cpp:467-479); A @ v is its matBVprod
    (SparseMatrix.cpp:497-519) without the |v| > 1e-20 skip"""
    A = dense_rect(si, vals, si.nRows)
    return A + np.triu(A, 1).T


def sparse_rect(si, vals, ncols):
    """dense_rect as a scipy CSR (models too large for dense arrays)"""
    n = si.nRows
    rows = np.repeat(np.arange(n), np.diff(si.iofeir.astype(np.int64)))
    return sp.csr_matrix((vals, (rows, si.columns.astype(np.int64) - 1)), shape=(n, ncols))


def sparse_sym(si, vals):
    """dense_sym as a scipy CSR"""
    up = sparse_rect(si, vals, si.nRows)
    return (up + sp.triu(up, 1).T).tocsr()


def sparse_sym_pattern(n, max_in_row=100):
    """the SparsityInfo of SparseSymMatrix(n) (SparseMatrix.cpp:423-431): every diagonal entry
    added by the constructor (BlockSparseSymMatrix's reinit does not, SparseMatrix.cpp:442-448)"""
    si = SparsityInfo(n, n, max_in_row)
    for i in range(1, n + 1):
        si.addEntry(i, i)
    return si


class OracleModel:
    """FEStorage + NonlinearFESolver restated for small models."""

    def __init__(self, xyz, elements, material=None, fixes=(), loads=(), rigid_bodies=(), elem_params=None, mpcs=(),
                 max_in_row=100):
        """elements: list of (kind, [1-based nodes]); fixes/loads: (node, dof, value);
        rigid_bodies: (master, [slaves], [dofs]); material: (type, MC); mpcs: constant
        linear MPCs ([(node, dof, coef), ...], b) as FEStorage::addMpc (FEStorage.cpp:175-179);
        max_in_row: the row capacity of BlockSparseSymMatrix<2> (BlockSparseMatrix.h:42, the
        reference's 100), 0 for none (a rigid-body master with more than 100 slave rows)."""
        self.max_in_row = max_in_row
        # SOLID81-only models assemble through the C restatement of the element loop
        # (identical additions in identical order; tests compare the two bit for bit)
        self.fast_solid81 = True
        self.xyz = np.asarray(xyz, dtype=np.float64).reshape(-1, 3)
        self.elements = [(k, list(map(int, n))) for k, n in elements]
        self.material = material
        self.fixes = list(fixes)
        self.loads = list(loads)
        self.rigid_bodies = [(int(m), list(map(int, s)), list(d)) for m, s, d in rigid_bodies]
        self.elem_params = elem_params
        self.ce_mpcs = [([(int(n), int(d), float(c)) for n, d, c in terms], float(b)) for terms, b in mpcs]

    # FEStorage::initDofs (FEStorage.cpp:338-379)
    def initDofs(self):
        nN, nE = len(self.xyz), len(self.elements)
        self.nodeMask = [set() for _ in range(nN)]
        self.elemMask = [set() for _ in range(nE)]
        for e, (kind, nodes) in enumerate(self.elements):
            if kind == SOLID81:
                for n in nodes:
                    self.nodeMask[n - 1] |= {UX, UY, UZ}
                self.elemMask[e] |= {HYDRO_PRESSURE}
            elif kind == PLANE41:  # PLANE41.cpp:15-29
                for n in nodes:
                    self.nodeMask[n - 1] |= {UX, UY}
                self.elemMask[e] |= {HYDRO_PRESSURE}
            else:
                for n in nodes:
                    self.nodeMask[n - 1] |= {TEMP}
        # [terms list of (node, dof, coef), b, eqNum]; the directly added MPCs come first
        # (main.cpp:207-212), then the rigid bodies' rows (MpcCollection::registerMpcsInStorage)
        self.mpcs = [[list(terms), b, 0] for terms, b in self.ce_mpcs]
        self.rb_first = []
        for master, slaves, dofs in self.rigid_bodies:
            self.rb_first.append(len(self.mpcs))
            for s in slaves:
                for d in dofs:
                    self.mpcs.append([[(s, d, 1.0), (master, d, -1.0), (master, ROTX, 0.0), (master, ROTY, 0.0),
                                       (master, ROTZ, 0.0)], 0.0, 0])
            self.nodeMask[master - 1] |= {UX, UY, UZ, ROTX, ROTY, ROTZ}
        self.nDofs = sum(len(m) for m in self.nodeMask) + sum(len(m) for m in self.elemMask)

    # FESolver::setConstrainedDofs + FEStorage::assignEquationNumbers (FEStorage.cpp:382-432)
    def assignEquationNumbers(self):
        constrained = set()
        for node, dof, _ in self.fixes:
            constrained.add((node, dof))
        self.nC = len(constrained)
        self.nU = self.nDofs - self.nC
        nxt_s, nxt_c = self.nC + 1, 1
        self.elemEq = [dict() for _ in self.elements]
        self.nodeEq = [dict() for _ in range(len(self.xyz))]
        for e in range(len(self.elements)):
            for t in range(NUM_DOF_TYPES):
                if t in self.elemMask[e]:
                    self.elemEq[e][t] = nxt_s
                    nxt_s += 1
        for n in range(len(self.xyz)):
            for t in range(NUM_DOF_TYPES):
                if t in self.nodeMask[n]:
                    if (n + 1, t) in constrained:
                        self.nodeEq[n][t] = nxt_c
                        nxt_c += 1
                    else:
                        self.nodeEq[n][t] = nxt_s
                        nxt_s += 1
        for m in self.mpcs:
            m[2] = nxt_s
            nxt_s += 1
        self.nMpc = len(self.mpcs)
        self.nAll = self.nDofs + self.nMpc

    def _block(self, i):
        return (1, i) if i <= self.nC else (2, i - self.nC)

    def _addEntry(self, i, j):  # BlockSparseSymMatrix::addEntry (BlockSparseMatrix.h:139-152)
        if i > j:
            i, j = j, i
        bi, pi = self._block(i)
        bj, pj = self._block(j)
        if bi == bj:
            self.blocks[bi].addEntry(pi, pj)
        else:
            self.blocks[12].addEntry(pi, pj)

    # FEStorage::initSolutionData (FEStorage.cpp:434-550)
    def initSolutionData(self):
        nC, nS = self.nC, self.nU + self.nMpc
        cap = self.max_in_row if self.max_in_row > 0 else max(nC, nS, 1)
        self.blocks = {1: SparsityInfo(nC, nC, cap), 12: SparsityInfo(nC, nS, cap), 2: SparsityInfo(nS, nS, cap)}
        topology = [set() for _ in range(len(self.xyz))]
        for e, (_, nodes) in enumerate(self.elements):
            for n in nodes:
                topology[n - 1].add(e + 1)
        for nn in range(1, len(self.xyz) + 1):
            nn_dofs = [self.nodeEq[nn - 1][t] for t in sorted(self.nodeMask[nn - 1])]
            for en in sorted(topology[nn - 1]):
                en_dofs = [self.elemEq[en - 1][t] for t in sorted(self.elemMask[en - 1])]
                for d1 in en_dofs:
                    for d2 in nn_dofs:
                        self._addEntry(d1, d2)
                for nn2 in self.elements[en - 1][1]:
                    if nn2 < nn:
                        continue
                    nn2_dofs = [self.nodeEq[nn2 - 1][t] for t in sorted(self.nodeMask[nn2 - 1])]
                    for d1 in nn2_dofs:
                        for d2 in nn_dofs:
                            self._addEntry(d1, d2)
        for en in range(1, len(self.elements) + 1):
            en_dofs = [self.elemEq[en - 1][t] for t in sorted(self.elemMask[en - 1])]
            for d1 in en_dofs:
                for d2 in en_dofs:
                    self._addEntry(d1, d2)
        for terms, _, eq in self.mpcs:
            for node, dof, _ in terms:
                self._addEntry(eq, self.nodeEq[node - 1][dof])
        for b in self.blocks.values():
            b.compress()

    def setup(self):
        self.initDofs()
        self.assignEquationNumbers()
        self.initSolutionData()
        self._p41 = None  # positions of the batched PLANE41 assembly, rebuilt on first use
        return self

    # BlockSparseSymMatrix::addValue (BlockSparseMatrix.h:155-168)
    def addValueK(self, i, j, v):
        if i > j:
            i, j = j, i
        bi, pi = self._block(i)
        bj, pj = self._block(j)
        blk = self.blocks[bi] if bi == bj else self.blocks[12]
        idx = blk.getIndex(pi, pj)
        if idx < 0:
            raise RuntimeError(f"The position({i}, {j}) is absent in the matrix")
        blk.values[idx] += v

    def solid81_elements_at(self, U):
        """per-element 350-vectors (assemble3 order) at solution U, SOLID81 elements in order"""
        Xe, Ue, pe = [], [], []
        for e, (kind, nodes) in enumerate(self.elements):
            if kind != SOLID81:
                continue
            Xe.append(self.xyz[np.array(nodes) - 1])
            Ue.append([U[self.nodeEq[n - 1][d] - 1] for n in nodes for d in range(3)])
            pe.append(U[self.elemEq[e][HYDRO_PRESSURE] - 1])
        if not Xe:
            return np.zeros((0, 350))
        return solid81_elements(np.array(Xe), np.array(Ue), np.array(pe), *self.material)

    # FEStorage::assembleGlobalEqMatrices (FEStorage.cpp:26-78) for SOLID81 + RigidBodyMpc
    def assemble(self, U):
        for b in self.blocks.values():
            b.values[:] = 0.0
        F = np.zeros(self.nAll)
        plane_only = all(kind != SOLID81 for kind, _ in self.elements)
        solid_only = all(kind == SOLID81 for kind, _ in self.elements)
        if solid_only and self.fast_solid81:
            self._assemble_solid81_c(U, F)
        else:
            ke = self.solid81_elements_at(U)
        s = 0
        if plane_only:
            self._assemble_plane41_batched(U, F)
        for e, (kind, nodes) in enumerate(self.elements):
            if solid_only and self.fast_solid81:
                break
            if kind == PLANE41 and plane_only:
                continue
            if kind == PLANE41:  # PLANE41::assembleK (PLANE41.h:55-89)
                eqn = [self.nodeEq[n - 1][d] for n in nodes for d in (UX, UY)] + [self.elemEq[e][HYDRO_PRESSURE]]
                out = plane41_element(self.xyz[np.array(nodes) - 1], [U[q - 1] for q in eqn[:8]], U[eqn[8] - 1],
                                      self.material[0], self.material[1])
                for k, (a, b) in enumerate(plane41_slot_pairs()):
                    self.addValueK(eqn[a], eqn[b], out[k])
                for a in range(9):
                    F[eqn[a] - 1] += out[45 + a]
                continue
            if kind != SOLID81:
                continue
            out = ke[s]
            s += 1
            eqn = [self.nodeEq[n - 1][d] for n in nodes for d in range(3)]
            ep = self.elemEq[e][HYDRO_PRESSURE]
            k = 0
            for i in range(8):  # SOLID81.h:62-74
                for di in range(3):
                    for j in range(i, 8):
                        for dj in range(3):
                            if i == j and dj < di:
                                continue
                            self.addValueK(eqn[3 * i + di], eqn[3 * j + dj], out[k])
                            k += 1
            for a in range(24):
                self.addValueK(eqn[a], ep, out[300 + a])
            self.addValueK(ep, ep, out[324])
            for a in range(24):
                F[eqn[a] - 1] += out[325 + a]
            F[ep - 1] += out[349]
        self.rigid_body_update(U)
        for terms, b, eq in self.mpcs:  # FEStorage.cpp:51-61
            F[eq - 1] = b
            for node, dof, coef in terms:
                self.addValueK(eq, self.nodeEq[node - 1][dof], coef)
        return {b: blk.values.copy() for b, blk in self.blocks.items()}, F

    def _assemble_solid81_c(self, U, F):
        """the same element loop as assemble()'s (SOLID81.h:54-94 through the binary-search
        addValue, element order) run by the C restatement solid81_assemble_ref, for models of
        SOLID81 elements only"""
        nE = len(self.elements)
        eq = np.zeros((nE, 25), dtype=np.uint32)
        for e, (_, nodes) in enumerate(self.elements):
            eq[e, :24] = [self.nodeEq[n - 1][d] for n in nodes for d in range(3)]
            eq[e, 24] = self.elemEq[e][HYDRO_PRESSURE]
        Xe = self.xyz[np.array([nodes for _, nodes in self.elements]) - 1].reshape(nE, 24)
        Ue = U[eq[:, :24].astype(np.int64) - 1]
        pe = U[eq[:, 24].astype(np.int64) - 1]
        pats = [(self.blocks[b].iofeir, self.blocks[b].columns) for b in (1, 12, 2)]
        vals, Fe = solid81_assemble_ref(Xe, Ue, pe, self.material[0], self.material[1], eq, self.nC, *pats, self.nAll)
        for b in (1, 12, 2):
            self.blocks[b].values[:] = vals[b]
        F += Fe

    def _assemble_plane41_batched(self, U, F):
        """PLANE41::assembleK for every PLANE41 element of a model without SOLID81: the same
        addValue calls as the per-element loop, with the positions looked up once
        (SparsityInfo::getIndex) and the values added by np.add.at, which adds repeated
        positions one after another in (element, slot) order, the loop's order."""
        if getattr(self, "_p41", None) is None:
            pairs = plane41_slot_pairs()
            els, eqs, blk, idx = [], [], {1: [], 12: [], 2: []}, {1: [], 12: [], 2: []}
            for e, (kind, nodes) in enumerate(self.elements):
                if kind != PLANE41:
                    continue
                eqn = [self.nodeEq[n - 1][d] for n in nodes for d in (UX, UY)] + [self.elemEq[e][HYDRO_PRESSURE]]
                els.append(e)
                eqs.append(eqn)
                for k, (a, b) in enumerate(pairs):
                    i, j = min(eqn[a], eqn[b]), max(eqn[a], eqn[b])
                    bi, pi = self._block(i)
                    bj, pj = self._block(j)
                    bb = bi if bi == bj else 12
                    q = self.blocks[bb].getIndex(pi, pj)
                    if q < 0:
                        raise RuntimeError(f"The position({i}, {j}) is absent in the matrix")
                    blk[bb].append(len(els) * 45 - 45 + k)
                    idx[bb].append(q)
            self._p41 = (np.array(els, dtype=np.int64), np.array(eqs, dtype=np.int64),
                         {b: (np.array(blk[b], dtype=np.int64), np.array(idx[b], dtype=np.int64)) for b in blk})
        els, eqs, where = self._p41
        out = np.empty((len(els), 54))
        for r, e in enumerate(els):
            nodes = self.elements[e][1]
            out[r] = plane41_element(self.xyz[np.array(nodes) - 1], U[eqs[r, :8] - 1], U[eqs[r, 8] - 1],
                                     self.material[0], self.material[1])
        kv = out[:, :45].ravel()
        for b, (src, q) in where.items():
            np.add.at(self.blocks[b].values, q, kv[src])
        np.add.at(F, (eqs - 1).ravel(), out[:, 45:].ravel())

    # RigidBodyMpc::update (Mpc.cpp:70-164)
    def rigid_body_update(self, U):
        LC = [[[0, 0, 0], [0, 0, 1], [0, -1, 0]], [[0, 0, -1], [0, 0, 0], [1, 0, 0]],
              [[0, 1, 0], [-1, 0, 0], [0, 0, 0]]]
        Id = [[1, 0, 0], [0, 1, 0], [0, 0, 1]]
        for r, (master, slaves, dofs) in enumerate(self.rigid_bodies):
            th = [U[self.nodeEq[master - 1][t] - 1] for t in (ROTX, ROTY, ROTZ)]
            w0 = [U[self.nodeEq[master - 1][t] - 1] for t in (UX, UY, UZ)]
            tn = math.sqrt(th[0] * th[0] + th[1] * th[1] + th[2] * th[2])
            if tn < 1.0e-3:
                c1, c2, c3, c4 = 1.0, 0.5, -1.0 / 3.0, -1.0 / 12.0
            else:
                c1 = math.sin(tn) / tn
                c2 = (1.0 - math.cos(tn)) / (tn * tn)
                c3 = (tn * math.cos(tn) - math.sin(tn)) / (tn * tn * tn)
                c4 = (tn * math.sin(tn) - 2.0 + 2.0 * math.cos(tn)) / (tn * tn * tn * tn)
            C = [[math.cos(tn) * Id[i][j] + c1 * (LC[i][0][j] * th[0] + LC[i][1][j] * th[1] + LC[i][2][j] * th[2])
                  + c2 * th[i] * th[j] for j in range(3)] for i in range(3)]
            mp = self.xyz[master - 1]
            for n, s in enumerate(slaves):
                pq = self.xyz[s - 1] - mp
                for dd, d in enumerate(dofs):
                    i = d
                    cD = [[Id[i][j] * (-c1) * th[l] + c3 * th[l] * (LC[i][0][j] * th[0] + LC[i][1][j] * th[1]
                                                                 + LC[i][2][j] * th[2]) + c1 * LC[i][l][j]
                           + c4 * th[l] * th[i] * th[j] + c2 * (Id[i][l] * th[j] + th[i] * Id[j][l])
                           for l in range(3)] for j in range(3)]
                    m = self.mpcs[self.rb_first[r] + n * len(dofs) + dd]
                    m[1] = (w0[i] + (C[i][0] - Id[i][0]) * pq[0] + (C[i][1] - Id[i][1]) * pq[1]
                            + (C[i][2] - Id[i][2]) * pq[2] - U[self.nodeEq[s - 1][d] - 1])
                    terms = m[0]
                    for l, t in enumerate((ROTX, ROTY, ROTZ)):
                        coef = -(cD[0][l] * pq[0] + cD[1][l] * pq[1] + cD[2][l] * pq[2])
                        terms[2 + l] = (terms[2 + l][0], t, coef)

    # ---------------- Newton-Raphson (NonlinearFESolver::solve, FESolver.cpp:407-521) ----------------
    def dense_blocks(self, K):
        """K values per block -> dense K12 (nC x n2) and symmetric K2 (n2 x n2)"""
        n2 = self.nU + self.nMpc
        return dense_rect(self.blocks[12], K[12], n2), dense_sym(self.blocks[2], K[2])

    @staticmethod
    def sparse_sym(b, vals, n):
        """symmetric scipy CSR from an upper 1-based CSR block"""
        return sparse_sym(b, vals)

    def sparse_blocks(self, K):
        """K values per block -> sparse K12 (nC x n2) and symmetric K2 (n2 x n2); large models"""
        n2 = self.nU + self.nMpc
        return sparse_rect(self.blocks[12], K[12], n2), sparse_sym(self.blocks[2], K[2])

    def dense_k11(self, K):
        """symmetric K11 (nC x nC) from the upper CSR of block 1"""
        return dense_sym(self.blocks[1], K[1])

    def reaction(self, node, dof):
        """FEStorage::getReaction(node, dof) (FEStorage.cpp:105-119): vecR at the DoF's equation"""
        if dof not in self.nodeMask[node - 1]:
            return 0.0
        return self.R[self.nodeEq[node - 1][dof] - 1]

    def newton_init(self):
        self.U = np.zeros(self.nAll)
        self.Ucprev = np.zeros(self.nC)

    def solve_load_step(self, t, criterion=0, convergence=1e-3, max_iterations=20, linear_solve=None):
        """one load step at normalised time t; a direct FP64 solve (dense, or SuperLU above
        SPARSE_ABOVE unknowns) stands in for PARDISO.  linear_solve(values2, rhs) -> dU, if
        given, replaces that solve (EquationSolver::solveEquations at FESolver.cpp:457 with
        K.block(2)'s upper CSR values on self.blocks[2]'s pattern): the hook through which a
        test runs this loop with another equation solver"""
        nC, nU = self.nC, self.nU
        U = self.U
        rhs1 = None
        for it in range(1, max_iterations + 1):
            K, F = self.assemble(U)
            R = np.zeros(self.nAll)
            for node, dof, val in self.loads:  # applyBoundaryConditions (FESolver.cpp:229-247)
                R[self.nodeEq[node - 1][dof] - 1] += val * t
            for node, dof, val in self.fixes:
                U[self.nodeEq[node - 1][dof] - 1] = val * t
            dUc = U[:nC] - self.Ucprev
            K12, K2 = self.sparse_blocks(K) if nU + self.nMpc > SPARSE_ABOVE else self.dense_blocks(K)
            rhs = F[nC:] + R[nC:] - K12.T @ dUc
            # equilibrium residual: lambda is a total, not an increment (FESolver.cpp:461), so the
            # multiplier forces K2 [0; lambda] are still in rhs and are taken out of the measure
            lam = np.zeros(rhs.size)
            lam[nU:] = U[nC + nU:]
            rn = np.linalg.norm(rhs - K2 @ lam)
            if it == 1:
                rhs1 = rn
            if criterion == 1 and it > 1 and rn <= convergence * rhs1:
                return it - 1
            if linear_solve is not None:
                dU = linear_solve(K[2], rhs)
            else:
                dU = spla.spsolve(K2.tocsc(), rhs) if sp.issparse(K2) else np.linalg.solve(K2, rhs)
            U[nC:nC + nU] += dU[:nU]
            U[nC + nU:] = dU[nU:]
            # constrained-DoF reactions: vecRc = K11 dUc + K12 dUsl - Fc (FESolver.cpp:464-467)
            self.R = R
            K11 = self.sparse_sym(self.blocks[1], K[1], nC) if sp.issparse(K2) else self.dense_k11(K)
            self.R[:nC] = K11 @ dUc + K12 @ dU - F[:nC]
            self.Ucprev = U[:nC].copy()
            if criterion == 0:
                crit = np.abs(dU[:nU]).sum() / nU / convergence
                if crit < convergence:
                    return it
                if crit > 1e6 or crit != crit:
                    raise RuntimeError("The solution is diverged!")
        raise RuntimeError("The solution is not converged")

    # ---------------- thermal (QUADTH / SurfaceLINETH) ----------------
    def assemble_thermal(self):
        """K, C, F of FEStorage::assembleGlobalEqMatrices with transient matrices (FEStorage.cpp:26-78):
        buildK / buildC of every element in element order (QUADTH.cpp:22-125)"""
        for b in self.blocks.values():
            b.values[:] = 0.0
        Cvals = {b: np.zeros_like(blk.values) for b, blk in self.blocks.items()}
        F = np.zeros(self.nAll)
        for e, (kind, nodes) in enumerate(self.elements):
            p = self.elem_params[e]
            X = self.xyz[np.array(nodes) - 1]
            eq = [self.nodeEq[n - 1][TEMP] for n in nodes]
            if kind == QUADTH:
                Ke, Fe, Ce = quadth_element(X, p[0], p[1], p[2], p[3])
                nn = 4
            elif kind == SURFACE_LINETH:
                Ke, Fe = lineth_element(X, p[0], p[1], p[2], p[3])
                Ce = None
                nn = 2
            else:
                continue
            k = 0
            for i in range(nn):  # Element::assembleK (element.h:356-386)
                for j in range(i, nn):
                    self.addValueK(eq[i], eq[j], Ke[k])
                    k += 1
            for i in range(nn):
                F[eq[i] - 1] += Fe[i]
            if Ce is not None:  # Element::assembleC (element.h:306-328)
                k = 0
                for i in range(nn):
                    for j in range(i, nn):
                        a, b = sorted((eq[i], eq[j]))
                        bi, pi = self._block(a)
                        bj, pj = self._block(b)
                        blk = 1 if (bi == bj == 1) else (2 if bi == bj == 2 else 12)
                        idx = self.blocks[blk].getIndex(pi, pj)
                        Cvals[blk][idx] += Ce[k]
                        k += 1
        return {b: blk.values.copy() for b, blk in self.blocks.items()}, Cvals, F

    def _apply_bc(self, U, R, t):
        for node, dof, val in self.loads:
            R[self.nodeEq[node - 1][dof] - 1] += val * t
        for node, dof, val in self.fixes:
            U[self.nodeEq[node - 1][dof] - 1] = val * t

    def solve_linear(self):
        """LinearFESolver::solve (FESolver.cpp:337-397) with a dense solve; returns vecU"""
        nC = self.nC
        K, C, F = self.assemble_thermal()
        U = np.zeros(self.nAll)
        R = np.zeros(self.nAll)
        self._apply_bc(U, R, 1.0)
        K12, K2 = self.dense_blocks(K)
        rhs = F[nC:] + R[nC:] - K12.T @ U[:nC]
        U[nC:] = np.linalg.solve(K2, rhs)
        return U

    def solve_transient(self, time0=0.0, time1=1.0, n_timesteps=100, init_value=0.0, alpha=0.25, delta=0.5,
                        dirichlet_coupling=False, probe=None):
        """LinearTransientFESolver::solve (FESolver.cpp:532-642) with dense solves.
        Returns (U after every step as a list of vectors if probe is None, else the probe eq history)."""
        nC = self.nC
        dt = (time1 - time0) / n_timesteps
        cur = time0 + dt
        a0 = 1.0 / (alpha * dt * dt); a1 = delta / (alpha * dt); a2 = 1.0 / (alpha * dt)
        a3 = 1.0 / (2.0 * alpha) - 1.0; a4 = delta / alpha - 1.0; a5 = dt / 2.0 * (delta / alpha - 2.0)
        a6 = dt * (1.0 - delta); a7 = delta * dt
        U = np.full(self.nAll, init_value)
        DU = np.zeros(self.nAll)
        DDU = np.zeros(self.nAll)
        K, Cv, F = self.assemble_thermal()
        R = np.zeros(self.nAll)
        self._apply_bc(U, R, 1.0)
        K12, K2 = self.dense_blocks(K)
        _, C2 = self.dense_blocks(Cv)
        Kmod = a1 * C2 + K2  # M = 0 for QUADTH
        out = []
        while cur <= time1:
            rhs = F[nC:] + R[nC:] + C2 @ (a1 * U[nC:] + a4 * DU[nC:] + a5 * DDU[nC:])
            if dirichlet_coupling:
                rhs = rhs - K12.T @ U[:nC]
            Un = U.copy()
            Un[nC:] = np.linalg.solve(Kmod, rhs)
            DDUn = a0 * (Un - U) - a2 * DU - a3 * DDU
            DUn = DU + a6 * DDUn + a7 * DDU
            U, DU, DDU = Un, DUn, DDUn
            out.append(U[probe] if probe is not None else U.copy())
            cur += dt
        return out
