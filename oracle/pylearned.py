"""CPU oracle of the learned block-Jacobi correction (test infrastructure only).

The reference has no counterpart (its solver is PARDISO, EquationSolver.cpp:158-249),
so this restates the definitions the GPU path implements, in FP64 numpy:

* block_structure: the block-Jacobi groups of block(2) (node DoF triples, element
  DoF and multiplier rows) and the learned nodes (free UX UY UZ triples);
* block_jacobi: inverse node blocks (Cholesky, |diag|^-1 when not positive
  definite) and the diagonal Schur estimate |K_ii| + sum_j K_ij^2/|K_jj| of the
  other rows;
* node_features: the 24 features of each learned node;
* mlp_forward / mlp_backward: the 24-128-128-128-128-6 GELU network;
* corrected_blocks: P_i = L^-T C C^T L^-1, K_ii = L L^T;
* minres_log_rate: mean log residual-reduction factor of k preconditioned MINRES
  steps (the Lanczos/plane-rotation recurrence of Ctx::minresSolve).

Pinning: mlp_backward against central differences of mlp_forward; the MINRES
recurrence against a dense solve (its residual estimate equals the true
residual of its iterate); minres_log_rate's gradient on the GPU is checked
against central differences of this function; node_features by its invariances
(unchanged under K -> sK, carried along by a renumbering of the nodes) and by
closed-form values on one undeformed unit cube (tests/test_oracle_pins.py).
"""
import math

import numpy as np
import scipy.sparse as sp

N_IN, WIDTH, N_OUT = 24, 128, 6
OFF_W = [0, 3200, 19712, 36224, 52736]
OFF_B = [3072, 19584, 36096, 52608, 53504]
N_PARAMS = 53510
SHAPES = [(WIDTH, N_IN), (WIDTH, WIDTH), (WIDTH, WIDTH), (WIDTH, WIDTH), (N_OUT, WIDTH)]


def unpack(w):
    w = np.asarray(w, dtype=np.float64)
    Ws = [w[OFF_W[l]:OFF_W[l] + SHAPES[l][0] * SHAPES[l][1]].reshape(SHAPES[l]) for l in range(5)]
    bs = [w[OFF_B[l]:OFF_B[l] + SHAPES[l][0]] for l in range(5)]
    return Ws, bs


def _erf(x):
    return np.vectorize(math.erf)(x)


def gelu(z):
    return 0.5 * z * (1.0 + _erf(z / math.sqrt(2.0)))


def gelu_grad(z):
    return 0.5 * (1.0 + _erf(z / math.sqrt(2.0))) + z * np.exp(-0.5 * z * z) / math.sqrt(2.0 * math.pi)


def mlp_forward(w, X, mean=None, istd=None):
    """returns (Y, cache)"""
    Ws, bs = unpack(w)
    X = np.asarray(X, dtype=np.float64)
    mean = np.zeros(N_IN) if mean is None else np.asarray(mean, dtype=np.float64)
    istd = np.ones(N_IN) if istd is None else np.asarray(istd, dtype=np.float64)
    a = (X - mean) * istd
    acts, zs = [a], []
    for l in range(4):
        z = a @ Ws[l].T + bs[l]
        zs.append(z)
        a = gelu(z)
        acts.append(a)
    Y = a @ Ws[4].T + bs[4]
    return Y, (acts, zs)


def mlp_backward(w, cache, dY):
    """gradient of sum(dY * Y) with respect to the flat parameters"""
    Ws, _ = unpack(w)
    acts, zs = cache
    g = np.zeros(N_PARAMS)
    dz = np.asarray(dY, dtype=np.float64)
    for l in range(4, -1, -1):
        gW = dz.T @ acts[l]
        g[OFF_W[l]:OFF_W[l] + gW.size] = gW.ravel()
        g[OFF_B[l]:OFF_B[l] + dz.shape[1]] = dz.sum(0)
        if l > 0:
            dz = (dz @ Ws[l]) * gelu_grad(zs[l - 1])
    return g


def block_structure(node_eq, elem_eq, nC, n2, mpc_eqs, K2):
    """groups [(r0, size, kind)] (kind 0 node block, 1 scalar Schur row), learned first rows,
    row kind (0 node, 1 element, 2 multiplier) and row node (-1 for non-node rows), from the
    1-based equation numbers (0 = no DoF)"""
    kind = np.full(n2, -1, dtype=np.int64)
    rtype = np.zeros(n2, dtype=np.int64)
    rnode = np.full(n2, -1, dtype=np.int64)
    for e in range(len(elem_eq)):
        for t in range(8):
            q = int(elem_eq[e, t])
            if q > nC:
                kind[q - nC - 1] = 1
    for n in range(len(node_eq)):
        for t in range(8):
            q = int(node_eq[n, t])
            if q > nC:
                kind[q - nC - 1] = 0
                rtype[q - nC - 1] = t
                rnode[q - nC - 1] = n
    for q in mpc_eqs:
        kind[q - nC - 1] = 2
    diag = K2.diagonal()
    has_diag = np.zeros(n2, dtype=bool)
    Kc = K2.tocsr()
    for i in range(n2):
        row = Kc.indices[Kc.indptr[i]:Kc.indptr[i + 1]]
        has_diag[i] = i in row
    groups, learned = [], []
    i = 0
    while i < n2:
        if kind[i] != 0 or not has_diag[i]:
            groups.append((i, 1, 1))
            i += 1
            continue
        sz = 1
        while (sz < 3 and i + sz < n2 and kind[i + sz] == 0 and has_diag[i + sz] and rnode[i + sz] == rnode[i]
               and rtype[i + sz] // 3 == rtype[i] // 3):
            sz += 1
        groups.append((i, sz, 0))
        if sz == 3 and list(rtype[i:i + 3]) == [0, 1, 2]:
            learned.append(i)
        i += sz
    return groups, np.array(learned, dtype=np.int64), kind, np.where(kind == 0, rnode, -1)


def block_jacobi(K2, groups):
    """list of (r0, size, inverse block) of bj_setup"""
    Kd = K2.tocsr()
    diag = np.abs(Kd.diagonal())
    out = []
    for r0, sz, kd in groups:
        if kd == 1:
            row = Kd.getrow(r0)
            d = 0.0
            for j, v in zip(row.indices, row.data):
                if j == r0:
                    d += abs(v)
                elif diag[j] > 0:
                    d += v * v / diag[j]
            out.append((r0, 1, np.array([[1.0 / d if d > 0 else 1.0]])))
            continue
        B = Kd[r0:r0 + sz, r0:r0 + sz].toarray()
        Bs = 0.5 * (B + B.T)
        try:
            L = np.linalg.cholesky(Bs)
            Li = np.linalg.inv(L)
            out.append((r0, sz, Li.T @ Li))
        except np.linalg.LinAlgError:
            dd = np.abs(np.diag(B))
            out.append((r0, sz, np.diag(np.where(dd > 0, 1.0 / np.where(dd > 0, dd, 1.0), 1.0))))
    return out


def correction_factor(o):
    c = np.exp(np.clip(np.asarray(o[:3], dtype=np.float64), -4.0, 4.0))
    return np.array([[c[0], 0, 0], [o[3], c[1], 0], [o[4], o[5], c[2]]], dtype=np.float64)


def corrected_blocks(K2, learned_rows, Y):
    Kd = K2.tocsr()
    out = []
    for r0, o in zip(learned_rows, Y):
        B = Kd[r0:r0 + 3, r0:r0 + 3].toarray()
        L = np.linalg.cholesky(0.5 * (B + B.T))
        Li = np.linalg.inv(L)
        C = correction_factor(o)
        out.append(Li.T @ C @ C.T @ Li)
    return np.array(out)


def precond_matrix(n2, bj, learned_rows=None, learned_blocks=None):
    rows, cols, vals = [], [], []
    over = {}
    if learned_rows is not None:
        over = {int(r): P for r, P in zip(learned_rows, learned_blocks)}
    for r0, sz, B in bj:
        B = over.get(r0, B)
        for a in range(sz):
            for b in range(sz):
                rows.append(r0 + a)
                cols.append(r0 + b)
                vals.append(B[a, b])
    return sp.csr_matrix((vals, (rows, cols)), shape=(n2, n2))


def node_features(K2, learned_rows, kind, rnode, k12_count, bj, U, nC, bulk, t):
    """features of each learned node (the definition the feature kernel evaluates)"""
    Kd = K2.tocsr()
    schur_inv = {r0: B[0, 0] for r0, sz, B in bj if sz == 1 and kind[r0] == 1}
    X = np.zeros((len(learned_rows), N_IN))
    for i, r0 in enumerate(learned_rows):
        node = rnode[r0]
        Kii = Kd[r0:r0 + 3, r0:r0 + 3].toarray()
        rs = np.zeros(3)
        offsq, offmax = 0.0, 0.0
        nbr = []
        pcols = {}
        nlam = 0
        for a in range(3):
            row = Kd.getrow(r0 + a)
            for c, v in sorted(zip(row.indices, row.data)):
                if kind[c] == 0:
                    rs[a] += v
                    if not (rnode[c] == node and r0 <= c < r0 + 3):
                        offsq += v * v
                        offmax = max(offmax, abs(v))
                        if a == 0 and rnode[c] != node and rnode[c] not in nbr:
                            nbr.append(rnode[c])
                elif kind[c] == 1:
                    pcols.setdefault(c, np.zeros(3))[a] = v
                else:
                    nlam += 1
        s = np.trace(Kii) / 3.0
        f = np.zeros(N_IN)
        f[0:6] = np.array([Kii[0, 0], 0.5 * (Kii[0, 1] + Kii[1, 0]), 0.5 * (Kii[0, 2] + Kii[2, 0]), Kii[1, 1],
                           0.5 * (Kii[1, 2] + Kii[2, 1]), Kii[2, 2]]) / s
        f[6] = math.sqrt(offsq) / s / 26.0
        f[7] = offmax / s
        f[8] = len(nbr) / 26.0
        f[9] = np.linalg.norm(rs) / s
        G = np.zeros((3, 3))
        pj = []
        cshare = 0.0
        for c, kp in pcols.items():
            kpp = abs(Kd[c, c])
            if kpp > 0:
                G += np.outer(kp, kp) / kpp
            pj.append(U[nC + c] / bulk)
            cshare += kpp * schur_inv[c]
        trG = np.trace(G)
        f[10:16] = np.array([G[0, 0], G[0, 1], G[0, 2], G[1, 1], G[1, 2], G[2, 2]]) / trG if trG > 0 else 0.0
        f[16] = math.log10(trG / (3.0 * s)) if trG > 0 else 0.0
        f[17] = len(pcols) / 8.0
        f[18] = nlam / 3.0
        f[19] = sum(k12_count[r0:r0 + 3]) / 81.0
        f[20] = np.mean(pj) if pj else 0.0
        f[21] = min(pj) if pj else 0.0
        f[22] = t
        f[23] = cshare / len(pcols) if pcols else 0.0
        X[i] = f
    return X


def minres_log_rate(K2, M, b, k):
    """mean over k steps of log(phibar_j / phibar_{j-1}) for MINRES on K2 x = b, preconditioner M"""
    r1 = np.asarray(b, dtype=np.float64).copy()
    y = M @ r1
    beta = math.sqrt(r1 @ y)
    oldb, dbar, cs, sn = 0.0, 0.0, -1.0, 0.0
    r2 = r1.copy()
    L = 0.0
    for itn in range(1, k + 1):
        v = y / beta
        y = K2 @ v
        if itn >= 2:
            y = y - (beta / oldb) * r1
        alfa = v @ y
        y = y - (alfa / beta) * r2
        r1, r2 = r2, y
        y = M @ r2
        oldb, beta = beta, math.sqrt(r2 @ y)
        gbar = sn * dbar - cs * alfa
        dbar = -cs * beta
        gamma = math.hypot(gbar, beta)
        cs, sn = gbar / gamma, beta / gamma
        L += math.log(sn)
    return L / k


def minres_solve(K2, M, b, tol, maxit):
    """plain preconditioned MINRES (the recurrence above carried to a solution); returns x, iterations"""
    n = len(b)
    x = np.zeros(n)
    r1 = np.asarray(b, dtype=np.float64).copy()
    y = M @ r1
    beta1 = math.sqrt(r1 @ y)
    beta, oldb, dbar, epsln, phibar, cs, sn = beta1, 0.0, 0.0, 0.0, beta1, -1.0, 0.0
    w = np.zeros(n)
    w2 = np.zeros(n)
    r2 = r1.copy()
    for itn in range(1, maxit + 1):
        v = y / beta
        y = K2 @ v
        if itn >= 2:
            y = y - (beta / oldb) * r1
        alfa = v @ y
        y = y - (alfa / beta) * r2
        r1, r2 = r2, y
        y = M @ r2
        oldb, beta = beta, math.sqrt(r2 @ y)
        oldeps = epsln
        delta = cs * dbar + sn * alfa
        gbar = sn * dbar - cs * alfa
        epsln = sn * beta
        dbar = -cs * beta
        gamma = max(math.hypot(gbar, beta), 1e-300)
        cs, sn = gbar / gamma, beta / gamma
        phi = cs * phibar
        phibar = sn * phibar
        w1, w2 = w2, w
        w = (v - oldeps * w1 - delta * w2) / gamma
        x = x + phi * w
        if phibar <= tol * beta1:
            return x, itn
    return x, maxit
