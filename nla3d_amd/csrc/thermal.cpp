// Thermal path on the device: QUADTH / SurfaceLINETH assembly of K, C and F,
// LinearFESolver::solve (FESolver.cpp:337-397) and
// LinearTransientFESolver::solve (FESolver.cpp:532-642) with CG in place of
// the direct solve.
#include <algorithm>
#include <chrono>
#include <climits>
#include <cmath>
#include <functional>
#include <mutex>

#include "../../include/nla3d_gpu.h"
#include "ctx.h"
#include "newton.h"
#include "thermal.h"

namespace nla3d_amd {

static double nowMsT() {
  return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// device data of the thermal elements, the [K | F | C] scatter map and the
// transient state; created by setupThermal() when the model has thermal elements
struct ThermalData {
  int nQuad = 0, nLine = 0;
  DevBuf quadConn, quadPar, lineConn, linePar, out;  // out: [nQuad*24 | nLine*5]
  DevBuf seg, src;
  int nTargets = 0;
  DevBuf full2C;  // C.block(2) in the grouped full-symmetric layout
  bool assembled = false;
  // transient
  nla3d_gpu_transient_opts o{};
  double a0 = 0, a1 = 0, a2 = 0, a3 = 0, a4 = 0, a5 = 0, a6 = 0, a7 = 0, dt = 0, curTime = 0;
  int curStep = 0;
  bool running = false;
  DevBuf DU, DDU, v, y, x;
  // the predictor: the last steps' increments of the unknowns (a ring, newest at histHead),
  // their products with the CG matrix, the residual of the current unknowns, a scratch vector
  static constexpr int kHist = 3;
  DevBuf hist[kHist], hy[kHist], r0, scratch;
  int histN = 0, histHead = 0;
  DevBuf cgw[5], cgPart2, cgSt, cgSt0;
  CgState* hst = nullptr;
  // scalar-row CG copy (built on the first solve when every row is its own group)
  bool sellTried = false, sellOk = false;
  DevBuf sOff, sBase, sCol16, sCol, sPos, sVal, sDinv;
  CgSell sell;
  ~ThermalData() {
    if (hst) (void)hipHostFree(hst);
  }
};

ThermalData& thermal(Ctx& c) {
  if (!c.thermal) c.thermal = std::make_shared<ThermalData>();
  return *c.thermal;
}

void setupThermal(Ctx& c) {
  FEStorage& m = c.model;
  ThermalData& T = thermal(c);
  std::vector<int> qc, lc, qIdx, lIdx;
  std::vector<double> qp, lp;
  for (uint32_t e = 0; e < m.nElements(); e++) {
    const uint64_t b = m.elemNodePtr[e];
    if (m.elemKind[e] == ELEM_QUADTH) {
      qIdx.push_back((int)e);
      for (int i = 0; i < 4; i++) qc.push_back((int)m.elemNodes[b + i] - 1);
      for (int i = 0; i < 4; i++) qp.push_back(m.elemParams[(size_t)e * 4 + i]);
    } else if (m.elemKind[e] == ELEM_SURFACE_LINETH) {
      lIdx.push_back((int)e);
      for (int i = 0; i < 2; i++) lc.push_back((int)m.elemNodes[b + i] - 1);
      for (int i = 0; i < 4; i++) lp.push_back(m.elemParams[(size_t)e * 4 + i]);
    }
  }
  T.nQuad = (int)qIdx.size();
  T.nLine = (int)lIdx.size();
  if (T.nQuad + T.nLine == 0) return;
  T.quadConn.upload(qc);
  T.quadPar.upload(qp);
  T.lineConn.upload(lc);
  T.linePar.upload(lp);
  T.out.alloc(((size_t)T.nQuad * kQuadthSlots + (size_t)T.nLine * kLinethSlots) * 8);
  // targets: [K (nnzAll) | F (nAll) | C (nnzAll)]
  const uint64_t nK = m.nnzAll(), offF = nK, offC = nK + m.nAll();
  const uint64_t nT = offC + nK;
  if (nT >= (1ull << 31)) throw ApiError(-1, "thermal model too large for 32-bit scatter indices");
  c.d_vals.alloc((size_t)nT * 8);
  hipCheck(hipMemset(c.d_vals.p, 0, c.d_vals.bytes), "memset vals");
  // contributions enumerated in global element order (the reference's element loop)
  std::vector<int> qpos(m.nElements(), -1), lpos(m.nElements(), -1);
  for (int i = 0; i < T.nQuad; i++) qpos[qIdx[i]] = i;
  for (int i = 0; i < T.nLine; i++) lpos[lIdx[i]] = i;
  std::vector<std::pair<int, int>> contrib;  // (target, src)
  contrib.reserve((size_t)T.nQuad * 24 + (size_t)T.nLine * 5);
  auto need = [&](int64_t t, uint32_t a, uint32_t b) {
    if (t < 0)
      throw ApiError(-4, "The position(" + std::to_string(std::min(a, b)) + ", " + std::to_string(std::max(a, b)) +
                             ") is absent in the matrix");
    return (int)t;
  };
  const size_t lineBase = (size_t)T.nQuad * kQuadthSlots;
  for (uint32_t e = 0; e < m.nElements(); e++) {
    const uint64_t b = m.elemNodePtr[e];
    if (qpos[e] >= 0) {
      uint32_t eq[4];
      for (int i = 0; i < 4; i++) eq[i] = m.nodeDofEq(m.elemNodes[b + i], TEMP);
      const int base = qpos[e] * kQuadthSlots;
      int k = 0;
      for (int i = 0; i < 4; i++)
        for (int j = i; j < 4; j++, k++) {
          const int t = need(m.findEntry(eq[i], eq[j]), eq[i], eq[j]);
          contrib.push_back({t, base + k});
          contrib.push_back({(int)(offC + t), base + 10 + k});
        }
      for (int i = 0; i < 4; i++) contrib.push_back({(int)(offF + eq[i] - 1), base + 20 + i});
    } else if (lpos[e] >= 0) {
      uint32_t eq[2];
      for (int i = 0; i < 2; i++) eq[i] = m.nodeDofEq(m.elemNodes[b + i], TEMP);
      const int base = (int)lineBase + lpos[e] * kLinethSlots;
      int k = 0;
      for (int i = 0; i < 2; i++)
        for (int j = i; j < 2; j++, k++) contrib.push_back({need(m.findEntry(eq[i], eq[j]), eq[i], eq[j]), base + k});
      for (int i = 0; i < 2; i++) contrib.push_back({(int)(offF + eq[i] - 1), base + 3 + i});
    }
  }
  std::vector<int> seg(nT + 1, 0), src(contrib.size());
  for (auto& p : contrib) seg[(size_t)p.first + 1]++;
  for (uint64_t t = 0; t < nT; t++) seg[t + 1] += seg[t];
  {
    std::vector<int> pos(seg.begin(), seg.end() - 1);
    for (auto& p : contrib) src[pos[p.first]++] = p.second;  // stable: contributions are in element order
  }
  T.seg.upload(seg);
  T.src.upload(src);
  T.nTargets = (int)nT;
  T.full2C.alloc(c.d_full2.bytes);
  hipCheck(hipMemset(T.full2C.p, 0, T.full2C.bytes), "memset C");
}

// FEStorage::assembleGlobalEqMatrices with transient = true (FEStorage.cpp:26-78):
// K and F by buildK, C by buildC (M is empty for QUADTH: QUADTH.h:33)
void assembleThermal(Ctx& c) {
  FEStorage& m = c.model;
  ThermalData& T = thermal(c);
  if (T.nQuad + T.nLine == 0) throw ApiError(-3, "model has no thermal elements");
  double* out = T.out.as<double>();
  launch_quadth(c.stream, T.nQuad, T.quadConn.as<int>(), c.d_xyz.as<double>(), T.quadPar.as<double>(), out);
  launch_lineth(c.stream, T.nLine, T.lineConn.as<int>(), c.d_xyz.as<double>(), T.linePar.as<double>(),
                out + (size_t)T.nQuad * kQuadthSlots);
  launch_segmented_gather(c.stream, T.nTargets, T.seg.as<int>(), T.src.as<int>(), out, c.d_vals.as<double>(), 0, 0,
                          nullptr, nullptr);
  hipCheck(hipGetLastError(), "thermal assembly");
  // block(2) of K and of C into the grouped full-symmetric layouts
  const int n2u = (int)m.block2.nnz();
  const double* K2 = c.d_vals.as<double>() + m.offB2();
  const double* C2 = c.d_vals.as<double>() + m.nnzAll() + m.nAll() + m.offB2();
  launch_combine_mirror(c.stream, n2u, 1.0, K2, 0.0, nullptr, c.d_mir2.as<int>(), c.d_full2.as<double>());
  launch_combine_mirror(c.stream, n2u, 1.0, C2, 0.0, nullptr, c.d_mir2.as<int>(), T.full2C.as<double>());
  c.refreshSpmvCopy(c.stream);
  hipCheck(hipGetLastError(), "thermal mirror");
  T.assembled = true;
}

// the sliced-ELLPACK copy of a matrix whose rows are all single-row groups (QUADTH: one
// TEMP DoF per node), from the grouped full-symmetric structure
static void buildSell(Ctx& c, ThermalData& T) {
  T.sellTried = true;
  const int n = c.n2, ng = c.nSpmvGroups;
  if (n == 0 || ng != n || c.nGroups != n) return;
  auto down = [](const DevBuf& d, size_t count) {
    std::vector<int> v(count);
    hipCheck(hipMemcpy(v.data(), d.p, count * 4, hipMemcpyDeviceToHost), "download");
    return v;
  };
  const std::vector<int> size = down(c.d_sgSize, ng), row0 = down(c.d_sgRow0, ng), cptr = down(c.d_sgColPtr, ng + 1),
                         vptr = down(c.d_sgValPtr, ng), info = down(c.d_gInfo, n);
  for (int g = 0; g < ng; g++)
    if (size[g] != 1 || row0[g] != g || (info[g] & 3) != 1) return;
  const std::vector<int> cols = down(c.d_sgCols, (size_t)cptr[ng]);
  const int nSl = (n + 63) / 64;
  std::vector<int> off(nSl + 1, 0), base(nSl, 0);
  bool narrow = true;
  for (int sl = 0; sl < nSl; sl++) {
    int w = 0, lo = INT_MAX, hi = -1;
    for (int i = sl * 64; i < std::min(n, sl * 64 + 64); i++) {
      w = std::max(w, cptr[i + 1] - cptr[i]);
      for (int k = cptr[i]; k < cptr[i + 1]; k++) {
        lo = std::min(lo, cols[k]);
        hi = std::max(hi, cols[k]);
      }
    }
    if (hi < 0) lo = hi = sl * 64;
    base[sl] = lo;
    if (hi - lo > 65535) narrow = false;
    if ((int64_t)off[sl] + 64ll * w >= (1ll << 31)) return;
    off[sl + 1] = off[sl] + 64 * w;
  }
  const int nSlots = off[nSl];
  std::vector<int> scol(nSlots), spos(nSlots);
  for (int sl = 0; sl < nSl; sl++) {
    const int w = (off[sl + 1] - off[sl]) / 64;
    for (int l = 0; l < 64; l++) {
      const int i = sl * 64 + l;
      for (int k = 0; k < w; k++) {
        const int j = off[sl] + k * 64 + l;
        const bool real = i < n && k < cptr[i + 1] - cptr[i];
        scol[j] = real ? cols[cptr[i] + k] : base[sl];
        spos[j] = real ? vptr[i] + k : -1;
      }
    }
  }
  T.sOff.upload(off);
  T.sPos.upload(spos);
  if (narrow) {
    std::vector<unsigned short> c16(nSlots);
    for (int sl = 0; sl < nSl; sl++)
      for (int j = off[sl]; j < off[sl + 1]; j++) c16[j] = (unsigned short)(scol[j] - base[sl]);
    T.sCol16.upload(c16);
    T.sBase.upload(base);
  } else {
    T.sCol.upload(scol);
  }
  T.sVal.alloc((size_t)std::max(nSlots, 1) * 8);
  T.sDinv.alloc((size_t)n * 8);
  CgSell& A = T.sell;
  A.n = n;
  A.nSlices = nSl;
  A.nSlots = nSlots;
  A.off = T.sOff.as<int>();
  A.base = narrow ? T.sBase.as<int>() : nullptr;
  A.col16 = narrow ? T.sCol16.as<unsigned short>() : nullptr;
  A.col = narrow ? nullptr : T.sCol.as<int>();
  A.pos = T.sPos.as<int>();
  A.val = T.sVal.as<double>();
  A.dinv = T.sDinv.as<double>();
  T.sellOk = true;
}

static CgWork cgWork(Ctx& c, ThermalData& T) {
  const size_t n = (size_t)c.n2 * 8;
  for (auto& b : T.cgw) b.alloc(n);
  T.cgPart2.alloc(c.d_partial.bytes);
  T.cgSt.alloc(sizeof(CgState));
  if (!T.cgSt0.p) {
    T.cgSt0.alloc(sizeof(CgState));
    hipCheck(hipMemset(T.cgSt0.p, 0, sizeof(CgState)), "memset");
  }
  if (!T.hst) hipCheck(hipHostMalloc((void**)&T.hst, sizeof(CgState), hipHostMallocDefault), "pinned");
  CgWork w;
  w.r = T.cgw[0].as<double>();
  w.z = T.cgw[1].as<double>();
  w.p = T.cgw[2].as<double>();
  w.q = T.cgw[3].as<double>();
  w.p2 = T.cgw[4].as<double>();
  w.partial = c.d_partial.as<double>();
  w.partial2 = T.cgPart2.as<double>();
  w.st = T.cgSt.as<CgState>();
  w.st0 = T.cgSt0.as<CgState>();
  return w;
}

static int cgSolve(Ctx& c, ThermalData& T, const double* b, double* x, double tol, int maxit,
                   nla3d_gpu_newton_stats& st) {
  CgWork w = cgWork(c, T);
  if (!T.sellTried) buildSell(c, T);
  const double avgCols = c.n2 > 0 ? (double)c.fullNnz2 / c.n2 : 0.0;
  const int it = cg_solve(c.stream, c.krylovMatrix(), c.precond(), w, b, x, tol, maxit, T.hst, avgCols,
                          T.sellOk ? &T.sell : nullptr);
  if (it < 0) throw ApiError(-2, "HIP error in CG");
  st.linear_solves++;
  st.krylov_iterations += it;
  st.last_krylov_iterations = it;
  if (T.hst->breakdown) st.krylov_breakdown = 1;
  if (!T.hst->converged) throw ApiError(-5, "CG did not converge in " + std::to_string(it) + " iterations");
  return it;
}

// LinearFESolver::solve (FESolver.cpp:337-397)
void solveLinearThermal(Ctx& c, double tol, int maxit, nla3d_gpu_newton_stats& st) {
  FEStorage& m = c.model;
  ThermalData& T = thermal(c);
  const int nC = (int)m.nConstrainedDofs;
  const double t0 = nowMsT();
  hipCheck(hipMemsetAsync(c.d_R.p, 0, c.d_R.bytes, c.stream), "zero R");
  assembleThermal(c);
  double* U = c.d_U.as<double>();
  newton_apply_bc(c.stream, (int)m.fixs.size(), c.d_fixEq.as<int>(), c.d_fixVal.as<double>(), (int)m.loads.size(),
                  c.d_loadEq.as<int>(), c.d_loadVal.as<double>(), 1.0, U, c.d_R.as<double>());
  // rhs = F_sl + R_sl - K12^T Uc
  launch_tr_rhs(c.stream, c.n2, nC, c.d_vals.as<double>() + m.nnzAll(), c.d_R.as<double>(), nullptr, U,
                c.d_k12tPtr.as<int>(), c.d_k12tPos.as<int>(), c.d_k12tRow.as<int>(),
                c.d_vals.as<double>() + m.offB12(), 1, c.d_rhs.as<double>());
  krylov_precond_setup(c.stream, c.krylovMatrix(), c.precond());
  T.x.alloc((size_t)c.n2 * 8);
  launch_fill(c.stream, c.n2, 0.0, T.x.as<double>());
  st.ms_assemble += nowMsT() - t0;
  const double t1 = nowMsT();
  cgSolve(c, T, c.d_rhs.as<double>(), T.x.as<double>(), tol, maxit, st);
  launch_copy(c.stream, c.n2, T.x.as<double>(), U + nC);
  hipCheck(hipStreamSynchronize(c.stream), "sync");
  st.ms_solve += nowMsT() - t1;
  st.load_steps++;
}

// LinearTransientFESolver::solve set-up part (FESolver.cpp:539-601)
void transientBegin(Ctx& c, const nla3d_gpu_transient_opts& o) {
  FEStorage& m = c.model;
  ThermalData& T = thermal(c);
  T.o = o;
  const double dt = (o.time1 - o.time0) / o.n_timesteps;
  T.dt = dt;
  T.a0 = 1.0 / (o.alpha * dt * dt);
  T.a1 = o.delta / (o.alpha * dt);
  T.a2 = 1.0 / (o.alpha * dt);
  T.a3 = 1.0 / (2.0 * o.alpha) - 1.0;
  T.a4 = o.delta / o.alpha - 1.0;
  T.a5 = dt / 2.0 * (o.delta / o.alpha - 2.0);
  T.a6 = dt * (1.0 - o.delta);
  T.a7 = o.delta * dt;
  T.curTime = o.time0 + dt;
  T.curStep = 1;
  const int nAll = (int)m.nAll();
  T.DU.alloc((size_t)nAll * 8);
  T.DDU.alloc((size_t)nAll * 8);
  T.v.alloc((size_t)c.n2 * 8);
  T.y.alloc((size_t)c.n2 * 8);
  T.x.alloc((size_t)c.n2 * 8);
  T.histN = T.histHead = 0;
  if (o.predictor) {
    for (int i = 0; i < ThermalData::kHist; i++) {
      T.hist[i].alloc((size_t)c.n2 * 8);
      T.hy[i].alloc((size_t)c.n2 * 8);
    }
    T.r0.alloc((size_t)c.n2 * 8);
    T.scratch.alloc((size_t)c.n2 * 8);
  }
  hipCheck(hipMemsetAsync(c.d_R.p, 0, c.d_R.bytes, c.stream), "zero R");
  double* U = c.d_U.as<double>();
  launch_fill(c.stream, nAll, o.init_value, U);  // vecU = initValue (FESolver.cpp:577-579)
  launch_fill(c.stream, nAll, 0.0, T.DU.as<double>());
  launch_fill(c.stream, nAll, 0.0, T.DDU.as<double>());
  assembleThermal(c);
  newton_apply_bc(c.stream, (int)m.fixs.size(), c.d_fixEq.as<int>(), c.d_fixVal.as<double>(), (int)m.loads.size(),
                  c.d_loadEq.as<int>(), c.d_loadVal.as<double>(), 1.0, U, c.d_R.as<double>());
  // Kmod = a0 M + a1 C + K on block(2), M = 0 (FESolver.cpp:594-598)
  const int n2u = (int)m.block2.nnz();
  const double* K2 = c.d_vals.as<double>() + m.offB2();
  const double* C2 = c.d_vals.as<double>() + m.nnzAll() + m.nAll() + m.offB2();
  launch_combine_mirror(c.stream, n2u, 1.0, K2, T.a1, C2, c.d_mir2.as<int>(), c.d_full2.as<double>());
  c.refreshSpmvCopy(c.stream);
  krylov_precond_setup(c.stream, c.krylovMatrix(), c.precond());
  hipCheck(hipGetLastError(), "transient set-up");
  hipCheck(hipStreamSynchronize(c.stream), "sync");
  T.running = true;
}

// T.x = Us + sum_i a_i d_i over the history (newest first), the a_i minimising ||b - A x||_2 for
// the CG matrix A; CG's tolerance is relative to ||b||_M, so a better start only saves
// iterations
static void transientPredict(Ctx& c, ThermalData& T, const double* Us) {
  const int n = c.n2, m0 = T.histN;
  const KrylovMatrix A = c.krylovMatrix();
  double* part = c.d_partial.as<double>();
  double* sc = c.d_scal.as<double>();
  auto dot = [&](const double* x, const double* y) {
    krylov_dot(c.stream, n, x, y, part, sc + 11);
    double v;
    hipCheck(hipMemcpyAsync(&v, sc + 11, 8, hipMemcpyDeviceToHost, c.stream), "dot");
    hipCheck(hipStreamSynchronize(c.stream), "sync");
    return v;
  };
  const double* d[ThermalData::kHist];
  double* y[ThermalData::kHist];
  for (int i = 0; i < m0; i++) {
    d[i] = T.hist[(T.histHead - i + ThermalData::kHist) % ThermalData::kHist].as<double>();
    y[i] = T.hy[i].as<double>();
    krylov_spmv(c.stream, A, d[i], y[i]);
  }
  double* r0 = T.r0.as<double>();
  krylov_spmv(c.stream, A, Us, r0);
  vec_rsub(c.stream, n, c.d_rhs.as<double>(), r0);  // r0 = b - A Us
  double G[kMaxLsBasis][kMaxLsBasis] = {}, g[kMaxLsBasis] = {}, a[kMaxLsBasis];
  for (int i = 0; i < m0; i++) {
    g[i] = dot(y[i], r0);
    for (int k = 0; k <= i; k++) G[i][k] = dot(y[i], y[k]);
  }
  const int m = lsNormalSolve(G, g, m0, a);
  // x = Us + sum a_i d_i (the residual update into a scratch vector is not needed)
  for (int i = 0; i < m; i++)
    newton_pred_axpy(c.stream, n, a[i], d[i], y[i], i == 0 ? Us : nullptr, i == 0 ? r0 : nullptr, T.x.as<double>(),
                     T.scratch.as<double>());
  hipCheck(hipGetLastError(), "transient predictor");
}

// one pass of the time loop (FESolver.cpp:604-636); returns false when curTime > time1
bool transientStep(Ctx& c, nla3d_gpu_newton_stats& st) {
  FEStorage& m = c.model;
  ThermalData& T = thermal(c);
  if (!T.running) throw ApiError(-3, "call nla3d_gpu_transient_begin first");
  if (!(T.curTime <= T.o.time1)) return false;
  const int nC = (int)m.nConstrainedDofs;
  double* U = c.d_U.as<double>();
  const double t0 = nowMsT();
  // rhs = F + R + C (a1 U + a4 DU + a5 DDU) on the s part [- K12^T Uc]
  launch_tr_comb(c.stream, c.n2, nC, T.a1, U, T.a4, T.DU.as<double>(), T.a5, T.DDU.as<double>(), T.v.as<double>());
  KrylovMatrix MC = c.krylovMatrix();
  MC.val = T.full2C.as<double>();
  MC.nDiaFit = 0;  // C has no node-diagonal copy of its own: all rows through the grouped CSR
  MC.csrList = nullptr;
  krylov_spmv(c.stream, MC, T.v.as<double>(), T.y.as<double>());
  launch_tr_rhs(c.stream, c.n2, nC, c.d_vals.as<double>() + m.nnzAll(), c.d_R.as<double>(), T.y.as<double>(), U,
                c.d_k12tPtr.as<int>(), c.d_k12tPos.as<int>(), c.d_k12tRow.as<int>(),
                c.d_vals.as<double>() + m.offB12(), T.o.dirichlet_coupling, c.d_rhs.as<double>());
  // initial guess: the current unknowns, plus with the predictor the least-squares combination
  // of the last steps' increments
  launch_copy(c.stream, c.n2, U + nC, T.x.as<double>());
  st.ms_assemble += nowMsT() - t0;
  const double t1 = nowMsT();
  if (T.o.predictor && T.histN > 0) transientPredict(c, T, U + nC);
  cgSolve(c, T, c.d_rhs.as<double>(), T.x.as<double>(), T.o.krylov_tol, T.o.krylov_max_iter, st);
  if (T.o.predictor) {  // this step's increment into the history
    const int slot = (T.histHead + 1) % ThermalData::kHist;
    double* d = T.hist[slot].as<double>();
    launch_copy(c.stream, c.n2, U + nC, d);
    vec_rsub(c.stream, c.n2, T.x.as<double>(), d);  // d = x - U_s
    T.histHead = slot;
    T.histN = std::min(ThermalData::kHist, T.histN + 1);
  }
  launch_tr_update(c.stream, (int)m.nAll(), nC, T.x.as<double>(), T.a0, T.a2, T.a3, T.a6, T.a7, U, T.DU.as<double>(),
                   T.DDU.as<double>());
  hipCheck(hipGetLastError(), "transient update");
  hipCheck(hipStreamSynchronize(c.stream), "sync");
  st.ms_solve += nowMsT() - t1;
  st.load_steps++;
  st.last_time = T.curTime;
  T.curStep++;
  T.curTime += T.dt;
  return true;
}

}  // namespace nla3d_amd

using namespace nla3d_amd;

static int thermalGuard(nla3d_gpu_t h, const std::function<void(Ctx&)>& fn) {
  if (!h) return -1;
  Ctx& c = *reinterpret_cast<Ctx*>(h);
  try {
    c.err.clear();
    c.requireDevice();
    if (!c.initialized) throw ApiError(-3, "call nla3d_gpu_init_solution first");
    hipCheck(hipSetDevice(c.device), "hipSetDevice");
    fn(c);
    return 0;
  } catch (const ApiError& e) {
    c.err = e.what();
    return e.code;
  } catch (const std::exception& e) {
    c.err = e.what();
    return -1;
  }
}

extern "C" {

int nla3d_gpu_assemble_thermal(nla3d_gpu_t h) {
  return thermalGuard(h, [&](Ctx& c) {
    assembleThermal(c);
    hipCheck(hipStreamSynchronize(c.stream), "sync");
  });
}

int nla3d_gpu_get_C(nla3d_gpu_t h, int block, double* values) {
  return thermalGuard(h, [&](Ctx& c) {
    if (!values) throw ApiError(-1, "NULL");
    FEStorage& m = c.model;
    uint64_t off, n;
    if (block == 1) { off = 0; n = m.block1.nnz(); }
    else if (block == 12) { off = m.offB12(); n = m.block12.nnz(); }
    else if (block == 2) { off = m.offB2(); n = m.block2.nnz(); }
    else throw ApiError(-1, "block must be 1, 12 or 2");
    if (c.d_vals.bytes < (m.nnzAll() * 2 + m.nAll()) * 8) throw ApiError(-3, "no C matrix (not a thermal model)");
    hipCheck(hipMemcpyAsync(values, c.d_vals.as<double>() + m.nnzAll() + m.nAll() + off, n * 8,
                            hipMemcpyDeviceToHost, c.stream), "copy C");
    hipCheck(hipStreamSynchronize(c.stream), "sync");
  });
}

int nla3d_gpu_solve_linear(nla3d_gpu_t h, double krylov_tol, int krylov_max_iter, nla3d_gpu_newton_stats* stats) {
  if (!stats) return -1;
  return thermalGuard(h, [&](Ctx& c) {
    const double t0 = nowMsT();
    solveLinearThermal(c, krylov_tol, krylov_max_iter, *stats);
    stats->ms_total += nowMsT() - t0;
  });
}

int nla3d_gpu_transient_defaults(nla3d_gpu_transient_opts* o) {
  if (!o) return -1;
  o->alpha = 0.25;  // FESolver.h:230-237
  o->delta = 0.5;
  o->time0 = 0.0;
  o->time1 = 1.0;
  o->n_timesteps = 100;
  o->init_value = 0.0;
  o->dirichlet_coupling = 0;
  o->krylov_tol = 1.0e-12;
  o->krylov_max_iter = 100000;
  o->predictor = 1;
  return 0;
}

int nla3d_gpu_transient_begin(nla3d_gpu_t h, const nla3d_gpu_transient_opts* opts) {
  if (!opts || opts->n_timesteps < 1 || !(opts->alpha > 0.0)) return -1;
  return thermalGuard(h, [&](Ctx& c) { transientBegin(c, *opts); });
}

int nla3d_gpu_transient_step(nla3d_gpu_t h, nla3d_gpu_newton_stats* stats, int* done) {
  if (!stats || !done) return -1;
  return thermalGuard(h, [&](Ctx& c) {
    const double t0 = nowMsT();
    *done = transientStep(c, *stats) ? 0 : 1;
    stats->ms_total += nowMsT() - t0;
  });
}

int nla3d_gpu_solve_transient(nla3d_gpu_t h, const nla3d_gpu_transient_opts* opts, nla3d_gpu_newton_stats* stats) {
  if (!opts || !stats || opts->n_timesteps < 1) return -1;
  return thermalGuard(h, [&](Ctx& c) {
    const double t0 = nowMsT();
    transientBegin(c, *opts);
    while (transientStep(c, *stats)) {
    }
    stats->ms_total += nowMsT() - t0;
  });
}

}  // extern "C"
