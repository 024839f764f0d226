// EquationSolver on a caller's matrix (math/EquationSolver.h:19-78): the drop-in for the
// solver nla3d's FESolver calls with K.block(2) (FESolver.cpp:377, 457, 601, 613), handed
// over as a SparseSymMatrix's raw 1-based upper CSR (getValuesArray / getIofeirArray /
// getColumnsArray, SparseMatrix.h:114-116).
//
// factorize() analyses the pattern on its first call (and again only when the arrays change)
// and then, for every call, copies the values to the device, mirrors them into the grouped
// full-symmetric copy the SpMV reads, inverts the block-Jacobi blocks and refreshes the
// Schur-row multigrid; substitute() runs preconditioned MINRES for one right-hand side.
// This is PARDISO's phase split (11 once, 22 per factorisation, 33 per right-hand side,
// EquationSolver.cpp:141-251), with an iterative solve in place of the factorisation.
//
// Rows: with a DoF map (set_dof_map) each row is a node DoF (node, type), an element DoF or a
// multiplier, as FEStorage numbers them.  Without one, rows are classified from the matrix:
// up to three consecutive rows with identical column lists are one node's DoFs; a row of its
// own with a positive diagonal is a one-DoF node block; every other row (a negative, zero or
// absent diagonal: element pressures, multipliers) is a Schur row.
#include <chrono>
#include <cstring>
#include <string>

#include "../../include/nla3d_gpu.h"
#include "ctx.h"

namespace nla3d_amd {

struct EqSolver {
  Ctx c;
  CsrPattern pat;  // the analysed pattern (1-based upper CSR)
  bool analysed = false, factorized = false, schurOk = false;
  std::vector<uint32_t> mapNode;  // optional DoF map, one entry per row
  std::vector<int> mapDof;
  nla3d_gpu_eqsolver_opts opts{};
  nla3d_gpu_eqsolver_stats stats{};
  DevBuf d_up;  // upper values [nnz]
  DevBuf d_x;   // solution [n]
  std::string err;
};

static double nowMsE() {
  return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// the reference's SparsityInfo invariants after compress(): 1-based, iofeir[0] = 1, rows
// sorted and strictly increasing, upper triangle (SparseMatrix.cpp:87-112)
static void checkPattern(uint32_t n, const uint32_t* iofeir, const uint32_t* columns) {
  if (n == 0) throw ApiError(NLA3D_GPU_EINVAL, "empty matrix");
  if (iofeir[0] != 1) throw ApiError(NLA3D_GPU_EPATTERN, "iofeir[0] must be 1 (1-based CSR)");
  for (uint32_t i = 0; i < n; i++) {
    if (iofeir[i + 1] < iofeir[i]) throw ApiError(NLA3D_GPU_EPATTERN, "iofeir is not nondecreasing");
    for (uint32_t k = iofeir[i] - 1; k < iofeir[i + 1] - 1; k++) {
      const uint32_t c = columns[k];
      if (c < i + 1 || c > n)
        throw ApiError(NLA3D_GPU_EPATTERN, "row " + std::to_string(i + 1) + ": column " + std::to_string(c) +
                                               " outside the upper triangle");
      if (k > iofeir[i] - 1 && columns[k - 1] >= c)
        throw ApiError(NLA3D_GPU_EPATTERN, "row " + std::to_string(i + 1) + ": columns not sorted");
    }
  }
}

static bool samePattern(const CsrPattern& p, uint32_t n, const uint32_t* iofeir, const uint32_t* columns) {
  if (p.nRows != n || p.iofeir.size() != (size_t)n + 1) return false;
  if (std::memcmp(p.iofeir.data(), iofeir, ((size_t)n + 1) * 4) != 0) return false;
  const size_t nnz = (size_t)iofeir[n] - 1;
  return p.columns.size() == nnz && (nnz == 0 || std::memcmp(p.columns.data(), columns, nnz * 4) == 0);
}

static void analyse(EqSolver& S, uint32_t n, const double* values, const uint32_t* iofeir, const uint32_t* columns) {
  checkPattern(n, iofeir, columns);
  CsrPattern& p = S.pat;
  p.nRows = p.nColumns = n;
  p.iofeir.assign(iofeir, iofeir + n + 1);
  p.columns.assign(columns, columns + (iofeir[n] - 1));
  std::vector<int8_t> kind3(n, 1);
  std::vector<uint8_t> rowType(n, 0);
  std::vector<uint32_t> rowNode(n, 0);
  if (!S.mapNode.empty()) {
    if (S.mapNode.size() != n) throw ApiError(NLA3D_GPU_EINVAL, "the DoF map has another row count than the matrix");
    for (uint32_t i = 0; i < n; i++) {
      if (S.mapNode[i] > 0) {
        kind3[i] = 0;
        rowType[i] = (uint8_t)S.mapDof[i];
        rowNode[i] = S.mapNode[i] - 1;
      } else {
        kind3[i] = S.mapDof[i] == 6 ? 1 : 2;  // HYDRO_PRESSURE (Dof.h:15-25) or a multiplier
      }
    }
  } else {
    std::vector<int64_t> rp;
    std::vector<int> col;
    fullSymmetricPattern(p, rp, col, nullptr);
    auto sameCols = [&](uint32_t a, uint32_t b) {
      return rp[a + 1] - rp[a] == rp[b + 1] - rp[b] &&
             std::equal(col.begin() + rp[a], col.begin() + rp[a + 1], col.begin() + rp[b]);
    };
    uint32_t node = 0;
    for (uint32_t i = 0; i < n;) {
      uint32_t sz = 1;
      while (sz < 3 && i + sz < n && sameCols(i, i + sz)) sz++;
      // a row of its own: a one-DoF node block only with a positive diagonal (its first upper entry)
      const bool diagPos = iofeir[i + 1] > iofeir[i] && columns[iofeir[i] - 1] == i + 1 && values[iofeir[i] - 1] > 0.0;
      if (sz > 1 || diagPos) {
        for (uint32_t a = 0; a < sz; a++) {
          kind3[i + a] = 0;
          rowType[i + a] = (uint8_t)a;
          rowNode[i + a] = node;
        }
        node++;
      } else {
        kind3[i] = 1;
      }
      i += sz;
    }
  }
  S.c.setupBlock2(p, kind3, rowType, rowNode);
  S.c.krylovReady = true;
  S.d_up.alloc(std::max<size_t>(p.nnz(), 1) * 8);
  S.d_x.alloc((size_t)n * 8);
  S.analysed = true;
}

static void factorize(EqSolver& S, uint32_t n, const double* values, const uint32_t* iofeir, const uint32_t* columns) {
  if (!values || !iofeir || !columns) throw ApiError(NLA3D_GPU_EINVAL, "null matrix array");
  const double t0 = nowMsE();
  S.factorized = false;
  if (!S.analysed || !samePattern(S.pat, n, iofeir, columns)) analyse(S, n, values, iofeir, columns);
  Ctx& c = S.c;
  const size_t nnz = S.pat.nnz();
  hipCheck(hipMemcpyAsync(S.d_up.p, values, nnz * 8, hipMemcpyHostToDevice, c.stream), "values");
  launch_mirror(c.stream, (int)nnz, S.d_up.as<double>(), nullptr, c.d_mir2.as<int>(), c.d_full2.as<double>());
  c.refreshSpmvCopy(c.stream);
  krylov_precond_setup(c.stream, c.krylovMatrix(), c.precond());
  hipCheck(hipGetLastError(), "factorize");
  S.schurOk = false;
  if (S.opts.preconditioner == 2) {
    std::string why;
    S.schurOk = schurUsable(c, &why) && schurRefresh(c, &why);
  }
  hipCheck(hipStreamSynchronize(c.stream), "factorize sync");
  S.factorized = true;
  S.stats.ms_factorize = nowMsE() - t0;
}

static void substitute(EqSolver& S, const double* rhs, double* x) {
  if (!rhs || !x) throw ApiError(NLA3D_GPU_EINVAL, "null rhs or unknowns");
  if (!S.factorized) throw ApiError(NLA3D_GPU_ESTATE, "call nla3d_gpu_eqsolver_factorize first");
  const double t0 = nowMsE();
  Ctx& c = S.c;
  const int n = c.n2;
  hipCheck(hipMemcpyAsync(c.d_rhs.p, rhs, (size_t)n * 8, hipMemcpyHostToDevice, c.stream), "rhs");
  nla3d_gpu_eqsolver_stats& st = S.stats;
  st.precond_fallbacks = 0;
  st.breakdown = 0;
  if (S.opts.preconditioner == 2 && !S.schurOk) st.precond_fallbacks = 1;
  PrecondHook hook;
  FusedStage fused;
  if (S.schurOk) {
    hook = schurHook(c);
    fused = schurFusedStage(c);
  }
  double relres = 0.0;
  bool brk = false;
  int it = c.minresSolve(c.d_rhs.as<double>(), S.d_x.as<double>(), S.opts.tol, S.opts.max_iter, &relres, &brk,
                         S.schurOk ? &hook : nullptr, 0.0, nullptr, 0.0, S.schurOk ? &fused : nullptr);
  int total = it;
  if (S.schurOk && (brk || !(relres <= S.opts.tol))) {  // the multigrid failed this solve: block Jacobi
    st.precond_fallbacks++;
    brk = false;
    it = c.minresSolve(c.d_rhs.as<double>(), S.d_x.as<double>(), S.opts.tol, S.opts.max_iter, &relres, &brk);
    total += it;
  }
  hipCheck(hipMemcpyAsync(x, S.d_x.p, (size_t)n * 8, hipMemcpyDeviceToHost, c.stream), "unknowns");
  hipCheck(hipStreamSynchronize(c.stream), "substitute sync");
  st.iterations = total;
  st.relres = relres;
  st.breakdown = brk ? 1 : 0;
  st.ms_substitute = nowMsE() - t0;
  if (brk) throw ApiError(NLA3D_GPU_ENOTCONV, "MINRES breakdown (preconditioner not positive definite)");
  if (!(relres <= S.opts.tol))
    throw ApiError(NLA3D_GPU_ENOTCONV, "MINRES did not reach the tolerance in " + std::to_string(it) + " iterations");
}

static int guard(nla3d_gpu_eqsolver_t h, const std::function<void(EqSolver&)>& fn) {
  if (!h) return NLA3D_GPU_EINVAL;
  EqSolver& S = *reinterpret_cast<EqSolver*>(h);
  try {
    S.err.clear();
    hipCheck(hipSetDevice(S.c.device), "hipSetDevice");
    fn(S);
    return NLA3D_GPU_OK;
  } catch (const ApiError& e) {
    S.err = e.what();
    return e.code;
  } catch (const std::exception& e) {
    S.err = e.what();
    return NLA3D_GPU_EINVAL;
  }
}

}  // namespace nla3d_amd

using namespace nla3d_amd;

extern "C" {

void nla3d_gpu_eqsolver_defaults(nla3d_gpu_eqsolver_opts* o) {
  if (!o) return;
  o->tol = 1.0e-12;
  o->max_iter = 20000;
  o->preconditioner = 0;
}

int nla3d_gpu_eqsolver_create(nla3d_gpu_eqsolver_t* s, int device) {
  if (!s) return NLA3D_GPU_EINVAL;
  *s = nullptr;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || device < 0 || device >= n) return NLA3D_GPU_EHIP;
  if (hipSetDevice(device) != hipSuccess) return NLA3D_GPU_EHIP;
  EqSolver* S = new (std::nothrow) EqSolver();
  if (!S) return NLA3D_GPU_EINVAL;
  S->c.device = device;
  nla3d_gpu_eqsolver_defaults(&S->opts);
  if (hipStreamCreateWithFlags(&S->c.stream, hipStreamNonBlocking) != hipSuccess) {
    delete S;
    return NLA3D_GPU_EHIP;
  }
  *s = reinterpret_cast<nla3d_gpu_eqsolver_t>(S);
  return NLA3D_GPU_OK;
}

int nla3d_gpu_eqsolver_destroy(nla3d_gpu_eqsolver_t s) {
  if (!s) return NLA3D_GPU_EINVAL;
  EqSolver* S = reinterpret_cast<EqSolver*>(s);
  (void)hipSetDevice(S->c.device);
  delete S;
  return NLA3D_GPU_OK;
}

const char* nla3d_gpu_eqsolver_last_error(nla3d_gpu_eqsolver_t s) {
  return s ? reinterpret_cast<EqSolver*>(s)->err.c_str() : "null solver handle";
}

int nla3d_gpu_eqsolver_set_options(nla3d_gpu_eqsolver_t s, const nla3d_gpu_eqsolver_opts* opts) {
  return guard(s, [&](EqSolver& S) {
    if (!opts || !(opts->tol > 0.0) || opts->max_iter < 1 || (opts->preconditioner != 0 && opts->preconditioner != 2))
      throw ApiError(NLA3D_GPU_EINVAL, "options: tol > 0, max_iter >= 1, preconditioner 0 or 2");
    const bool pcChanged = opts->preconditioner != S.opts.preconditioner;
    S.opts = *opts;
    if (pcChanged) S.factorized = false;  // the multigrid is refreshed by the next factorize
  });
}

int nla3d_gpu_eqsolver_set_dof_map(nla3d_gpu_eqsolver_t s, uint32_t n, const uint32_t* row_node, const int* row_dof) {
  return guard(s, [&](EqSolver& S) {
    if (n == 0) {  // back to classifying rows from the matrix
      S.mapNode.clear();
      S.mapDof.clear();
    } else {
      if (!row_node || !row_dof) throw ApiError(NLA3D_GPU_EINVAL, "null DoF map");
      for (uint32_t i = 0; i < n; i++)
        if (row_dof[i] < -1 || row_dof[i] > 7 || (row_node[i] > 0 && (row_dof[i] < 0 || row_dof[i] == 6)))
          throw ApiError(NLA3D_GPU_EINVAL, "DoF map row " + std::to_string(i + 1) + ": bad DoF type");
      S.mapNode.assign(row_node, row_node + n);
      S.mapDof.assign(row_dof, row_dof + n);
    }
    S.analysed = S.factorized = false;
  });
}

int nla3d_gpu_eqsolver_factorize(nla3d_gpu_eqsolver_t s, uint32_t n, const double* values, const uint32_t* iofeir,
                                 const uint32_t* columns) {
  return guard(s, [&](EqSolver& S) { factorize(S, n, values, iofeir, columns); });
}

int nla3d_gpu_eqsolver_substitute(nla3d_gpu_eqsolver_t s, const double* rhs, double* unknowns) {
  return guard(s, [&](EqSolver& S) { substitute(S, rhs, unknowns); });
}

int nla3d_gpu_eqsolver_solve(nla3d_gpu_eqsolver_t s, uint32_t n, const double* values, const uint32_t* iofeir,
                             const uint32_t* columns, const double* rhs, double* unknowns) {
  return guard(s, [&](EqSolver& S) {
    factorize(S, n, values, iofeir, columns);
    substitute(S, rhs, unknowns);
  });
}

int nla3d_gpu_eqsolver_get_stats(nla3d_gpu_eqsolver_t s, nla3d_gpu_eqsolver_stats* stats) {
  if (!s || !stats) return NLA3D_GPU_EINVAL;
  EqSolver& S = *reinterpret_cast<EqSolver*>(s);
  *stats = S.stats;
  stats->spmv_dia_nodes = S.analysed ? S.c.krylovMatrix().nDiaFit : 0;
  return NLA3D_GPU_OK;
}

}  // extern "C"
