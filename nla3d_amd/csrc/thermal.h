// QUADTH / SurfaceLINETH kernels and the CG solve (thermal.hip).
#pragma once

#include <hip/hip_runtime.h>

#include "krylov.h"

namespace nla3d_amd {

constexpr int kQuadthSlots = 24;  // Ke 10, Ce 10, Fe 4
constexpr int kLinethSlots = 5;   // Ke 3, Fe 2

struct CgState {
  double rz, alpha, beta, rnorm, bnorm;
  int itn, converged, breakdown, skip;
};

struct CgWork {
  double *r, *z, *p, *q;
  double* p2;  // second search-direction buffer (scalar path: p is rebuilt into it every iteration)
  double *partial, *partial2;
  CgState* st;
  CgState* st0;  // an always-active state, for the initial residual product
};

// Scalar-row matrices (one DoF per node, every row its own group): the matrix in sliced
// ELLPACK over 64-row slices (FP64 values; columns as 16-bit offsets from the slice's
// base when every slice spans < 65,536 columns, else full indices; padding: value 0 at
// the base column) and the Jacobi inverse diagonal.  val is refreshed from the grouped
// copy through pos before each solve.
struct CgSell {
  int n = 0, nSlices = 0, nSlots = 0;
  const int* off = nullptr;
  const int* base = nullptr;
  const unsigned short* col16 = nullptr;
  const int* col = nullptr;
  const int* pos = nullptr;  // [nSlots] index into the grouped values, -1 padding
  double* val = nullptr;
  double* dinv = nullptr;
};

void launch_quadth(hipStream_t s, int n, const int* conn, const double* xyz, const double* par, double* out);
void launch_lineth(hipStream_t s, int n, const int* conn, const double* xyz, const double* par, double* out);
void launch_combine_mirror(hipStream_t s, int n, double a, const double* A, double b, const double* B, const int* mir,
                           double* full);
// preconditioned CG on M x = b from the x passed in; returns iterations (negative on HIP error).
// avgCols: mean entries per row of the full matrix (picks the lanes per row group of the SpMV)
int cg_solve(hipStream_t s, const KrylovMatrix& M, const Precond& P, CgWork& w, const double* b, double* x, double tol,
             int maxit, CgState* hstate, double avgCols, const CgSell* sell = nullptr);
void launch_tr_comb(hipStream_t s, int n, int off, double a, const double* U, double b, const double* DU, double c,
                    const double* DDU, double* v);
void launch_tr_rhs(hipStream_t s, int n2, int nC, const double* F, const double* R, const double* y, const double* U,
                   const int* cptr, const int* cpos, const int* crow, const double* K12, int coupling, double* rhs);
void launch_tr_update(hipStream_t s, int nAll, int nC, const double* x, double a0, double a2, double a3, double a6,
                      double a7, double* U, double* DU, double* DDU);
void launch_fill(hipStream_t s, int n, double v, double* x);
void launch_copy(hipStream_t s, int n, const double* a, double* b);

}  // namespace nla3d_amd
