// Vector kernels of the Newton loop (newton.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>

namespace nla3d_amd {

void newton_apply_bc(hipStream_t s, int nFix, const int* fixEq, const double* fixVal, int nLoad, const int* loadEq,
                     const double* loadVal, double t, double* U, double* R);
void newton_rhs(hipStream_t s, int n2, int nC, const double* F, const double* R, const double* U, const double* Ucprev,
                const int* cptr, const int* cpos, const int* crow, const double* K12, double* rhs);
void newton_update(hipStream_t s, int nC, int nU, int nMpc, const double* dU, double* U, double* Ucprev);
// constrained-DoF reactions into R[0, nC): K11 dUc + K12 dUsl - Fc
void newton_reactions(hipStream_t s, int nC, const int* k11Ptr, const int* k11Col, const int* k11Pos, const double* K11,
                      const int* k12Ptr, const int* k12Col, const double* K12, const double* U, const double* Ucprev,
                      const double* dUsl, const double* F, double* R);
void vec_abssum(hipStream_t s, int n, const double* x, double* partial, double* out);
void newton_lambda_vec(hipStream_t s, int nC, int nU, int nMpc, const double* U, double* v);
void vec_diffsq(hipStream_t s, int n, const double* a, const double* b, double* partial, double* out);
void vec_sumsq(hipStream_t s, int n, const double* x, double* partial, double* out);
// the first solve's starting point at a new load step: x0 = [scale (U_s - Ustart) ; U_l] over the
// block(2) unknowns (free DoFs, then the multipliers, which the solve replaces), then
// Ustart = U[nC, nC + n2); with predict = false only the copy
void newton_predictor(hipStream_t s, int nC, int nU, int nMpc, const double* U, double* Ustart, bool record,
                      double* inc);
// y += x over n entries; y = b - y
void vec_add(hipStream_t s, int n, const double* x, double* y);
// the predictors' pieces: P = [p_u ; 0] and (Q non-null) Q = [0 ; U_l]; then x0 += a P and
// bc -= a Y per basis vector, the first call (Q, R0 non-null) starting from x0 = Q, bc = R0
void newton_pred_split(hipStream_t s, int nC, int nU, int nMpc, const double* p, const double* U, double* P,
                       double* Q);
void newton_pred_axpy(hipStream_t s, int n, double a, const double* P, const double* Y, const double* Q,
                      const double* R0, double* x0, double* bc);
void vec_rsub(hipStream_t s, int n, const double* b, double* y);

}  // namespace nla3d_amd
