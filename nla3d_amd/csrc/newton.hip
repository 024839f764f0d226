// Small vector kernels of the Newton-Raphson loop (FESolver.cpp:441-490):
// boundary conditions, the block(2) right-hand side with the eliminated
// constrained DoFs, the increment update and the convergence measures.
#include "newton.h"

namespace nla3d_amd {
namespace {

__device__ __forceinline__ double wsum(double v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

__device__ __forceinline__ double bsum256(double v, double* sh) {
  v = wsum(v);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
  __syncthreads();
  double r = 0.0;
  if (threadIdx.x == 0)
    for (int i = 0; i < 4; i++) r += sh[i];
  return r;
}

// vecUc[eq] = fix * t (FESolver.cpp:241-246); fixes are distinct DoFs
__global__ void fix_kernel(int n, const int* __restrict__ eq, const double* __restrict__ val, double t,
                           double* __restrict__ U) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) U[eq[i]] = val[i] * t;
}

// vecR[eq] += load * t in list order (FESolver.cpp:235-237)
__global__ void load_kernel(int n, const int* __restrict__ eq, const double* __restrict__ val, double t,
                            double* __restrict__ R) {
  if (blockIdx.x == 0 && threadIdx.x == 0)
    for (int i = 0; i < n; i++) R[eq[i]] += val[i] * t;
}

// rhs_j = F_sl + R_sl - (K12^T dUc)_j, dUc = Uc - Ucprev (FESolver.cpp:449-454)
__global__ __launch_bounds__(256) void rhs_kernel(int n2, int nC, const double* __restrict__ F,
                                                 const double* __restrict__ R, const double* __restrict__ U,
                                                 const double* __restrict__ Ucprev, const int* __restrict__ cptr,
                                                 const int* __restrict__ cpos, const int* __restrict__ crow,
                                                 const double* __restrict__ K12, double* __restrict__ rhs) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n2) return;
  double v = F[nC + j] + R[nC + j];
  double acc = 0.0;
  for (int k = cptr[j]; k < cptr[j + 1]; k++) {
    const int i = crow[k];
    acc += K12[cpos[k]] * (U[i] - Ucprev[i]);
  }
  rhs[j] = v - acc;
}

// vecRc = K11 dUc + K12 dUsl - Fc (FESolver.cpp:464-467), dUc = Uc - Ucprev; K11 through
// full-symmetric row lists (positions into its upper values), K12 by rows.  One 64-lane wave per
// constrained row: lane l sums the row's entries l, l + 64, ... and a butterfly adds the lanes (a
// rigid-body master's row holds every slave row of its MPC, tens of thousands of entries)
__global__ __launch_bounds__(256) void reaction_kernel(int nC, const int* __restrict__ k11Ptr,
                                                      const int* __restrict__ k11Col, const int* __restrict__ k11Pos,
                                                      const double* __restrict__ K11, const int* __restrict__ k12Ptr,
                                                      const int* __restrict__ k12Col, const double* __restrict__ K12,
                                                      const double* __restrict__ U, const double* __restrict__ Ucprev,
                                                      const double* __restrict__ dUsl, const double* __restrict__ F,
                                                      double* __restrict__ R) {
  const int i = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
  if (i >= nC) return;  // the 64 lanes of a row leave together
  double r = 0.0;
  for (int k = k11Ptr[i] + lane; k < k11Ptr[i + 1]; k += 64) {
    const int j = k11Col[k];
    r += K11[k11Pos[k]] * (U[j] - Ucprev[j]);
  }
  for (int k = k12Ptr[i] + lane; k < k12Ptr[i + 1]; k += 64) r += K12[k] * dUsl[k12Col[k]];
  r = wsum(r);
  if (lane == 0) R[i] = r - F[i];
}

// vecUs += dUs ; vecUl = dUl ; Ucprev = Uc (FESolver.cpp:460-469)
__global__ __launch_bounds__(256) void update_kernel(int nC, int nU, int nMpc, const double* __restrict__ dU,
                                                    double* __restrict__ U, double* __restrict__ Ucprev) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < nU) U[nC + i] += dU[i];
  else if (i < nU + nMpc) U[nC + i] = dU[i];
  if (i < nC) Ucprev[i] = U[i];
}

__global__ __launch_bounds__(256) void abssum_partial_kernel(int n, const double* __restrict__ x,
                                                            double* __restrict__ partial) {
  __shared__ double sh[4];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const double s = bsum256(i < n ? fabs(x[i]) : 0.0, sh);
  if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

__global__ __launch_bounds__(256) void sumsq_partial_kernel(int n, const double* __restrict__ x,
                                                           double* __restrict__ partial) {
  __shared__ double sh[4];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const double v = i < n ? x[i] : 0.0;
  const double s = bsum256(v * v, sh);
  if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

__global__ __launch_bounds__(256) void finish_kernel(const double* __restrict__ partial, int n,
                                                    double* __restrict__ out) {
  __shared__ double sh[4];
  double v = 0.0;
  for (int i = threadIdx.x; i < n; i += 256) v += partial[i];
  const double s = bsum256(v, sh);
  if (threadIdx.x == 0) *out = s;
}

// v = [0 (nU) ; U_l (nMpc)]
__global__ __launch_bounds__(256) void lambda_vec_kernel(int nC, int nU, int nMpc, const double* __restrict__ U,
                                                        double* __restrict__ v) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < nU) v[i] = 0.0;
  else if (i < nU + nMpc) v[i] = U[nC + i];
}

// partial sums of (a - b)^2
__global__ __launch_bounds__(256) void diffsq_partial_kernel(int n, const double* __restrict__ a,
                                                            const double* __restrict__ b,
                                                            double* __restrict__ partial) {
  __shared__ double sh[4];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const double v = i < n ? a[i] - b[i] : 0.0;
  const double s = bsum256(v * v, sh);
  if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

}  // namespace

static inline int cdiv(int a, int b) { return (a + b - 1) / b; }

void newton_apply_bc(hipStream_t s, int nFix, const int* fixEq, const double* fixVal, int nLoad, const int* loadEq,
                     const double* loadVal, double t, double* U, double* R) {
  if (nFix > 0) hipLaunchKernelGGL(fix_kernel, dim3(cdiv(nFix, 256)), dim3(256), 0, s, nFix, fixEq, fixVal, t, U);
  if (nLoad > 0) hipLaunchKernelGGL(load_kernel, dim3(1), dim3(64), 0, s, nLoad, loadEq, loadVal, t, R);
}

void newton_rhs(hipStream_t s, int n2, int nC, const double* F, const double* R, const double* U, const double* Ucprev,
                const int* cptr, const int* cpos, const int* crow, const double* K12, double* rhs) {
  hipLaunchKernelGGL(rhs_kernel, dim3(cdiv(n2, 256)), dim3(256), 0, s, n2, nC, F, R, U, Ucprev, cptr, cpos, crow, K12,
                     rhs);
}

void newton_reactions(hipStream_t s, int nC, const int* k11Ptr, const int* k11Col, const int* k11Pos, const double* K11,
                      const int* k12Ptr, const int* k12Col, const double* K12, const double* U, const double* Ucprev,
                      const double* dUsl, const double* F, double* R) {
  if (nC <= 0) return;
  hipLaunchKernelGGL(reaction_kernel, dim3(cdiv(nC, 4)), dim3(256), 0, s, nC, k11Ptr, k11Col, k11Pos, K11, k12Ptr,
                     k12Col, K12, U, Ucprev, dUsl, F, R);
}

void newton_update(hipStream_t s, int nC, int nU, int nMpc, const double* dU, double* U, double* Ucprev) {
  const int n = std::max(nU + nMpc, nC);
  hipLaunchKernelGGL(update_kernel, dim3(cdiv(std::max(n, 1), 256)), dim3(256), 0, s, nC, nU, nMpc, dU, U, Ucprev);
}

void vec_abssum(hipStream_t s, int n, const double* x, double* partial, double* out) {
  const int nb = cdiv(std::max(n, 1), 256);
  hipLaunchKernelGGL(abssum_partial_kernel, dim3(nb), dim3(256), 0, s, n, x, partial);
  hipLaunchKernelGGL(finish_kernel, dim3(1), dim3(256), 0, s, partial, nb, out);
}

void newton_lambda_vec(hipStream_t s, int nC, int nU, int nMpc, const double* U, double* v) {
  hipLaunchKernelGGL(lambda_vec_kernel, dim3(cdiv(std::max(nU + nMpc, 1), 256)), dim3(256), 0, s, nC, nU, nMpc, U, v);
}

// the load step's start: inc = U_sl - Ustart on the node and pressure rows (0 on the
// multipliers) when record, then Ustart = U_sl
__global__ __launch_bounds__(256) void predictor_kernel(int nC, int nU, int nMpc, const double* __restrict__ U,
                                                       double* __restrict__ Ustart, int record,
                                                       double* __restrict__ inc) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nU + nMpc) return;
  const double u = U[nC + i];
  if (record) inc[i] = i < nU ? u - Ustart[i] : 0.0;
  Ustart[i] = u;
}

// predictor basis: P = [p_u ; 0] and, with Q, Q = [0 ; U_l] (the multipliers' current values)
__global__ __launch_bounds__(256) void pred_split_kernel(int nC, int nU, int nMpc, const double* __restrict__ p,
                                                        const double* __restrict__ U, double* __restrict__ P,
                                                        double* __restrict__ Q) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nU + nMpc) return;
  P[i] = i < nU ? p[i] : 0.0;
  if (Q) Q[i] = i < nU ? 0.0 : U[nC + i];
}

// x0 += a P ; bc -= a Y (one basis vector of the predictor at a time; the first call starts
// from x0 = Q, bc = R0)
__global__ __launch_bounds__(256) void pred_axpy_kernel(int n, double a, const double* __restrict__ P,
                                                       const double* __restrict__ Y, const double* __restrict__ Q,
                                                       const double* __restrict__ R0, double* __restrict__ x0,
                                                       double* __restrict__ bc) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  x0[i] = (Q ? Q[i] : x0[i]) + a * P[i];
  bc[i] = (R0 ? R0[i] : bc[i]) - a * Y[i];
}

__global__ __launch_bounds__(256) void add_kernel(int n, const double* __restrict__ x, double* __restrict__ y) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] += x[i];
}

void newton_predictor(hipStream_t s, int nC, int nU, int nMpc, const double* U, double* Ustart, bool record,
                      double* inc) {
  hipLaunchKernelGGL(predictor_kernel, dim3(cdiv(std::max(nU + nMpc, 1), 256)), dim3(256), 0, s, nC, nU, nMpc, U,
                     Ustart, record ? 1 : 0, inc);
}

__global__ __launch_bounds__(256) void rsub_kernel(int n, const double* __restrict__ b, double* __restrict__ y) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = b[i] - y[i];
}

void vec_rsub(hipStream_t s, int n, const double* b, double* y) {
  hipLaunchKernelGGL(rsub_kernel, dim3(cdiv(std::max(n, 1), 256)), dim3(256), 0, s, n, b, y);
}

void vec_add(hipStream_t s, int n, const double* x, double* y) {
  hipLaunchKernelGGL(add_kernel, dim3(cdiv(std::max(n, 1), 256)), dim3(256), 0, s, n, x, y);
}

void vec_diffsq(hipStream_t s, int n, const double* a, const double* b, double* partial, double* out) {
  const int nb = cdiv(std::max(n, 1), 256);
  hipLaunchKernelGGL(diffsq_partial_kernel, dim3(nb), dim3(256), 0, s, n, a, b, partial);
  hipLaunchKernelGGL(finish_kernel, dim3(1), dim3(256), 0, s, partial, nb, out);
}

void vec_sumsq(hipStream_t s, int n, const double* x, double* partial, double* out) {
  const int nb = cdiv(std::max(n, 1), 256);
  hipLaunchKernelGGL(sumsq_partial_kernel, dim3(nb), dim3(256), 0, s, n, x, partial);
  hipLaunchKernelGGL(finish_kernel, dim3(1), dim3(256), 0, s, partial, nb, out);
}

}  // namespace nla3d_amd

namespace nla3d_amd {

void newton_pred_split(hipStream_t s, int nC, int nU, int nMpc, const double* p, const double* U, double* P,
                       double* Q) {
  hipLaunchKernelGGL(pred_split_kernel, dim3(cdiv(std::max(nU + nMpc, 1), 256)), dim3(256), 0, s, nC, nU, nMpc, p, U,
                     P, Q);
}

void newton_pred_axpy(hipStream_t s, int n, double a, const double* P, const double* Y, const double* Q,
                      const double* R0, double* x0, double* bc) {
  hipLaunchKernelGGL(pred_axpy_kernel, dim3(cdiv(std::max(n, 1), 256)), dim3(256), 0, s, n, a, P, Y, Q, R0, x0, bc);
}

}  // namespace nla3d_amd
