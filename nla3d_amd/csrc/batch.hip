// Batched Krylov solve: one persistent launch solves many independent block(2)
// systems.  Each workgroup (a "team") takes whole systems from a queue (an
// atomic counter) and runs preconditioned MINRES on one system to its own
// convergence with block-level reductions only, so no synchronisation crosses
// workgroups; every team leaves the loop when the queue is exhausted.  The
// recurrence is the one of the single-system solver (krylov.hip), with the
// reductions in a fixed order inside the team.
#include <hip/hip_runtime.h>

#include "batch.h"

namespace nla3d_amd {
namespace {

__device__ __forceinline__ double wave_sum(double v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// sum over the team in a fixed order, result returned in every thread
template <int NT>
__device__ __forceinline__ double team_sum(double v, double* red) {
  v = wave_sum(v);
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) red[w] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    double r = 0.0;
    for (int i = 0; i < NT / 64; i++) r += red[i];
    red[NT / 64] = r;
  }
  __syncthreads();
  const double r = red[NT / 64];
  __syncthreads();  // red is reused by the next reduction
  return r;
}

template <int LPG>
__device__ __forceinline__ void team_rowsums(const KrylovMatrix& M, int g, int lane, const double* __restrict__ x,
                                             double& a0, double& a1, double& a2, int& r0, int& sz) {
  const int c0 = M.gColPtr[g], L = M.gColPtr[g + 1] - c0;
  sz = M.gSize[g];
  r0 = M.gRow0[g];
  a0 = a1 = a2 = 0.0;
  const int* __restrict__ cols = M.gCols + c0;
  const double* __restrict__ v = M.val + M.gValPtr[g];
  for (int k = lane; k < L; k += LPG) {
    const double xv = x[cols[k]];
    a0 += v[k] * xv;
    if (sz > 1) a1 += v[L + k] * xv;
    if (sz > 2) a2 += v[2 * L + k] * xv;
  }
  for (int o = LPG / 2; o > 0; o >>= 1) {
    a0 += __shfl_xor(a0, o, LPG);
    a1 += __shfl_xor(a1, o, LPG);
    a2 += __shfl_xor(a2, o, LPG);
  }
}

// z = P r over all block-Jacobi groups; returns this thread's part of r . z
template <int NT>
__device__ __forceinline__ double team_precond(const Precond& P, const double* __restrict__ r, double* __restrict__ z) {
  double part = 0.0;
  for (int g = threadIdx.x; g < P.nGroups; g += NT) {
    const int r0 = P.gStart[g], size = P.gInfo[g] & 3;
    const double* B = P.binv + (size_t)g * 9;
    double t[3] = {0, 0, 0};
    for (int i = 0; i < size; i++) t[i] = r[r0 + i];
    for (int i = 0; i < size; i++) {
      double s = 0.0;
      for (int j = 0; j < size; j++) s += B[i * 3 + j] * t[j];
      z[r0 + i] = s;
      part += t[i] * s;
    }
  }
  return part;
}

template <int NT, int LPG>
__device__ void team_minres(const TeamSolve& ts, double* red) {
  const KrylovMatrix& M = ts.M;
  const int n = M.n;
  double* R1 = ts.v[0];
  double* R2 = ts.v[1];
  double* T = ts.v[2];
  double* Yp = ts.v[3];
  double* Z = ts.v[4];
  double* Wa = ts.v[5];  // w
  double* Wb = ts.v[6];  // w2
  double* Wc = ts.v[7];  // w1
  double* X = ts.x;
  for (int i = threadIdx.x; i < n; i += NT) {
    const double bi = ts.b[i];
    R1[i] = bi;
    R2[i] = bi;
    X[i] = 0.0;
    Wa[i] = Wb[i] = Wc[i] = 0.0;
  }
  __syncthreads();
  const double bb = team_sum<NT>(team_precond<NT>(ts.P, R1, Yp), red);
  int itn = 0, flag = 0;
  double beta1 = bb > 0.0 ? sqrt(bb) : 0.0, phibar = beta1;
  if (!(bb > 0.0)) flag = bb < 0.0 ? 2 : 1;  // indefinite preconditioner / zero rhs
  double beta = beta1, oldb = 0.0, dbar = 0.0, epsln = 0.0, cs = -1.0, sn = 0.0;
  const int groupsPerPass = NT / LPG;
  const int lane = threadIdx.x % LPG;
  while (!flag && itn < ts.maxit) {
    itn++;
    const double s = 1.0 / beta;
    const double c1 = itn >= 2 ? beta / oldb : 0.0;
    // T = s K Yp - c1 R1 ; alfa = s Yp . T
    double part = 0.0;
    for (int g = threadIdx.x / LPG; g < M.nGroups; g += groupsPerPass) {
      double a0, a1, a2;
      int r0, sz;
      team_rowsums<LPG>(M, g, lane, Yp, a0, a1, a2, r0, sz);
      if (lane < sz) {
        const int row = r0 + lane;
        const double t = s * (lane == 0 ? a0 : (lane == 1 ? a1 : a2)) - c1 * R1[row];
        T[row] = t;
        part += s * Yp[row] * t;
      }
    }
    __syncthreads();
    const double alfa = team_sum<NT>(part, red);
    const double c2 = alfa / beta;
    for (int i = threadIdx.x; i < n; i += NT) T[i] -= c2 * R2[i];
    // R1 <- R2, R2 <- T, T <- old R1
    double* tmp = R1;
    R1 = R2;
    R2 = T;
    T = tmp;
    __syncthreads();
    const double bn = team_sum<NT>(team_precond<NT>(ts.P, R2, Z), red);
    if (bn < 0.0) {
      flag = 2;
      break;
    }
    oldb = beta;
    beta = sqrt(bn);
    const double oldeps = epsln;
    const double delta = cs * dbar + sn * alfa;
    const double gbar = sn * dbar - cs * alfa;
    epsln = sn * beta;
    dbar = -cs * beta;
    const double gamma = fmax(sqrt(gbar * gbar + beta * beta), 1e-300);
    cs = gbar / gamma;
    sn = beta / gamma;
    const double phi = cs * phibar;
    phibar = sn * phibar;
    // w1 <- w2, w2 <- w, w <- (v - oldeps w1 - delta w2) / gamma ; x += phi w   (v = s Yp)
    const double ig = 1.0 / gamma;
    for (int i = threadIdx.x; i < n; i += NT) {
      const double wn = (s * Yp[i] - oldeps * Wb[i] - delta * Wa[i]) * ig;
      Wc[i] = wn;
      X[i] += phi * wn;
    }
    tmp = Wc;
    Wc = Wb;
    Wb = Wa;
    Wa = tmp;
    tmp = Yp;
    Yp = Z;
    Z = tmp;
    __syncthreads();
    if (phibar <= ts.tol * beta1 || beta == 0.0) break;
  }
  if (threadIdx.x == 0) {
    ts.out[0] = itn;
    ts.out[1] = flag;
    ts.relres[0] = beta1 > 0.0 ? phibar / beta1 : 0.0;
  }
  __syncthreads();
}

template <int NT, int LPG>
__global__ __launch_bounds__(NT) void minres_team_kernel(int nQ, const TeamSolve* __restrict__ solves,
                                                       const int* __restrict__ queue, int* __restrict__ counter) {
  __shared__ double red[NT / 64 + 1];
  __shared__ int sq;
  for (;;) {
    __syncthreads();
    if (threadIdx.x == 0) sq = atomicAdd(counter, 1);
    __syncthreads();
    const int q = sq;
    if (q >= nQ) break;
    team_minres<NT, LPG>(solves[queue[q]], red);
  }
}

__global__ void zero_int_kernel(int* p) { *p = 0; }

template <int NT, int LPG>
void launchT(hipStream_t s, int grid, int nQ, const TeamSolve* solves, const int* queue, int* counter) {
  hipLaunchKernelGGL((minres_team_kernel<NT, LPG>), dim3(grid), dim3(NT), 0, s, nQ, solves, queue, counter);
}

}  // namespace

bool team_config_valid(int threads, int lpg) {
  return (threads == 256 || threads == 512 || threads == 1024) && (lpg == 4 || lpg == 8 || lpg == 16);
}

void launch_minres_team(hipStream_t s, int threads, int lpg, int grid, int nQ, const TeamSolve* solves,
                        const int* queue, int* counter) {
  if (nQ <= 0) return;
  hipLaunchKernelGGL(zero_int_kernel, dim3(1), dim3(1), 0, s, counter);
#define NLA3D_TEAM(NT, L) \
  if (threads == NT && lpg == L) return launchT<NT, L>(s, grid, nQ, solves, queue, counter);
  NLA3D_TEAM(256, 4) NLA3D_TEAM(256, 8) NLA3D_TEAM(256, 16)
  NLA3D_TEAM(512, 4) NLA3D_TEAM(512, 8) NLA3D_TEAM(512, 16)
  NLA3D_TEAM(1024, 4) NLA3D_TEAM(1024, 8) NLA3D_TEAM(1024, 16)
#undef NLA3D_TEAM
}

}  // namespace nla3d_amd
