// QUADTH / SurfaceLINETH element kernels, preconditioned CG, and the vector
// kernels of the linear and transient thermal solution procedures.
//
// Elements: ElementQUADTH::buildK / buildC (elements/QUADTH.cpp:22-67) and
// SurfaceLINETH::buildK (QUADTH.cpp:94-125), one thread per element, FP64.
// Outputs per QUADTH element (24 doubles): Ke packed upper 4x4 (10), Ce packed
// upper 4x4 (10), Fe (4); per SurfaceLINETH element (5): Ke packed upper 2x2
// (3), Fe (2).  The deterministic scatter of assemble.hip sums them into K, C
// and F.
//
// CG: Hestenes-Stiefel with the block-Jacobi preconditioner of krylov.hip (for
// TEMP DoFs the blocks are 1x1, i.e. Jacobi).  As for MINRES, the scalar
// recurrences run on the device and inner products are reduced in a fixed
// order.
#include <algorithm>

#include "internal.h"
#include "krylov.h"
#include "thermal.h"

namespace nla3d_amd {
namespace {

constexpr double kG = 0.5773502691896258;  // isoparametric.h:49-52, 82-87

__global__ __launch_bounds__(256) void quadth_kernel(int n, const int* __restrict__ conn,
                                                    const double* __restrict__ xyz,
                                                    const double* __restrict__ par, double* __restrict__ out) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  double X[4][2];
  for (int i = 0; i < 4; i++) {
    const int nd = conn[(size_t)e * 4 + i];
    X[i][0] = xyz[(size_t)nd * 3 + 0];
    X[i][1] = xyz[(size_t)nd * 3 + 1];
  }
  const double k = par[(size_t)e * 4 + 0], rho = par[(size_t)e * 4 + 1], c = par[(size_t)e * 4 + 2],
               q = par[(size_t)e * 4 + 3];
  double Ke[10] = {0}, Ce[10] = {0}, Fe[4] = {0};
  for (int g = 0; g < 4; g++) {
    const double r = (g & 1) ? kG : -kG, s = (g & 2) ? kG : -kG;
    // formFuncDeriv (isoparametric.cpp:130-135) and the Jacobian (isoparametric.cpp:55-106)
    const double dN[4][2] = {{-0.25 * (1.0 - s), -0.25 * (1.0 - r)},
                             {+0.25 * (1.0 - s), -0.25 * (1.0 + r)},
                             {+0.25 * (1.0 + s), +0.25 * (1.0 + r)},
                             {-0.25 * (1.0 + s), +0.25 * (1.0 - r)}};
    double J00 = 0, J01 = 0, J10 = 0, J11 = 0;
    for (int i = 0; i < 4; i++) {
      J00 += dN[i][0] * X[i][0];
      J01 += dN[i][0] * X[i][1];
      J10 += dN[i][1] * X[i][0];
      J11 += dN[i][1] * X[i][1];
    }
    const double det = J00 * J11 - J01 * J10;
    const double rd = 1.0 / det;
    double N[4][2];
    for (int i = 0; i < 4; i++) {
      N[i][0] = (J11 * dN[i][0] - J01 * dN[i][1]) * rd;
      N[i][1] = (-J10 * dN[i][0] + J00 * dN[i][1]) * rd;
    }
    const double w = det;  // Gauss weight 1
    const double ff[4] = {0.25 * (1.0 - r) * (1.0 - s), 0.25 * (1.0 + r) * (1.0 - s), 0.25 * (1.0 + r) * (1.0 + s),
                          0.25 * (1.0 - r) * (1.0 + s)};
    int idx = 0;
    for (int i = 0; i < 4; i++)
      for (int j = i; j < 4; j++) {
        Ke[idx] += k * w * (N[i][0] * N[j][0] + N[i][1] * N[j][1]);  // B^T (k I) B
        Ce[idx] += ff[i] * ff[j] * rho * c * w;                        // consistent capacity
        idx++;
      }
    if (q != 0.0)
      for (int i = 0; i < 4; i++) Fe[i] += ff[i] * (q * w);
  }
  double* o = out + (size_t)e * kQuadthSlots;
  for (int i = 0; i < 10; i++) o[i] = Ke[i];
  for (int i = 0; i < 10; i++) o[10 + i] = Ce[i];
  for (int i = 0; i < 4; i++) o[20 + i] = Fe[i];
}

__global__ __launch_bounds__(256) void lineth_kernel(int n, const int* __restrict__ conn,
                                                    const double* __restrict__ xyz,
                                                    const double* __restrict__ par, double* __restrict__ out) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  const int a = conn[(size_t)e * 2], b = conn[(size_t)e * 2 + 1];
  const double dx = xyz[(size_t)b * 3] - xyz[(size_t)a * 3], dy = xyz[(size_t)b * 3 + 1] - xyz[(size_t)a * 3 + 1],
               dz = xyz[(size_t)b * 3 + 2] - xyz[(size_t)a * 3 + 2];
  const double det = sqrt(dx * dx + dy * dy + dz * dz) * 0.5;  // isoparametric.cpp:13-27
  const double flux = par[(size_t)e * 4 + 0], htc = par[(size_t)e * 4 + 1];
  const double e0 = par[(size_t)e * 4 + 2];
  const double e1 = par[(size_t)e * 4 + 3] == 0.0 ? e0 : par[(size_t)e * 4 + 3];  // QUADTH.cpp:87-90
  double K0 = 0, K1 = 0, K2 = 0, F0 = 0, F1 = 0;
  for (int g = 0; g < 2; g++) {
    const double r = g ? kG : -kG;
    const double N0 = 0.5 * (1.0 - r), N1 = 0.5 * (1.0 + r);
    if (flux != 0.0) {
      F0 += N0 * (flux * det);
      F1 += N1 * (flux * det);
    }
    if (htc != 0.0) {
      const double t00 = N0 * N0 * htc * det, t01 = N0 * N1 * htc * det, t11 = N1 * N1 * htc * det;
      K0 += t00;
      K1 += t01;
      K2 += t11;
      F0 += t00 * e0 + t01 * e1;
      F1 += t01 * e0 + t11 * e1;
    }
  }
  double* o = out + (size_t)e * kLinethSlots;
  o[0] = K0;
  o[1] = K1;
  o[2] = K2;
  o[3] = F0;
  o[4] = F1;
}

// full[mir[2k]] = full[mir[2k+1]] = a * A[k] + b * B[k]  (B may be null)
__global__ __launch_bounds__(256) void combine_mirror_kernel(int n, double a, const double* __restrict__ A, double b,
                                                            const double* __restrict__ B, const int* __restrict__ mir,
                                                            double* __restrict__ full) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const double v = a * A[k] + (B ? b * B[k] : 0.0);
  full[mir[2 * k]] = v;
  full[mir[2 * k + 1]] = v;
}

__device__ __forceinline__ double wsum(double v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}
__device__ __forceinline__ double bsum(double v, double* sh) {
  v = wsum(v);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
  __syncthreads();
  double r = 0.0;
  if (threadIdx.x == 0)
    for (int i = 0; i < 4; i++) r += sh[i];
  return r;
}
__device__ double finish(const double* partial, int n, double* sh) {
  double v = 0.0;
  for (int i = threadIdx.x; i < n; i += 256) v += partial[i];
  return bsum(v, sh);
}

// ---- CG ------------------------------------------------------------------
// q = A p ; partial p.q
template <int LPG>
__global__ __launch_bounds__(256) void cg_spmv_kernel(KrylovMatrix M, const double* __restrict__ p,
                                                     double* __restrict__ q, const CgState* __restrict__ st,
                                                     double* __restrict__ partial) {
  __shared__ double sh[4];
  if (st->converged) return;
  const int lane = threadIdx.x % LPG, gpb = blockDim.x / LPG;
  double contrib = 0.0;
  for (int g = blockIdx.x * gpb + threadIdx.x / LPG; g < M.nGroups; g += gridDim.x * gpb) {
    const int c0 = M.gColPtr[g], L = M.gColPtr[g + 1] - c0, vb = M.gValPtr[g], sz = M.gSize[g], r0 = M.gRow0[g];
    double a0 = 0, a1 = 0, a2 = 0;
    for (int k = lane; k < L; k += LPG) {
      const double xv = p[M.gCols[c0 + k]];
      a0 += M.val[vb + k] * xv;
      if (sz > 1) a1 += M.val[vb + L + k] * xv;
      if (sz > 2) a2 += M.val[vb + 2 * L + k] * xv;
    }
    for (int o = LPG / 2; o > 0; o >>= 1) {
      a0 += __shfl_xor(a0, o, LPG);
      a1 += __shfl_xor(a1, o, LPG);
      a2 += __shfl_xor(a2, o, LPG);
    }
    if (lane < sz) {
      const double v = lane == 0 ? a0 : (lane == 1 ? a1 : a2);
      q[r0 + lane] = v;
      contrib += p[r0 + lane] * v;
    }
  }
  const double s = bsum(contrib, sh);
  if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

__global__ __launch_bounds__(256) void cg_alpha_kernel(CgState* st, const double* partial, int n) {
  __shared__ double sh[4];
  if (st->converged) return;
  const double pq = finish(partial, n, sh);
  if (threadIdx.x == 0) st->alpha = pq != 0.0 ? st->rz / pq : 0.0;
}

// x += alpha p ; r -= alpha q ; z = M r ; partial r.z   (one thread per preconditioner group)
__global__ __launch_bounds__(256) void cg_update_kernel(int nGroups, const int* __restrict__ gStart,
                                                       const int* __restrict__ gInfo, const double* __restrict__ binv,
                                                       double* __restrict__ x, double* __restrict__ r,
                                                       const double* __restrict__ p, const double* __restrict__ q,
                                                       double* __restrict__ z, const CgState* __restrict__ st,
                                                       double* __restrict__ partial) {
  __shared__ double sh[4];
  if (st->converged) return;
  const double a = st->alpha;
  double contrib = 0.0;
  for (int g = blockIdx.x * blockDim.x + threadIdx.x; g < nGroups; g += gridDim.x * blockDim.x) {
    const int r0 = gStart[g], size = gInfo[g] & 3;
    double t[3];
    for (int i = 0; i < size; i++) {
      x[r0 + i] += a * p[r0 + i];
      t[i] = r[r0 + i] - a * q[r0 + i];
      r[r0 + i] = t[i];
    }
    const double* B = binv + (size_t)g * 9;
    for (int i = 0; i < size; i++) {
      double zz = 0.0;
      for (int j = 0; j < size; j++) zz += B[i * 3 + j] * t[j];
      z[r0 + i] = zz;
      contrib += t[i] * zz;
    }
  }
  const double s = bsum(contrib, sh);
  if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

__global__ __launch_bounds__(256) void cg_beta_kernel(CgState* st, const double* partial, int n, double tol) {
  __shared__ double sh[4];
  if (st->converged) {
    if (threadIdx.x == 0) st->skip = 1;
    return;
  }
  const double rz = finish(partial, n, sh);
  if (threadIdx.x != 0) return;
  st->itn += 1;
  st->beta = st->rz != 0.0 ? rz / st->rz : 0.0;
  st->rz = rz;
  st->rnorm = rz > 0.0 ? sqrt(rz) : 0.0;
  st->skip = 0;
  if (!(rz == rz) || rz < 0.0) {
    st->breakdown = 1;
    st->converged = 1;
  } else if (st->rnorm <= tol * st->bnorm) {
    st->converged = 1;
  }
}

// p = z + beta p
__global__ __launch_bounds__(256) void cg_dir_kernel(int n, const double* __restrict__ z, double* __restrict__ p,
                                                    const CgState* __restrict__ st) {
  if (st->skip || st->converged) return;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = z[i] + st->beta * p[i];
}

// start: r = b - A x (Ax precomputed in q), z = M r, p = z; partials r.z and b.Mb
__global__ __launch_bounds__(256) void cg_start_kernel(int nGroups, const int* __restrict__ gStart,
                                                      const int* __restrict__ gInfo, const double* __restrict__ binv,
                                                      const double* __restrict__ b, const double* __restrict__ Ax,
                                                      double* __restrict__ r, double* __restrict__ z,
                                                      double* __restrict__ p, double* __restrict__ partial,
                                                      double* __restrict__ partial2) {
  __shared__ double sh[4];
  double c1 = 0.0, c2 = 0.0;
  for (int g = blockIdx.x * blockDim.x + threadIdx.x; g < nGroups; g += gridDim.x * blockDim.x) {
    const int r0 = gStart[g], size = gInfo[g] & 3;
    double t[3], bb[3];
    for (int i = 0; i < size; i++) {
      bb[i] = b[r0 + i];
      t[i] = bb[i] - Ax[r0 + i];
      r[r0 + i] = t[i];
    }
    const double* B = binv + (size_t)g * 9;
    for (int i = 0; i < size; i++) {
      double zz = 0.0, zb = 0.0;
      for (int j = 0; j < size; j++) {
        zz += B[i * 3 + j] * t[j];
        zb += B[i * 3 + j] * bb[j];
      }
      z[r0 + i] = zz;
      p[r0 + i] = zz;
      c1 += t[i] * zz;
      c2 += bb[i] * zb;
    }
  }
  const double s1 = bsum(c1, sh);
  __syncthreads();
  const double s2 = bsum(c2, sh);
  if (threadIdx.x == 0) {
    partial[blockIdx.x] = s1;
    partial2[blockIdx.x] = s2;
  }
}

__global__ __launch_bounds__(256) void cg_init_kernel(CgState* st, const double* partial, const double* partial2,
                                                     int n, double tol) {
  __shared__ double sh[4];
  const double rz = finish(partial, n, sh);
  __syncthreads();
  const double bmb = finish(partial2, n, sh);
  if (threadIdx.x != 0) return;
  CgState z{};
  z.rz = rz;
  z.rnorm = rz > 0.0 ? sqrt(rz) : 0.0;
  z.bnorm = bmb > 0.0 ? sqrt(bmb) : 0.0;
  z.converged = (z.rnorm <= tol * z.bnorm) ? 1 : 0;
  if (!(rz == rz) || rz < 0.0) z.breakdown = z.converged = 1;
  *st = z;
}

// ---- transient vectors ----------------------------------------------------
// v = a1 U + a4 DU + a5 DDU on the s part
__global__ __launch_bounds__(256) void tr_comb_kernel(int n, int off, double a, const double* __restrict__ U, double b,
                                                     const double* __restrict__ DU, double c,
                                                     const double* __restrict__ DDU, double* __restrict__ v) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) v[i] = a * U[off + i] + b * DU[off + i] + c * DDU[off + i];
}

// rhs = F_sl + R_sl + y  [- (K12^T Uc)]   (LinearTransientFESolver, FESolver.cpp:606-608)
__global__ __launch_bounds__(256) void tr_rhs_kernel(int n2, int nC, const double* __restrict__ F,
                                                    const double* __restrict__ R, const double* __restrict__ y,
                                                    const double* __restrict__ U, const int* __restrict__ cptr,
                                                    const int* __restrict__ cpos, const int* __restrict__ crow,
                                                    const double* __restrict__ K12, int coupling,
                                                    double* __restrict__ rhs) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n2) return;
  double v = F[nC + j] + R[nC + j] + (y ? y[j] : 0.0);
  if (coupling) {
    double acc = 0.0;
    for (int k = cptr[j]; k < cptr[j + 1]; k++) acc += K12[cpos[k]] * U[crow[k]];
    v -= acc;
  }
  rhs[j] = v;
}

// Unext = [Uc ; x]; DDUn = a0 (Un - U) - a2 DU - a3 DDU; DUn = DU + a6 DDUn + a7 DDU (FESolver.cpp:617-624)
__global__ __launch_bounds__(256) void tr_update_kernel(int nAll, int nC, const double* __restrict__ x, double a0,
                                                       double a2, double a3, double a6, double a7,
                                                       double* __restrict__ U, double* __restrict__ DU,
                                                       double* __restrict__ DDU) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nAll) return;
  const double un = i < nC ? U[i] : x[i - nC];
  const double ddun = a0 * (un - U[i]) - a2 * DU[i] - a3 * DDU[i];
  const double dun = DU[i] + a6 * ddun + a7 * DDU[i];
  U[i] = un;
  DU[i] = dun;
  DDU[i] = ddun;
}

// ---- scalar-row CG on the sliced-ELLPACK copy --------------------------------
__global__ __launch_bounds__(256) void cg_sell_fill_kernel(int nSlots, const int* __restrict__ pos,
                                                         const double* __restrict__ full, double* __restrict__ val,
                                                         int nGroups, const int* __restrict__ gStart,
                                                         const double* __restrict__ binv, double* __restrict__ dinv) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j < nSlots) {
    const int p = pos[j];
    val[j] = p >= 0 ? full[p] : 0.0;
  }
  if (j < nGroups) dinv[gStart[j]] = binv[(size_t)j * 9];
}

// pnew = z + beta pold (rows of this slice), q = A pnew (the gathered entries of pnew are
// formed on the fly from z and pold), partial pnew . q; one thread per row, one wave per
// 64-row slice, grid-stride over slices on the fixed grid
template <bool kNarrow>
__global__ __launch_bounds__(256) void cg_sell_spmv_kernel(CgSell A, const double* __restrict__ z,
                                                          const double* __restrict__ pold, double* __restrict__ pnew,
                                                          double* __restrict__ q, const CgState* __restrict__ st,
                                                          double* __restrict__ partial) {
  __shared__ double sh[4];
  if (st->converged) return;
  const double beta = st->beta;
  const int lane = threadIdx.x & 63;
  double contrib = 0.0;
  for (int sl = blockIdx.x * 4 + (threadIdx.x >> 6); sl < A.nSlices; sl += gridDim.x * 4) {
    const int i = sl * 64 + lane;
    if (i >= A.n) continue;
    const int b = A.off[sl], w = (A.off[sl + 1] - b) >> 6;
    const double* __restrict__ v = A.val + b + lane;
    double acc = 0.0;
    if constexpr (kNarrow) {
      const unsigned short* __restrict__ c = A.col16 + b + lane;
      const int cb = A.base[sl];
      for (int k = 0; k < w; k++) {
        const int cc = cb + c[k * 64];
        acc += v[k * 64] * (z[cc] + beta * pold[cc]);
      }
    } else {
      const int* __restrict__ c = A.col + b + lane;
      for (int k = 0; k < w; k++) {
        const int cc = c[k * 64];
        acc += v[k * 64] * (z[cc] + beta * pold[cc]);
      }
    }
    const double pi = z[i] + beta * pold[i];
    pnew[i] = pi;
    q[i] = acc;
    contrib += pi * acc;
  }
  const double s = bsum(contrib, sh);
  if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

// alpha = rz / (p . q) from the SpMV's partials (every workgroup sums them in the same
// order, so all hold the same alpha); x += alpha p ; r -= alpha q ; z = D^-1 r ; partial r . z
// into a second partial buffer (the first is still being read by other workgroups)
__global__ __launch_bounds__(256) void cg_jacobi_update_kernel(int n, const double* __restrict__ dinv,
                                                              double* __restrict__ x, double* __restrict__ r,
                                                              const double* __restrict__ p,
                                                              const double* __restrict__ q, double* __restrict__ z,
                                                              const CgState* __restrict__ st,
                                                              const double* __restrict__ pqPartial, int nPart,
                                                              double* __restrict__ partial) {
  __shared__ double sh[4];
  __shared__ double alpha;
  if (st->converged) return;
  const double pq = finish(pqPartial, nPart, sh);
  if (threadIdx.x == 0) alpha = pq != 0.0 ? st->rz / pq : 0.0;
  __syncthreads();
  const double a = alpha;
  double contrib = 0.0;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    x[i] += a * p[i];
    const double t = r[i] - a * q[i];
    r[i] = t;
    const double zz = dinv[i] * t;
    z[i] = zz;
    contrib += t * zz;
  }
  const double s = bsum(contrib, sh);
  if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

__global__ __launch_bounds__(256) void fill_kernel(int n, double v, double* __restrict__ x) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) x[i] = v;
}

__global__ __launch_bounds__(256) void copy_kernel(int n, const double* __restrict__ a, double* __restrict__ b) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) b[i] = a[i];
}

}  // namespace

static inline int cdiv(int a, int b) { return (a + b - 1) / b; }

void launch_quadth(hipStream_t s, int n, const int* conn, const double* xyz, const double* par, double* out) {
  if (n > 0) hipLaunchKernelGGL(quadth_kernel, dim3(cdiv(n, 256)), dim3(256), 0, s, n, conn, xyz, par, out);
}
void launch_lineth(hipStream_t s, int n, const int* conn, const double* xyz, const double* par, double* out) {
  if (n > 0) hipLaunchKernelGGL(lineth_kernel, dim3(cdiv(n, 256)), dim3(256), 0, s, n, conn, xyz, par, out);
}
void launch_combine_mirror(hipStream_t s, int n, double a, const double* A, double b, const double* B, const int* mir,
                           double* full) {
  if (n > 0)
    hipLaunchKernelGGL(combine_mirror_kernel, dim3(cdiv(n, 256)), dim3(256), 0, s, n, a, A, b, B, mir, full);
}

template <int LPG>
static int cg_solve_t(hipStream_t s, const KrylovMatrix& M, const Precond& P, CgWork& w, const double* b, double* x,
                      double tol, int maxit, CgState* hstate) {
  const int nb = kKrylovBlocks;
  // r = b - A x0
  hipLaunchKernelGGL(cg_spmv_kernel<LPG>, dim3(nb), dim3(256), 0, s, M, x, w.q, w.st0, w.partial);
  hipLaunchKernelGGL(cg_start_kernel, dim3(nb), dim3(256), 0, s, P.nGroups, P.gStart, P.gInfo, P.binv, b, w.q, w.r,
                     w.z, w.p, w.partial, w.partial2);
  hipLaunchKernelGGL(cg_init_kernel, dim3(1), dim3(256), 0, s, w.st, w.partial, w.partial2, nb, tol);
  int it = 0;
  const int check = 16;
  for (;;) {
    for (int k = 0; k < check && it < maxit; k++, it++) {
      hipLaunchKernelGGL(cg_spmv_kernel<LPG>, dim3(nb), dim3(256), 0, s, M, w.p, w.q, w.st, w.partial);
      hipLaunchKernelGGL(cg_alpha_kernel, dim3(1), dim3(256), 0, s, w.st, w.partial, nb);
      hipLaunchKernelGGL(cg_update_kernel, dim3(nb), dim3(256), 0, s, P.nGroups, P.gStart, P.gInfo, P.binv, x, w.r,
                         w.p, w.q, w.z, w.st, w.partial);
      hipLaunchKernelGGL(cg_beta_kernel, dim3(1), dim3(256), 0, s, w.st, w.partial, nb, tol);
      hipLaunchKernelGGL(cg_dir_kernel, dim3(cdiv(M.n, 256)), dim3(256), 0, s, M.n, w.z, w.p, w.st);
    }
    if (hipMemcpyAsync(hstate, w.st, sizeof(CgState), hipMemcpyDeviceToHost, s) != hipSuccess) return -1;
    if (hipStreamSynchronize(s) != hipSuccess) return -1;
    if (hstate->converged || it >= maxit) break;
  }
  return hstate->itn;
}

// scalar rows: two kernels per iteration besides one one-block scalar step (the search
// direction is formed inside the SpMV, alpha inside the Jacobi pass, which reads one inverse
// per row)
static int cg_solve_sell(hipStream_t s, const KrylovMatrix& M, const Precond& P, CgWork& w, const double* b, double* x,
                         double tol, int maxit, CgState* hstate, const CgSell& A) {
  const int nb = kKrylovBlocks;
  hipLaunchKernelGGL(cg_sell_fill_kernel, dim3(cdiv(std::max(A.nSlots, P.nGroups), 256)), dim3(256), 0, s, A.nSlots,
                     A.pos, M.val, A.val, P.nGroups, P.gStart, P.binv, A.dinv);
  // r = b - A x0, z = M r, p = z (the grouped SpMV with the always-active state)
  hipLaunchKernelGGL(cg_spmv_kernel<4>, dim3(nb), dim3(256), 0, s, M, x, w.q, w.st0, w.partial);
  hipLaunchKernelGGL(cg_start_kernel, dim3(nb), dim3(256), 0, s, P.nGroups, P.gStart, P.gInfo, P.binv, b, w.q, w.r,
                     w.z, w.p, w.partial, w.partial2);
  hipLaunchKernelGGL(cg_init_kernel, dim3(1), dim3(256), 0, s, w.st, w.partial, w.partial2, nb, tol);
  double* pa = w.p;   // direction of the previous iteration (beta = 0 at the first)
  double* pb = w.p2;  // this iteration's direction
  int it = 0;
  const int check = 16;
  for (;;) {
    for (int k = 0; k < check && it < maxit; k++, it++) {
      if (A.col16)
        hipLaunchKernelGGL(cg_sell_spmv_kernel<true>, dim3(nb), dim3(256), 0, s, A, w.z, pa, pb, w.q, w.st,
                           w.partial);
      else
        hipLaunchKernelGGL(cg_sell_spmv_kernel<false>, dim3(nb), dim3(256), 0, s, A, w.z, pa, pb, w.q, w.st,
                           w.partial);
      hipLaunchKernelGGL(cg_jacobi_update_kernel, dim3(nb), dim3(256), 0, s, A.n, A.dinv, x, w.r, pb, w.q, w.z, w.st,
                         w.partial, nb, w.partial2);
      hipLaunchKernelGGL(cg_beta_kernel, dim3(1), dim3(256), 0, s, w.st, w.partial2, nb, tol);
      std::swap(pa, pb);
    }
    if (hipMemcpyAsync(hstate, w.st, sizeof(CgState), hipMemcpyDeviceToHost, s) != hipSuccess) return -1;
    if (hipStreamSynchronize(s) != hipSuccess) return -1;
    if (hstate->converged || it >= maxit) break;
  }
  return hstate->itn;
}

int cg_solve(hipStream_t s, const KrylovMatrix& M, const Precond& P, CgWork& w, const double* b, double* x, double tol,
             int maxit, CgState* hstate, double avgCols, const CgSell* sell) {
  if (sell) return cg_solve_sell(s, M, P, w, b, x, tol, maxit, hstate, *sell);
  // lanes per row group sized to the row length (scalar thermal rows hold about 9 entries)
  if (avgCols <= 12.0) return cg_solve_t<4>(s, M, P, w, b, x, tol, maxit, hstate);
  if (avgCols <= 40.0) return cg_solve_t<8>(s, M, P, w, b, x, tol, maxit, hstate);
  return cg_solve_t<16>(s, M, P, w, b, x, tol, maxit, hstate);
}

void launch_tr_comb(hipStream_t s, int n, int off, double a, const double* U, double b, const double* DU, double c,
                    const double* DDU, double* v) {
  hipLaunchKernelGGL(tr_comb_kernel, dim3(cdiv(n, 256)), dim3(256), 0, s, n, off, a, U, b, DU, c, DDU, v);
}
void launch_tr_rhs(hipStream_t s, int n2, int nC, const double* F, const double* R, const double* y, const double* U,
                   const int* cptr, const int* cpos, const int* crow, const double* K12, int coupling, double* rhs) {
  hipLaunchKernelGGL(tr_rhs_kernel, dim3(cdiv(n2, 256)), dim3(256), 0, s, n2, nC, F, R, y, U, cptr, cpos, crow, K12,
                     coupling, rhs);
}
void launch_tr_update(hipStream_t s, int nAll, int nC, const double* x, double a0, double a2, double a3, double a6,
                      double a7, double* U, double* DU, double* DDU) {
  hipLaunchKernelGGL(tr_update_kernel, dim3(cdiv(nAll, 256)), dim3(256), 0, s, nAll, nC, x, a0, a2, a3, a6, a7, U, DU,
                     DDU);
}
void launch_fill(hipStream_t s, int n, double v, double* x) {
  if (n > 0) hipLaunchKernelGGL(fill_kernel, dim3(cdiv(n, 256)), dim3(256), 0, s, n, v, x);
}
void launch_copy(hipStream_t s, int n, const double* a, double* b) {
  if (n > 0) hipLaunchKernelGGL(copy_kernel, dim3(cdiv(n, 256)), dim3(256), 0, s, n, a, b);
}

}  // namespace nla3d_amd
