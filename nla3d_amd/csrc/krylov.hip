// Krylov solve of K.block(2) x = b on the device: preconditioned MINRES with
// a block-Jacobi preconditioner, replacing EquationSolver::solveEquations
// (math/EquationSolver.h:22; the PARDISO call of FESolver.cpp:457).
//
// Matrix: block(2) is symmetric and is kept on the device twice: the
// reference's upper-triangle CSR (for export) and a full-symmetric CSR
// (0-based, both triangles) that the SpMV reads.  The assembly's segmented
// reduction writes both (see assemble.hip).
//
// MINRES follows Paige & Saunders (SIAM J. Numer. Anal. 12, 1975), in the
// preconditioned form of their SYMMLQ/MINRES code: per iteration one SpMV,
// one preconditioner application and two inner products.  All scalar
// recurrences run in a one-thread kernel on device-resident state, so an
// iteration needs no host round-trip; inner products are reduced
// deterministically (fixed-order per-block partials, then a fixed-order
// single-block finish), so results do not depend on timing.
//
// Preconditioner (SPD, as MINRES requires): 3x3 blocks of the node DoFs
// (inverted once per Newton iteration; a block that is not positive definite
// falls back to |diag|^-1), and for the pressure and Lagrange-multiplier rows
// the diagonal Schur estimate  d_i = |K_ii| + sum_{j != i} K_ij^2 / |K_jj|.

#include "internal.h"
#include "krylov.h"

namespace nla3d_amd {
namespace {

constexpr int kRedThreads = 256;

__device__ __forceinline__ double warp_sum(double v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// block-level sum in a fixed order (wave sums, then wave 0 adds the waves in order)
template <int NT>
__device__ __forceinline__ double block_sum(double v, double* sh) {
  v = warp_sum(v);
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  if (l == 0) sh[w] = v;
  __syncthreads();
  double r = 0.0;
  if (threadIdx.x == 0) {
    for (int i = 0; i < NT / 64; i++) r += sh[i];
  }
  return r;  // valid in thread 0
}

// ---- preconditioner set-up: one thread per group --------------------------
__global__ __launch_bounds__(256) void bj_setup_kernel(int nGroups, const int* __restrict__ gStart,
                                                      const int* __restrict__ gInfo,
                                                      const int* __restrict__ gDiagPos, KrylovMatrix M,
                                                      double* __restrict__ binv) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= nGroups) return;
  const int info = gInfo[g];
  const int size = info & 3, kind = info >> 2;
  const double* A = M.val;
  double* out = binv + (size_t)g * 9;
  for (int k = 0; k < 9; k++) out[k] = 0.0;
  if (kind == 1) {  // pressure / multiplier row: diagonal Schur estimate
    const int r = gStart[g];
    const int sg = M.rowGroup[r];
    const int c0 = M.gColPtr[sg], L = M.gColPtr[sg + 1] - c0;
    const int vb = M.gValPtr[sg] + (r - M.gRow0[sg]) * L;
    double d = 0.0;
    for (int k = 0; k < L; k++) {
      const int c = M.gCols[c0 + k];
      const double a = A[vb + k];
      if (c == r) {
        d += fabs(a);
      } else {
        const int dp = M.diagpos[c];
        const double ajj = dp >= 0 ? fabs(A[dp]) : 0.0;
        if (ajj > 0.0) d += a * a / ajj;
      }
    }
    out[0] = d > 0.0 ? 1.0 / d : 1.0;
    return;
  }
  // node block of `size` consecutive equations
  double B[3][3];
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) {
      const int p = (i < size && j < size) ? gDiagPos[(size_t)g * 9 + i * 3 + j] : -1;
      B[i][j] = p >= 0 ? A[p] : 0.0;
    }
  // Cholesky of the symmetric part; not positive definite -> |diag|^-1
  bool spd = true;
  double L[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
  for (int i = 0; i < size && spd; i++) {
    for (int j = 0; j <= i; j++) {
      double s = 0.5 * (B[i][j] + B[j][i]);
      for (int k = 0; k < j; k++) s -= L[i][k] * L[j][k];
      if (i == j) {
        if (!(s > 0.0)) { spd = false; break; }
        L[i][i] = sqrt(s);
      } else {
        L[i][j] = s / L[j][j];
      }
    }
  }
  if (!spd) {
    for (int i = 0; i < size; i++) out[i * 3 + i] = fabs(B[i][i]) > 0.0 ? 1.0 / fabs(B[i][i]) : 1.0;
    return;
  }
  // inverse via L^-1 (lower) then B^-1 = L^-T L^-1
  double Li[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
  for (int i = 0; i < size; i++) {
    Li[i][i] = 1.0 / L[i][i];
    for (int j = 0; j < i; j++) {
      double s = 0.0;
      for (int k = j; k < i; k++) s -= L[i][k] * Li[k][j];
      Li[i][j] = s / L[i][i];
    }
  }
  for (int i = 0; i < size; i++)
    for (int j = 0; j < size; j++) {
      double s = 0.0;
      for (int k = (i > j ? i : j); k < size; k++) s += Li[k][i] * Li[k][j];
      out[i * 3 + j] = s;
    }
}

// ---- grouped SpMV -----------------------------------------------------------
// LPG lanes per group, grid-stride over groups.  Each lane walks the group's
// column list with stride LPG, reads x[col] once and multiplies it into the
// (up to 3) rows of the group; the row sums are reduced over the LPG lanes.
template <int LPG>
__device__ __forceinline__ void group_rowsums(const KrylovMatrix& M, int g, int lane, const double* __restrict__ x,
                                              double& a0, double& a1, double& a2, int& r0, int& sz) {
  const int c0 = M.gColPtr[g], L = M.gColPtr[g + 1] - c0;
  const int vb = M.gValPtr[g];
  sz = M.gSize[g];
  r0 = M.gRow0[g];
  a0 = a1 = a2 = 0.0;
  const int* __restrict__ cols = M.gCols + c0;
  const double* __restrict__ v0 = M.val + vb;
  // kU entries of the lane at a time: their column indices, then the x values, then the
  // products (one dependent load round trip per kU entries instead of per entry); each lane
  // still sums its entries in increasing k
  constexpr int kU = 4;
  for (int k0 = lane; k0 < L; k0 += kU * LPG) {
    int c[kU];
    double xv[kU];
#pragma unroll
    for (int u = 0; u < kU; u++) {
      const int k = k0 + u * LPG;
      c[u] = k < L ? cols[k] : -1;
    }
#pragma unroll
    for (int u = 0; u < kU; u++) xv[u] = c[u] >= 0 ? x[c[u]] : 0.0;
#pragma unroll
    for (int u = 0; u < kU; u++) {
      const int k = k0 + u * LPG;
      if (k < L) {
        a0 += v0[k] * xv[u];
        if (sz > 1) a1 += v0[L + k] * xv[u];
        if (sz > 2) a2 += v0[2 * L + k] * xv[u];
      }
    }
  }
  for (int o = LPG / 2; o > 0; o >>= 1) {
    a0 += __shfl_xor(a0, o, LPG);
    a1 += __shfl_xor(a1, o, LPG);
    a2 += __shfl_xor(a2, o, LPG);
  }
}

// the three rows of node-diagonal node i (see KrylovMatrix): the diagonal block, the upper
// blocks (t, t + d) and the transposes of the upper blocks (t - d, t) of the neighbours, then
// the ELL columns; consecutive i read consecutive addresses of every diagonal
__device__ __forceinline__ void dia_rows(const KrylovMatrix& M, int i, int t, const double* __restrict__ x,
                                         double& a0, double& a1, double& a2, int& row) {
  const size_t ld = (size_t)M.diaNodes;
  const double* __restrict__ xr = x + M.diaRow0;
  const double* __restrict__ D = M.dia + t;
  {
    const double x0 = xr[3 * t], x1 = xr[3 * t + 1], x2 = xr[3 * t + 2];
    a0 = D[0] * x0 + D[ld] * x1 + D[2 * ld] * x2;
    a1 = D[3 * ld] * x0 + D[4 * ld] * x1 + D[5 * ld] * x2;
    a2 = D[6 * ld] * x0 + D[7 * ld] * x1 + D[8 * ld] * x2;
  }
#pragma unroll 1
  for (int k = 1; k < M.nd; k++) {
    const int d = M.diaOff[k];
    const double* __restrict__ U = D + (size_t)k * 9 * ld;  // block (t, t + d)
    const double* __restrict__ B = U - d;                   // block (t - d, t)
    const double u0 = xr[3 * (t + d)], u1 = xr[3 * (t + d) + 1], u2 = xr[3 * (t + d) + 2];
    const double l0 = xr[3 * (t - d)], l1 = xr[3 * (t - d) + 1], l2 = xr[3 * (t - d) + 2];
    a0 += U[0] * u0 + U[ld] * u1 + U[2 * ld] * u2;
    a1 += U[3 * ld] * u0 + U[4 * ld] * u1 + U[5 * ld] * u2;
    a2 += U[6 * ld] * u0 + U[7 * ld] * u1 + U[8 * ld] * u2;
    a0 += B[0] * l0 + B[3 * ld] * l1 + B[6 * ld] * l2;
    a1 += B[ld] * l0 + B[4 * ld] * l1 + B[7 * ld] * l2;
    a2 += B[2 * ld] * l0 + B[5 * ld] * l1 + B[8 * ld] * l2;
  }
  const int nf = M.nDiaFit;
  for (int j = 0; j < M.remW; j++) {
    const int c = M.remCol[j * nf + i];
    if (c >= 0) {
      const double xv = x[c];
      const double* __restrict__ v = M.remVal + (size_t)3 * j * nf + i;
      a0 += v[0] * xv;
      a1 += v[nf] * xv;
      a2 += v[2 * nf] * xv;
    }
  }
  row = M.diaRow0 + 3 * t;
}

// f(i, t) for every node-diagonal node (t = fit[i]): the run in chunks of M.diaChunk nodes,
// chunk c walked by the workgroups of XCD c % 8 (workgroup b runs on XCD b % 8; gridDim.x is
// a multiple of 8), each XCD's chunks in order, all XCDs at about the same place of the run
template <class F>
__device__ __forceinline__ void dia_walk(const KrylovMatrix& M, F&& f) {
  const int x = (int)blockIdx.x & 7, q = (int)blockIdx.x >> 3, nq = (int)gridDim.x >> 3;
  const int C = M.diaChunk;
  for (int l = q * (int)blockDim.x + (int)threadIdx.x;; l += nq * (int)blockDim.x) {
    const int t = (x + 8 * (l / C)) * C + l % C;
    if (t >= M.diaNodes) break;
    const int i = M.diaRank[t];
    if (i >= 0) f(i, t);
  }
}

// the group of the j-th grouped-row item after the leading single rows
__device__ __forceinline__ int csr_group(const KrylovMatrix& M, int j) {
  return M.csrList ? M.csrList[j] : M.nLead1 + j;
}
__device__ __forceinline__ int csr_count(const KrylovMatrix& M) {
  return M.csrList ? M.nCsrList : M.nGroups - M.nLead1;
}

// T = s * A Yp - c1 * R1 over grouped-row items [0, n) (group of item j: LIST ? csr_group :
// j), LPG lanes per group, grid-stride; returns this thread's part of s * sum Yp . T
template <int LPG, bool LIST>
__device__ __forceinline__ double spmv_minres_range(const KrylovMatrix& M, int n, const double* __restrict__ Yp,
                                                    const double* __restrict__ R1, double* __restrict__ T, double s,
                                                    double c1) {
  const int lane = threadIdx.x % LPG;
  const int gpb = blockDim.x / LPG;
  double contrib = 0.0;
  for (int j = blockIdx.x * gpb + threadIdx.x / LPG; j < n; j += gridDim.x * gpb) {
    const int g = LIST ? csr_group(M, j) : j;
    double a0, a1, a2;
    int r0, sz;
    group_rowsums<LPG>(M, g, lane, Yp, a0, a1, a2, r0, sz);
    if (lane < sz) {
      const int row = r0 + lane;
      const double acc = lane == 0 ? a0 : (lane == 1 ? a1 : a2);
      const double t = s * acc - c1 * R1[row];
      T[row] = t;
      contrib += s * Yp[row] * t;
    }
  }
  return contrib;
}

// T = s * A Yp - c1 * R1 ; partial alfa = s * sum Yp . T over the grouped rows: the leading
// single-row groups (the element pressures: 25 entries each) take LPG1 lanes each, the others
// LPG lanes
template <int LPG1, int LPG>
__global__ __launch_bounds__(256) void spmv_minres_kernel(KrylovMatrix M, const double* __restrict__ Yp,
                                                         const double* __restrict__ R1, double* __restrict__ T,
                                                         MinresState* __restrict__ st, double* __restrict__ partial) {
  __shared__ double sh[4];
  if (st->converged) return;
  const double s = st->s, c1 = st->c_r1;
  double contrib = spmv_minres_range<LPG1, false>(M, M.nLead1, Yp, R1, T, s, c1);
  contrib += spmv_minres_range<LPG, true>(M, csr_count(M), Yp, R1, T, s, c1);
  const double b = block_sum<256>(contrib, sh);
  if (threadIdx.x == 0) partial[blockIdx.x] = b;
}

// the same over the node-diagonal nodes, one thread each, launched after spmv_minres_kernel on
// the same grid: each block adds its sum to that kernel's partial of the same block
__global__ __launch_bounds__(256) void spmv_minres_dia_kernel(KrylovMatrix M, const double* __restrict__ Yp,
                                                             const double* __restrict__ R1, double* __restrict__ T,
                                                             MinresState* __restrict__ st,
                                                             double* __restrict__ partial) {
  __shared__ double sh[4];
  if (st->converged) return;
  const double s = st->s, c1 = st->c_r1;
  double contrib = 0.0;
  dia_walk(M, [&](int i, int t) {
    double a0, a1, a2;
    int row;
    dia_rows(M, i, t, Yp, a0, a1, a2, row);
    const double t0 = s * a0 - c1 * R1[row], t1 = s * a1 - c1 * R1[row + 1], t2 = s * a2 - c1 * R1[row + 2];
    T[row] = t0;
    T[row + 1] = t1;
    T[row + 2] = t2;
    contrib += s * Yp[row] * t0;
    contrib += s * Yp[row + 1] * t1;
    contrib += s * Yp[row + 2] * t2;
  });
  const double b = block_sum<256>(contrib, sh);
  if (threadIdx.x == 0) partial[blockIdx.x] += b;
}

// plain y = A x over grouped-row items [0, n)
template <int LPG, bool LIST>
__device__ __forceinline__ void spmv_range(const KrylovMatrix& M, int n, const double* __restrict__ x,
                                           double* __restrict__ y) {
  const int lane = threadIdx.x % LPG;
  const int gpb = blockDim.x / LPG;
  for (int j = blockIdx.x * gpb + threadIdx.x / LPG; j < n; j += gridDim.x * gpb) {
    const int g = LIST ? csr_group(M, j) : j;
    double a0, a1, a2;
    int r0, sz;
    group_rowsums<LPG>(M, g, lane, x, a0, a1, a2, r0, sz);
    if (lane < sz) y[r0 + lane] = lane == 0 ? a0 : (lane == 1 ? a1 : a2);
  }
}

template <int LPG1, int LPG>
__global__ __launch_bounds__(256) void spmv_kernel(KrylovMatrix M, const double* __restrict__ x,
                                                  double* __restrict__ y) {
  spmv_range<LPG1, false>(M, M.nLead1, x, y);
  spmv_range<LPG, true>(M, csr_count(M), x, y);
}

__global__ __launch_bounds__(256) void spmv_dia_kernel(KrylovMatrix M, const double* __restrict__ x,
                                                      double* __restrict__ y) {
  dia_walk(M, [&](int i, int t) {
    double a0, a1, a2;
    int row;
    dia_rows(M, i, t, x, a0, a1, a2, row);
    y[row] = a0;
    y[row + 1] = a1;
    y[row + 2] = a2;
  });
}

// the node-diagonal copy from the grouped values: upper blocks of every run node (absent
// blocks are zero), then the ELL values of the node-diagonal nodes.  The diagonal (ELL slot)
// is the fastest index, so the threads of a wave read neighbouring entries of a few nodes'
// row blocks, each node's rows once, and the writes of one diagonal row merge in L2.
__global__ __launch_bounds__(256) void dia_fill_kernel(KrylovMatrix M) {
  const long long nD = (long long)M.nd * M.diaNodes, nR = (long long)M.remW * M.nDiaFit;
  const size_t ld = (size_t)M.diaNodes;
  for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < nD + nR;
       e += (long long)gridDim.x * blockDim.x) {
    if (e < nD) {
      const int t = (int)(e / M.nd), k = (int)(e % M.nd);
      const int p = M.diaPos[(size_t)k * M.diaNodes + t];
      const int g = M.diaG0 + t;
      const int L = M.gColPtr[g + 1] - M.gColPtr[g];
      const double* v = M.val + M.gValPtr[g];
      double* D = M.dia + (size_t)k * 9 * ld + t;
#pragma unroll
      for (int r = 0; r < 3; r++)
#pragma unroll
        for (int c = 0; c < 3; c++) D[(3 * r + c) * ld] = p >= 0 ? v[r * L + p + c] : 0.0;
    } else {
      const long long q = e - nD;
      const int i = (int)(q / M.remW), j = (int)(q % M.remW);
      const int p = M.remPos[(size_t)j * M.nDiaFit + i];
      const int g = M.diaG0 + M.fit[i];
      const int L = M.gColPtr[g + 1] - M.gColPtr[g];
      const double* v = M.val + M.gValPtr[g];
      double* R = M.remVal + (size_t)3 * j * M.nDiaFit + i;
#pragma unroll
      for (int r = 0; r < 3; r++) R[r * M.nDiaFit] = p >= 0 ? v[r * L + p] : 0.0;
    }
  }
}

// ---- fixed-order finish of a reduction + MINRES scalar steps --------------
// Each thread issues its partials' loads 8 at a time before adding any: they were just written by
// other compute units, so every load misses this one's caches, and a loop that adds as it loads
// waits out that latency once per partial.  Fixed order: slot k of each thread's group of 8, then
// the 8 slots pairwise, then the block.
__device__ double finish_sum(const double* partial, int nPart, double* sh) {
  double acc[8] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  for (int base = threadIdx.x; base < nPart; base += 8 * kRedThreads) {
    double a[8];
#pragma unroll
    for (int k = 0; k < 8; k++) {
      const int i = base + k * kRedThreads;
      a[k] = i < nPart ? partial[i] : 0.0;
    }
#pragma unroll
    for (int k = 0; k < 8; k++) acc[k] += a[k];
  }
  const double v = ((acc[0] + acc[1]) + (acc[2] + acc[3])) + ((acc[4] + acc[5]) + (acc[6] + acc[7]));
  return block_sum<kRedThreads>(v, sh);
}

// after the SpMV: alfa, and the coefficient of R2 for the preconditioner pass
__global__ __launch_bounds__(kRedThreads) void minres_scalar1_kernel(MinresState* st, const double* partial,
                                                                   int nPart) {
  __shared__ double sh[kRedThreads / 64];
  if (st->converged) return;
  const double alfa = finish_sum(partial, nPart, sh);
  if (threadIdx.x == 0) {
    st->alfa = alfa;
    st->c_r2 = alfa / st->beta;
  }
}

// after the preconditioner pass: beta, the plane rotation, phi and the
// coefficients of the W / X update; convergence test on the residual estimate
__global__ __launch_bounds__(kRedThreads) void minres_scalar2_kernel(MinresState* st, const double* partial,
                                                                   int nPart, double tol) {
  __shared__ double sh[kRedThreads / 64];
  if (st->converged) {
    if (threadIdx.x == 0) st->skip_update = 1;
    return;
  }
  const double beta2 = finish_sum(partial, nPart, sh);
  if (threadIdx.x != 0) return;
  st->itn += 1;
  const double oldb = st->beta;
  const double beta = beta2 > 0.0 ? sqrt(beta2) : 0.0;
  const double alfa = st->alfa;
  st->tnorm2 += alfa * alfa + oldb * oldb + beta * beta;
  const double oldeps = st->epsln;
  const double delta = st->cs * st->dbar + st->sn * alfa;
  const double gbar = st->sn * st->dbar - st->cs * alfa;
  st->epsln = st->sn * beta;
  st->dbar = -st->cs * beta;
  double gamma = sqrt(gbar * gbar + beta * beta);
  if (gamma < 1e-300) gamma = 1e-300;
  st->cs = gbar / gamma;
  st->sn = beta / gamma;
  const double phi = st->cs * st->phibar;
  st->phibar = st->sn * st->phibar;
  st->oldeps = oldeps;
  st->delta = delta;
  st->denom = 1.0 / gamma;
  st->phi = phi;
  st->s_update = st->s;  // the v = s * Yp of this iteration
  st->oldb = oldb;
  st->beta = beta;
  st->rnorm = st->phibar;
  // next iteration's SpMV coefficients
  st->s = beta > 0.0 ? 1.0 / beta : 0.0;
  st->c_r1 = oldb > 0.0 ? beta / oldb : 0.0;
  st->skip_update = 0;
  if (!(beta2 >= 0.0) || beta2 != beta2) {  // indefinite preconditioner or NaN
    st->breakdown = 1;
    st->converged = 1;
  } else if (st->phibar <= tol * st->beta1 || beta == 0.0) {
    st->converged = 1;
  }
}

// T -= c_r2 * R2 ; Z = M T ; partial beta2 = T . Z   (one thread per group, grid-stride)
// (skipQ: the scalar Schur groups are left alone, a fused stage takes them)
__global__ __launch_bounds__(256) void precond_kernel(int nGroups, const int* __restrict__ gStart,
                                                     const int* __restrict__ gInfo, const double* __restrict__ binv,
                                                     double* __restrict__ T, const double* __restrict__ R2,
                                                     double* __restrict__ Z, MinresState* __restrict__ st,
                                                     double* __restrict__ partial, int skipQ) {
  __shared__ double sh[4];
  __shared__ double sB[256 * 9];  // this pass's 256 inverse blocks, loaded with coalesced reads
  if (st->converged) return;
  const double c2 = st->c_r2;
  double contrib = 0.0;
  for (int g0 = blockIdx.x * 256; g0 < nGroups; g0 += gridDim.x * 256) {
    const int ng = min(256, nGroups - g0);
    const int g = g0 + threadIdx.x;
    const int info = g < nGroups ? gInfo[g] : 0;
    const int r0 = g < nGroups ? gStart[g] : 0;
    const bool mine = g < nGroups && !(skipQ && (info >> 2) == 1);
    // a run of 256 groups all left to the fused stage (the element pressures come first, the
    // multipliers last) is skipped whole, its inverse blocks unread
    if (!__syncthreads_or(mine)) continue;
    // every load of the pass is issued before any is used (two rounds: the inverse blocks with
    // the group's rows of T and R2, whose addresses need gStart): a loop that used each load as it
    // came would wait out the memory latency once per entry
    const int size = info & 3;
    double b[9], tv[3], rv[3];
#pragma unroll
    for (int k = 0; k < 9; k++) {
      const int e = threadIdx.x + k * 256;
      b[k] = e < ng * 9 ? binv[(size_t)g0 * 9 + e] : 0.0;
    }
#pragma unroll
    for (int i = 0; i < 3; i++) {  // unconditional loads of rows in range (the group's first row, or
                                   // row 0 past the last group), the unused ones zeroed after
      const bool in = mine && i < size;
      const int q = in ? r0 + i : r0;
      const double tq = T[q], rq = R2[q];
      tv[i] = in ? tq : 0.0;
      rv[i] = in ? rq : 0.0;
    }
#pragma unroll
    for (int k = 0; k < 9; k++) {
      const int e = threadIdx.x + k * 256;
      if (e < ng * 9) sB[e] = b[k];
    }
    __syncthreads();
    if (mine) {
      double t[3];
#pragma unroll
      for (int i = 0; i < 3; i++) t[i] = tv[i] - c2 * rv[i];
      const double* B = sB + threadIdx.x * 9;
#pragma unroll
      for (int i = 0; i < 3; i++) {
        if (i < size) {
          double z = 0.0;
#pragma unroll
          for (int j = 0; j < 3; j++)
            if (j < size) z += B[i * 3 + j] * t[j];
          T[r0 + i] = t[i];
          Z[r0 + i] = z;
          contrib += t[i] * z;
        }
      }
    }
    __syncthreads();
  }
  const double b = block_sum<256>(contrib, sh);
  if (threadIdx.x == 0) partial[blockIdx.x] = b;
}

// Wn = (s Yp - oldeps W1 - delta W2) * denom ; X += phi Wn   (Wn overwrites W1); two
// neighbouring entries per thread, each vector read and written 16 bytes at a time
struct Dbl2u {
  double a, b;
};
__global__ __launch_bounds__(256) void minres_update_kernel(int n, const double* __restrict__ Yp,
                                                           double* __restrict__ W1, const double* __restrict__ W2,
                                                           double* __restrict__ X, const MinresState* __restrict__ st) {
  if (st->skip_update) return;
  const int i = 2 * (blockIdx.x * blockDim.x + threadIdx.x);
  if (i >= n) return;
  const double su = st->s_update, oe = st->oldeps, de = st->delta, dn = st->denom, ph = st->phi;
  if (i + 1 < n) {
    Dbl2u y, w1, w2, x, wn;
    __builtin_memcpy(&y, Yp + i, 16);
    __builtin_memcpy(&w1, W1 + i, 16);
    __builtin_memcpy(&w2, W2 + i, 16);
    __builtin_memcpy(&x, X + i, 16);
    wn.a = (su * y.a - oe * w1.a - de * w2.a) * dn;
    wn.b = (su * y.b - oe * w1.b - de * w2.b) * dn;
    x.a += ph * wn.a;
    x.b += ph * wn.b;
    __builtin_memcpy(W1 + i, &wn, 16);
    __builtin_memcpy(X + i, &x, 16);
    return;
  }
  const double wn = (su * Yp[i] - oe * W1[i] - de * W2[i]) * dn;
  W1[i] = wn;
  X[i] += ph * wn;
}

// start: R1 = R2 = b (x0 = 0), Z = M b, partial = b . Z   (grid-stride)
__global__ __launch_bounds__(256) void precond_start_kernel(int nGroups, const int* __restrict__ gStart,
                                                           const int* __restrict__ gInfo,
                                                           const double* __restrict__ binv,
                                                           const double* __restrict__ b, double* __restrict__ R1,
                                                           double* __restrict__ R2, double* __restrict__ Z,
                                                           double* __restrict__ partial) {
  __shared__ double sh[4];
  double contrib = 0.0;
  for (int g = blockIdx.x * blockDim.x + threadIdx.x; g < nGroups; g += gridDim.x * blockDim.x) {
    const int r0 = gStart[g], size = gInfo[g] & 3;
    double t[3] = {0, 0, 0};
    for (int i = 0; i < size; i++) {
      t[i] = b[r0 + i];
      R1[r0 + i] = t[i];
      R2[r0 + i] = t[i];
    }
    const double* B = binv + (size_t)g * 9;
    for (int i = 0; i < size; i++) {
      double z = 0.0;
      for (int j = 0; j < size; j++) z += B[i * 3 + j] * t[j];
      Z[r0 + i] = z;
      contrib += t[i] * z;
    }
  }
  const double s = block_sum<256>(contrib, sh);
  if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

__global__ __launch_bounds__(kRedThreads) void minres_init_kernel(MinresState* st, const double* partial, int nPart) {
  __shared__ double sh[kRedThreads / 64];
  const double b2 = finish_sum(partial, nPart, sh);
  if (threadIdx.x != 0) return;
  const double beta1 = b2 > 0.0 ? sqrt(b2) : 0.0;
  MinresState z{};
  z.beta1 = beta1;
  z.beta = beta1;
  z.oldb = 0.0;
  z.phibar = beta1;
  z.rnorm = beta1;
  z.cs = -1.0;
  z.sn = 0.0;
  z.s = beta1 > 0.0 ? 1.0 / beta1 : 0.0;
  z.c_r1 = 0.0;
  z.converged = (beta1 == 0.0) ? 1 : 0;
  z.breakdown = (b2 < 0.0 || b2 != b2) ? 1 : 0;
  if (z.breakdown) z.converged = 1;
  *st = z;
}

__global__ __launch_bounds__(256) void zero3_kernel(int n, double* __restrict__ a, double* __restrict__ b,
                                                   double* __restrict__ c) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    a[i] = 0.0;
    b[i] = 0.0;
    c[i] = 0.0;
  }
}

// generic dot partials: sum a_i b_i
__global__ __launch_bounds__(256) void dot_partial_kernel(int n, const double* __restrict__ a,
                                                         const double* __restrict__ b, double* __restrict__ partial) {
  __shared__ double sh[4];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const double v = (i < n) ? a[i] * b[i] : 0.0;
  const double s = block_sum<256>(v, sh);
  if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

__global__ __launch_bounds__(kRedThreads) void finish_to_kernel(const double* partial, int nPart, double* out) {
  __shared__ double sh[kRedThreads / 64];
  const double v = finish_sum(partial, nPart, sh);
  if (threadIdx.x == 0) *out = v;
}

}  // namespace

// ---------------------------------------------------------------------------
// host-side launchers
// ---------------------------------------------------------------------------
static inline int cdiv(int a, int b) { return (a + b - 1) / b; }

void krylov_precond_setup(hipStream_t s, const KrylovMatrix& M, const Precond& P) {
  hipLaunchKernelGGL(bj_setup_kernel, dim3(cdiv(P.nGroups, 256)), dim3(256), 0, s, P.nGroups, P.gStart, P.gInfo,
                     P.gDiagPos, M, P.binv);
}

constexpr int kLPG = 16, kLPG1 = 8;
static_assert(kKrylovBlocks % 8 == 0, "the node-diagonal walk splits the grid over the 8 XCDs");

void krylov_refresh(hipStream_t s, const KrylovMatrix& M) {
  if (M.nDiaFit == 0) return;
  hipLaunchKernelGGL(dia_fill_kernel, dim3(kKrylovBlocks), dim3(256), 0, s, M);
}

void krylov_spmv(hipStream_t s, const KrylovMatrix& M, const double* x, double* y) {
  hipLaunchKernelGGL((spmv_kernel<kLPG1, kLPG>), dim3(kKrylovBlocks), dim3(256), 0, s, M, x, y);
  if (M.nDiaFit > 0) hipLaunchKernelGGL(spmv_dia_kernel, dim3(kKrylovBlocks), dim3(256), 0, s, M, x, y);
}

void krylov_dot(hipStream_t s, int n, const double* a, const double* b, double* partial, double* out) {
  const int nb = cdiv(n, 256);
  hipLaunchKernelGGL(dot_partial_kernel, dim3(nb), dim3(256), 0, s, n, a, b, partial);
  hipLaunchKernelGGL(finish_to_kernel, dim3(1), dim3(kRedThreads), 0, s, partial, nb, out);
}

void minres_start(hipStream_t s, const KrylovMatrix& M, const Precond& P, MinresWork& w, const double* b,
                  const PrecondHook* hook) {
  hipLaunchKernelGGL(precond_start_kernel, dim3(kKrylovBlocks), dim3(256), 0, s, P.nGroups, P.gStart, P.gInfo,
                     P.binv, b, w.R1, w.R2, w.Yp, w.partial);
  if (hook) (*hook)(s, w.R1, w.Yp, w.partial, nullptr);
  hipLaunchKernelGGL(minres_init_kernel, dim3(1), dim3(kRedThreads), 0, s, w.st, w.partial, kKrylovBlocks);
  // x0 = 0, w = w2 = 0 (a fill kernel rather than memset nodes, so the sequence can be graph-captured)
  hipLaunchKernelGGL(zero3_kernel, dim3(cdiv(M.n, 256)), dim3(256), 0, s, M.n, w.X, w.W1, w.W2);
}

// one MINRES iteration; buffers rotate on the host (pure pointer swaps)
void minres_iteration(hipStream_t s, const KrylovMatrix& M, const Precond& P, MinresWork& w, double tol,
                      const PrecondHook* hook, const FusedStage* fused) {
  hipLaunchKernelGGL((spmv_minres_kernel<kLPG1, kLPG>), dim3(kKrylovBlocks), dim3(256), 0, s, M, w.Yp, w.R1, w.T,
                     w.st, w.partial);
  if (M.nDiaFit > 0)
    hipLaunchKernelGGL(spmv_minres_dia_kernel, dim3(kKrylovBlocks), dim3(256), 0, s, M, w.Yp, w.R1, w.T, w.st,
                       w.partial);
  hipLaunchKernelGGL(minres_scalar1_kernel, dim3(1), dim3(kRedThreads), 0, s, w.st, w.partial, kKrylovBlocks);
  if (fused) {
    // the block-Jacobi pass of the other rows, then the stage on the Schur rows
    hipLaunchKernelGGL(precond_kernel, dim3(kKrylovBlocks), dim3(256), 0, s, P.nGroups, P.gStart, P.gInfo, P.binv,
                       w.T, w.R2, w.Z, w.st, w.partial, 1);
    fused->run(s, w.R2, w.T, w.Z, w.partial + kKrylovBlocks, w.st);
    hipLaunchKernelGGL(minres_scalar2_kernel, dim3(1), dim3(kRedThreads), 0, s, w.st, w.partial,
                       kKrylovBlocks + fused->partials, tol);
  } else {
    hipLaunchKernelGGL(precond_kernel, dim3(kKrylovBlocks), dim3(256), 0, s, P.nGroups, P.gStart, P.gInfo, P.binv,
                       w.T, w.R2, w.Z, w.st, w.partial, P.skipQ);
    if (hook) (*hook)(s, w.T, w.Z, w.partial, w.st);
    hipLaunchKernelGGL(minres_scalar2_kernel, dim3(1), dim3(kRedThreads), 0, s, w.st, w.partial, kKrylovBlocks,
                       tol);
  }
  hipLaunchKernelGGL(minres_update_kernel, dim3(cdiv(M.n, 512)), dim3(256), 0, s, M.n, w.Yp, w.W1, w.W2, w.X, w.st);
  // rotations: R1 <- R2, R2 <- T, T <- old R1 ; Yp <- Z ; (W1, W2) <- (W2, Wn)
  double* t = w.R1;
  w.R1 = w.R2;
  w.R2 = w.T;
  w.T = t;
  t = w.Yp;
  w.Yp = w.Z;
  w.Z = t;
  t = w.W2;
  w.W2 = w.W1;  // W1 now holds Wn
  w.W1 = t;
}

}  // namespace nla3d_amd
