// Kernels of the Schur-row multigrid preconditioner (see schur.h).
#include <hip/hip_runtime.h>

#include "schur.h"

namespace nla3d_amd {
namespace {

inline int cdiv(long long a, int b) { return (int)((a + b - 1) / b); }
constexpr int kLanes = 8;  // lanes per CSR row (S rows hold about 27 entries)
constexpr int kGrid = kKrylovBlocks;

// one row of A times x, spread over kLanes lanes and reduced (every lane returns it)
__device__ __forceinline__ double row_dot(const AmgCsr& A, int i, int lane, const double* __restrict__ x) {
  double s = 0.0;
  for (int k = A.rowPtr[i] + lane; k < A.rowPtr[i + 1]; k += kLanes) s += A.val[k] * x[A.col[k]];
  for (int o = kLanes / 2; o > 0; o >>= 1) s += __shfl_xor(s, o, kLanes);
  return s;
}

// one row of the sliced-ELLPACK copy times x, one thread per row: the 64 rows of a slice
// are the 64 lanes of one wave, so every load of values and columns is coalesced and the
// loop trip count (the slice width) is uniform over the wave
// (kNarrow: 16-bit column offsets from the slice's base column)
// (pattern slices, cOff[s] < 0: the symmetric diagonals, 64 consecutive rows per wave, so every
// value and x load of the wave is one contiguous segment)
template <bool kNarrow>
__device__ __forceinline__ double sell_dot(const AmgSell& A, int i, const double* __restrict__ x) {
  const int sl = i >> 6;
  const int b = A.off[sl], w = (A.off[sl + 1] - b) >> 6;
  const float* __restrict__ v = A.val + b + (i & 63);
  double acc = 0.0;
  int cb = b;
  if (A.cOff) {
    cb = A.cOff[sl];
    if (cb < 0) {
      const float* __restrict__ u = A.dia + i;
      const int* __restrict__ dOff = A.dOff;
      acc = (double)u[0] * x[i];
#pragma unroll 2
      for (int k = 1; k < A.nd; k++) {
        const int d = dOff[k];
        const float* __restrict__ uk = u + (size_t)k * A.ldd;
        acc += (double)uk[0] * x[i + d] + (double)uk[-d] * x[i - d];
      }
      return acc;
    }
  }
  cb += i & 63;
  if constexpr (kNarrow) {
    const unsigned short* __restrict__ c = A.col16 + cb;
    const double* __restrict__ xb = x + A.base[sl];
#pragma unroll 4
    for (int k = 0; k < w; k++) acc += (double)v[k * 64] * xb[c[k * 64]];
  } else {
    const int* __restrict__ c = A.col + cb;
#pragma unroll 4
    for (int k = 0; k < w; k++) acc += (double)v[k * 64] * x[c[k * 64]];
  }
  return acc;
}

// ---- the finest level's half-precision smoother (AmgHalf) ------------------------------
// FP64 -> FP16 in one correctly rounded step (round to nearest even): no FP32 intermediate,
// whose double rounding would turn values just off an FP16 midpoint into ties
__device__ __forceinline__ unsigned short to_half_bits(double v) {
  return __builtin_bit_cast(unsigned short, static_cast<_Float16>(v));
}
__device__ __forceinline__ double from_half_bits(unsigned short b) {
  return (double)__builtin_bit_cast(_Float16, b);
}

// row i of S^' = H + diag(1 + Delta) times x, i in slice sl (the layout of sell_dot); sl is
// the same for every lane of the wave, so the slice's descriptors are scalar loads
template <bool kNarrow>
__device__ __forceinline__ double half_dot(const AmgHalf& H, int i, int sl, const double* __restrict__ x) {
  double acc = (double)H.diag[i] * x[i];
  if (H.cOff && H.cOff[sl] < 0) {
    const unsigned short* __restrict__ u = H.dia + i;
#pragma unroll 2
    for (int k = 1; k < H.nd; k++) {
      const int d = H.dOff[k];
      const unsigned short* __restrict__ uk = u + (size_t)k * H.ldd;
      acc += from_half_bits(uk[0]) * x[i + d] + from_half_bits(uk[-d]) * x[i - d];
    }
    return acc;
  }
  const int b = H.off[sl], w = (H.off[sl + 1] - b) >> 6;
  const unsigned short* __restrict__ v = H.val + b + (i & 63);
  const int cb = (H.cOff ? H.cOff[sl] : b) + (i & 63);
  if constexpr (kNarrow) {
    const unsigned short* __restrict__ c = H.col16 + cb;
    const double* __restrict__ xb = x + H.base[sl];
#pragma unroll 4
    for (int k = 0; k < w; k++) acc += from_half_bits(v[k * 64]) * xb[c[k * 64]];
  } else {
    const int* __restrict__ c = H.col + cb;
#pragma unroll 4
    for (int k = 0; k < w; k++) acc += from_half_bits(v[k * 64]) * x[c[k * 64]];
  }
  return acc;
}

// rows j0 (slice sA) and j1 (slice sA + 1, which exists) of a wave whose two slices are both
// in ELLPACK form: one loop over the wider slice's slots with the two rows' loads interleaved,
// so the two dependent chains (column index, then x) run side by side instead of one after the
// other (the ELLPACK waves were the step's longest); each row summed in half_dot's order
template <bool kNarrow>
__device__ __forceinline__ void half_dot_ell2(const AmgHalf& H, int sA, int lane, int j0, int j1,
                                              const double* __restrict__ x, double& a0, double& a1) {
  const int bA = H.off[sA], bB = H.off[sA + 1];
  const int wA = (bB - bA) >> 6, wB = (H.off[sA + 2] - bB) >> 6;
  const unsigned short* __restrict__ vA = H.val + bA + lane;
  const unsigned short* __restrict__ vB = H.val + bB + lane;
  const int cA = (H.cOff ? H.cOff[sA] : bA) + lane, cB = (H.cOff ? H.cOff[sA + 1] : bB) + lane;
  double acc0 = (double)H.diag[j0] * x[j0];
  double acc1 = j1 >= 0 ? (double)H.diag[j1] * x[j1] : 0.0;
  // the slots both rows have (no branch inside, so the loads of an unrolled group issue
  // together), then the longer row's rest
  const int wm = wA < wB ? wA : wB;
  if constexpr (kNarrow) {
    const unsigned short* __restrict__ kA = H.col16 + cA;
    const unsigned short* __restrict__ kB = H.col16 + cB;
    const double* __restrict__ xA = x + H.base[sA];
    const double* __restrict__ xB = x + H.base[sA + 1];
#pragma unroll 4
    for (int k = 0; k < wm; k++) {
      acc0 += from_half_bits(vA[k * 64]) * xA[kA[k * 64]];
      acc1 += from_half_bits(vB[k * 64]) * xB[kB[k * 64]];
    }
#pragma unroll 4
    for (int k = wm; k < wA; k++) acc0 += from_half_bits(vA[k * 64]) * xA[kA[k * 64]];
#pragma unroll 4
    for (int k = wm; k < wB; k++) acc1 += from_half_bits(vB[k * 64]) * xB[kB[k * 64]];
  } else {
    const int* __restrict__ kA = H.col + cA;
    const int* __restrict__ kB = H.col + cB;
#pragma unroll 4
    for (int k = 0; k < wm; k++) {
      acc0 += from_half_bits(vA[k * 64]) * x[kA[k * 64]];
      acc1 += from_half_bits(vB[k * 64]) * x[kB[k * 64]];
    }
#pragma unroll 4
    for (int k = wm; k < wA; k++) acc0 += from_half_bits(vA[k * 64]) * x[kA[k * 64]];
#pragma unroll 4
    for (int k = wm; k < wB; k++) acc1 += from_half_bits(vB[k * 64]) * x[kB[k * 64]];
  }
  a0 = acc0;
  a1 = acc1;
}

// unaligned loads of 2 FP16 entries or 2 doubles (the hardware takes them as one access)
struct Pair64 {
  double a, b;
};
template <typename T>
__device__ __forceinline__ T load_any(const void* p) {
  T v;
  __builtin_memcpy(&v, p, sizeof(T));
  return v;
}

// Two rows of S^' x per thread (256-thread workgroups), a wave covering the 128 rows of two
// slices: where both slices are in diagonal form (and kND > 0), lane l takes rows 2l and
// 2l + 1, so each FP16 load takes two neighbouring entries of a diagonal and each x load two
// neighbouring values (half the loads of one row per thread); otherwise lane l takes row l of
// each slice, one at a time, in that slice's own layout.  Rows j0, j1 (-1 past the matrix)
// and their products a0, a1.
template <bool kNarrow, int kND>
__device__ __forceinline__ void half_dot_wave(const AmgHalf& H, const double* __restrict__ x, int& j0, double& a0,
                                              int& j1, double& a1) {
  // the workgroup's dispatch word (AmgHalf::blk: its 512-row block and, per wave, whether both
  // slices are in diagonal form) is the one load before the rows are known; block and wave
  // index are uniform, so they live in scalar registers
  const int word = __builtin_amdgcn_readfirstlane(H.blk[blockIdx.x]);
  const int v = __builtin_amdgcn_readfirstlane((int)threadIdx.x >> 6);
  const int w = (word & 0xffffff) * 4 + v;
  const int lane = threadIdx.x & 63;
  const int r0 = w * 128;
  const int sA = 2 * w;
  const bool pairForm = kND > 0 && ((word >> (24 + v)) & 1);
  if (pairForm) {
    const int i = r0 + 2 * lane;
    const Pair64 xc = load_any<Pair64>(x + i);
    a0 = (double)H.diag[i] * xc.a;
    a1 = (double)H.diag[i + 1] * xc.b;
#pragma unroll
    for (int k = 1; k < (kND > 0 ? kND : 1); k++) {
      const int d = H.dOff[k];
      const unsigned short* __restrict__ uk = H.dia + (size_t)k * H.ldd + i;
      const unsigned int up = load_any<unsigned int>(uk);
      const unsigned int lo = load_any<unsigned int>(uk - d);
      const Pair64 xu = load_any<Pair64>(x + i + d), xl = load_any<Pair64>(x + i - d);
      a0 += from_half_bits(up & 0xffffu) * xu.a + from_half_bits(lo & 0xffffu) * xl.a;
      a1 += from_half_bits(up >> 16) * xu.b + from_half_bits(lo >> 16) * xl.b;
    }
    j0 = i;
    j1 = i + 1;
    return;
  }
  j0 = r0 + lane < H.n ? r0 + lane : -1;
  j1 = r0 + 64 + lane < H.n ? r0 + 64 + lane : -1;
  if (r0 + 64 < H.n && (!H.cOff || (H.cOff[sA] >= 0 && H.cOff[sA + 1] >= 0))) {
    half_dot_ell2<kNarrow>(H, sA, lane, j0, j1, x, a0, a1);  // (every row of slice sA exists)
    return;
  }
  a0 = j0 >= 0 ? half_dot<kNarrow>(H, j0, sA, x) : 0.0;
  a1 = j1 >= 0 ? half_dot<kNarrow>(H, j1, sA + 1, x) : 0.0;
}

__global__ __launch_bounds__(256) void half_scale_kernel(int n, const double* __restrict__ dinv,
                                                       double* __restrict__ sc) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) sc[i] = sqrt(dinv[i]);
}

// S^(i, c) = S(i, c) (s_i s_c) rounded to FP16 for the ELLPACK rows (the diagonal slot: 0)
template <bool kNarrow>
__global__ __launch_bounds__(256) void half_fill_ell_kernel(AmgHalf H, const double* __restrict__ csrVal,
                                                          const int* __restrict__ sPos,
                                                          unsigned short* __restrict__ val) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= H.n) return;
  const int sl = i >> 6;
  if (H.cOff && H.cOff[sl] < 0) return;
  const int b = H.off[sl], w = (H.off[sl + 1] - b) >> 6;
  const int cb = (H.cOff ? H.cOff[sl] : b) + (i & 63);
  const double si = H.s[i];
  for (int k = 0; k < w; k++) {
    const int slot = b + k * 64 + (i & 63);
    const int p = sPos[slot];
    const int c = kNarrow ? H.base[sl] + H.col16[cb + k * 64] : H.col[cb + k * 64];
    const double v = (p >= 0 && c != i) ? csrVal[p] * (si * H.s[c]) : 0.0;
    val[slot] = to_half_bits(v);
  }
}

// dia[k][j] = S^(j, j + d_k) rounded to FP16, k >= 1, every row j (0 where absent)
__global__ __launch_bounds__(256) void half_fill_dia_kernel(AmgHalf H, const double* __restrict__ csrVal,
                                                          const int* __restrict__ dPos,
                                                          unsigned short* __restrict__ dia) {
  const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (long long)(H.nd - 1) * H.ldd) return;
  const int k = 1 + (int)(idx / H.ldd), j = (int)(idx - (long long)(k - 1) * H.ldd);
  double v = 0.0;
  if (j < H.n) {
    const int p = dPos[(size_t)k * H.ldd + j];
    if (p >= 0) v = csrVal[p] * (H.s[j] * H.s[j + H.dOff[k]]);
  }
  dia[(size_t)k * H.ldd + j] = to_half_bits(v);
}

// diag_i = 1 + Delta_i rounded up to FP32, Delta_i = |1 - S^(i,i)| + the absolute FP16 rounding
// errors of the row's other entries, summed in the row's CSR order
__global__ __launch_bounds__(256) void half_diag_kernel(AmgCsr A, const double* __restrict__ sc,
                                                      float* __restrict__ diag) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= A.n) return;
  const double si = sc[i];
  double delta = 0.0;
  for (int k = A.rowPtr[i]; k < A.rowPtr[i + 1]; k++) {
    const int c = A.col[k];
    const double v = A.val[k] * (si * sc[c]);
    delta += c == i ? fabs(1.0 - v) : fabs(from_half_bits(to_half_bits(v)) - v);
  }
  const double d = 1.0 + delta;
  float f = (float)d;
  if ((double)f < d) f = nextafterf(f, INFINITY);
  diag[i] = f;
}

// r^ = s r ; x^_1 = c0 r^
__global__ __launch_bounds__(256) void half_start_kernel(int n, const int* __restrict__ qrow,
                                                       const double* __restrict__ rsrc, const double* __restrict__ sc,
                                                       double c0, double* __restrict__ rhat, double* __restrict__ x1,
                                                       const MinresState* __restrict__ st) {
  if (st && st->converged) return;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const double rh = sc[i] * rsrc[qrow ? qrow[i] : i];
    rhat[i] = rh;
    x1[i] = c0 * rh;
  }
}

// the same start when the stage runs fused with MINRES's preconditioner pass: first the
// pass's own update of the Schur rows, t = T - c_r2 R2 (T[qrow] = t), then r^ = s t
__global__ __launch_bounds__(256) void half_start_fused_kernel(int n, const int* __restrict__ qrow,
                                                             const double* __restrict__ R2, double* __restrict__ T,
                                                             const double* __restrict__ sc, double c0,
                                                             double* __restrict__ rhat, double* __restrict__ x1,
                                                             const MinresState* __restrict__ st) {
  if (st->converged) return;
  const double c2 = st->c_r2;
  // two neighbouring Schur rows per thread, 16-byte accesses where their block(2) rows are
  // neighbours too (the element pressures)
  for (int i = 2 * (blockIdx.x * blockDim.x + threadIdx.x); i < n; i += 2 * gridDim.x * blockDim.x) {
    if (i + 1 < n) {
      const int r0 = qrow[i], r1 = qrow[i + 1];
      const Pair64 s2 = load_any<Pair64>(sc + i);
      Pair64 t2, rh, xo;
      if (r1 == r0 + 1) {
        const Pair64 tv = load_any<Pair64>(T + r0), rv = load_any<Pair64>(R2 + r0);
        t2.a = tv.a - c2 * rv.a;
        t2.b = tv.b - c2 * rv.b;
        __builtin_memcpy(T + r0, &t2, sizeof(Pair64));
      } else {
        t2.a = T[r0] - c2 * R2[r0];
        t2.b = T[r1] - c2 * R2[r1];
        T[r0] = t2.a;
        T[r1] = t2.b;
      }
      rh.a = s2.a * t2.a;
      rh.b = s2.b * t2.b;
      xo.a = c0 * rh.a;
      xo.b = c0 * rh.b;
      __builtin_memcpy(rhat + i, &rh, sizeof(Pair64));
      __builtin_memcpy(x1 + i, &xo, sizeof(Pair64));
    } else {
      const int r = qrow[i];
      const double t = T[r] - c2 * R2[r];
      T[r] = t;
      const double rh1 = sc[i] * t;
      rhat[i] = rh1;
      x1[i] = c0 * rh1;
    }
  }
}

// The three product kernels take two rows per thread (half_dot_wave).
template <bool kNarrow, int kND>
__global__ __launch_bounds__(256) void half_start_x_kernel(AmgHalf H, const double* __restrict__ x0,
                                                         const double* __restrict__ rhat, double c0,
                                                         double* __restrict__ x1, const MinresState* __restrict__ st) {
  if (st && st->converged) return;
  int j[2];
  double a[2];
  half_dot_wave<kNarrow, kND>(H, x0, j[0], a[0], j[1], a[1]);
#pragma unroll
  for (int r = 0; r < 2; r++)
    if (j[r] >= 0) x1[j[r]] = x0[j[r]] + c0 * (rhat[j[r]] - a[r]);
}

// x^_(k+1) = x^_k + c1 (x^_k - x^_(k-1)) + c2 (r^ - S^' x^_k); xout may be xprev (each thread
// reads its own entries of xprev before writing them)
template <bool kNarrow, int kND>
__global__ __launch_bounds__(256) void half_step_kernel(AmgHalf H, const double* __restrict__ xk, const double* xprev,
                                                      const double* __restrict__ rhat, double c1, double c2,
                                                      double* xout, const int* __restrict__ qrow,
                                                      double* __restrict__ z, const MinresState* __restrict__ st,
                                                      double* __restrict__ partial, double zs) {
  // the state is tested before the stores only, so that no load waits for it
  const bool done = st && st->converged;
  int j[2];
  double a[2];
  half_dot_wave<kNarrow, kND>(H, xk, j[0], a[0], j[1], a[1]);
  if (done) return;
  double contrib = 0.0;
  if (kND > 0 && j[1] == j[0] + 1) {
    // the paired form's two neighbouring rows: one 16-byte access per vector (the same
    // arithmetic as the loop below)
    const int i = j[0];
    const Pair64 xc = load_any<Pair64>(xk + i);
    const Pair64 xp = xprev ? load_any<Pair64>(xprev + i) : Pair64{0.0, 0.0};
    const Pair64 rh = load_any<Pair64>(rhat + i);
    Pair64 xn;
    xn.a = xc.a + c1 * (xc.a - xp.a) + c2 * (rh.a - a[0]);
    xn.b = xc.b + c1 * (xc.b - xp.b) + c2 * (rh.b - a[1]);
    if (z) {
      z[qrow ? qrow[i] : i] = H.s[i] * (zs * xn.a);
      z[qrow ? qrow[i + 1] : i + 1] = H.s[i + 1] * (zs * xn.b);
      contrib += rh.a * (zs * xn.a);
      contrib += rh.b * (zs * xn.b);
    } else {
      __builtin_memcpy(xout + i, &xn, sizeof(Pair64));
    }
  } else {
#pragma unroll
    for (int r = 0; r < 2; r++) {
      const int q = j[r];
      if (q < 0) continue;
      const double xq = xk[q];
      const double xp = xprev ? xprev[q] : 0.0;
      const double rq = rhat[q];
      const double xn = xq + c1 * (xq - xp) + c2 * (rq - a[r]);
      if (z)
        z[qrow ? qrow[q] : q] = H.s[q] * (zs * xn);
      else
        xout[q] = xn;
      contrib += rq * (zs * xn);  // r . z on this row: (s t) x^ = t (s x^)
    }
  }
  if (partial) {  // the workgroup's sum of r . z (the fused stage's partials)
    __shared__ double sh[4];
    for (int o = 32; o > 0; o >>= 1) contrib += __shfl_down(contrib, o, 64);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = contrib;
    __syncthreads();
    if (threadIdx.x == 0) partial[blockIdx.x] = (sh[0] + sh[1]) + (sh[2] + sh[3]);
  }
}

template <bool kNarrow, int kND>
__global__ __launch_bounds__(256) void half_residual_kernel(AmgHalf H, const double* __restrict__ x,
                                                          const double* __restrict__ rhat, double* __restrict__ t,
                                                          const MinresState* __restrict__ st) {
  if (st && st->converged) return;
  int j[2];
  double a[2];
  half_dot_wave<kNarrow, kND>(H, x, j[0], a[0], j[1], a[1]);
#pragma unroll
  for (int r = 0; r < 2; r++) {
    const int q = j[r];
    if (q < 0) continue;
    const double sq = H.s[q];
    t[q] = sq > 0.0 ? (rhat[q] - a[r]) / sq : 0.0;
  }
}

// power-iteration step on the stored operator: y = S^' x / ||x||, ||x||^2 read from the device
template <bool kNarrow, int kND>
__global__ __launch_bounds__(256) void half_power_kernel(AmgHalf H, const double* __restrict__ x,
                                                       const double* __restrict__ nsq, double* __restrict__ y) {
  int j[2];
  double a[2];
  half_dot_wave<kNarrow, kND>(H, x, j[0], a[0], j[1], a[1]);
  const double scale = *nsq > 0.0 ? 1.0 / sqrt(*nsq) : 0.0;
#pragma unroll
  for (int r = 0; r < 2; r++)
    if (j[r] >= 0) y[j[r]] = scale * a[r];
}

__global__ __launch_bounds__(256) void half_prolong_kernel(int n, const int* __restrict__ agg,
                                                         const double* __restrict__ xc, const double* __restrict__ dc,
                                                         const double* __restrict__ sc, double* __restrict__ x,
                                                         const MinresState* __restrict__ st) {
  if (st && st->converged) return;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const double si = sc[i];
    const int a = agg[i];
    if (si > 0.0) x[i] += (dc ? 1.0 * dc[a] + 1.0 * xc[a] : xc[a]) / si;
  }
}

__global__ __launch_bounds__(256) void sell_fill_kernel(int nSlots, const int* __restrict__ pos,
                                                      const double* __restrict__ csrVal, float* __restrict__ val) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= nSlots) return;
  const int p = pos[j];
  val[j] = p >= 0 ? (float)csrVal[p] : 0.0f;
}

// contribution B(:,a)^T D^-1 B(:,b) of node n to S(q_a, q_b), one thread per (node, a)
__global__ __launch_bounds__(256) void schur_contrib_kernel(int nPairs, const int* __restrict__ pairNode,
                                                          const int* __restrict__ nodeGroup,
                                                          const int* __restrict__ nodeSize,
                                                          const int* __restrict__ qPtr, const int* __restrict__ cPtr,
                                                          const int* __restrict__ posB, const double* __restrict__ kval,
                                                          const double* __restrict__ binv, double alpha,
                                                          double* __restrict__ contrib) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nPairs) return;
  const int n = pairNode[k];
  const int q0 = qPtr[n], m = qPtr[n + 1] - q0, a = k - q0, sz = nodeSize[n];
  const double* D = binv + (size_t)nodeGroup[n] * 9;
  double ba[3] = {0, 0, 0};
  for (int r = 0; r < sz; r++) {
    const int p = posB[(size_t)(q0 + a) * 3 + r];
    ba[r] = p >= 0 ? kval[p] : 0.0;
  }
  double wa[3] = {0, 0, 0};  // D^-1 B(:,a)
  for (int r = 0; r < sz; r++)
    for (int c = 0; c < sz; c++) wa[r] += D[r * 3 + c] * ba[c];
  double* out = contrib + cPtr[n] + (size_t)a * m;
  for (int b = 0; b < m; b++) {
    double s = 0.0;
    for (int r = 0; r < sz; r++) {
      const int p = posB[(size_t)(q0 + b) * 3 + r];
      if (p >= 0) s += wa[r] * kval[p];
    }
    out[b] = alpha * s;
  }
}

__global__ __launch_bounds__(256) void schur_sub_c_kernel(int nq, const int* __restrict__ sDiag,
                                                        const int* __restrict__ cPos, const double* __restrict__ kval,
                                                        double* __restrict__ sval) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= nq) return;
  if (cPos[q] >= 0) sval[sDiag[q]] -= kval[cPos[q]];
}

__global__ __launch_bounds__(256) void gather_sum_kernel(int n, const int* __restrict__ seg,
                                                       const int* __restrict__ src, const double* __restrict__ in,
                                                       double* __restrict__ out) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  double acc = 0.0;
  for (int s = seg[t]; s < seg[t + 1]; s++) acc += in[src[s]];
  out[t] = acc;
}

__global__ __launch_bounds__(256) void inv_diag_kernel(AmgCsr A, double* __restrict__ dinv) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= A.n) return;
  double d = 0.0;
  for (int k = A.rowPtr[i]; k < A.rowPtr[i + 1]; k++)
    if (A.col[k] == i) d = A.val[k];
  dinv[i] = d > 0.0 ? 1.0 / d : 0.0;
}

// res = r ; d = c0 D^-1 r ; x = 0
__global__ __launch_bounds__(256) void cheb_start_kernel(int n, const double* __restrict__ r,
                                                       const double* __restrict__ dinv, double c0,
                                                       double* __restrict__ res, double* __restrict__ d,
                                                       double* __restrict__ x) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const double ri = r[i];
    res[i] = ri;
    d[i] = c0 * dinv[i] * ri;
    x[i] = 0.0;
  }
}

// the MINRES stage's entry: r = rext[qrow] (the Schur rows of the Krylov vector, kept for the
// post-smoother), then cheb_start: res = r ; d = c0 D^-1 r ; x = 0
__global__ __launch_bounds__(256) void cheb_start_gather_kernel(int n, const int* __restrict__ qrow,
                                                              const double* __restrict__ rext,
                                                              const double* __restrict__ dinv, double c0,
                                                              double* __restrict__ r, double* __restrict__ res,
                                                              double* __restrict__ d, double* __restrict__ x) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const double ri = rext[qrow[i]];
    r[i] = ri;
    res[i] = ri;
    d[i] = c0 * dinv[i] * ri;
    x[i] = 0.0;
  }
}

// the same start fused with MINRES's preconditioner pass on the Schur rows (FusedStage): first
// the pass's own update t = T[qrow] - c_r2 R2[qrow] (written back to T), then r = res = t,
// d = c0 D^-1 t, x = 0; nothing once the solve has converged
__global__ __launch_bounds__(256) void cheb_start_fused_kernel(int n, const int* __restrict__ qrow,
                                                             const double* __restrict__ R2, double* __restrict__ T,
                                                             const double* __restrict__ dinv, double c0,
                                                             double* __restrict__ r, double* __restrict__ res,
                                                             double* __restrict__ d, double* __restrict__ x,
                                                             const MinresState* __restrict__ st) {
  if (st->converged) return;
  const double c2 = st->c_r2;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const int q = qrow[i];
    const double t = T[q] - c2 * R2[q];
    T[q] = t;
    r[i] = t;
    res[i] = t;
    d[i] = c0 * dinv[i] * t;
    x[i] = 0.0;
  }
}

// the exit of the fused stage: z[qrow] = d + x as finish_scatter, and workgroup b's sum of
// r . z over its rows in partial[b] (one workgroup per 512 rows, two rows per thread)
__global__ __launch_bounds__(256) void finish_fused_kernel(int n, const int* __restrict__ qrow,
                                                         const double* __restrict__ d, const double* __restrict__ x,
                                                         const double* __restrict__ r,
                                                         const MinresState* __restrict__ st, double* __restrict__ z,
                                                         double* __restrict__ partial, double zs) {
  if (st->converged) return;
  double contrib = 0.0;
  for (int k = 0; k < 2; k++) {
    const int i = blockIdx.x * 512 + k * 256 + (int)threadIdx.x;
    if (i < n) {
      const double zi = zs * (1.0 * d[i] + 1.0 * x[i]);
      z[qrow[i]] = zi;
      contrib += r[i] * zi;
    }
  }
  __shared__ double sh[4];
  for (int o = 32; o > 0; o >>= 1) contrib += __shfl_down(contrib, o, 64);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = contrib;
  __syncthreads();
  if (threadIdx.x == 0) partial[blockIdx.x] = (sh[0] + sh[1]) + (sh[2] + sh[3]);
}

// the MINRES stage's exit: z[qrow] = d + x (the last smoothing increment added on the way out),
// nothing once the solve has converged
__global__ __launch_bounds__(256) void finish_scatter_kernel(int n, const int* __restrict__ qrow,
                                                           const double* __restrict__ d,
                                                           const double* __restrict__ x,
                                                           const MinresState* __restrict__ st,
                                                           double* __restrict__ z, double zs) {
  if (st && st->converged) return;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    z[qrow[i]] = zs * (1.0 * d[i] + 1.0 * x[i]);
}

// res = r - A x ; d = c0 D^-1 res   (x is read, not written)
template <bool kNarrow>
__global__ __launch_bounds__(256) void cheb_start_x_kernel(AmgSell A, const double* __restrict__ r,
                                                         const double* __restrict__ dinv, double c0,
                                                         double* __restrict__ res, double* __restrict__ d,
                                                         const double* __restrict__ x) {
  const int i = xcd_block() * blockDim.x + threadIdx.x;
  if (i >= A.n) return;
  const double ri = r[i] - sell_dot<kNarrow>(A, i, x);
  res[i] = ri;
  d[i] = c0 * dinv[i] * ri;
}

// res -= A din ; dout = c1 din + c2 D^-1 res ; then by xmode: x += din (kChebPlain),
// x += din + dout (kChebPair: the next step, kChebSkip, leaves x alone), or nothing
template <bool kNarrow>
__global__ __launch_bounds__(256) void cheb_step_kernel(AmgSell A, const double* __restrict__ din,
                                                      const double* __restrict__ dinv, double c1, double c2,
                                                      double* __restrict__ res, double* __restrict__ dout,
                                                      double* __restrict__ x, int xmode) {
  const int i = xcd_block() * blockDim.x + threadIdx.x;
  if (i >= A.n) return;
  const double ri = res[i] - sell_dot<kNarrow>(A, i, din);
  res[i] = ri;
  const double di = din[i];
  const double dn = c1 * di + c2 * dinv[i] * ri;
  dout[i] = dn;
  if (xmode == kChebPlain)
    x[i] += di;
  else if (xmode == kChebPair)
    x[i] += di + dn;
}

// t = res - A dlast (the residual of x + dlast) ; x += dlast
template <bool kNarrow>
__global__ __launch_bounds__(256) void residual_kernel(AmgSell A, const double* __restrict__ res,
                                                     const double* __restrict__ dlast, double* __restrict__ t,
                                                     double* __restrict__ x) {
  const int i = xcd_block() * blockDim.x + threadIdx.x;
  if (i >= A.n) return;
  t[i] = res[i] - sell_dot<kNarrow>(A, i, dlast);
  x[i] += dlast[i];
}

// rc[I] = sum of t over the members of aggregate I: 16 lanes per aggregate, each summing every
// 16th member in member order, then a fixed butterfly over the lanes (an aggregate holds ~25
// rows, so one thread per aggregate left the launch latency-bound with few threads)
constexpr int kRestrictLanes = 16;
// (with nx.res: also the coarse level's cheb_start, res = r, d = c0 D^-1 r, x = 0)
__global__ __launch_bounds__(256) void restrict_kernel(int nc, const int* __restrict__ memPtr,
                                                     const int* __restrict__ mem, const double* __restrict__ t,
                                                     double* __restrict__ rc, ChebStartNext nx) {
  const int lane = threadIdx.x & (kRestrictLanes - 1);
  const int I = (blockIdx.x * blockDim.x + threadIdx.x) / kRestrictLanes;
  if (I >= nc) return;  // the 16 lanes of an aggregate leave together
  double s = 0.0;
  for (int k = memPtr[I] + lane; k < memPtr[I + 1]; k += kRestrictLanes) s += t[mem[k]];
  for (int o = kRestrictLanes / 2; o > 0; o >>= 1) s += __shfl_xor(s, o, kRestrictLanes);
  if (lane == 0) {
    rc[I] = s;
    if (nx.res) {
      nx.res[I] = s;
      nx.d[I] = nx.c0 * nx.dinv[I] * s;
      nx.x[I] = 0.0;
    }
  }
}

// x += xc[agg] (+ dc[agg]: the coarse level's last smoothing increment, not yet added to xc)
__global__ __launch_bounds__(256) void prolong_kernel(int n, const int* __restrict__ agg, const double* __restrict__ xc,
                                                    const double* __restrict__ dc, double* __restrict__ x) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const int a = agg[i];
    x[i] += dc ? 1.0 * dc[a] + 1.0 * xc[a] : xc[a];
  }
}

// x = Ainv r, one 64-lane wave per row (4 rows per block)
__global__ __launch_bounds__(256) void dense_matvec_kernel(int n, const double* __restrict__ Ainv,
                                                         const double* __restrict__ r, double* __restrict__ x) {
  const int i = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
  if (i >= n) return;
  double s = 0.0;
  for (int j = lane; j < n; j += 64) s += Ainv[(size_t)i * n + j] * r[j];
  for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
  if (lane == 0) x[i] = s;
}

__global__ __launch_bounds__(256) void q_gather_kernel(int nq, const int* __restrict__ qrow,
                                                     const double* __restrict__ v, const MinresState* __restrict__ st,
                                                     double* __restrict__ out) {
  if (st && st->converged) return;
  for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < nq; q += gridDim.x * blockDim.x) out[q] = v[qrow[q]];
}

__global__ __launch_bounds__(256) void q_scatter_kernel(int nq, const int* __restrict__ qrow,
                                                      const double* __restrict__ v,
                                                      const MinresState* __restrict__ st, double* __restrict__ out) {
  if (st && st->converged) return;
  for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < nq; q += gridDim.x * blockDim.x) out[qrow[q]] = v[q];
}

__device__ __forceinline__ double wsum(double v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// partial[b] = sum over block b's grid-stride rows of a . b, the layout of the MINRES partials
__global__ __launch_bounds__(256) void dot_partials_kernel(int n, const double* __restrict__ a,
                                                         const double* __restrict__ b,
                                                         const MinresState* __restrict__ st,
                                                         double* __restrict__ partial) {
  __shared__ double sh[4];
  if (st && st->converged) return;
  double v = 0.0;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) v += a[i] * b[i];
  v = wsum(v);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) partial[blockIdx.x] = sh[0] + sh[1] + sh[2] + sh[3];
}

__global__ __launch_bounds__(256) void dinv_spmv_kernel(AmgCsr A, const double* __restrict__ dinv,
                                                      const double* __restrict__ x, const double* __restrict__ nsq,
                                                      double* __restrict__ y) {
  const double scale = *nsq > 0.0 ? 1.0 / sqrt(*nsq) : 0.0;
  const int lane = threadIdx.x % kLanes, rpb = blockDim.x / kLanes;
  for (int i = blockIdx.x * rpb + threadIdx.x / kLanes; i < A.n; i += gridDim.x * rpb) {
    const double ax = row_dot(A, i, lane, x);
    if (lane == 0) y[i] = scale * dinv[i] * ax;
  }
}

inline int rowGrid(int n) { return std::min(kGrid, std::max(1, cdiv(n, 256 / kLanes))); }
inline int elemGrid(int n) { return std::min(kGrid, std::max(1, cdiv(n, 256))); }

// ---- the coarse levels of the V-cycle in one workgroup --------------------------------
// The same per-row arithmetic as the one-level kernels above (cheb_start, cheb_step,
// residual, restrict, dense_matvec, prolong, cheb_start_x, the final x += d), phase by phase
// with a workgroup barrier between phases: for small levels the cycle is bound by launch
// latency, not by the rows.
constexpr int kCoarseThreads = 1024;

__device__ __forceinline__ double sell_dot_any(const AmgSell& A, int i, const double* __restrict__ x) {
  return A.col16 ? sell_dot<true>(A, i, x) : sell_dot<false>(A, i, x);
}

__device__ void coarse_smooth(const FusedLevel& L, bool fromZero) {
  double* din = L.da;
  double* dout = L.db;
  for (int i = threadIdx.x; i < L.n; i += blockDim.x) {
    const double ri = fromZero ? L.r[i] : L.r[i] - sell_dot_any(L.A, i, L.x);
    L.res[i] = ri;
    din[i] = L.c0 * L.dinv[i] * ri;
    if (fromZero) L.x[i] = 0.0;
  }
  __syncthreads();
  for (int k = 1; k < L.steps; k++) {
    for (int i = threadIdx.x; i < L.n; i += blockDim.x) {
      const double ri = L.res[i] - sell_dot_any(L.A, i, din);
      L.res[i] = ri;
      const double di = din[i];
      dout[i] = L.c1[k] * di + L.c2[k] * L.dinv[i] * ri;
      L.x[i] += di;
    }
    __syncthreads();
    double* tmp = din;
    din = dout;
    dout = tmp;
  }
}

// the buffer holding the last smoothing increment after coarse_smooth (da or db by parity)
__device__ __forceinline__ double* last_increment(const FusedLevel& L) { return ((L.steps - 1) & 1) ? L.db : L.da; }

__global__ __launch_bounds__(kCoarseThreads) void coarse_cycle_kernel(const FusedLevel* __restrict__ lv, int nl) {
  // descend
  for (int l = 0; l + 1 < nl; l++) {
    const FusedLevel& L = lv[l];
    coarse_smooth(L, true);
    const double* dlast = last_increment(L);
    for (int i = threadIdx.x; i < L.n; i += blockDim.x) {
      L.t[i] = L.res[i] - sell_dot_any(L.A, i, dlast);
      L.x[i] += dlast[i];
    }
    __syncthreads();
    const FusedLevel& C = lv[l + 1];
    for (int I = threadIdx.x; I < L.nc; I += blockDim.x) {
      double sum = 0.0;
      for (int k = L.memPtr[I]; k < L.memPtr[I + 1]; k++) sum += L.t[L.mem[k]];
      C.r[I] = sum;
    }
    __syncthreads();
  }
  {  // coarsest: x = Ainv r, one wave per row
    const FusedLevel& Z = lv[nl - 1];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
    for (int i = wave; i < Z.n; i += nw) {
      double sum = 0.0;
      for (int j = lane; j < Z.n; j += 64) sum += Z.cinv[(size_t)i * Z.n + j] * Z.r[j];
      for (int o = 32; o > 0; o >>= 1) sum += __shfl_xor(sum, o, 64);
      if (lane == 0) Z.x[i] = sum;
    }
    __syncthreads();
  }
  // ascend
  for (int l = nl - 2; l >= 0; l--) {
    const FusedLevel& L = lv[l];
    const FusedLevel& C = lv[l + 1];
    for (int i = threadIdx.x; i < L.n; i += blockDim.x) L.x[i] += C.x[L.agg[i]];
    __syncthreads();
    coarse_smooth(L, false);
    const double* dlast = last_increment(L);
    for (int i = threadIdx.x; i < L.n; i += blockDim.x) L.x[i] = 1.0 * dlast[i] + 1.0 * L.x[i];
    __syncthreads();
  }
}

}  // namespace

void launch_schur_contrib(hipStream_t s, int nPairs, const int* pairNode, const int* nodeGroup, const int* nodeSize,
                          const int* qPtr, const int* cPtr, const int* posB, const double* kval, const double* binv,
                          double alpha, double* contrib) {
  if (nPairs <= 0) return;
  hipLaunchKernelGGL(schur_contrib_kernel, dim3(cdiv(nPairs, 256)), dim3(256), 0, s, nPairs, pairNode, nodeGroup,
                     nodeSize, qPtr, cPtr, posB, kval, binv, alpha, contrib);
}

void launch_schur_sub_c(hipStream_t s, int nq, const int* sDiag, const int* cPos, const double* kval, double* sval) {
  if (nq <= 0) return;
  hipLaunchKernelGGL(schur_sub_c_kernel, dim3(cdiv(nq, 256)), dim3(256), 0, s, nq, sDiag, cPos, kval, sval);
}

void launch_gather_sum(hipStream_t s, int n, const int* seg, const int* src, const double* in, double* out) {
  if (n <= 0) return;
  hipLaunchKernelGGL(gather_sum_kernel, dim3(cdiv(n, 256)), dim3(256), 0, s, n, seg, src, in, out);
}

void launch_inv_diag(hipStream_t s, const AmgCsr& A, double* dinv) {
  if (A.n <= 0) return;
  hipLaunchKernelGGL(inv_diag_kernel, dim3(cdiv(A.n, 256)), dim3(256), 0, s, A, dinv);
}

void launch_half_fill(hipStream_t s, const AmgHalf& H, const AmgCsr& A, const int* sPos, const int* dPos,
                      const double* dinv, double* sOut, unsigned short* val, unsigned short* dia, float* diag) {
  if (H.n <= 0) return;
  hipLaunchKernelGGL(half_scale_kernel, dim3(cdiv(H.n, 256)), dim3(256), 0, s, H.n, dinv, sOut);
  AmgHalf W = H;
  W.s = sOut;
  if (W.col16)
    hipLaunchKernelGGL(half_fill_ell_kernel<true>, dim3(cdiv(W.n, 256)), dim3(256), 0, s, W, A.val, sPos, val);
  else
    hipLaunchKernelGGL(half_fill_ell_kernel<false>, dim3(cdiv(W.n, 256)), dim3(256), 0, s, W, A.val, sPos, val);
  if (dia && W.nd > 1)
    hipLaunchKernelGGL(half_fill_dia_kernel, dim3(cdiv((long long)(W.nd - 1) * W.ldd, 256)), dim3(256), 0, s, W,
                       A.val, dPos, dia);
  hipLaunchKernelGGL(half_diag_kernel, dim3(cdiv(A.n, 256)), dim3(256), 0, s, A, sOut, diag);
}

void launch_half_start(hipStream_t s, int n, const int* qrow, const double* rsrc, const double* sc, double c0,
                       double* rhat, double* x1, const MinresState* st) {
  hipLaunchKernelGGL(half_start_kernel, dim3(elemGrid(n)), dim3(256), 0, s, n, qrow, rsrc, sc, c0, rhat, x1, st);
}

void launch_half_start_fused(hipStream_t s, int n, const int* qrow, const double* R2, double* T, const double* sc,
                             double c0, double* rhat, double* x1, const MinresState* st) {
  hipLaunchKernelGGL(half_start_fused_kernel, dim3(elemGrid((n + 1) / 2)), dim3(256), 0, s, n, qrow, R2, T, sc, c0,
                     rhat, x1, st);
}

// the diagonal-form width the kernels unroll: 14 offsets (a hexahedral mesh's element-pressure
// Schur rows: the element and its 13 upper neighbours), 5 (quadrilaterals), else a loop
template <template <bool, int> class Launch, typename... Args>
static void half_dispatch(const AmgHalf& H, Args... args) {
  const bool narrow = H.col16 != nullptr;
  const int nd = H.cOff ? H.nd : 0;
  if (nd == 14) narrow ? Launch<true, 14>::go(H, args...) : Launch<false, 14>::go(H, args...);
  else if (nd == 5) narrow ? Launch<true, 5>::go(H, args...) : Launch<false, 5>::go(H, args...);
  else narrow ? Launch<true, 0>::go(H, args...) : Launch<false, 0>::go(H, args...);
}

template <bool kNarrow, int kND>
struct StartXLaunch {
  static void go(const AmgHalf& H, hipStream_t s, const double* x0, const double* rhat, double c0, double* x1,
                 const MinresState* st) {
    hipLaunchKernelGGL((half_start_x_kernel<kNarrow, kND>), dim3(cdiv(H.n, 512)), dim3(256), 0, s, H, x0, rhat, c0,
                       x1, st);
  }
};
template <bool kNarrow, int kND>
struct PowerLaunch {
  static void go(const AmgHalf& H, hipStream_t s, const double* x, const double* nsq, double* y) {
    hipLaunchKernelGGL((half_power_kernel<kNarrow, kND>), dim3(cdiv(H.n, 512)), dim3(256), 0, s, H, x, nsq, y);
  }
};
template <bool kNarrow, int kND>
struct StepLaunch {
  static void go(const AmgHalf& H, hipStream_t s, const double* xk, const double* xprev, const double* rhat, double c1,
                 double c2, double* xout, const int* qrow, double* z, const MinresState* st, double* partial,
                 double zs) {
    hipLaunchKernelGGL((half_step_kernel<kNarrow, kND>), dim3(cdiv(H.n, 512)), dim3(256), 0, s, H, xk, xprev, rhat,
                       c1, c2, xout, qrow, z, st, partial, zs);
  }
};
template <bool kNarrow, int kND>
struct ResidualLaunch {
  static void go(const AmgHalf& H, hipStream_t s, const double* x, const double* rhat, double* t,
                 const MinresState* st) {
    hipLaunchKernelGGL((half_residual_kernel<kNarrow, kND>), dim3(cdiv(H.n, 512)), dim3(256), 0, s, H, x, rhat, t,
                       st);
  }
};

void launch_half_start_x(hipStream_t s, const AmgHalf& H, const double* x0, const double* rhat, double c0, double* x1,
                         const MinresState* st) {
  half_dispatch<StartXLaunch>(H, s, x0, rhat, c0, x1, st);
}

void launch_half_power(hipStream_t s, const AmgHalf& H, const double* x, const double* nsq, double* y) {
  half_dispatch<PowerLaunch>(H, s, x, nsq, y);
}

void launch_half_step(hipStream_t s, const AmgHalf& H, const double* xk, const double* xprev, const double* rhat,
                      double c1, double c2, double* xout, const int* qrow, double* z, const MinresState* st,
                      double* partial, double zs) {
  half_dispatch<StepLaunch>(H, s, xk, xprev, rhat, c1, c2, xout, qrow, z, st, partial, zs);
}

void launch_half_residual_restrict(hipStream_t s, const AmgHalf& H, const double* x, const double* rhat, double* t,
                                   int nc, const int* memPtr, const int* mem, double* rc, const MinresState* st,
                                   const ChebStartNext& nx) {
  half_dispatch<ResidualLaunch>(H, s, x, rhat, t, st);
  hipLaunchKernelGGL(restrict_kernel, dim3(cdiv((long long)nc * kRestrictLanes, 256)), dim3(256), 0, s, nc, memPtr,
                     mem, t, rc, nx);
}

void launch_half_prolong(hipStream_t s, int n, const int* agg, const double* xc, const double* dc, const double* sc,
                         double* x, const MinresState* st) {
  hipLaunchKernelGGL(half_prolong_kernel, dim3(elemGrid(n)), dim3(256), 0, s, n, agg, xc, dc, sc, x, st);
}

void launch_cheb_start(hipStream_t s, int n, const double* r, const double* dinv, double c0, double* res, double* d,
                       double* x) {
  hipLaunchKernelGGL(cheb_start_kernel, dim3(elemGrid(n)), dim3(256), 0, s, n, r, dinv, c0, res, d, x);
}

void launch_cheb_start_gather(hipStream_t s, int n, const int* qrow, const double* rext, const double* dinv, double c0,
                              double* r, double* res, double* d, double* x) {
  hipLaunchKernelGGL(cheb_start_gather_kernel, dim3(elemGrid(n)), dim3(256), 0, s, n, qrow, rext, dinv, c0, r, res, d,
                     x);
}

void launch_finish_scatter(hipStream_t s, int n, const int* qrow, const double* d, const double* x,
                           const MinresState* st, double* z, double zs) {
  hipLaunchKernelGGL(finish_scatter_kernel, dim3(elemGrid(n)), dim3(256), 0, s, n, qrow, d, x, st, z, zs);
}

void launch_cheb_start_fused(hipStream_t s, int n, const int* qrow, const double* R2, double* T, const double* dinv,
                             double c0, double* r, double* res, double* d, double* x, const MinresState* st) {
  hipLaunchKernelGGL(cheb_start_fused_kernel, dim3(elemGrid(n)), dim3(256), 0, s, n, qrow, R2, T, dinv, c0, r, res, d,
                     x, st);
}

void launch_finish_fused(hipStream_t s, int n, const int* qrow, const double* d, const double* x, const double* r,
                         const MinresState* st, double* z, double* partial, double zs) {
  hipLaunchKernelGGL(finish_fused_kernel, dim3(cdiv(n, 512)), dim3(256), 0, s, n, qrow, d, x, r, st, z, partial,
                     zs);
}

void launch_sell_fill(hipStream_t s, int nSlots, const int* pos, const double* csrVal, float* sellVal) {
  if (nSlots <= 0) return;
  hipLaunchKernelGGL(sell_fill_kernel, dim3(cdiv(nSlots, 256)), dim3(256), 0, s, nSlots, pos, csrVal, sellVal);
}

void launch_cheb_start_x(hipStream_t s, const AmgSell& A, const double* r, const double* dinv, double c0, double* res,
                         double* d, const double* x) {
  if (A.col16)
    hipLaunchKernelGGL(cheb_start_x_kernel<true>, dim3(cdiv(A.n, 256)), dim3(256), 0, s, A, r, dinv, c0, res, d, x);
  else
    hipLaunchKernelGGL(cheb_start_x_kernel<false>, dim3(cdiv(A.n, 256)), dim3(256), 0, s, A, r, dinv, c0, res, d, x);
}

void launch_cheb_step(hipStream_t s, const AmgSell& A, const double* din, const double* dinv, double c1, double c2,
                      double* res, double* dout, double* x, int xmode) {
  if (A.col16)
    hipLaunchKernelGGL(cheb_step_kernel<true>, dim3(cdiv(A.n, 256)), dim3(256), 0, s, A, din, dinv, c1, c2, res, dout,
                       x, xmode);
  else
    hipLaunchKernelGGL(cheb_step_kernel<false>, dim3(cdiv(A.n, 256)), dim3(256), 0, s, A, din, dinv, c1, c2, res, dout,
                       x, xmode);
}

void launch_residual_restrict(hipStream_t s, const AmgSell& A, const double* res, const double* dlast, double* t,
                              int nc, const int* memPtr, const int* mem, double* rc, double* x,
                              const ChebStartNext& nx) {
  if (A.col16)
    hipLaunchKernelGGL(residual_kernel<true>, dim3(cdiv(A.n, 256)), dim3(256), 0, s, A, res, dlast, t, x);
  else
    hipLaunchKernelGGL(residual_kernel<false>, dim3(cdiv(A.n, 256)), dim3(256), 0, s, A, res, dlast, t, x);
  hipLaunchKernelGGL(restrict_kernel, dim3(cdiv((long long)nc * kRestrictLanes, 256)), dim3(256), 0, s, nc, memPtr,
                     mem, t, rc, nx);
}

void launch_prolong(hipStream_t s, int n, const int* agg, const double* xc, const double* dc, double* x) {
  hipLaunchKernelGGL(prolong_kernel, dim3(elemGrid(n)), dim3(256), 0, s, n, agg, xc, dc, x);
}

void launch_coarse_cycle(hipStream_t s, const FusedLevel* lv, int nl) {
  hipLaunchKernelGGL(coarse_cycle_kernel, dim3(1), dim3(kCoarseThreads), 0, s, lv, nl);
}

void launch_dense_matvec(hipStream_t s, int n, const double* Ainv, const double* r, double* x) {
  hipLaunchKernelGGL(dense_matvec_kernel, dim3(cdiv(n, 4)), dim3(256), 0, s, n, Ainv, r, x);
}

void launch_q_gather(hipStream_t s, int nq, const int* qrow, const double* v, const MinresState* st, double* out) {
  hipLaunchKernelGGL(q_gather_kernel, dim3(elemGrid(nq)), dim3(256), 0, s, nq, qrow, v, st, out);
}

void launch_q_scatter(hipStream_t s, int nq, const int* qrow, const double* v, const MinresState* st, double* out) {
  hipLaunchKernelGGL(q_scatter_kernel, dim3(elemGrid(nq)), dim3(256), 0, s, nq, qrow, v, st, out);
}

void launch_dot_partials(hipStream_t s, int n, const double* a, const double* b, const MinresState* st,
                         double* partial) {
  hipLaunchKernelGGL(dot_partials_kernel, dim3(kGrid), dim3(256), 0, s, n, a, b, st, partial);
}

void launch_dinv_spmv(hipStream_t s, const AmgCsr& A, const double* dinv, const double* x, const double* nsq,
                      double* y) {
  hipLaunchKernelGGL(dinv_spmv_kernel, dim3(rowGrid(A.n)), dim3(256), 0, s, A, dinv, x, nsq, y);
}

}  // namespace nla3d_amd
