// Learned block-Jacobi correction on the device.
//
//  * node_features_kernel: 24 scale-free features of each node's rows of the
//    tangent (one thread per node; runs once per Newton iteration).
//  * mlp_bf16_kernel: the per-node MLP 24-128-128-128-128-6 (GELU) on MFMA
//    (v_mfma_f32_16x16x16_bf16, FP32 accumulation), 64 nodes per workgroup,
//    activations ping-ponged through LDS.
//  * dense_* kernels: the FP32 training passes (forward with stored
//    pre-activations, input and weight gradients with fixed-order reductions).
//  * learned_blocks_kernel / learned_grad_kernel: the corrected block
//    P_i = L^-T C C^T L^-1 (K_ii = L L^T, C lower triangular from the six
//    outputs, exp on its diagonal) and its adjoint.
#include <hip/hip_runtime.h>

#include <cmath>

#include "learned.h"

namespace nla3d_amd {
namespace {

inline int cdiv(long long a, int b) { return (int)((a + b - 1) / b); }

// ---------------------------------------------------------------------------
// features
// ---------------------------------------------------------------------------
constexpr int kMaxP = 16;  // element columns tracked per node

__global__ __launch_bounds__(256) void node_features_kernel(KrylovMatrix M, FeatureInputs in, float* __restrict__ X) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= in.nL) return;
  const int r0 = in.lRow0[i];
  const int node = in.rowNode[r0];
  double Kii[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
  double rs[3] = {0, 0, 0};
  double offSq = 0.0, offMax = 0.0;
  int nNbr = 0, nLam = 0;
  int pCol[kMaxP];
  double pVal[kMaxP][3];
  int nP = 0;
  for (int a = 0; a < 3; a++) {
    const int r = r0 + a;
    const int sg = M.rowGroup[r];
    const int c0 = M.gColPtr[sg], L = M.gColPtr[sg + 1] - c0;
    const int vb = M.gValPtr[sg] + (r - M.gRow0[sg]) * L;
    int lastNode = -1;
    for (int k = 0; k < L; k++) {
      const int c = M.gCols[c0 + k];
      const double v = M.val[vb + k];
      const int kind = in.rowKind[c];
      if (kind == 0) {
        rs[a] += v;
        const int cn = in.rowNode[c];
        if (cn == node && c >= r0 && c < r0 + 3) {
          Kii[a][c - r0] = v;
        } else {
          offSq += v * v;
          offMax = fmax(offMax, fabs(v));
          if (a == 0 && cn != lastNode && cn != node) nNbr++;
        }
        if (a == 0) lastNode = cn;
      } else if (kind == 1) {
        int q = 0;
        while (q < nP && pCol[q] != c) q++;
        if (q == nP && nP < kMaxP) {
          pCol[nP] = c;
          pVal[nP][0] = pVal[nP][1] = pVal[nP][2] = 0.0;
          nP++;
        }
        if (q < kMaxP) pVal[q][a] = v;
      } else {
        nLam++;
      }
    }
  }
  const double s = (Kii[0][0] + Kii[1][1] + Kii[2][2]) / 3.0;
  const double is = s > 0.0 ? 1.0 / s : 0.0;
  float f[kMlpIn];
  f[0] = (float)(Kii[0][0] * is);
  f[1] = (float)(0.5 * (Kii[0][1] + Kii[1][0]) * is);
  f[2] = (float)(0.5 * (Kii[0][2] + Kii[2][0]) * is);
  f[3] = (float)(Kii[1][1] * is);
  f[4] = (float)(0.5 * (Kii[1][2] + Kii[2][1]) * is);
  f[5] = (float)(Kii[2][2] * is);
  f[6] = (float)(sqrt(offSq) * is / 26.0);
  f[7] = (float)(offMax * is);
  f[8] = (float)(nNbr / 26.0);
  f[9] = (float)(sqrt(rs[0] * rs[0] + rs[1] * rs[1] + rs[2] * rs[2]) * is);
  // augmented block G = sum_p k_p k_p^T / |K_pp| (the element-pressure coupling)
  double G[6] = {0, 0, 0, 0, 0, 0};
  double pj = 0.0, pjMin = 0.0, cshare = 0.0;
  for (int q = 0; q < nP; q++) {
    const int c = pCol[q];
    const int dp = M.diagpos[c];
    const double kpp = dp >= 0 ? fabs(M.val[dp]) : 0.0;
    if (kpp > 0.0) {
      const double w = 1.0 / kpp;
      G[0] += pVal[q][0] * pVal[q][0] * w;
      G[1] += pVal[q][0] * pVal[q][1] * w;
      G[2] += pVal[q][0] * pVal[q][2] * w;
      G[3] += pVal[q][1] * pVal[q][1] * w;
      G[4] += pVal[q][1] * pVal[q][2] * w;
      G[5] += pVal[q][2] * pVal[q][2] * w;
    }
    const double ratio = in.bulk > 0.0 ? in.U[in.nC + c] / in.bulk : 0.0;
    pj += ratio;
    pjMin = q == 0 ? ratio : fmin(pjMin, ratio);
    cshare += kpp * in.binv[(size_t)in.rowPGroup[c] * 9];
  }
  const double trG = G[0] + G[3] + G[5];
  const double ig = trG > 0.0 ? 1.0 / trG : 0.0;
  for (int k = 0; k < 6; k++) f[10 + k] = (float)(G[k] * ig);
  f[16] = (float)(trG > 0.0 && s > 0.0 ? log10(trG / (3.0 * s)) : 0.0);
  f[17] = (float)(nP / 8.0);
  f[18] = (float)(nLam / 3.0);
  int ncon = 0;
  for (int a = 0; a < 3; a++) ncon += in.k12Ptr[r0 + a + 1] - in.k12Ptr[r0 + a];
  f[19] = (float)(ncon / 81.0);
  f[20] = (float)(nP ? pj / nP : 0.0);
  f[21] = (float)pjMin;
  f[22] = (float)in.t;
  f[23] = (float)(nP ? cshare / nP : 0.0);
  for (int k = 0; k < kMlpIn; k++) X[(size_t)i * kMlpIn + k] = f[k];
}

// ---------------------------------------------------------------------------
// corrected blocks
// ---------------------------------------------------------------------------
__device__ inline double diag_factor(float o) { return exp(fmin(fmax((double)o, -4.0), 4.0)); }

// Cholesky of the symmetric part of B; returns false when not positive definite
__device__ inline bool chol3(const double B[3][3], double L[3][3]) {
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) L[i][j] = 0.0;
  for (int i = 0; i < 3; i++)
    for (int j = 0; j <= i; j++) {
      double s = 0.5 * (B[i][j] + B[j][i]);
      for (int k = 0; k < j; k++) s -= L[i][k] * L[j][k];
      if (i == j) {
        if (!(s > 0.0)) return false;
        L[i][i] = sqrt(s);
      } else {
        L[i][j] = s / L[j][j];
      }
    }
  return true;
}

__device__ inline void lower_inverse(const double L[3][3], double Li[3][3]) {
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) Li[i][j] = 0.0;
  for (int i = 0; i < 3; i++) {
    Li[i][i] = 1.0 / L[i][i];
    for (int j = 0; j < i; j++) {
      double s = 0.0;
      for (int k = j; k < i; k++) s -= L[i][k] * Li[k][j];
      Li[i][j] = s / L[i][i];
    }
  }
}

__device__ inline void correction_factor(const float* o, double C[3][3]) {
  C[0][0] = diag_factor(o[0]);
  C[1][1] = diag_factor(o[1]);
  C[2][2] = diag_factor(o[2]);
  C[0][1] = C[0][2] = C[1][2] = 0.0;
  C[1][0] = o[3];
  C[2][0] = o[4];
  C[2][1] = o[5];
}

__global__ __launch_bounds__(256) void learned_blocks_kernel(int nL, const int* __restrict__ lGroup,
                                                           const int* __restrict__ gDiagPos,
                                                           const double* __restrict__ val, const float* __restrict__ Y,
                                                           double* __restrict__ binv, double* __restrict__ Rinv,
                                                           int* __restrict__ ok) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nL) return;
  const int g = lGroup[i];
  double B[3][3], L[3][3], Li[3][3], C[3][3];
  for (int a = 0; a < 3; a++)
    for (int b = 0; b < 3; b++) {
      const int p = gDiagPos[(size_t)g * 9 + a * 3 + b];
      B[a][b] = p >= 0 ? val[p] : 0.0;
    }
  const bool spd = chol3(B, L);
  if (ok) ok[i] = spd ? 1 : 0;
  if (!spd) return;  // the block-Jacobi fallback already written for this group stays
  lower_inverse(L, Li);
  correction_factor(Y + (size_t)i * kMlpOut, C);
  // Mx = C^T Li ; P = Mx^T Mx = Li^T C C^T Li
  double Mx[3][3];
  for (int a = 0; a < 3; a++)
    for (int b = 0; b < 3; b++) {
      double s = 0.0;
      for (int k = 0; k < 3; k++) s += C[k][a] * Li[k][b];
      Mx[a][b] = s;
    }
  double* out = binv + (size_t)g * 9;
  for (int a = 0; a < 3; a++)
    for (int b = 0; b < 3; b++) {
      double s = 0.0;
      for (int k = 0; k < 3; k++) s += Mx[k][a] * Mx[k][b];
      out[a * 3 + b] = s;
    }
  if (Rinv) {
    double* r = Rinv + (size_t)i * 6;
    r[0] = Li[0][0];
    r[1] = Li[1][0];
    r[2] = Li[1][1];
    r[3] = Li[2][0];
    r[4] = Li[2][1];
    r[5] = Li[2][2];
  }
}

__global__ __launch_bounds__(256) void learned_grad_kernel(int nL, const double* __restrict__ Rinv,
                                                         const int* __restrict__ ok, const float* __restrict__ Y,
                                                         const double* __restrict__ Pbar, float* __restrict__ dY) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nL) return;
  float* d = dY + (size_t)i * kMlpOut;
  if (!ok[i]) {
    for (int k = 0; k < kMlpOut; k++) d[k] = 0.0f;
    return;
  }
  const double* r = Rinv + (size_t)i * 6;
  const double Li[3][3] = {{r[0], 0, 0}, {r[1], r[2], 0}, {r[3], r[4], r[5]}};
  const float* o = Y + (size_t)i * kMlpOut;
  double C[3][3];
  correction_factor(o, C);
  const double* pb = Pbar + (size_t)i * 9;
  double Ps[3][3];
  for (int a = 0; a < 3; a++)
    for (int b = 0; b < 3; b++) Ps[a][b] = 0.5 * (pb[a * 3 + b] + pb[b * 3 + a]);
  // dL/dC = 2 Li Ps Li^T C
  double T1[3][3], T[3][3];
  for (int a = 0; a < 3; a++)
    for (int b = 0; b < 3; b++) {
      double s = 0.0;
      for (int k = 0; k < 3; k++) s += Li[a][k] * Ps[k][b];
      T1[a][b] = s;
    }
  for (int a = 0; a < 3; a++)
    for (int b = 0; b < 3; b++) {
      double s = 0.0;
      for (int k = 0; k < 3; k++) s += T1[a][k] * Li[b][k];
      T[a][b] = s;
    }
  double G[3][3];
  for (int a = 0; a < 3; a++)
    for (int b = 0; b < 3; b++) {
      double s = 0.0;
      for (int k = 0; k < 3; k++) s += T[a][k] * C[k][b];
      G[a][b] = 2.0 * s;
    }
  for (int k = 0; k < 3; k++) {
    const bool inside = o[k] > -4.0f && o[k] < 4.0f;
    d[k] = inside ? (float)(G[k][k] * C[k][k]) : 0.0f;
  }
  d[3] = (float)G[1][0];
  d[4] = (float)G[2][0];
  d[5] = (float)G[2][1];
}

// ---------------------------------------------------------------------------
// MLP: BF16 MFMA inference
// ---------------------------------------------------------------------------
typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float floatx4 __attribute__((ext_vector_type(4)));

__device__ inline unsigned short f2bf(float f) {
  unsigned int u = __float_as_uint(f);
  if ((u & 0x7f800000u) == 0x7f800000u) return (unsigned short)((u >> 16) | ((u & 0xffffu) ? 0x40u : 0u));
  u += 0x7fffu + ((u >> 16) & 1u);  // round to nearest even
  return (unsigned short)(u >> 16);
}

__device__ inline float gelu(float z) { return 0.5f * z * (1.0f + erff(z * 0.70710678118654752f)); }
__device__ inline float gelu_grad(float z) {
  return 0.5f * (1.0f + erff(z * 0.70710678118654752f)) + z * 0.39894228040143268f * expf(-0.5f * z * z);
}

__global__ __launch_bounds__(256) void mlp_pack_kernel(const float* __restrict__ w, unsigned short* __restrict__ wp,
                                                     float* __restrict__ bp) {
  const int offW[5] = {kMlpOffW[0], kMlpOffW[1], kMlpOffW[2], kMlpOffW[3], kMlpOffW[4]};
  const int offB[5] = {kMlpOffB[0], kMlpOffB[1], kMlpOffB[2], kMlpOffB[3], kMlpOffB[4]};
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < kMlpPackedW) {
    float v = 0.0f;
    if (t < 128 * 32) {  // W1 [128][24] -> [128][32]
      const int o = t / 32, k = t % 32;
      v = k < kMlpIn ? w[offW[0] + o * kMlpIn + k] : 0.0f;
    } else if (t < 128 * 32 + 3 * 128 * 128) {
      const int q = t - 128 * 32;
      const int l = 1 + q / (128 * 128);
      v = w[offW[l] + q % (128 * 128)];
    } else {  // W5 [6][128] -> [16][128]
      const int q = t - 128 * 32 - 3 * 128 * 128;
      const int o = q / 128, k = q % 128;
      v = o < kMlpOut ? w[offW[4] + o * 128 + k] : 0.0f;
    }
    wp[t] = f2bf(v);
  }
  if (t < kMlpPackedB) {
    float v = 0.0f;
    if (t < 4 * 128) v = w[offB[t / 128] + t % 128];
    else if (t - 512 < kMlpOut) v = w[offB[4] + t - 512];
    bp[t] = v;
  }
}

constexpr int kTile = 64;          // nodes per workgroup
constexpr int kActStride = 136;    // LDS row stride in bf16 elements (128 + 8 padding)

__global__ __launch_bounds__(256) void mlp_bf16_kernel(int n, const float* __restrict__ X,
                                                     const float* __restrict__ mean, const float* __restrict__ istd,
                                                     const unsigned short* __restrict__ wp,
                                                     const float* __restrict__ bp, float* __restrict__ Y) {
  __shared__ __attribute__((aligned(16))) unsigned short act[2][kTile][kActStride];
  const int tile0 = blockIdx.x * kTile;
  for (int idx = threadIdx.x; idx < kTile * 32; idx += 256) {
    const int m = idx / 32, k = idx % 32;
    const int node = tile0 + m;
    float v = 0.0f;
    if (k < kMlpIn && node < n) v = (X[(size_t)node * kMlpIn + k] - mean[k]) * istd[k];
    act[0][m][k] = f2bf(v);
  }
  __syncthreads();
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int lr = lane & 15, lk = (lane >> 4) * 4;
  int cur = 0;
  for (int l = 0; l < 4; l++) {
    const int K = l == 0 ? 32 : 128;
    const unsigned short* W = wp + (l == 0 ? 0 : 128 * 32 + (l - 1) * 128 * 128);
    const float* b = bp + l * 128;
    floatx4 acc[4][2];
    for (int mt = 0; mt < 4; mt++)
      for (int t = 0; t < 2; t++) acc[mt][t] = floatx4{0.0f, 0.0f, 0.0f, 0.0f};
    for (int k0 = 0; k0 < K; k0 += 16) {
      bf16x4 bf[2];
      for (int t = 0; t < 2; t++) {
        const int o = wave * 32 + t * 16 + lr;
        bf[t] = *reinterpret_cast<const bf16x4*>(W + o * K + k0 + lk);
      }
      for (int mt = 0; mt < 4; mt++) {
        const bf16x4 af = *reinterpret_cast<const bf16x4*>(&act[cur][mt * 16 + lr][k0 + lk]);
        for (int t = 0; t < 2; t++)
          acc[mt][t] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(af, bf[t], acc[mt][t], 0, 0, 0);
      }
    }
    for (int mt = 0; mt < 4; mt++)
      for (int t = 0; t < 2; t++) {
        const int o = wave * 32 + t * 16 + lr;
        const float bo = b[o];
        for (int j = 0; j < 4; j++) act[cur ^ 1][mt * 16 + lk + j][o] = f2bf(gelu(acc[mt][t][j] + bo));
      }
    __syncthreads();
    cur ^= 1;
  }
  // output layer: 16 padded outputs, wave w takes node rows [16w, 16w+16)
  const unsigned short* W = wp + 128 * 32 + 3 * 128 * 128;
  floatx4 acc = floatx4{0.0f, 0.0f, 0.0f, 0.0f};
  for (int k0 = 0; k0 < 128; k0 += 16) {
    const bf16x4 bf = *reinterpret_cast<const bf16x4*>(W + lr * 128 + k0 + lk);
    const bf16x4 af = *reinterpret_cast<const bf16x4*>(&act[cur][wave * 16 + lr][k0 + lk]);
    acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(af, bf, acc, 0, 0, 0);
  }
  if (lr < kMlpOut)
    for (int j = 0; j < 4; j++) {
      const int node = tile0 + wave * 16 + lk + j;
      if (node < n) Y[(size_t)node * kMlpOut + lr] = acc[j] + bp[512 + lr];
    }
}

// ---------------------------------------------------------------------------
// MLP: FP32 training passes
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void normalize_kernel(int n, const float* __restrict__ X,
                                                      const float* __restrict__ mean, const float* __restrict__ istd,
                                                      float* __restrict__ X0) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long long)n * kMlpIn) return;
  const int k = (int)(t % kMlpIn);
  X0[t] = (X[t] - mean[k]) * istd[k];
}

__global__ __launch_bounds__(256) void dense_fwd_kernel(int n, int in, int out, const float* __restrict__ X,
                                                      const float* __restrict__ W, const float* __restrict__ b,
                                                      float* __restrict__ Z, float* __restrict__ A, int act) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long long)n * out) return;
  const int node = (int)(t / out), o = (int)(t % out);
  const float* x = X + (size_t)node * in;
  const float* w = W + (size_t)o * in;
  float z = b[o];
  for (int k = 0; k < in; k++) z = fmaf(x[k], w[k], z);
  Z[t] = z;
  if (A) A[t] = act ? gelu(z) : z;
}

__global__ __launch_bounds__(256) void dense_bwd_input_kernel(int n, int in, int out, const float* __restrict__ dZ,
                                                            const float* __restrict__ W,
                                                            const float* __restrict__ Zprev,
                                                            float* __restrict__ dZprev) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long long)n * in) return;
  const int node = (int)(t / in), k = (int)(t % in);
  const float* dz = dZ + (size_t)node * out;
  float s = 0.0f;
  for (int o = 0; o < out; o++) s = fmaf(dz[o], W[(size_t)o * in + k], s);
  dZprev[t] = s * gelu_grad(Zprev[t]);
}

// partial[chunk][o*(in+1) + k] = sum over the chunk's nodes of dZ[node][o] * X[node][k] (k == in: bias)
__global__ __launch_bounds__(256) void dense_bwd_weight_kernel(int n, int in, int out, const float* __restrict__ dZ,
                                                             const float* __restrict__ X,
                                                             float* __restrict__ partial) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int m = out * (in + 1);
  if (t >= m) return;
  const int chunk = blockIdx.y;
  const int o = t / (in + 1), k = t % (in + 1);
  const int n0 = chunk * kGradChunk, n1 = min(n, n0 + kGradChunk);
  float s = 0.0f;
  for (int node = n0; node < n1; node++) {
    const float x = k < in ? X[(size_t)node * in + k] : 1.0f;
    s = fmaf(dZ[(size_t)node * out + o], x, s);
  }
  partial[(size_t)chunk * m + t] = s;
}

__global__ __launch_bounds__(256) void reduce_chunks_kernel(int nChunks, int in, int out,
                                                          const float* __restrict__ partial, float* __restrict__ dW,
                                                          float* __restrict__ db) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int m = out * (in + 1);
  if (t >= m) return;
  float s = 0.0f;
  for (int c = 0; c < nChunks; c++) s += partial[(size_t)c * m + t];
  const int o = t / (in + 1), k = t % (in + 1);
  if (k < in) dW[(size_t)o * in + k] = s;
  else db[o] = s;
}

__global__ __launch_bounds__(256) void adam_kernel(int n, float lr, float b1, float b2, float eps, float c1, float c2,
                                                 const float* __restrict__ g, float* __restrict__ w,
                                                 float* __restrict__ m, float* __restrict__ v) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const float gt = g[t];
  const float mt = b1 * m[t] + (1.0f - b1) * gt;
  const float vt = b2 * v[t] + (1.0f - b2) * gt * gt;
  m[t] = mt;
  v[t] = vt;
  w[t] -= lr * (mt / c1) / (sqrtf(vt / c2) + eps);
}

// ---------------------------------------------------------------------------
// unrolled MINRES vector pieces
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void bj_apply_kernel(int nGroups, const int* __restrict__ gStart,
                                                     const int* __restrict__ gInfo, const double* __restrict__ binv,
                                                     const double* __restrict__ r, double* __restrict__ z) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= nGroups) return;
  const int size = gInfo[g] & 3;
  const int r0 = gStart[g];
  const double* B = binv + (size_t)g * 9;
  double x[3] = {0, 0, 0};
  for (int a = 0; a < size; a++) x[a] = r[r0 + a];
  for (int a = 0; a < size; a++) {
    double s = 0.0;
    for (int b = 0; b < size; b++) s += B[a * 3 + b] * x[b];
    z[r0 + a] = s;
  }
}

__global__ __launch_bounds__(256) void axpby_kernel(int n, double a, const double* x, double b,
                                                  const double* y, double* out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = a * x[i] + b * y[i];
}

__global__ __launch_bounds__(256) void outer_acc_kernel(int nL, const int* __restrict__ lRow0,
                                                      const double* __restrict__ ybar, const double* __restrict__ r,
                                                      double* __restrict__ Pbar) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nL) return;
  const int r0 = lRow0[i];
  double* p = Pbar + (size_t)i * 9;
  for (int a = 0; a < 3; a++)
    for (int b = 0; b < 3; b++) p[a * 3 + b] += ybar[r0 + a] * r[r0 + b];
}

__global__ __launch_bounds__(256) void zero_d_kernel(size_t n, double* __restrict__ x) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) x[i] = 0.0;
}

}  // namespace

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
void launch_node_features(hipStream_t s, const KrylovMatrix& M, const FeatureInputs& in, float* X) {
  if (in.nL <= 0) return;
  hipLaunchKernelGGL(node_features_kernel, dim3(cdiv(in.nL, 256)), dim3(256), 0, s, M, in, X);
}

void launch_learned_blocks(hipStream_t s, int nL, const int* lGroup, const Precond& P, const KrylovMatrix& M,
                           const float* Y, double* Rinv, int* ok) {
  if (nL <= 0) return;
  hipLaunchKernelGGL(learned_blocks_kernel, dim3(cdiv(nL, 256)), dim3(256), 0, s, nL, lGroup, P.gDiagPos, M.val, Y,
                     P.binv, Rinv, ok);
}

void launch_learned_grad(hipStream_t s, int nL, const double* Rinv, const int* ok, const float* Y, const double* Pbar,
                         float* dY) {
  if (nL <= 0) return;
  hipLaunchKernelGGL(learned_grad_kernel, dim3(cdiv(nL, 256)), dim3(256), 0, s, nL, Rinv, ok, Y, Pbar, dY);
}

void launch_mlp_pack(hipStream_t s, const float* w, unsigned short* wp, float* bp) {
  hipLaunchKernelGGL(mlp_pack_kernel, dim3(cdiv(kMlpPackedW, 256)), dim3(256), 0, s, w, wp, bp);
}

void launch_mlp_bf16(hipStream_t s, int n, const float* X, const float* mean, const float* istd,
                     const unsigned short* wp, const float* bp, float* Y) {
  if (n <= 0) return;
  hipLaunchKernelGGL(mlp_bf16_kernel, dim3(cdiv(n, kTile)), dim3(256), 0, s, n, X, mean, istd, wp, bp, Y);
}

void launch_normalize(hipStream_t s, int n, const float* X, const float* mean, const float* istd, float* X0) {
  if (n <= 0) return;
  hipLaunchKernelGGL(normalize_kernel, dim3(cdiv((long long)n * kMlpIn, 256)), dim3(256), 0, s, n, X, mean, istd, X0);
}

void launch_dense_fwd(hipStream_t s, int n, int in, int out, const float* X, const float* W, const float* b, float* Z,
                      float* A, int act) {
  if (n <= 0) return;
  hipLaunchKernelGGL(dense_fwd_kernel, dim3(cdiv((long long)n * out, 256)), dim3(256), 0, s, n, in, out, X, W, b, Z,
                     A, act);
}

void launch_dense_bwd_input(hipStream_t s, int n, int in, int out, const float* dZ, const float* W, const float* Zprev,
                            float* dZprev) {
  if (n <= 0) return;
  hipLaunchKernelGGL(dense_bwd_input_kernel, dim3(cdiv((long long)n * in, 256)), dim3(256), 0, s, n, in, out, dZ, W,
                     Zprev, dZprev);
}

void launch_dense_bwd_weight(hipStream_t s, int n, int in, int out, const float* dZ, const float* X, float* partial,
                             float* dW, float* db) {
  const int m = out * (in + 1);
  const int nChunks = cdiv(n, kGradChunk);  // n >= 1 (checked by the caller)
  hipLaunchKernelGGL(dense_bwd_weight_kernel, dim3(cdiv(m, 256), nChunks), dim3(256), 0, s, n, in, out, dZ, X, partial);
  hipLaunchKernelGGL(reduce_chunks_kernel, dim3(cdiv(m, 256)), dim3(256), 0, s, nChunks, in, out, partial, dW, db);
}

void launch_adam(hipStream_t s, int n, float lr, float b1, float b2, float eps, int step, const float* g, float* w,
                 float* m, float* v) {
  const float c1 = 1.0f - powf(b1, (float)step), c2 = 1.0f - powf(b2, (float)step);
  hipLaunchKernelGGL(adam_kernel, dim3(cdiv(n, 256)), dim3(256), 0, s, n, lr, b1, b2, eps, c1, c2, g, w, m, v);
}

void launch_bj_apply(hipStream_t s, const Precond& P, const double* r, double* z) {
  hipLaunchKernelGGL(bj_apply_kernel, dim3(cdiv(P.nGroups, 256)), dim3(256), 0, s, P.nGroups, P.gStart, P.gInfo,
                     P.binv, r, z);
}

void launch_axpby(hipStream_t s, int n, double a, const double* x, double b, const double* y, double* out) {
  hipLaunchKernelGGL(axpby_kernel, dim3(cdiv(n, 256)), dim3(256), 0, s, n, a, x, b, y, out);
}

void launch_outer_acc(hipStream_t s, int nL, const int* lRow0, const double* ybar, const double* r, double* Pbar) {
  if (nL <= 0) return;
  hipLaunchKernelGGL(outer_acc_kernel, dim3(cdiv(nL, 256)), dim3(256), 0, s, nL, lRow0, ybar, r, Pbar);
}

void launch_zero_d(hipStream_t s, size_t n, double* x) {
  if (!n) return;
  hipLaunchKernelGGL(zero_d_kernel, dim3(cdiv((long long)n, 256)), dim3(256), 0, s, n, x);
}

}  // namespace nla3d_amd
