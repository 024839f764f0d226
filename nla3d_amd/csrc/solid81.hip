// SOLID81 fused update() + buildK() on gfx950, FP64.
//
// Reference: ElementSOLID81::update (elements/SOLID81.cpp:85-115) followed by
// ElementSOLID81::buildK (elements/SOLID81.cpp:29-82) with the material
// Mat_Hyper_Isotrop_General::getS_UP / getDdDp_UP
// (materials/materials_hyperelastic.cpp:30-129) and the hexahedron Jacobian
// ElementIsoParamHEXAHEDRON::makeJacob (elements/isoparametric.cpp:140-188).
// update() followed by buildK() at the same U is a pure function of (X, U, p):
// the Gauss-point state (O, C, S) is recomputed here rather than stored.
//
// Lane mapping (one 64-lane wave = 7 elements x 9 lanes, lane 63 idle):
//   phase A  lane r < 8 of an element evaluates Gauss point r: Jacobian,
//            shape-function gradients dN/dX, displacement gradient O, C,
//            invariants, PK2 stress S, tangent Dd, Dp, and from those the
//            9x9 symmetric "gradient-space" tangent
//               At[(d,p),(e,r)] = 2w sum_{q,s} F_dq Dd[v(p,q)][v(r,s)] F_es
//                                 + w delta_de S[v(p,r)]
//            so that Kuu[(i,d),(j,e)] = sum_gp sum_{p,r} dN_i,p At dN_j,r.
//            This equals the reference's 1/2 w B^T Dd B + w B_NL^T S B_NL
//            with B = B_L + 2 Omega B_NL (SOLID81.cpp:63-71): B_k(i,d) =
//            sum_p dN_i,p * 2 F_dq for Voigt k = v(p,q).
//            Results go to LDS (89 doubles per Gauss point).
//   phase B  lane (d,e) in 0..8 of an element accumulates, over the 8 Gauss
//            points, the (d,e) component of all 36 node-pair blocks of Kuu,
//            plus Fu (lanes e==0), Kup (lanes e==1), Fp / Kpp (lanes 0 / 1).
// The work in phase B is the same for every lane (no triangular imbalance):
// per Gauss point 8 3x3 mat-vecs (72 FMA) + 36 dot3 (108 FMA) + 24 FMA.
//
// Output slot order per element (350 doubles), chosen so each store of the
// 9 lanes of an element is contiguous:
//   [0,252)   off-diagonal node pairs i<j (28 pairs x 9 (d,e))
//   [252,300) diagonal node pairs i (8 x 6, d<=e in Voigt order)
//   [300,324) Kup (3i+d), 324 Kpp, [325,349) Fu (3i+d), 349 Fp
// solid81_slot_to_assemble3() gives the permutation to the reference's
// assemble3 order (SOLID81.h:54-94).

#include "hyper.h"
#include "internal.h"

namespace nla3d_amd {
namespace {

constexpr int kEPB = 7;   // elements per 64-lane block
constexpr int kLPE = 9;   // lanes per element
constexpr int kGPW = 89;  // doubles per Gauss point in LDS
constexpr double kGP2 = 0.5773502691896258;  // isoparametric.h:128-136

// LDS layout of one Gauss point record
constexpr int kOffA = 0;    // 45: At packed upper 9x9
constexpr int kOffN = 45;   // 24: dN/dX, node-major
constexpr int kOffP = 69;   // 9:  -w F S        (Fu)
constexpr int kOffQ = 78;   // 9:   w F J C^-1   (Kup)
constexpr int kOffFp = 87;  // -(J - 1 - p/k) w
constexpr int kOffKpp = 88; // -w/k

// Voigt index of the symmetric pair (p,q): XX XY XZ YY YZ ZZ (solidmech.h:12-19)
__host__ __device__ constexpr int voigt(int p, int q) {
  return (p > q) ? (q == 0 ? p : (q == 1 ? 2 + p : 5)) : (p == 0 ? q : (p == 1 ? 2 + q : 5));
}
// packed upper index of a symmetric 9x9
__host__ __device__ constexpr int pk9u(int a, int b) { return 9 * a - (a * (a - 1)) / 2 + (b - a); }
__host__ __device__ constexpr int pk9(int a, int b) { return (a > b) ? pk9u(b, a) : pk9u(a, b); }
// off-diagonal node pair (i<j) -> 0..27, enumerated j-major: (0,1) (0,2) (1,2) (0,3) ...
__host__ __device__ constexpr int offpair(int i, int j) { return (j * (j - 1)) / 2 + i; }

// node sign pattern of the trilinear hexahedron (isoparametric.cpp:197-208)
__device__ constexpr int kSx[8] = {-1, 1, 1, -1, -1, 1, 1, -1};
__device__ constexpr int kSy[8] = {-1, -1, 1, 1, -1, -1, 1, 1};
__device__ constexpr int kSz[8] = {-1, -1, -1, -1, 1, 1, 1, 1};

// Voigt component -> tensor indices
__device__ constexpr int kVi[6] = {0, 0, 0, 1, 1, 2};
__device__ constexpr int kVj[6] = {0, 1, 2, 1, 2, 2};

// Phase A: one Gauss point of one element.
__device__ __forceinline__ void gauss_point(const double (*X)[3], const double (*U)[3], double p_e, int g,
                                            const MaterialParams& mat, double* out,
                                            double* __restrict__ state) {
  const double gx = (g & 1) ? kGP2 : -kGP2;
  const double gy = (g & 2) ? kGP2 : -kGP2;
  const double gz = (g & 4) ? kGP2 : -kGP2;

  // dN/dxi at this Gauss point (formFuncDeriv, isoparametric.cpp:211-221)
  double dN[8][3];
#pragma unroll
  for (int n = 0; n < 8; n++) {
    const double ax = 1.0 + kSx[n] * gx, ay = 1.0 + kSy[n] * gy, az = 1.0 + kSz[n] * gz;
    dN[n][0] = (0.125 * kSx[n]) * ay * az;
    dN[n][1] = (0.125 * kSy[n]) * ax * az;
    dN[n][2] = (0.125 * kSz[n]) * ax * ay;
  }
  // Jacobian J[i][j] = sum_n dN_n,i X_n,j ; det ; inverse (Mat.cpp:25-46)
  double J[3][3] = {{0.0, 0.0, 0.0}, {0.0, 0.0, 0.0}, {0.0, 0.0, 0.0}};
#pragma unroll
  for (int n = 0; n < 8; n++)
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
      for (int j = 0; j < 3; j++) J[i][j] += dN[n][i] * X[n][j];
  const double det = J[0][0] * (J[1][1] * J[2][2] - J[1][2] * J[2][1]) -
                     J[0][1] * (J[1][0] * J[2][2] - J[1][2] * J[2][0]) +
                     J[0][2] * (J[1][0] * J[2][1] - J[1][1] * J[2][0]);
  const double rdet = 1.0 / det;
  double Ji[3][3];
  Ji[0][0] = (J[1][1] * J[2][2] - J[1][2] * J[2][1]) * rdet;
  Ji[0][1] = (J[0][2] * J[2][1] - J[0][1] * J[2][2]) * rdet;
  Ji[0][2] = (J[0][1] * J[1][2] - J[0][2] * J[1][1]) * rdet;
  Ji[1][0] = (J[2][0] * J[1][2] - J[1][0] * J[2][2]) * rdet;
  Ji[1][1] = (J[0][0] * J[2][2] - J[0][2] * J[2][0]) * rdet;
  Ji[1][2] = (J[1][0] * J[0][2] - J[0][0] * J[1][2]) * rdet;
  Ji[2][0] = (J[1][0] * J[2][1] - J[1][1] * J[2][0]) * rdet;
  Ji[2][1] = (J[0][1] * J[2][0] - J[0][0] * J[2][1]) * rdet;
  Ji[2][2] = (J[0][0] * J[1][1] - J[0][1] * J[1][0]) * rdet;

  // dN/dX (NiXj) and displacement gradient O[3d+q] = du_d/dX_q (SOLID81.cpp:97-101)
  double N[8][3];
  double O[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
#pragma unroll
  for (int n = 0; n < 8; n++) {
#pragma unroll
    for (int i = 0; i < 3; i++) N[n][i] = Ji[i][0] * dN[n][0] + Ji[i][1] * dN[n][1] + Ji[i][2] * dN[n][2];
#pragma unroll
    for (int d = 0; d < 3; d++)
#pragma unroll
      for (int q = 0; q < 3; q++) O[3 * d + q] += N[n][q] * U[n][d];
  }
  // X and U are not read past this point: the record `out` written below overlays them in
  // LDS (the block is one wave, so every lane has passed this point before any lane
  // writes); the fence keeps the compiler from moving those writes above the reads
  __atomic_signal_fence(__ATOMIC_SEQ_CST);
  double F[3][3];
#pragma unroll
  for (int d = 0; d < 3; d++)
#pragma unroll
    for (int q = 0; q < 3; q++) F[d][q] = (d == q ? 1.0 : 0.0) + O[3 * d + q];

  // right Cauchy-Green tensor, SOLID81.cpp:102-107
  double C[6];
  C[0] = 1.0 + 2 * O[0] + O[0] * O[0] + O[3] * O[3] + O[6] * O[6];
  C[3] = 1.0 + 2 * O[4] + O[1] * O[1] + O[4] * O[4] + O[7] * O[7];
  C[5] = 1.0 + 2 * O[8] + O[2] * O[2] + O[5] * O[5] + O[8] * O[8];
  C[1] = O[1] + O[3] + O[0] * O[1] + O[3] * O[4] + O[6] * O[7];
  C[4] = O[5] + O[7] + O[1] * O[2] + O[4] * O[5] + O[7] * O[8];
  C[2] = O[2] + O[6] + O[0] * O[2] + O[3] * O[5] + O[6] * O[8];

  // invariants and C^-1 (solidmech.cpp:10-42)
  const double I1 = C[0] + C[3] + C[5];
  const double I2 = C[0] * C[3] + C[3] * C[5] + C[0] * C[5] - C[1] * C[1] - C[4] * C[4] - C[2] * C[2];
  const double Jv = sqrt(C[0] * (C[3] * C[5] - C[4] * C[4]) - C[1] * (C[1] * C[5] - C[4] * C[2]) +
                         C[2] * (C[1] * C[4] - C[3] * C[2]));
  const double oo = 1.0 / (Jv * Jv);
  double Ci[6];
  Ci[0] = oo * (C[3] * C[5] - C[4] * C[4]);
  Ci[1] = oo * (C[2] * C[4] - C[1] * C[5]);
  Ci[2] = oo * (C[1] * C[4] - C[2] * C[3]);
  Ci[3] = oo * (C[0] * C[5] - C[2] * C[2]);
  Ci[4] = oo * (C[1] * C[2] - C[0] * C[4]);
  Ci[5] = oo * (C[0] * C[3] - C[1] * C[1]);

  const double pp = pow(Jv, -2.0 / 3.0);
  const double pppp = pp * pp;
  double al[5];
  material_alpha(mat, I1 * pp, al);

  const double third_I1 = (1.0 / 3.0) * I1;
  const double two3_I2 = (2.0 / 3.0) * I2;
  double A[6], B[6];
#pragma unroll
  for (int i = 0; i < 6; i++) {
    const double Ii = (i == 0 || i == 3 || i == 5) ? 1.0 : 0.0;
    A[i] = Ii - third_I1 * Ci[i];
    B[i] = I1 * Ii - C[i] - two3_I2 * Ci[i];
  }
  // PK2 stress, getS_UP (materials_hyperelastic.cpp:30-62)
  const double ko1 = 2.0 * al[0] * pp;
  const double ko2 = 2.0 * al[1] * pppp;
  const double pJ = p_e * Jv;
  double S[6];
#pragma unroll
  for (int i = 0; i < 6; i++) S[i] = ko1 * A[i] + ko2 * B[i] + pJ * Ci[i];

  if (state) {
#pragma unroll
    for (int i = 0; i < 9; i++) state[i] = O[i];
#pragma unroll
    for (int i = 0; i < 6; i++) state[9 + i] = C[i];
#pragma unroll
    for (int i = 0; i < 6; i++) state[15 + i] = S[i];
  }

  // deviatoric tangent Dd, getDdDp_UP (materials_hyperelastic.cpp:65-129); the
  // symmetrised -C^-1 (x) C^-1 term is IIt_(ij)(kl) = -(Ci_ik Ci_jl + Ci_il Ci_jk)/2.
  const double c11 = 4.0 * al[2] * pppp;
  const double c12 = 4.0 * al[3] * oo;
  const double c22 = 4.0 * al[4] * pppp * pppp;
  const double c1 = (4.0 / 3.0) * al[0] * pp;
  const double c2 = (8.0 / 3.0) * al[1] * pppp;
  double Cf[3][3];
#pragma unroll
  for (int v = 0; v < 6; v++) {
    Cf[kVi[v]][kVj[v]] = Ci[v];
    Cf[kVj[v]][kVi[v]] = Ci[v];
  }
  double Df[6][6];
#pragma unroll
  for (int i = 0; i < 6; i++) {
#pragma unroll
    for (int j = i; j < 6; j++) {
      const int a = kVi[i], b = kVj[i], c = kVi[j], d = kVj[j];
      const double IIt = -0.5 * (Cf[a][c] * Cf[b][d] + Cf[a][d] * Cf[b][c]);
      const double Ii = (i == 0 || i == 3 || i == 5) ? 1.0 : 0.0;
      const double Ij = (j == 0 || j == 3 || j == 5) ? 1.0 : 0.0;
      const double IIij = (i == j) ? ((i == 0 || i == 3 || i == 5) ? 1.0 : 0.5) : 0.0;
      const double CC = Ci[i] * Ci[j];
      double v = c11 * A[i] * A[j] + c12 * (B[i] * A[j] + A[i] * B[j]) + c22 * B[i] * B[j] -
                 c1 * (Ci[i] * A[j] + A[i] * Ci[j] + third_I1 * CC + I1 * IIt) -
                 c2 * (Ci[i] * B[j] + B[i] * Ci[j] + two3_I2 * CC - 1.5 * Ii * Ij + 1.5 * IIij + I2 * IIt) +
                 pJ * (CC + 2.0 * IIt);
      v *= 0.5;
      Df[i][j] = v;
      Df[j][i] = v;
    }
  }

  const double w = det;  // Gauss weight 1 (2x2x2 rule) times det, intWeight (isoparametric.h:303-308)

  // H[k][(e,r)] = sum_s Dd[k][v(r,s)] F[e][s]
  double H[6][9];
#pragma unroll
  for (int k = 0; k < 6; k++)
#pragma unroll
    for (int e = 0; e < 3; e++)
#pragma unroll
      for (int r = 0; r < 3; r++)
        H[k][3 * e + r] = Df[k][voigt(r, 0)] * F[e][0] + Df[k][voigt(r, 1)] * F[e][1] + Df[k][voigt(r, 2)] * F[e][2];

  // At[(d,p),(e,r)] = 2w sum_q F[d][q] H[v(p,q)][(e,r)] + w delta_de S[v(p,r)]
  const double w2 = 2.0 * w;
#pragma unroll
  for (int a = 0; a < 9; a++) {
#pragma unroll
    for (int b = a; b < 9; b++) {
      const int d = a / 3, p = a % 3, e = b / 3, r = b % 3;
      double gsum = F[d][0] * H[voigt(p, 0)][b] + F[d][1] * H[voigt(p, 1)][b] + F[d][2] * H[voigt(p, 2)][b];
      double v = w2 * gsum;
      if (d == e) v += w * S[voigt(p, r)];
      out[kOffA + pk9(a, b)] = v;
    }
  }
#pragma unroll
  for (int n = 0; n < 8; n++)
#pragma unroll
    for (int i = 0; i < 3; i++) out[kOffN + 3 * n + i] = N[n][i];
  // Fu: -1/2 w B^T S ; Kup: 1/2 w B^T Dp with Dp = J C^-1  (SOLID81.cpp:73-77)
#pragma unroll
  for (int d = 0; d < 3; d++)
#pragma unroll
    for (int p = 0; p < 3; p++) {
      out[kOffP + 3 * d + p] =
          -w * (F[d][0] * S[voigt(p, 0)] + F[d][1] * S[voigt(p, 1)] + F[d][2] * S[voigt(p, 2)]);
      out[kOffQ + 3 * d + p] =
          w * Jv * (F[d][0] * Ci[voigt(p, 0)] + F[d][1] * Ci[voigt(p, 1)] + F[d][2] * Ci[voigt(p, 2)]);
    }
  // Fp, Kpp (SOLID81.cpp:79-80)
  out[kOffFp] = -(Jv - 1 - p_e / mat.k) * w;
  out[kOffKpp] = -1.0 / mat.k * w;
}

// The element coordinates and DoF values are needed only until phase A has formed each
// Gauss point's deformation gradient, so the Gauss-point records overlay them: 39.9 KB a
// block instead of 42.6 KB, four blocks (waves) per CU instead of three.
struct Solid81XU {
  double X[kEPB][8][3];
  double U[kEPB][8][3];
};
struct Solid81Shared {
  union {
    Solid81XU xu;
    double G[kEPB][8][kGPW];
  };
  double P[kEPB];
};
static_assert(sizeof(Solid81Shared) <= 40960, "four blocks per CU need at most 40 KB of LDS each");

// One 64-lane block of kEPB elements starting at element e0 (of nE): gather, Gauss
// points (phase A), node-pair blocks (phase B).  Every lane of the block calls it; it
// returns with a barrier passed, so the shared arrays can be reused by the next call.
__device__ __forceinline__ void solid81_block(Solid81Shared& sh, int e0, int nE, const double* __restrict__ xyz,
                                              const int* __restrict__ conn, const int* __restrict__ node_eq,
                                              const int* __restrict__ elem_eq, const double* __restrict__ Uall,
                                              const MaterialParams& mat, double* __restrict__ Ke,
                                              double* __restrict__ gp_state) {
  const int lane = threadIdx.x;
  const int el = lane / kLPE;
  const int r = lane - el * kLPE;
  const int e = e0 + el;
  const bool valid = (el < kEPB) && (e < nE);

  // phase 0: gather coordinates and the element's DoF values
  if (valid) {
    if (r < 8) {
      const int n = conn[(size_t)e * 8 + r];
#pragma unroll
      for (int d = 0; d < 3; d++) {
        sh.xu.X[el][r][d] = xyz[(size_t)n * 3 + d];
        const int q = node_eq[(size_t)n * 3 + d];
        sh.xu.U[el][r][d] = (q >= 0) ? Uall[q] : 0.0;
      }
    } else {
      const int q = elem_eq[e];
      sh.P[el] = (q >= 0) ? Uall[q] : 0.0;
    }
  }
  __syncthreads();

  // phase A: lane r evaluates Gauss point r
  if (valid && r < 8) {
    gauss_point(sh.xu.X[el], sh.xu.U[el], sh.P[el], r, mat, sh.G[el][r],
                gp_state ? gp_state + ((size_t)e * 8 + r) * kSolid81GpState : nullptr);
  }
  __syncthreads();

  // phase B: lane (d,e) accumulates component (d,e) of every node-pair block
  if (valid) {
    const int d = r / 3, ee = r - 3 * (r / 3);
    int ia[3][3];
#pragma unroll
    for (int p = 0; p < 3; p++)
#pragma unroll
      for (int q = 0; q < 3; q++) ia[p][q] = pk9(3 * d + p, 3 * ee + q);
    const int vsel = (ee == 0) ? kOffP : kOffQ;  // lanes e==2 compute a discarded copy of Kup

    double kij[36];
#pragma unroll
    for (int i = 0; i < 36; i++) kij[i] = 0.0;
    double kf[8];
#pragma unroll
    for (int i = 0; i < 8; i++) kf[i] = 0.0;
    double ksc = 0.0;

#pragma unroll 1
    for (int g = 0; g < 8; g++) {
      const double* G = sh.G[el][g];
      double Ab[3][3];
#pragma unroll
      for (int p = 0; p < 3; p++)
#pragma unroll
        for (int q = 0; q < 3; q++) Ab[p][q] = G[ia[p][q]];
      double N[8][3];
#pragma unroll
      for (int n = 0; n < 8; n++)
#pragma unroll
        for (int i = 0; i < 3; i++) N[n][i] = G[kOffN + 3 * n + i];
      const double v0 = G[vsel + 3 * d + 0], v1 = G[vsel + 3 * d + 1], v2 = G[vsel + 3 * d + 2];
      ksc += G[kOffFp + (r & 1)];
#pragma unroll
      for (int j = 0; j < 8; j++) {
        const double w0 = Ab[0][0] * N[j][0] + Ab[0][1] * N[j][1] + Ab[0][2] * N[j][2];
        const double w1 = Ab[1][0] * N[j][0] + Ab[1][1] * N[j][1] + Ab[1][2] * N[j][2];
        const double w2 = Ab[2][0] * N[j][0] + Ab[2][1] * N[j][1] + Ab[2][2] * N[j][2];
#pragma unroll
        for (int i = 0; i <= j; i++) {
          const int slot = (i < j) ? offpair(i, j) : 28 + i;
          kij[slot] += N[i][0] * w0 + N[i][1] * w1 + N[i][2] * w2;
        }
      }
#pragma unroll
      for (int i = 0; i < 8; i++) kf[i] += N[i][0] * v0 + N[i][1] * v1 + N[i][2] * v2;
    }

    double* out = Ke + (size_t)e * kSolid81Slots;
#pragma unroll
    for (int s = 0; s < 28; s++) out[s * 9 + r] = kij[s];
    if (d <= ee) {
      const int up = voigt(d, ee);
#pragma unroll
      for (int i = 0; i < 8; i++) out[252 + 6 * i + up] = kij[28 + i];
    }
    if (ee < 2) {
      const int base = (ee == 0) ? 325 : 300;
#pragma unroll
      for (int i = 0; i < 8; i++) out[base + 3 * i + d] = kf[i];
    }
    if (r == 0) out[349] = ksc;
    if (r == 1) out[324] = ksc;
  }
  __syncthreads();
}

__global__ __launch_bounds__(64) void solid81_kernel(int nE, const double* __restrict__ xyz,
                                                     const int* __restrict__ conn,
                                                     const int* __restrict__ node_eq,
                                                     const int* __restrict__ elem_eq,
                                                     const double* __restrict__ Uall, MaterialParams mat,
                                                     double* __restrict__ Ke, double* __restrict__ gp_state) {
  __shared__ Solid81Shared sh;
  solid81_block(sh, blockIdx.x * kEPB, nE, xyz, conn, node_eq, elem_eq, Uall, mat, Ke, gp_state);
}

// Batched assembly: one persistent grid-stride launch over the element blocks of
// every model of a batch; block b belongs to the model whose block range holds it.
__global__ __launch_bounds__(64) void solid81_batched_kernel(int nModels, const Solid81Batch* __restrict__ mods,
                                                             const int* __restrict__ blockBegin, int nBlocks) {
  __shared__ Solid81Shared sh;
  for (int b = blockIdx.x; b < nBlocks; b += gridDim.x) {
    int lo = 0, hi = nModels - 1;  // last model with blockBegin <= b
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (blockBegin[mid] <= b) lo = mid;
      else hi = mid - 1;
    }
    const Solid81Batch& md = mods[lo];
    solid81_block(sh, (b - blockBegin[lo]) * kEPB, md.nE, md.xyz, md.conn, md.node_eq, md.elem_eq, md.U, md.mat,
                  md.Ke, nullptr);
  }
}

}  // namespace

void launch_solid81(hipStream_t stream, int nE, const double* xyz, const int* conn, const int* node_eq,
                    const int* elem_eq, const double* U, const MaterialParams& mat, double* Ke,
                    double* gp_state) {
  if (nE <= 0) return;
  const int blocks = (nE + kEPB - 1) / kEPB;
  hipLaunchKernelGGL(solid81_kernel, dim3(blocks), dim3(64), 0, stream, nE, xyz, conn, node_eq, elem_eq, U,
                     mat, Ke, gp_state);
}

int solid81_blocks(int nE) { return (nE + kEPB - 1) / kEPB; }

void launch_solid81_batched(hipStream_t stream, int grid, int nModels, const Solid81Batch* mods, const int* blockBegin,
                            int nBlocks) {
  if (nBlocks <= 0) return;
  hipLaunchKernelGGL(solid81_batched_kernel, dim3(grid), dim3(64), 0, stream, nModels, mods, blockBegin, nBlocks);
}

void solid81_slot_to_assemble3(int* map) {
  // packed MatSym<24> index of (a,b), a<=b (Mat.h:591-598)
  auto pk24 = [](int a, int b) { return 24 * a - (a * (a - 1)) / 2 + (b - a); };
  for (int j = 1; j < 8; j++)
    for (int i = 0; i < j; i++)
      for (int d = 0; d < 3; d++)
        for (int e = 0; e < 3; e++) map[offpair(i, j) * 9 + 3 * d + e] = pk24(3 * i + d, 3 * j + e);
  for (int i = 0; i < 8; i++)
    for (int d = 0; d < 3; d++)
      for (int e = d; e < 3; e++) map[252 + 6 * i + voigt(d, e)] = pk24(3 * i + d, 3 * i + e);
  for (int k = 300; k < 350; k++) map[k] = k;
}

}  // namespace nla3d_amd
