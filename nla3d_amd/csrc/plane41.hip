// PLANE41 fused update() + buildK() on gfx950, FP64: the plane-strain u/p quadrilateral.
//
// Reference: ElementPLANE41::update (elements/PLANE41.cpp:136-149) followed by
// ElementPLANE41::buildK (elements/PLANE41.cpp:31-103), the material
// Mat_Hyper_Isotrop_General::getS_UP / getDdDp_UP (materials/materials_hyperelastic.cpp:30-129)
// evaluated at C = [[Cxx, Cxy, 0], [Cxy, Cyy, 0], [0, 0, 1]], of which the element keeps the
// components XX, YY, XY (PLANE41.cpp:11), and the quadrilateral Jacobian
// ElementIsoParamQUAD::makeJacob (elements/isoparametric.cpp:55-106).  As for SOLID81,
// update() then buildK() at the same U is a pure function of (X, U, p), so the Gauss-point
// state is recomputed rather than stored.
//
// One thread per element: a plane element is small (8 + 1 DoFs, 4 Gauss points), so its
// 8x8 Kuu, Kup, Qe, Kpp and Fp stay in registers and the element needs no cross-lane
// exchange.  Outputs are slot-major, Ke[s * nE + e], so that each of the 54 stores of a wave is
// one contiguous 512-byte line.  Slot order is the reference's assembleK order
// (PLANE41.h:55-89): node pairs i <= j with (di, dj) over UX UY, skipping dj > di when
// i == j (36 slots), then Kup of the 8 node DoFs, Kpp, then Fe = [-Qe, -Fp] (9 slots).
//
// The tangent is Kuu = sum_gp (2 B^T E B + Bw^T S Bw) det with B = B_L + Omega Bw
// (PLANE41.cpp:63-71); written out per node pair (a, b) and DoF pair (d, e) with
// N = dN/dX, F = I + grad u (2x2) this is
//   Kuu[(a,d),(b,e)] = det [ 2 sum_{k,l} B_k(a,d) E_kl B_l(b,e) + delta_de (N_a . S N_b) ]
// where B_XX(a,d) = F_d0 N_a0, B_YY(a,d) = F_d1 N_a1, B_XY(a,d) = F_d0 N_a1 + F_d1 N_a0,
// which is the reference's B row by row.

#include "hyper.h"
#include "internal.h"

namespace nla3d_amd {
namespace {

constexpr double kGP2 = 0.5773502691896258;  // isoparametric.h:128-136
// packed upper index of a symmetric 8x8
__host__ __device__ constexpr int pk8(int a, int b) {
  return (a > b) ? (8 * b - (b * (b - 1)) / 2 + (a - b)) : (8 * a - (a * (a - 1)) / 2 + (b - a));
}

__global__ __launch_bounds__(256) void plane41_kernel(int nE, const double* __restrict__ xyz,
                                                      const int* __restrict__ conn, const int* __restrict__ node_eq,
                                                      const int* __restrict__ elem_eq, const double* __restrict__ Uall,
                                                      MaterialParams mat, double* __restrict__ Ke) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nE) return;
  double X[4][2], u[4][2];
#pragma unroll
  for (int a = 0; a < 4; a++) {
    const int n = conn[(size_t)e * 4 + a];
#pragma unroll
    for (int d = 0; d < 2; d++) {
      X[a][d] = xyz[(size_t)n * 3 + d];
      const int q = node_eq[(size_t)n * 3 + d];
      u[a][d] = q >= 0 ? Uall[q] : 0.0;
    }
  }
  const int qp = elem_eq[e];
  const double p_e = qp >= 0 ? Uall[qp] : 0.0;

  double K[36], Kup[8], Qe[8];
#pragma unroll
  for (int i = 0; i < 36; i++) K[i] = 0.0;
#pragma unroll
  for (int i = 0; i < 8; i++) Kup[i] = Qe[i] = 0.0;
  double Kpp = 0.0, Fp = 0.0;

#pragma unroll 1
  for (int g = 0; g < 4; g++) {
    // Gauss point order r fastest (isoparametric.h:82-87)
    const double r = (g & 1) ? kGP2 : -kGP2;
    const double s = (g & 2) ? kGP2 : -kGP2;
    // dN/dxi of the bilinear quadrilateral (isoparametric.cpp:123-128)
    const double dN[4][2] = {{-0.25 * (1.0 - s), -0.25 * (1.0 - r)},
                             {0.25 * (1.0 - s), -0.25 * (1.0 + r)},
                             {0.25 * (1.0 + s), 0.25 * (1.0 + r)},
                             {-0.25 * (1.0 + s), 0.25 * (1.0 - r)}};
    double J00 = 0.0, J01 = 0.0, J10 = 0.0, J11 = 0.0;
#pragma unroll
    for (int a = 0; a < 4; a++) {
      J00 += dN[a][0] * X[a][0];
      J01 += dN[a][0] * X[a][1];
      J10 += dN[a][1] * X[a][0];
      J11 += dN[a][1] * X[a][1];
    }
    const double det = J00 * J11 - J01 * J10;
    const double rd = 1.0 / det;
    const double i00 = J11 * rd, i01 = -J01 * rd, i10 = -J10 * rd, i11 = J00 * rd;
    double N[4][2];
    double O[4] = {0.0, 0.0, 0.0, 0.0};  // dux/dx dux/dy duy/dx duy/dy (PLANE41.cpp:140-143)
#pragma unroll
    for (int a = 0; a < 4; a++) {
      N[a][0] = i00 * dN[a][0] + i01 * dN[a][1];
      N[a][1] = i10 * dN[a][0] + i11 * dN[a][1];
      O[0] += N[a][0] * u[a][0];
      O[1] += N[a][1] * u[a][0];
      O[2] += N[a][0] * u[a][1];
      O[3] += N[a][1] * u[a][1];
    }
    const double F[2][2] = {{1.0 + O[0], O[1]}, {O[2], 1.0 + O[3]}};
    // C (PLANE41.cpp:144-146) in the 6-component Voigt order XX XY XZ YY YZ ZZ
    const double Cxx = 1.0 + 2 * O[0] + (O[0] * O[0] + O[2] * O[2]);
    const double Cyy = 1.0 + 2 * O[3] + (O[3] * O[3] + O[1] * O[1]);
    const double Cxy = O[1] + O[2] + O[0] * O[1] + O[2] * O[3];
    const double C[6] = {Cxx, Cxy, 0.0, Cyy, 0.0, 1.0};

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

    // the three kept components, Voigt indices XX = 0, YY = 3, XY = 1, and their tensor indices
    constexpr int kc[3] = {0, 3, 1};
    constexpr int ti[3] = {0, 1, 0}, tj[3] = {0, 1, 1};
    double A[3], Bv[3], Cik[3], S[3];
    const double ko1 = 2.0 * al[0] * pp;
    const double ko2 = 2.0 * al[1] * pppp;
    const double pJ = p_e * Jv;
#pragma unroll
    for (int k = 0; k < 3; k++) {
      const double Ii = (k < 2) ? 1.0 : 0.0;
      Cik[k] = Ci[kc[k]];
      A[k] = Ii - third_I1 * Cik[k];
      Bv[k] = I1 * Ii - C[kc[k]] - two3_I2 * Cik[k];
      S[k] = ko1 * A[k] + ko2 * Bv[k] + pJ * Cik[k];  // getS_UP (materials_hyperelastic.cpp:30-62)
    }
    // in-plane C^-1 as a 2x2 tensor (its XZ and YZ terms are zero)
    const double Cf[2][2] = {{Ci[0], Ci[1]}, {Ci[1], Ci[3]}};
    // Dd restricted to the kept components (materials_hyperelastic.cpp:65-129)
    const double c11 = 4.0 * al[2] * pppp;
    const double c12 = 4.0 * al[3] * oo;
    const double c22 = 4.0 * al[4] * pppp * pppp;
    const double c1 = (4.0 / 3.0) * al[0] * pp;
    const double c2 = (8.0 / 3.0) * al[1] * pppp;
    double E[3][3];
#pragma unroll
    for (int k = 0; k < 3; k++)
#pragma unroll
      for (int l = k; l < 3; l++) {
        const double IIt = -0.5 * (Cf[ti[k]][ti[l]] * Cf[tj[k]][tj[l]] + Cf[ti[k]][tj[l]] * Cf[tj[k]][ti[l]]);
        const double Ik = (k < 2) ? 1.0 : 0.0, Il = (l < 2) ? 1.0 : 0.0;
        const double IIkl = (k == l) ? ((k < 2) ? 1.0 : 0.5) : 0.0;
        const double CC = Cik[k] * Cik[l];
        double v = c11 * A[k] * A[l] + c12 * (Bv[k] * A[l] + A[k] * Bv[l]) + c22 * Bv[k] * Bv[l] -
                   c1 * (Cik[k] * A[l] + A[k] * Cik[l] + third_I1 * CC + I1 * IIt) -
                   c2 * (Cik[k] * Bv[l] + Bv[k] * Cik[l] + two3_I2 * CC - 1.5 * Ik * Il + 1.5 * IIkl + I2 * IIt) +
                   pJ * (CC + 2.0 * IIt);
        v *= 0.5;
        E[k][l] = v;
        E[l][k] = v;
      }
    const double w = det;  // Gauss weight 1 (2x2 rule) times det

    // B columns: Bc[a][d][k], k over XX YY XY (make_B + Omega make_Bomega, PLANE41.cpp:106-118)
    double Bc[4][2][3];
#pragma unroll
    for (int a = 0; a < 4; a++)
#pragma unroll
      for (int d = 0; d < 2; d++) {
        Bc[a][d][0] = F[d][0] * N[a][0];
        Bc[a][d][1] = F[d][1] * N[a][1];
        Bc[a][d][2] = F[d][0] * N[a][1] + F[d][1] * N[a][0];
      }
    // G = E B (3 x 8), then Kuu += 2w B^T G + w delta_de N_a . S N_b
    double G[8][3];
#pragma unroll
    for (int c = 0; c < 8; c++)
#pragma unroll
      for (int k = 0; k < 3; k++)
        G[c][k] = E[k][0] * Bc[c / 2][c % 2][0] + E[k][1] * Bc[c / 2][c % 2][1] + E[k][2] * Bc[c / 2][c % 2][2];
    double SN[4][2];  // S N_b with S = [[Sxx, Sxy], [Sxy, Syy]]
#pragma unroll
    for (int b = 0; b < 4; b++) {
      SN[b][0] = S[0] * N[b][0] + S[2] * N[b][1];
      SN[b][1] = S[2] * N[b][0] + S[1] * N[b][1];
    }
    const double w2 = 2.0 * w;
#pragma unroll
    for (int r = 0; r < 8; r++)
#pragma unroll
      for (int c = r; c < 8; c++) {
        const int a = r / 2, d = r % 2, b = c / 2, ee = c % 2;
        double v = w2 * (Bc[a][d][0] * G[c][0] + Bc[a][d][1] * G[c][1] + Bc[a][d][2] * G[c][2]);
        if (d == ee) v += w * (N[a][0] * SN[b][0] + N[a][1] * SN[b][1]);
        K[pk8(r, c)] += v;
      }
    // Qe = B^T S w, Kup = B^T Dp w with Dp = J C^-1 (PLANE41.cpp:79-83)
#pragma unroll
    for (int r = 0; r < 8; r++) {
      const int a = r / 2, d = r % 2;
      Qe[r] += (Bc[a][d][0] * S[0] + Bc[a][d][1] * S[1] + Bc[a][d][2] * S[2]) * w;
      Kup[r] += (Bc[a][d][0] * Cik[0] + Bc[a][d][1] * Cik[1] + Bc[a][d][2] * Cik[2]) * Jv * w;
    }
    Fp += (Jv - 1 - p_e / mat.k) * w;
    Kpp -= 1.0 / mat.k * w;
  }

  // assembleK order (PLANE41.h:55-89)
  double* out = Ke + e;
  const size_t st = (size_t)nE;
  int slot = 0;
#pragma unroll
  for (int i = 0; i < 4; i++)
#pragma unroll
    for (int j = i; j < 4; j++)
#pragma unroll
      for (int di = 0; di < 2; di++)
#pragma unroll
        for (int dj = 0; dj < 2; dj++) {
          if (i == j && dj > di) continue;
          out[st * slot++] = K[pk8(2 * i + di, 2 * j + dj)];
        }
#pragma unroll
  for (int r = 0; r < 8; r++) out[st * (36 + r)] = Kup[r];
  out[st * 44] = Kpp;
#pragma unroll
  for (int r = 0; r < 8; r++) out[st * (45 + r)] = -Qe[r];
  out[st * 53] = -Fp;
}

}  // namespace

void launch_plane41(hipStream_t stream, int nE, const double* xyz, const int* conn, const int* node_eq,
                    const int* elem_eq, const double* U, const MaterialParams& mat, double* Ke) {
  if (nE <= 0) return;
  hipLaunchKernelGGL(plane41_kernel, dim3((nE + 255) / 256), dim3(256), 0, stream, nE, xyz, conn, node_eq, elem_eq, U,
                     mat, Ke);
}

void plane41_slot_pairs(int* la, int* lb) {
  int s = 0;
  for (int i = 0; i < 4; i++)
    for (int j = i; j < 4; j++)
      for (int di = 0; di < 2; di++)
        for (int dj = 0; dj < 2; dj++) {
          if (i == j && dj > di) continue;
          la[s] = 2 * i + di;
          lb[s] = 2 * j + dj;
          s++;
        }
  for (int r = 0; r < 8; r++, s++) { la[s] = r; lb[s] = 8; }
  la[s] = 8; lb[s] = 8; s++;
  for (int r = 0; r < 8; r++, s++) { la[s] = r; lb[s] = -1; }
  la[s] = 8; lb[s] = -1;
}

}  // namespace nla3d_amd
