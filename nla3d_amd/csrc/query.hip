// Element result queries of the u/p solids on the device: Element::getScalar / getVector /
// getTensor (elements/element.h:122-131) of SOLID81 (elements/SOLID81.cpp:191-345) and
// PLANE41 (elements/PLANE41.cpp:153-300), at one Gauss point or averaged over the element
// (GP_MEAN: each Gauss point weighted by intWeight / volume, element.h:122-131).
//
// The Gauss-point state the reference keeps from update() (O, C, S) is recomputed from the
// current solution, as the element kernels do.  The results follow the reference's
// arithmetic, including what it does beyond the mechanics:
//   * SOLID81 COUCHY and PK2 add p J C^-1 to S, which already holds it (SOLID81.cpp:306-323),
//   * SOLID81 WU evaluates W at the invariants of C itself, I3 = J (SOLID81.cpp:215-219),
//   * PLANE41 COUCHY forms F^T S F / J with F = I + grad u and ignores the weight, so GP_MEAN
//     returns the sum over the four Gauss points (PLANE41.cpp:245-265).
// One thread per element; a query is output, not part of the Newton iteration.

#include "hyper.h"
#include "internal.h"

namespace nla3d_amd {
namespace {

constexpr double kGP = 0.5773502691896258;  // isoparametric.h:128-136
__device__ constexpr int kHx[8] = {-1, 1, 1, -1, -1, 1, 1, -1};  // isoparametric.cpp:197-208
__device__ constexpr int kHy[8] = {-1, -1, 1, 1, -1, -1, 1, 1};
__device__ constexpr int kHz[8] = {-1, -1, -1, -1, 1, 1, 1, 1};

// Gauss point g of the hexahedron: det J and dN/dX (ElementIsoParamHEXAHEDRON::makeJacob)
__device__ void hex_point(const double (*X)[3], int g, double* det, double (*N)[3]) {
  const double gx = (g & 1) ? kGP : -kGP, gy = (g & 2) ? kGP : -kGP, gz = (g & 4) ? kGP : -kGP;
  double dN[8][3], J[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
  for (int n = 0; n < 8; n++) {
    const double ax = 1.0 + kHx[n] * gx, ay = 1.0 + kHy[n] * gy, az = 1.0 + kHz[n] * gz;
    dN[n][0] = (0.125 * kHx[n]) * ay * az;
    dN[n][1] = (0.125 * kHy[n]) * ax * az;
    dN[n][2] = (0.125 * kHz[n]) * ax * ay;
    for (int i = 0; i < 3; i++)
      for (int j = 0; j < 3; j++) J[i][j] += dN[n][i] * X[n][j];
  }
  const double d = J[0][0] * (J[1][1] * J[2][2] - J[1][2] * J[2][1]) - J[0][1] * (J[1][0] * J[2][2] - J[1][2] * J[2][0]) +
                   J[0][2] * (J[1][0] * J[2][1] - J[1][1] * J[2][0]);
  const double r = 1.0 / d;
  double Ji[3][3];
  Ji[0][0] = (J[1][1] * J[2][2] - J[1][2] * J[2][1]) * r;
  Ji[0][1] = (J[0][2] * J[2][1] - J[0][1] * J[2][2]) * r;
  Ji[0][2] = (J[0][1] * J[1][2] - J[0][2] * J[1][1]) * r;
  Ji[1][0] = (J[2][0] * J[1][2] - J[1][0] * J[2][2]) * r;
  Ji[1][1] = (J[0][0] * J[2][2] - J[0][2] * J[2][0]) * r;
  Ji[1][2] = (J[1][0] * J[0][2] - J[0][0] * J[1][2]) * r;
  Ji[2][0] = (J[1][0] * J[2][1] - J[1][1] * J[2][0]) * r;
  Ji[2][1] = (J[0][1] * J[2][0] - J[0][0] * J[2][1]) * r;
  Ji[2][2] = (J[0][0] * J[1][1] - J[0][1] * J[1][0]) * r;
  for (int n = 0; n < 8; n++)
    for (int i = 0; i < 3; i++) N[n][i] = Ji[i][0] * dN[n][0] + Ji[i][1] * dN[n][1] + Ji[i][2] * dN[n][2];
  *det = d;
}

// Gauss point g of the quadrilateral (ElementIsoParamQUAD::makeJacob, isoparametric.cpp:55-106)
__device__ void quad_point(const double (*X)[3], int g, double* det, double (*N)[3]) {
  const double r = (g & 1) ? kGP : -kGP, s = (g & 2) ? kGP : -kGP;
  const double dN[4][2] = {{-0.25 * (1.0 - s), -0.25 * (1.0 - r)},
                           {0.25 * (1.0 - s), -0.25 * (1.0 + r)},
                           {0.25 * (1.0 + s), 0.25 * (1.0 + r)},
                           {-0.25 * (1.0 + s), 0.25 * (1.0 - r)}};
  double J00 = 0, J01 = 0, J10 = 0, J11 = 0;
  for (int a = 0; a < 4; a++) {
    J00 += dN[a][0] * X[a][0];
    J01 += dN[a][0] * X[a][1];
    J10 += dN[a][1] * X[a][0];
    J11 += dN[a][1] * X[a][1];
  }
  const double d = J00 * J11 - J01 * J10, rd = 1.0 / d;
  for (int a = 0; a < 4; a++) {
    N[a][0] = (J11 * rd) * dN[a][0] + (-J01 * rd) * dN[a][1];
    N[a][1] = (-J10 * rd) * dN[a][0] + (J00 * rd) * dN[a][1];
    N[a][2] = 0.0;
  }
  *det = d;
}

// acc += c * B^T D B for a 3x3 B and symmetric D (Voigt order), upper result in Voigt order
// (matBTDBprod, math/Mat.h:687-741)
__device__ void btdb3(const double (*B)[3], const double* D, double c, double* acc) {
  constexpr int vi[6] = {0, 0, 0, 1, 1, 2}, vj[6] = {0, 1, 2, 1, 2, 2};
  double Df[3][3];
  for (int v = 0; v < 6; v++) Df[vi[v]][vj[v]] = Df[vj[v]][vi[v]] = D[v];
  for (int v = 0; v < 6; v++) {
    const int i = vi[v], j = vj[v];
    double sum = 0.0;
    for (int k = 0; k < 3; k++)
      for (int l = 0; l < 3; l++) sum += B[k][i] * Df[k][l] * B[l][j];
    acc[v] += sum * c;
  }
}

template <bool kHex>
__global__ __launch_bounds__(128) void element_query_kernel(int nE, const double* __restrict__ xyz,
                                                            const int* __restrict__ conn,
                                                            const int* __restrict__ node_eq,
                                                            const int* __restrict__ elem_eq,
                                                            const double* __restrict__ Uall, MaterialParams mat,
                                                            int kind, int query, int gp, double* __restrict__ out) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nE) return;
  constexpr int NN = kHex ? 8 : 4;  // nodes = Gauss points
  constexpr int ND = kHex ? 3 : 2;
  double X[NN][3], u[NN][3];
  for (int a = 0; a < NN; a++) {
    const int n = conn[(size_t)e * NN + a];
    for (int d = 0; d < 3; d++) {
      X[a][d] = xyz[(size_t)n * 3 + d];
      const int q = d < ND ? node_eq[(size_t)n * 3 + d] : -1;
      u[a][d] = q >= 0 ? Uall[q] : 0.0;
    }
  }
  const int qp = elem_eq[e];
  const double p = qp >= 0 ? Uall[qp] : 0.0;

  double det[NN], N[NN][3], vol = 0.0;
  for (int g = 0; g < NN; g++) {
    if (kHex)
      hex_point(X, g, &det[g], N);
    else
      quad_point(X, g, &det[g], N);
    vol += det[g];  // volume(): sum of intWeight, Gauss weights 1
  }
  const int width = kind == 0 ? 1 : (kind == 1 ? 3 : 6);
  double acc[6] = {0, 0, 0, 0, 0, 0};
  const int g0 = gp < 0 ? 0 : gp, g1 = gp < 0 ? NN : gp + 1;
  for (int g = g0; g < g1; g++) {
    const double scale = gp < 0 ? det[g] / vol : 1.0;
    double dg;
    if (kHex)
      hex_point(X, g, &dg, N);
    else
      quad_point(X, g, &dg, N);
    // F[d][q] = delta + du_d / dX_q
    double F[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
    for (int a = 0; a < NN; a++)
      for (int d = 0; d < ND; d++)
        for (int q = 0; q < ND; q++) F[d][q] += N[a][q] * u[a][d];
    // C = F^T F (SOLID81.cpp:102-107, PLANE41.cpp:144-146; C_ZZ = 1 and C_XZ = C_YZ = 0 in the plane)
    double C[6];
    {
      constexpr int vi[6] = {0, 0, 0, 1, 1, 2}, vj[6] = {0, 1, 2, 1, 2, 2};
      for (int v = 0; v < 6; v++) C[v] = F[0][vi[v]] * F[0][vj[v]] + F[1][vi[v]] * F[1][vj[v]] + F[2][vi[v]] * F[2][vj[v]];
    }
    double I1, I2, J, Ci[6];
    c_invariants(C, I1, I2, J, Ci);
    if (kind == 0) {
      if (query == 1) acc[0] += p * scale;                                          // SP
      else if (query == 3) acc[0] += strain_energy(mat, I1, I2) * scale;            // WU
      else if (query == 4) acc[0] += 0.5 * mat.k * (J - 1.0) * (J - 1.0) * scale;   // WP
      else if (query == 5) acc[0] += vol * scale;                                   // VOL
    } else if (kind == 1) {  // IC
      acc[0] += I1 * scale;
      acc[1] += I2 * scale;
      acc[2] += J * scale;
    } else if (query == 3 || query == 4) {  // E, C
      for (int v = 0; v < 6; v++) {
        const double dia = (v == 0 || v == 3 || v == 5) ? 1.0 : 0.0;
        acc[v] += (query == 4 ? C[v] : (C[v] - dia) * 0.5) * scale;
      }
    } else {  // COUCHY, PK2
      double S[6];
      pk2_up(mat, C, p, S);
      if (kHex)
        for (int v = 0; v < 6; v++) S[v] += p * J * Ci[v];
      if (query == 2) {
        for (int v = 0; v < 6; v++) acc[v] += S[v] * scale;
      } else if (kHex) {  // sigma = F S F^T / J: matBTDBprod with B = F^T
        double Ft[3][3];
        for (int i = 0; i < 3; i++)
          for (int j = 0; j < 3; j++) Ft[i][j] = F[j][i];
        btdb3(Ft, S, scale / J, acc);
      } else {  // F^T S F / J, weight not applied
        btdb3(F, S, 1.0 / J, acc);
      }
    }
  }
  for (int k = 0; k < width; k++) out[(size_t)e * width + k] = acc[k];
}

}  // namespace

void launch_element_query(hipStream_t stream, bool hex, int nE, const double* xyz, const int* conn, const int* node_eq,
                          const int* elem_eq, const double* U, const MaterialParams& mat, int kind, int query, int gp,
                          double* out) {
  if (nE <= 0) return;
  const dim3 grid((nE + 127) / 128), block(128);
  if (hex)
    hipLaunchKernelGGL(element_query_kernel<true>, grid, block, 0, stream, nE, xyz, conn, node_eq, elem_eq, U, mat,
                       kind, query, gp, out);
  else
    hipLaunchKernelGGL(element_query_kernel<false>, grid, block, 0, stream, nE, xyz, conn, node_eq, elem_eq, U, mat,
                       kind, query, gp, out);
}

}  // namespace nla3d_amd
