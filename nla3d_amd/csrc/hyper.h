// Device helpers shared by the u/p solid element kernels (solid81.hip, plane41.hip).
#pragma once

#include "internal.h"

namespace nla3d_amd {

// W_first/second_derivatives of the three materials (materials_hyperelastic.cpp:135-198).
// alpha = {a1, a2, a11, a12, a22}
__device__ __forceinline__ void material_alpha(const MaterialParams& m, double I1bar, double* al) {
  if (m.type == 1) {  // neo-Hookean: G, K
    al[0] = 0.5 * m.mc[0];
    al[1] = 0.0;
    al[2] = 0.0;
  } else if (m.type == 2) {  // Biderman: C10 C20 C30 C01 K
    double x = I1bar - 3.0;
    al[0] = m.mc[0] + 2.0 * m.mc[1] * x + 3.0 * m.mc[2] * x * x;
    al[1] = m.mc[3];
    al[2] = 2.0 * m.mc[1] + 6.0 * m.mc[2] * x;
  } else {  // Mooney-Rivlin: C10 C01 K
    al[0] = m.mc[0];
    al[1] = m.mc[1];
    al[2] = 0.0;
  }
  al[3] = 0.0;
  al[4] = 0.0;
}

// invariants and inverse of a symmetric C (Voigt XX XY XZ YY YZ ZZ; solidmech.cpp:10-42):
// I1, I2, J = sqrt(det C) and C^-1
__device__ __forceinline__ void c_invariants(const double* C, double& I1, double& I2, double& J, double* Ci) {
  I1 = C[0] + C[3] + C[5];
  I2 = C[0] * C[3] + C[3] * C[5] + C[0] * C[5] - C[1] * C[1] - C[4] * C[4] - C[2] * C[2];
  J = sqrt(C[0] * (C[3] * C[5] - C[4] * C[4]) - C[1] * (C[1] * C[5] - C[4] * C[2]) +
           C[2] * (C[1] * C[4] - C[3] * C[2]));
  const double oo = 1.0 / (J * J);
  Ci[0] = oo * (C[3] * C[5] - C[4] * C[4]);
  Ci[1] = oo * (C[2] * C[4] - C[1] * C[5]);
  Ci[2] = oo * (C[1] * C[4] - C[2] * C[3]);
  Ci[3] = oo * (C[0] * C[5] - C[2] * C[2]);
  Ci[4] = oo * (C[1] * C[2] - C[0] * C[4]);
  Ci[5] = oo * (C[0] * C[3] - C[1] * C[1]);
}

// PK2 stress of the u/p law, all six components (getS_UP, materials_hyperelastic.cpp:30-62)
__device__ __forceinline__ void pk2_up(const MaterialParams& m, const double* C, double p, double* S) {
  double I1, I2, J, Ci[6];
  c_invariants(C, I1, I2, J, Ci);
  const double pp = pow(J, -2.0 / 3.0);
  double al[5];
  material_alpha(m, I1 * pp, al);
  const double ko1 = 2.0 * al[0] * pp, ko2 = 2.0 * al[1] * pp * pp;
#pragma unroll
  for (int i = 0; i < 6; i++) {
    const double Ii = (i == 0 || i == 3 || i == 5) ? 1.0 : 0.0;
    S[i] = ko1 * (Ii - (1.0 / 3.0) * I1 * Ci[i]) + ko2 * (I1 * Ii - C[i] - (2.0 / 3.0) * I2 * Ci[i]) + p * J * Ci[i];
  }
}

// strain energy W(I1, I2, I3) of the three materials (materials_hyperelastic.cpp:150-200)
__device__ __forceinline__ double strain_energy(const MaterialParams& m, double I1, double I2) {
  const double x = I1 - 3.0;
  if (m.type == 1) return 0.5 * m.mc[0] * x;
  if (m.type == 2) return m.mc[0] * x + m.mc[1] * x * x + m.mc[2] * x * x * x + m.mc[3] * (I2 - 3.0);
  return m.mc[0] * x + m.mc[1] * (I2 - 3.0);
}

}  // namespace nla3d_amd
