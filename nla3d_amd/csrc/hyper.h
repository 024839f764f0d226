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

}  // namespace nla3d_amd
