// Internal declarations shared by the HIP sources of libnla3d_gpu.so.
// Not part of the C ABI (that is include/nla3d_gpu.h).
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <string>
#include <vector>

namespace nla3d_amd {

// Material constants as they cross into device code: a type tag (the
// MaterialFactory id) and the MC[] array of the reference Material.
struct MaterialParams {
  int type;       // 1 neo-Hookean (G, K), 2 Biderman (C10 C20 C30 C01 K), 3 Mooney-Rivlin (C10 C01 K)
  double mc[5];
  double k;       // bulk modulus, getK()
};

// Number of per-element values produced by the SOLID81 kernel, in the
// library's slot order (see solid81_slot_to_assemble3).
constexpr int kSolid81Slots = 350;
constexpr int kSolid81GpState = 21;  // O[9], C[6], S[6] per Gauss point

// Kernels (launchers). All run on `stream`.
// xyz[nN*3], conn[nE*8] 0-based, node_eq[nN*3] / elem_eq[nE] 0-based
// equation index (-1 = DoF not present), U[nAll].
void launch_solid81(hipStream_t stream, int nE, const double* xyz, const int* conn, const int* node_eq,
                    const int* elem_eq, const double* U, const MaterialParams& mat, double* Ke,
                    double* gp_state);

// PLANE41 (plane-strain u/p quadrilateral): 54 values per element, slot-major
// (Ke[s * nE + e]) in the reference's assembleK order; conn[nE*4], node_eq as above (the UX
// and UY entries are used).  plane41_slot_pairs gives each slot's local DoF pair (0..7 node
// DoFs 2a+d, 8 the pressure; lb = -1 for the 9 force slots).
constexpr int kPlane41Slots = 54;
void launch_plane41(hipStream_t stream, int nE, const double* xyz, const int* conn, const int* node_eq,
                    const int* elem_eq, const double* U, const MaterialParams& mat, double* Ke);
void plane41_slot_pairs(int* la, int* lb);

// Element result queries (query.hip): per u/p element, kind 0 scalar / 1 vector / 2 tensor with
// the reference's query numbers; gp = Gauss point or -1 (GP_MEAN); out[nE * width],
// width 1 / 3 / 6.  conn holds 8 (hex) or 4 nodes per element.
void launch_element_query(hipStream_t stream, bool hex, int nE, const double* xyz, const int* conn, const int* node_eq,
                          const int* elem_eq, const double* U, const MaterialParams& mat, int kind, int query, int gp,
                          double* out);

// One model of a batched assembly (device pointers of that model's buffers).
struct Solid81Batch {
  int nE;
  const double* xyz;
  const int* conn;
  const int* node_eq;
  const int* elem_eq;
  const double* U;
  MaterialParams mat;
  double* Ke;
};
int solid81_blocks(int nE);  // 64-lane blocks of a model's elements
// persistent grid-stride element kernel over the blocks of every model: blockBegin[m] is
// model m's first block, nBlocks the total; grid = number of resident workgroups
void launch_solid81_batched(hipStream_t stream, int grid, int nModels, const Solid81Batch* mods, const int* blockBegin,
                            int nBlocks);

struct GatherBatch {
  int nTargets;
  const int* seg;
  const int* src;
  const double* src_vals;
  double* out;
  int mirBegin, mirEnd;
  const int* mir;
  double* full;
};
// the segmented gather of every model in one launch; targetBegin[m] is model m's first target
void launch_gather_batched(hipStream_t stream, int threads, int nModels, const GatherBatch* mods,
                           const int* targetBegin, int nTargets);

// Segmented reduction: out[t] = sum_{s in [seg[t], seg[t+1])} src_vals[src[s]], summed in s order
// (starting from +0.0).  int32 offsets: the host checks the total count fits.
// Targets in [mirBegin, mirEnd) are also stored at full[mir[2k]] and full[mir[2k+1]],
// k = t - mirBegin (full may be null).
void launch_segmented_gather(hipStream_t stream, int nTargets, const int* seg, const int* src,
                             const double* src_vals, double* out, int mirBegin, int mirEnd, const int* mir,
                             double* full);
// full[mir[2k]] = full[mir[2k+1]] = upper2[k] for k = idx[0..n) (idx null: k = 0..n)
void launch_mirror(hipStream_t stream, int n, const double* upper2, const int* idx, const int* mir, double* full);
// constant MPC rows (FEStorage.cpp:51-61): row r adds coef[k] at vals[tgt[k]] for its terms
// k in [termPtr[r], termPtr[r+1]) in order, and sets vals[fTgt[r]] = b[r]
void launch_ce_mpc(hipStream_t stream, int nRows, const int* termPtr, const int* tgt, const double* coef,
                   const int* fTgt, const double* b, double* vals);

// One RigidBodyMpc collection on the device (Mpc.h:49-60).
struct RigidBodyDev {
  int masterEq[6];      // 0-based equation index of the master UX..ROTZ (-1 absent)
  double masterPos[3];
  int nRows;            // slaves * dofs
  int nDofs;
  int dofs[3];          // 0..2 = UX..UZ
};
void launch_rigid_body_mpc(hipStream_t stream, const RigidBodyDev& rb, const double* xyz, const int* slaveNode,
                           const int* slaveEq, const double* U, const int* tgt, const int* fTgt, double* vals);

// plain copy b[i] = a[i], i < n (byte-counter calibration)
void launch_calibration_copy(hipStream_t s, size_t n, const double* a, double* b);

// Library slot k (0..349) of the SOLID81 kernel output -> index in the
// reference assemble3 order (SOLID81.h:54-94).
void solid81_slot_to_assemble3(int* map350);

// Least squares by the normal equations for the predicted starts: given G[i][k] = y_i . y_k
// (k <= i read) and g[i] = y_i . r for m <= kMaxLsBasis vectors, the a minimising
// ||r - sum_i a_i y_i||_2.  Cholesky of G row by row, stopping at the first vector numerically
// dependent on the earlier ones (so the earlier ones should be the more important); returns how
// many vectors are used, the a_i beyond them are 0.
constexpr int kMaxLsBasis = 4;
inline int lsNormalSolve(const double G[kMaxLsBasis][kMaxLsBasis], const double* g, int m, double* a) {
  double L[kMaxLsBasis][kMaxLsBasis] = {};
  int r = 0;
  for (; r < m; r++) {
    double d = G[r][r];
    for (int k = 0; k < r; k++) {
      double t = G[r][k];
      for (int q = 0; q < k; q++) t -= L[r][q] * L[k][q];
      L[r][k] = t / L[k][k];
      d -= L[r][k] * L[r][k];
    }
    if (!(d > 1e-12 * G[r][r])) break;
    L[r][r] = std::sqrt(d);
  }
  for (int i = 0; i < kMaxLsBasis; i++) a[i] = 0.0;
  for (int i = 0; i < r; i++) {  // L y = g, then L^T a = y
    double t = g[i];
    for (int k = 0; k < i; k++) t -= L[i][k] * a[k];
    a[i] = t / L[i][i];
  }
  for (int i = r - 1; i >= 0; i--) {
    double t = a[i];
    for (int k = i + 1; k < r; k++) t -= L[k][i] * a[k];
    a[i] = t / L[i][i];
  }
  return r;
}

}  // namespace nla3d_amd
