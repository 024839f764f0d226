// Learned block-Jacobi correction (learned.hip): per-node features of the
// tangent, the per-node MLP (BF16 MFMA inference, FP32 training passes), the
// corrected 3x3 blocks and the pieces of the unrolled-MINRES adjoint.
#pragma once

#include <hip/hip_runtime.h>

#include "krylov.h"

namespace nla3d_amd {

constexpr int kMlpIn = 24;     // features per node
constexpr int kMlpWidth = 128; // hidden width (4 hidden layers, GELU)
constexpr int kMlpOut = 6;     // lower-triangular factor of the block correction
// flat FP32 parameter layout: W1[128x24] b1 W2[128x128] b2 W3 b3 W4 b4 W5[6x128] b5 (row-major W[out][in])
constexpr int kMlpOffW[5] = {0, 3200, 19712, 36224, 52736};
constexpr int kMlpOffB[5] = {3072, 19584, 36096, 52608, 53504};
constexpr int kMlpParams = 53510;
// packed BF16 weights for the MFMA kernel: W1 padded to [128][32], W2..W4 [128][128], W5 padded to [16][128]
constexpr int kMlpPackedW = 128 * 32 + 3 * 128 * 128 + 16 * 128;
constexpr int kMlpPackedB = 4 * 128 + 16;

// what the feature kernel reads besides the full-symmetric block(2) matrix
struct FeatureInputs {
  int nL;                      // learned nodes
  const int* lRow0;            // [nL] first block(2) row of the node's UX UY UZ triple
  const signed char* rowKind;  // [n2] 0 node DoF, 1 element DoF, 2 multiplier
  const int* rowNode;          // [n2] node index of a node-DoF row, -1 otherwise
  const int* k12Ptr;           // [n2+1] block(1,2) entries per column (constrained couplings)
  const int* rowPGroup;        // [n2] block-Jacobi group of each row
  const double* binv;          // block-Jacobi inverse blocks (the Schur estimates of element rows)
  const double* U;             // [nAll] solution, element pressures at nC + row
  int nC;
  double bulk;                 // material bulk modulus
  double t;                    // normalised load-step time
};

void launch_node_features(hipStream_t s, const KrylovMatrix& M, const FeatureInputs& in, float* X);
// corrected blocks P = L^-T C C^T L^-1 (K_ii = L L^T) written over the block-Jacobi
// inverse of each learned node whose diagonal block is positive definite; Rinv
// (nullable) receives L^-1 (6 doubles per node) for the adjoint
void launch_learned_blocks(hipStream_t s, int nL, const int* lGroup, const Precond& P, const KrylovMatrix& M,
                           const float* Y, double* Rinv, int* ok);
// dL/dY from the block adjoint Pbar (9 per node, accumulated over the unrolled iterations)
void launch_learned_grad(hipStream_t s, int nL, const double* Rinv, const int* ok, const float* Y, const double* Pbar,
                         float* dY);

// MLP
void launch_mlp_pack(hipStream_t s, const float* w, unsigned short* wp, float* bp);
void launch_mlp_bf16(hipStream_t s, int n, const float* X, const float* mean, const float* istd,
                     const unsigned short* wp, const float* bp, float* Y);
void launch_normalize(hipStream_t s, int n, const float* X, const float* mean, const float* istd, float* X0);
void launch_dense_fwd(hipStream_t s, int n, int in, int out, const float* X, const float* W, const float* b, float* Z,
                      float* A, int gelu);
void launch_dense_bwd_input(hipStream_t s, int n, int in, int out, const float* dZ, const float* W, const float* Zprev,
                            float* dZprev);
// weight and bias gradient of one layer into dW (row-major [out][in]) and db; partial: scratch of
// ceil(n/kChunk) * out * (in+1) floats
constexpr int kGradChunk = 1024;
void launch_dense_bwd_weight(hipStream_t s, int n, int in, int out, const float* dZ, const float* X, float* partial,
                             float* dW, float* db);
void launch_adam(hipStream_t s, int n, float lr, float b1, float b2, float eps, int step, const float* g, float* w,
                 float* m, float* v);

// vector pieces of the unrolled MINRES recurrence
void launch_bj_apply(hipStream_t s, const Precond& P, const double* r, double* z);
void launch_axpby(hipStream_t s, int n, double a, const double* x, double b, const double* y, double* out);
void launch_outer_acc(hipStream_t s, int nL, const int* lRow0, const double* ybar, const double* r, double* Pbar);
void launch_zero_d(hipStream_t s, size_t n, double* x);

}  // namespace nla3d_amd
