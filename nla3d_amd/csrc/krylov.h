// Device data of the block(2) Krylov solve (see krylov.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <functional>

namespace nla3d_amd {

// K.block(2) in full-symmetric "grouped CSR": consecutive rows with identical
// column lists (the 3 DoF rows of a node) form a group that stores its column
// list once and its values row by row (group value block = size x L).
struct KrylovMatrix {
  int n = 0;
  int nGroups = 0;
  const int* gRow0 = nullptr;    // [nGroups] first row of the group
  const int* gSize = nullptr;    // [nGroups] rows in the group, 1..3
  const int* gColPtr = nullptr;  // [nGroups+1] start of the group's columns in gCols
  const int* gCols = nullptr;    // 0-based column indices
  const int* gValPtr = nullptr;  // [nGroups] start of the group's values in val
  const double* val = nullptr;
  const int* diagpos = nullptr;  // [n] position of (i,i) in val, or -1 (multiplier rows have none)
  const int* rowGroup = nullptr; // [n] group of row i
  int nLead1 = 0;                // groups [0, nLead1) are single rows (the element pressures come first)

  // Node-diagonal copy, read by the SpMV in place of the grouped rows of most nodes of a
  // structured block (absent when nDiaFit == 0).  The run: diaNodes consecutive 3-row groups
  // diaG0 + t with rows diaRow0 + 3t.  Node t's upper node couplings are the 3x3 blocks
  // (t, t + diaOff[k]), k < nd, diaOff[0] = 0, stored once each (symmetric storage):
  // dia[(9k + 3r + c) * diaNodes + t] = K(row(t) + r, row(t + diaOff[k]) + c).  The nDiaFit
  // nodes fit[i] whose node couplings are exactly (t, t +- diaOff[k]) are read from there
  // (the lower blocks as the transposes of the neighbours' upper blocks), with their other
  // columns (element pressures, multipliers) in a width-remW ELL: column remCol[j * nDiaFit + i]
  // (-1 = padding), values remVal[(3j + r) * nDiaFit + i].  All other groups after the leading
  // single rows are csrList[0, nCsrList).  The grouped rows stay complete; the copy is refreshed
  // from val by krylov_refresh (diaPos / remPos: positions in each group's column list).
  static constexpr int kMaxDiaOffsets = 32;
  int nDiaFit = 0, diaNodes = 0, diaG0 = 0, diaRow0 = 0, nd = 0, remW = 0, nCsrList = 0;
  int diaOff[kMaxDiaOffsets] = {};
  const int* fit = nullptr;
  double* dia = nullptr;
  const int* remCol = nullptr;
  double* remVal = nullptr;
  const int* csrList = nullptr;
  const int* diaPos = nullptr;  // [k * diaNodes + t] or -1
  const int* remPos = nullptr;  // [j * nDiaFit + i] or -1
  // the SpMV walks the run in chunks of diaChunk nodes, chunk c on XCD c % 8 (a chunk size
  // that puts most pairs (t, t - diaOff[k]) on one XCD); diaRank[t] = i for fit[i] = t, else -1
  int diaChunk = 0;
  const int* diaRank = nullptr;
};

// block-Jacobi groups: each row belongs to exactly one group
//   gInfo = size | kind << 2 ; kind 0 = node block (size 1..3), 1 = Schur-estimate scalar row
struct Precond {
  int nGroups = 0;
  int skipQ = 0;                  // 1: the pass leaves the scalar Schur groups to a fused stage
  const int* gStart = nullptr;    // first row of the group
  const int* gInfo = nullptr;
  const int* gDiagPos = nullptr;  // [nGroups*9] positions of the block entries in the full CSR (-1 absent)
  double* binv = nullptr;         // [nGroups*9] inverse blocks (row-major, size x size used)
};

// scalar state of one MINRES solve, resident on the device
struct MinresState {
  double beta1, beta, oldb, alfa, dbar, epsln, phibar, cs, sn, tnorm2, rnorm;
  double s, c_r1, c_r2;                      // coefficients of the next SpMV / preconditioner pass
  double oldeps, delta, denom, phi, s_update;  // coefficients of the W / X update
  int itn, converged, breakdown, skip_update;
};

struct MinresWork {
  double *X, *R1, *R2, *T, *Yp, *Z, *W1, *W2;
  double* partial;  // per-block partial sums (kKrylovBlocks of them)
  MinresState* st;
};

// The logical index of this workgroup for one-row-block-per-workgroup and chunked launches.
// Workgroups go to the 8 XCDs in turn (workgroup b on XCD b % 8); renumbering them so that
// each XCD takes one contiguous range of logical blocks keeps what a workgroup reads around
// its rows (the lower entries i - d of symmetric diagonals, the neighbours' x) in its own
// XCD's L2 instead of another XCD's.  A bijection of [0, gridDim.x).
__device__ __forceinline__ int xcd_block() {
  constexpr int kXcd = 8;
  const int nb = (int)gridDim.x, b = (int)blockIdx.x;
  const int x = b % kXcd, q = b / kXcd, base = nb / kXcd, rem = nb % kXcd;
  return x * base + (x < rem ? x : rem) + q;
}

// fixed grid of the grid-stride Krylov kernels (also the number of reduction partials)
constexpr int kKrylovBlocks = 2048;

void krylov_precond_setup(hipStream_t s, const KrylovMatrix& M, const Precond& P);
// the node-diagonal copy from the current values (nothing when M has none); every write of
// M.val must be followed by it before the next SpMV
void krylov_refresh(hipStream_t s, const KrylovMatrix& M);
void krylov_spmv(hipStream_t s, const KrylovMatrix& M, const double* x, double* y);
void krylov_dot(hipStream_t s, int n, const double* a, const double* b, double* partial, double* out);
// A preconditioner stage after the block-Jacobi pass: given r and z = P_bj r (device, n), it
// may replace part of z and must then rewrite the kKrylovBlocks partials of r . z (the
// layout of the block-Jacobi pass); every kernel it launches must do nothing once
// st->converged is set.
using PrecondHook = std::function<void(hipStream_t, const double* r, double* z, double* partial,
                                       const MinresState* st)>;
// A stage on the scalar Schur groups' rows that takes over their part of the preconditioner
// pass: the block-Jacobi pass leaves those rows alone (Precond::skipQ), then `run` does
// T -= c_r2 R2 on them, writes Z there and `partials` sums of T . Z over them to
// partial[kKrylovBlocks ..), so no separate pass over r . z follows; every kernel it launches
// must do nothing once st->converged is set.  (A second stream running the block-Jacobi pass
// beside the stage was slower: the pass took the compute units the stage's short,
// latency-bound launches wait for.)
struct FusedStage {
  std::function<void(hipStream_t, const double* R2, double* T, double* Z, double* partial,
                     const MinresState* st)>
      run;
  int partials = 0;
};
void minres_start(hipStream_t s, const KrylovMatrix& M, const Precond& P, MinresWork& w, const double* b,
                  const PrecondHook* hook = nullptr);
void minres_iteration(hipStream_t s, const KrylovMatrix& M, const Precond& P, MinresWork& w, double tol,
                      const PrecondHook* hook = nullptr, const FusedStage* fused = nullptr);

}  // namespace nla3d_amd
