// Preconditioner stage for the Schur rows of block(2) (schur.hip, schur.cpp).
//
// The rows of block(2) that block Jacobi treats by a diagonal Schur estimate (element
// pressures, MPC multipliers) are preconditioned instead through the sparse Schur
// approximation
//     S = alpha B^T D^-1 B - C,
// B the couplings of those rows to the node DoFs, D^-1 the block-Jacobi inverse node
// blocks, C their own block (the pressure rows' -V/k).  S is SPD when the node blocks
// are.  alpha is empirical: the block-Jacobi blocks misjudge the displacement block's
// smooth modes, and shrinking their Schur term cuts MINRES iterations (0.3 for the default
// stage, 1/2 for the V-cycle; 1/4 was best for a small case under the V-cycle).  The default
// stage is one Jacobi-Chebyshev sweep on S from zero, its products on an FP32 or
// half-precision copy of S.  The V-cycle form: piecewise-constant aggregation of S's graph
// (greedy, distance one), Galerkin coarse matrices P^T A P (FP64), Jacobi-Chebyshev smoothing
// (the same polynomial before and after, so the cycle is symmetric; a long one on the finest
// level) whose products use an FP32 sliced-ELLPACK copy of each level, an explicit dense FP64
// inverse on the coarsest level.  All structure is built once per model; values, smoother
// intervals (power iteration) and the coarsest inverse are refreshed with every tangent.
#pragma once

#include <hip/hip_runtime.h>

#include "krylov.h"

namespace nla3d_amd {

// device view of one level's CSR (full symmetric)
struct AmgCsr {
  int n;
  const int* rowPtr;
  const int* col;
  const double* val;
};

// the smoother's copy of a level's matrix: sliced ELLPACK over 64-row slices with FP32
// values.  Row i = 64 s + l holds (off[s+1] - off[s]) / 64 slots at off[s] + k * 64 + l;
// padding slots hold value 0.  Columns are 16-bit offsets from the slice's base column
// (col16, base) when every slice spans fewer than 65,536 columns, else full indices (col).
// The values are the FP64 matrix rounded once per tangent, so the smoother stays one fixed
// symmetric operator.
// A level's matrix as 64-row sliced ELLPACK: slice s holds its rows' values slot by slot,
// val[off[s] + 64 k + lane], and their columns from cOff[s] (off[s] when cOff is null), as full
// indices or as 16-bit offsets from base[s].  A "pattern slice" (cOff[s] < 0) stores nothing
// there: its rows' columns are i + o for the offsets o of one symmetric list (0, +-d_1 ..
// +-d_m), and the matrix being symmetric, row i reads the entry (i, i + d_k) from diagonal k,
// dia[k * ldd + i], and the entry (i, i - d_k) as (i - d_k, i) from dia[k * ldd + i - d_k]:
// each entry of the pattern rows is stored once.  Absent entries are zeros.
struct AmgSell {
  int n;
  const int* off;
  const int* col;                // full column indices (col16 == nullptr)
  const unsigned short* col16;   // offsets from base[s]
  const int* base;               // [slices] smallest column of the slice
  const float* val;
  const int* cOff;               // [slices] start of the slice's stored columns, -1 for a pattern slice; or null
  const int* dOff;               // [nd] 0, d_1 .. d_m (d_k > 0)
  const float* dia;              // [nd][ldd] dia[k * ldd + j] = A(j, j + d_k), over all rows j
  int nd, ldd;
};

// One level of the fused coarse cycle (schur.hip coarse_cycle_kernel): the level's smoother
// matrix, diagonal, Chebyshev coefficients (c1[i], c2[i] for i = 1 .. steps-1) and buffers; the
// coarsest level (steps = 0) holds its dense inverse instead.
constexpr int kFusedMaxSteps = 4;
struct FusedLevel {
  int n, nc, steps;
  AmgSell A;
  const double* dinv;
  double c0, c1[kFusedMaxSteps], c2[kFusedMaxSteps];
  double *r, *x, *res, *da, *db, *t;
  const int *memPtr, *mem, *agg;
  const double* cinv;  // coarsest level only
};
// the V-cycle below and including level 0 of `lv` (nl levels, the last the coarsest) in one
// workgroup: every level must be small (one workgroup walks all its rows per pass)
void launch_coarse_cycle(hipStream_t s, const FusedLevel* lv, int nl);

// The finest level's smoother operator in half precision.  With s = D^-1/2 (D the diagonal
// of S), the symmetrically scaled matrix S^ = diag(s) S diag(s) has a unit diagonal and
// off-diagonal entries of magnitude below one; those are stored as FP16 (the FP64 value
// rounded once, to nearest even) in the slots of the level's ELLPACK and diagonal layout
// above (AmgSell: the same off, cOff, columns, dOff, nd, ldd), the ELLPACK rows' diagonal slots
// holding zero.  The diagonal is stored apart in FP32 as 1 + Delta_i rounded up, Delta_i the
// sum over row i of the absolute rounding errors of its stored entries, so the stored
// operator S^' = H + diag(1 + Delta) satisfies S^' - S^ >= 0 (E + diag(Delta), E the
// rounding errors, is diagonally dominant with a non-negative diagonal): S^' is SPD whenever
// S is, and the smoother stays one fixed symmetric positive-definite operator.  The
// smoother runs in the scaled variables (x^ = x / s, r^ = s r), where the Jacobi step
// D^-1 (r - S x) becomes r^ - S^' x^: no diagonal is read per step.
struct AmgHalf {
  int n;
  const int* off;
  const int* col;
  const unsigned short* col16;
  const int* base;
  const unsigned short* val;  // FP16 bits of the ELLPACK slots
  const int* cOff;
  const int* dOff;
  const unsigned short* dia;  // FP16 bits, dia[k * ldd + j] = S^(j, j + d_k) for k >= 1
  const float* diag;          // 1 + Delta_i
  const double* s;            // D^-1/2
  int nd, ldd;
  // dispatch order of the product kernels' 512-row blocks: hardware workgroup b (on XCD b mod 8)
  // takes block blk[b] & 0xffffff; bit 24 + v of blk[b] is set when both slices of the block's
  // wave v (rows 128 v .. 128 v + 127) are in diagonal form.  The blocks holding ELLPACK slices
  // (the rows slower per row) come first on every XCD, spread over the eight, then each XCD
  // takes one contiguous range of the rest
  const int* blk;
};

// the half-precision copy of a level from its FP64 CSR values: s = sqrt(dinv), the FP16
// slots (ELLPACK rows, diagonals k >= 1 of every row), then the compensated diagonal
void launch_half_fill(hipStream_t s, const AmgHalf& H, const AmgCsr& A, const int* sPos, const int* dPos,
                      const double* dinv, double* sOut, unsigned short* val, unsigned short* dia, float* diag);
// Chebyshev smoothing of the scaled system in three-term form: with x^_0 = 0 (pre-smoothing)
// or the prolongated iterate (post-smoothing),
//   x^_1 = x^_0 + c0 (r^ - S^' x^_0),
//   x^_(k+1) = x^_k + c1_k (x^_k - x^_(k-1)) + c2_k (r^ - S^' x^_k),
// the same polynomial as the two-term Jacobi-Chebyshev sweep (d_k = x_(k+1) - x_k).  Every
// kernel returns at once when st says the MINRES solve has converged.
// start from zero: r^ = s r (r = rsrc[qrow[i]], or rsrc[i] without qrow) ; x^_1 = c0 r^
void launch_half_start(hipStream_t s, int n, const int* qrow, const double* rsrc, const double* sc, double c0,
                       double* rhat, double* x1, const MinresState* st);
// the start fused with MINRES's preconditioner pass on the Schur rows: t = T[qrow] - c_r2 R2[qrow]
// written back to T[qrow], then r^ = s t, x^_1 = c0 r^
void launch_half_start_fused(hipStream_t s, int n, const int* qrow, const double* R2, double* T, const double* sc,
                             double c0, double* rhat, double* x1, const MinresState* st);
// start from x^_0 (post-smoothing): x^_1 = x^_0 + c0 (r^ - S^' x^_0)
void launch_half_start_x(hipStream_t s, const AmgHalf& H, const double* x0, const double* rhat, double c0, double* x1,
                         const MinresState* st);
// one step: xout = x^_(k+1) from xk = x^_k and xprev = x^_(k-1) (null: zero; xout may be xprev).
// With z, the step is the sweep's last: z[qrow[i]] (z[i] without qrow) = s_i x^_(k+1) instead;
// with partial too, workgroup b also writes partial[b] = its rows' sum of r^_i zs x^_(k+1),i (the
// rows' share of r . z), one per workgroup of cdiv(n, 512); zs scales z (the stage's weight).
void launch_half_step(hipStream_t s, const AmgHalf& H, const double* xk, const double* xprev, const double* rhat,
                      double c1, double c2, double* xout, const int* qrow, double* z, const MinresState* st,
                      double* partial = nullptr, double zs = 1.0);
// The coarse level's cheb_start done by the restriction that feeds it (res = r, d = c0 D^-1 r,
// x = 0 for every coarse row; res == nullptr: none, the coarse level starts itself)
struct ChebStartNext {
  const double* dinv = nullptr;
  double c0 = 0.0;
  double* res = nullptr;
  double* d = nullptr;
  double* x = nullptr;
};
// t = (r^ - S^' x^) / s, the unscaled residual, then the restriction rc = P^T t
void launch_half_residual_restrict(hipStream_t s, const AmgHalf& H, const double* x, const double* rhat, double* t,
                                   int nc, const int* memPtr, const int* mem, double* rc, const MinresState* st,
                                   const ChebStartNext& nx);
// x^ += (xc + dc)[agg] / s (the prolongated coarse correction in the scaled variables; dc: the
// coarse level's last smoothing increment not yet added to xc, or null)
void launch_half_prolong(hipStream_t s, int n, const int* agg, const double* xc, const double* dc, const double* sc,
                         double* x, const MinresState* st);

// level-0 assembly of S: one thread per (node, Schur column a) writes alpha B(:,a)^T D^-1 B(:,b)
void launch_schur_contrib(hipStream_t s, int nPairs, const int* pairNode, const int* nodeGroup, const int* nodeSize,
                          const int* qPtr, const int* cPtr, const int* posB, const double* kval, const double* binv,
                          double alpha, double* contrib);
void launch_schur_sub_c(hipStream_t s, int nq, const int* sDiag, const int* cPos, const double* kval, double* sval);
void launch_gather_sum(hipStream_t s, int n, const int* seg, const int* src, const double* in, double* out);
void launch_inv_diag(hipStream_t s, const AmgCsr& A, double* dinv);

// sellVal[j] = csrVal[pos[j]] rounded to FP32 (0 where pos[j] < 0)
void launch_sell_fill(hipStream_t s, int nSlots, const int* pos, const double* csrVal, float* sellVal);

// V-cycle pieces (c0, c1, c2: Chebyshev coefficients of a step)
void launch_cheb_start(hipStream_t s, int n, const double* r, const double* dinv, double c0, double* res, double* d,
                       double* x);
void launch_cheb_start_x(hipStream_t s, const AmgSell& A, const double* r, const double* dinv, double c0, double* res,
                         double* d, const double* x);
// the MINRES stage's ends: the first pass with the Schur rows gathered from the Krylov vector
// (r = rext[qrow], kept in r), and the last with z[qrow] = d + x (skipped once converged)
void launch_cheb_start_gather(hipStream_t s, int n, const int* qrow, const double* rext, const double* dinv, double c0,
                              double* r, double* res, double* d, double* x);
void launch_finish_scatter(hipStream_t s, int n, const int* qrow, const double* d, const double* x,
                           const MinresState* st, double* z, double zs = 1.0);
// the same pair fused with MINRES's preconditioner pass: the start also does the pass's
// T[qrow] -= c_r2 R2[qrow] (r = the result), the exit also writes partial[b] = sum of r . z over
// rows 512 b .. 512 b + 511 (cdiv(n, 512) partials)
void launch_cheb_start_fused(hipStream_t s, int n, const int* qrow, const double* R2, double* T, const double* dinv,
                             double c0, double* r, double* res, double* d, double* x, const MinresState* st);
void launch_finish_fused(hipStream_t s, int n, const int* qrow, const double* d, const double* x, const double* r,
                         const MinresState* st, double* z, double* partial, double zs = 1.0);
// how Chebyshev step i of a sweep of `steps` terms (i = 1 .. steps-1) adds the increments to x:
// odd steps with a step after them add their input and output increments (kChebPair), the even
// steps after them add nothing (kChebSkip), and a last odd step adds its input (kChebPlain), so
// the increment a sweep ends with is always still to be added, as with one addition per step
// (x then takes 5 of the 7 vector passes of a step every other step)
enum { kChebPlain = 0, kChebPair = 1, kChebSkip = 2 };
inline int chebXMode(int i, int steps) { return (i & 1) ? (i + 1 < steps ? kChebPair : kChebPlain) : kChebSkip; }
void launch_cheb_step(hipStream_t s, const AmgSell& A, const double* din, const double* dinv, double c1, double c2,
                      double* res, double* dout, double* x, int xmode);
void launch_residual_restrict(hipStream_t s, const AmgSell& A, const double* res, const double* dlast, double* t,
                              int nc, const int* memPtr, const int* mem, double* rc, double* x,
                              const ChebStartNext& nx);
// x += (xc + dc)[agg] (dc: the coarse level's pending last increment, or null)
void launch_prolong(hipStream_t s, int n, const int* agg, const double* xc, const double* dc, double* x);
void launch_dense_matvec(hipStream_t s, int n, const double* Ainv, const double* r, double* x);
void launch_q_gather(hipStream_t s, int nq, const int* qrow, const double* v, const MinresState* st, double* out);
void launch_q_scatter(hipStream_t s, int nq, const int* qrow, const double* v, const MinresState* st, double* out);
void launch_dot_partials(hipStream_t s, int n, const double* a, const double* b, const MinresState* st,
                         double* partial);
// power-iteration steps, the scale 1 / sqrt(*nsq) read on the device (no host round trip per
// step): y = D^-1 A x / ||x|| on the FP64 values, and y = S^' x / ||x|| on the stored
// half-precision operator (its largest eigenvalue is the sweep's, in the scaled variables)
void launch_dinv_spmv(hipStream_t s, const AmgCsr& A, const double* dinv, const double* x, const double* nsq,
                      double* y);
void launch_half_power(hipStream_t s, const AmgHalf& H, const double* x, const double* nsq, double* y);

}  // namespace nla3d_amd
