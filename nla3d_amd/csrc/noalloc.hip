// nla3d_gpu_solve_block2_noalloc: the block(2) MINRES solve of nla3d_gpu_solve_block2 for callers
// that cannot allocate device memory (code under stream capture, tight serving loops, harnesses
// that forbid it).  The solve's own work space (seven MINRES vectors, the reduction partials
// and the scalar state) lives in `scratch`, a file-scope device array that every GPU the library
// runs on holds until the process ends; the matrix and the block-Jacobi blocks are the handle's
// (allocated by nla3d_gpu_init_solution), the right-hand side and the unknowns the caller's.
//
// `scratch` names a host-side stand-in in host code: its device address is looked up once per
// device with hipGetSymbolAddress (the device current) and only that address is handed to
// kernels.  Calls on one GPU share its copy; the kernels rewrite every word they read before
// reading it (minres_start fills R1, R2, the first preconditioned vector and the state, and
// zeroes X, W1 and W2; T and Z are written before use), so a call never sees an earlier call's
// values, but two calls that overlap on one GPU race.
#include <mutex>

#include "../../include/nla3d_gpu.h"
#include "ctx.h"

namespace nla3d_amd {
namespace {

constexpr size_t kVec = (size_t)NLA3D_GPU_NOALLOC_MAX_ROWS;  // doubles per vector (a multiple of 32)
constexpr size_t kStateDoubles = 64;
static_assert(sizeof(MinresState) <= kStateDoubles * 8, "MINRES state larger than its scratch slot");
static_assert((kVec * 8) % 256 == 0, "vectors must keep 256-byte alignment");
constexpr size_t kScratchDoubles = 7 * kVec + kKrylovBlocks + kStateDoubles;  // 7,342,144 doubles, 56.0 MiB

static __device__ __align__(256) double scratch[kScratchDoubles];

constexpr int kMaxDevices = 64;
std::mutex scratchMutex;
double* scratchPtr[kMaxDevices] = {};

// the device address of this device's scratch (the device must be current)
double* deviceScratch(int device) {
  if (device < 0 || device >= kMaxDevices) throw ApiError(NLA3D_GPU_EINVAL, "device index out of range");
  std::lock_guard<std::mutex> lock(scratchMutex);
  if (!scratchPtr[device]) {
    void* p = nullptr;
    hipCheck(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
    scratchPtr[device] = static_cast<double*>(p);
  }
  return scratchPtr[device];
}

}  // namespace
}  // namespace nla3d_amd

using namespace nla3d_amd;

extern "C" int nla3d_gpu_solve_block2_noalloc(nla3d_gpu_t h, const double* rhs, double* x, double tol, int max_iter,
                                              int* iters, double* relres) {
  if (!h || !rhs || !x || max_iter < 1) return NLA3D_GPU_EINVAL;
  Ctx& c = *reinterpret_cast<Ctx*>(h);
  try {
    c.err.clear();
    c.requireSolid();
    if (!c.krylovReady) throw ApiError(NLA3D_GPU_ESTATE, "the handle's block(2) solver data are not set up");
    if ((size_t)c.n2 > kVec)
      throw ApiError(NLA3D_GPU_EINVAL, "block(2) has more rows than the scratch holds (NLA3D_GPU_NOALLOC_MAX_ROWS)");
    hipCheck(hipSetDevice(c.device), "hipSetDevice");
    double* base = deviceScratch(c.device);
    MinresWork w;
    w.X = x;
    w.R1 = base;
    w.R2 = base + kVec;
    w.T = base + 2 * kVec;
    w.Yp = base + 3 * kVec;
    w.Z = base + 4 * kVec;
    w.W1 = base + 5 * kVec;
    w.W2 = base + 6 * kVec;
    w.partial = base + 7 * kVec;
    w.st = reinterpret_cast<MinresState*>(base + 7 * kVec + kKrylovBlocks);
    const KrylovMatrix M = c.krylovMatrix();
    const Precond P = c.precond();
    krylov_precond_setup(c.stream, M, P);
    minres_start(c.stream, M, P, w, rhs);
    // the launches of nla3d_gpu_solve_block2's solve, issued directly (instantiating a graph
    // allocates); the state is read back every 12 iterations into the handle's pinned mirror
    int it = 0;
    for (;;) {
      for (int k = 0; k < 12 && it < max_iter; k++, it++) minres_iteration(c.stream, M, P, w, tol);
      hipCheck(hipGetLastError(), "minres iteration");
      hipCheck(hipMemcpyAsync(c.h_mstate, w.st, sizeof(MinresState), hipMemcpyDeviceToHost, c.stream), "state");
      hipCheck(hipStreamSynchronize(c.stream), "minres sync");
      if (c.h_mstate->converged || it >= max_iter) break;
    }
    if (iters) *iters = c.h_mstate->itn;
    if (relres) *relres = c.h_mstate->beta1 > 0 ? c.h_mstate->phibar / c.h_mstate->beta1 : 0.0;
    if (c.h_mstate->breakdown) throw ApiError(NLA3D_GPU_ENOTCONV, "MINRES breakdown (preconditioner not positive definite)");
    return NLA3D_GPU_OK;
  } catch (const ApiError& e) {
    c.err = e.what();
    return e.code;
  } catch (const std::exception& e) {
    c.err = e.what();
    return NLA3D_GPU_EINVAL;
  }
}
