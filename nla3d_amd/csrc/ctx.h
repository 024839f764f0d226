// The object behind an nla3d_gpu_t handle: model store, material, device
// buffers and the assembly pipeline.
#pragma once

#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "internal.h"
#include "krylov.h"
#include "model.h"

namespace nla3d_amd {

struct ApiError : std::runtime_error {
  int code;
  ApiError(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

void hipCheck(hipError_t e, const char* what);

struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  DevBuf(DevBuf&& o) noexcept : p(o.p), bytes(o.bytes) { o.p = nullptr; o.bytes = 0; }
  ~DevBuf() { free_(); }
  void alloc(size_t nbytes);
  void free_();
  template <class T>
  T* as() const { return reinterpret_cast<T*>(p); }
  template <class T>
  void upload(const std::vector<T>& v) {
    alloc(v.size() * sizeof(T));
    if (!v.empty()) hipCheck(hipMemcpy(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice), "upload");
  }
};

// A launch sequence captured from a stream into a graph, and its executable.  Both are
// destroyed together (the graph outlives the executable instantiated from it); a capture
// still open when the holder goes away, e.g. after a throw inside the captured calls, is
// ended first, so the stream leaves capture mode.
struct StreamGraph {
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  hipStream_t capturing = nullptr;
  StreamGraph() = default;
  StreamGraph(const StreamGraph&) = delete;
  StreamGraph& operator=(const StreamGraph&) = delete;
  ~StreamGraph() { reset(); }
  void begin(hipStream_t s) {
    reset();
    hipCheck(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal), "begin capture");
    capturing = s;
  }
  void end() {
    hipStream_t s = capturing;
    capturing = nullptr;
    hipCheck(hipStreamEndCapture(s, &graph), "end capture");
    hipCheck(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0), "instantiate");
  }
  void launch(hipStream_t s) const { hipCheck(hipGraphLaunch(exec, s), "graph launch"); }
  void reset() {
    if (capturing) {
      hipGraph_t g = nullptr;
      (void)hipStreamEndCapture(capturing, &g);
      if (g) (void)hipGraphDestroy(g);
      capturing = nullptr;
    }
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    exec = nullptr;
    graph = nullptr;
  }
  explicit operator bool() const { return exec != nullptr; }
};

struct ThermalData;  // thermal.cpp
struct LearnedWork;  // learned.cpp
struct SchurAmg;     // schur.cpp
struct Mlp;          // learned.cpp

struct Ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  std::string err;
  bool initialized = false;

  FEStorage model;
  MaterialParams mat{};
  bool hasMaterial = false;
  uint32_t maxInRow = 100;  // BlockSparseSymMatrix<2> default (BlockSparseMatrix.h:42)

  // device copies of the mesh and DoF maps (u/p element range: SOLID81 or PLANE41)
  DevBuf d_xyz;        // double [nN*3]
  DevBuf d_conn;       // int [nSolid*nn] (nn = 8 SOLID81, 4 PLANE41), 0-based node index
  DevBuf d_nodeEq3;    // int [nN*3], 0-based equation index of UX UY UZ or -1
  DevBuf d_elemEqP;    // int [nSolid], 0-based equation index of HYDRO_PRESSURE or -1
  DevBuf d_U;          // double [nAll]
  DevBuf d_Ke;         // element outputs: SOLID81 [nSolid][350] slot order, PLANE41 [54][nSolid]
  DevBuf d_vals;       // double [nnz1 + nnz12 + nnz2 + nAll]: K blocks then F
  DevBuf d_seg;        // int [nTargets+1]
  DevBuf d_src;        // int [nContrib]
  int nTargets = 0;
  int64_t nContrib = 0;

  // ---- block(2) Krylov data (solver.cpp) ----
  int n2 = 0;                 // nUnknown + nMpc
  // full-symmetric block 2 in grouped CSR (krylov.h)
  DevBuf d_sgRow0, d_sgSize, d_sgColPtr, d_sgCols, d_sgValPtr, d_rowGroup, d_diagpos2, d_full2;
  int nSpmvGroups = 0, nLead1 = 0;  // groups, and the leading single-row ones
  // node-diagonal copy of the SpMV (KrylovMatrix): values, node list, ELL columns / values,
  // the grouped-row list, fill positions; shape in spmvDia (nDiaFit == 0: no copy)
  DevBuf d_dia, d_diaFit, d_remCol, d_remVal, d_csrList, d_diaPos, d_remPos, d_diaRank;
  KrylovMatrix spmvDia;
  DevBuf d_mir2;              // int [2*nnz2]: full positions of each upper entry of block 2
  int64_t fullNnz2 = 0;
  DevBuf d_gStart, d_gInfo, d_gDiagPos, d_binv;  // block-Jacobi groups
  int nGroups = 0;
  DevBuf d_k12tPtr, d_k12tPos, d_k12tRow;       // block(1,2) by columns: K12^T products
  DevBuf d_k11Ptr, d_k11Col, d_k11Pos;          // block(1) full-symmetric rows (reactions)
  DevBuf d_k12rPtr, d_k12rCol;                  // block(1,2) by rows (reactions)
  // Newton / Krylov vectors
  DevBuf d_R;                 // vecR [nAll]
  DevBuf d_Rsave;             // reactions of the last solve while a criterion-1 check runs
  DevBuf d_Ucprev, d_fixEq, d_fixVal, d_loadEq, d_loadVal;
  DevBuf d_rhs, d_dU;         // rhs [n2], increment [nAll]
  DevBuf d_mw[8], d_partial, d_mstate, d_scal;
  DevBuf d_trueY;  // K2 x of minresSolve's true-residual check (allocated on first use)
  // the first-solve predictor (nla3d_gpu_newton_opts.first_solve_predictor): U[nC, nC + n2) at
  // the start of the last load step and the times of that state and of its target; the
  // starting point and the right-hand side of the correction
  DevBuf d_lsStart, d_x0, d_rhsC;
  bool lsValid = false;
  // the predictors' history: for the first (j = 0) and second (j = 1) solve of a load step, the
  // increments of the last load steps (whole load step for j = 0, the second Newton increment
  // for j = 1) in a ring of kPredDepth, newest at predHead[j]; work space of the least-squares
  // fit.  The fit uses the newest predDepth() of them (NLA3D_GPU_PREDICTOR_DEPTH, 1 .. kPredDepth,
  // default 3)
  static constexpr int kPredDepth = 4;
  DevBuf d_pred[2][kPredDepth], d_q, d_kq, d_pw[kPredDepth], d_yw[kPredDepth];
  int predN[2] = {0, 0}, predHead[2] = {0, 0};
  double lsTargetT = 0.0;  // the time of the last load step the predictors saw
  DevBuf d_mpcMirIdx;         // block(2) upper indices of MPC entries
  int nMpcMir = 0;
  // learned block-Jacobi correction (learned.cpp)
  DevBuf d_lGroup, d_lRow0;   // [nLearned] block-Jacobi group and first row of each UX UY UZ node triple
  DevBuf d_rowKind2;          // int8 [n2]: 0 node DoF, 1 element DoF, 2 multiplier
  DevBuf d_rowNode2, d_rowPG; // int [n2]: node of a node-DoF row (-1 otherwise), block-Jacobi group of each row
  int nLearned = 0;
  std::shared_ptr<LearnedWork> lw;  // unrolled-MINRES and feature buffers
  Mlp* mlp = nullptr;               // attached network (not owned)
  MinresState* h_mstate = nullptr;  // pinned host mirror of the MINRES state
  bool krylovReady = false;

  std::shared_ptr<ThermalData> thermal;  // QUADTH / SurfaceLINETH data (thermal.cpp)

  // rigid-body MPC collections
  struct RB {
    RigidBodyDev dev{};
    DevBuf slaveNode, slaveEq, tgt, fTgt;
  };
  std::vector<RB> rbs;
  // constant MPC rows: per row its terms' value positions and coefficients, F position, b
  int nCeRows = 0;
  DevBuf ceTermPtr, ceTgt, ceCoef, ceFTgt, ceB;

  ~Ctx() {
    if (h_mstate) (void)hipHostFree(h_mstate);
    if (stream) (void)hipStreamDestroy(stream);
  }

  void requireDevice() const {
    if (device < 0) throw ApiError(-3, "host-only handle (device -1) has no device storage");
  }
  void requireSolid() const {
    requireDevice();
    if (!initialized) throw ApiError(-3, "call nla3d_gpu_init_solution first");
    if (model.solidCount == 0) throw ApiError(-3, "model has no SOLID81 or PLANE41 elements");
    if (!hasMaterial) throw ApiError(-3, "material is not defined");
  }
  void requireSolid81() const {
    requireSolid();
    if (model.upKind != ELEM_SOLID81) throw ApiError(-3, "this entry point needs SOLID81 elements");
  }
  void setupDevice();
  // the u/p element kernel (SOLID81 or PLANE41) into d_Ke; gp_state (SOLID81 only) may be null
  void runElements(double* gp_state);
  void scatter();
  void applyMpc();
  void assemble();
  // the SpMV's node-diagonal copy from d_full2 (after every write of d_full2)
  void refreshSpmvCopy(hipStream_t s) const;

  // solver.cpp
  void setupKrylov();
  void setupBlock2(const CsrPattern& U2, const std::vector<int8_t>& kind3, const std::vector<uint8_t>& rowType,
                   const std::vector<uint32_t>& rowNode);
  void setupModelRows();
  void setupNodeDia(const std::vector<int>& sgRow0, const std::vector<int>& sgSize, const std::vector<int>& sgColPtr,
                    const std::vector<int>& sgCols);
  KrylovMatrix krylovMatrix() const;
  Precond precond();
  // solve block(2) x = b with the current K; b, x device pointers [n2]; returns iterations
  // trueAbs > 0 also stops, at the end of a 12-iteration group, once ||b - K2 x||_2 <= trueAbs
  // (*trueStop then set); the preconditioned test stays in force.  beta1Ref > 0: tol is taken
  // relative to beta1Ref instead of b's own preconditioned norm beta1 = sqrt(b^T M b) (the
  // relative tolerance becomes tol * beta1Ref / beta1, at most 0.1)
  // fused: the iterations run the Schur rows' stage fused with the preconditioner pass instead
  // of hook (FusedStage; the start still runs hook)
  int minresSolve(const double* b, double* x, double tol, int maxit, double* relres_est, bool* breakdown,
                  const PrecondHook* hook = nullptr, double trueAbs = 0.0, bool* trueStop = nullptr,
                  double beta1Ref = 0.0, const FusedStage* fused = nullptr);
  // sqrt(b^T M b) for the current preconditioner M (block Jacobi, and the hook's stage)
  double minresBeta1(const double* b, const PrecondHook* hook = nullptr);
  // multigrid on the Schur rows (schur.cpp)
  std::shared_ptr<SchurAmg> schur;
};

// full-symmetric rows of a 1-based upper CSR (solver.cpp)
void fullSymmetricPattern(const CsrPattern& U2, std::vector<int64_t>& rp, std::vector<int>& col,
                          std::vector<int64_t>* mirFull);
// multigrid on the Schur rows (schur.cpp): structure from the full block(2) pattern and the
// block-Jacobi groups; refresh for the current tangent (block Jacobi current); the MINRES stage
void setupSchur(Ctx& c, const std::vector<int64_t>& rp, const std::vector<int>& col,
                const std::vector<int64_t>& fullToVal, const std::vector<int>& gStart, const std::vector<int>& gInfo);
bool schurUsable(Ctx& c, std::string* why);
bool schurRefresh(Ctx& c, std::string* why);
PrecondHook schurHook(Ctx& c);
// the stage as a FusedStage (the one-sweep stage, on its half-precision or FP32 copy; partials
// = 0 when the stage cannot run fused, the V-cycle or a dense one-level inverse, and schurHook
// must be used alone)
FusedStage schurFusedStage(Ctx& c);
// constrained-DoF reactions of the last solve into vecR (solver.cpp), on `stream` (null: the handle's)
void computeReactions(Ctx& c, hipStream_t stream = nullptr);
// block-Jacobi blocks with the attached network's corrections (learned.cpp); t = load-step time
void learnedPrecondSetup(Ctx& c, double t);
// the attached network's corrections over the current block-Jacobi blocks (learned.cpp)
void learnedCorrect(Ctx& c, double t);

}  // namespace nla3d_amd
