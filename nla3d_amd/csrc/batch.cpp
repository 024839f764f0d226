// Many independent SOLID81 models on one device (the batched path).
//
// Assembly: one persistent grid-stride launch of the element kernel over the
// element blocks of every model, then one launch of the segmented gather over
// every model's CSR targets.  Boundary conditions, right-hand sides, residual
// norms and block-Jacobi set-up of every model follow as a chain of small
// per-model kernels; the whole pre-solve sequence is captured once per load
// step into a hipGraph and replayed at each Newton iteration.
//
// Solve: the systems of the still-unconverged models go to one persistent
// team-MINRES launch (batch.hip), except the largest, which use the
// whole-device solver on their own streams, concurrently.  Each model has its
// own Newton convergence mask and its own Krylov convergence.
//
// The autotuner times candidate launch shapes on the batch itself: the element
// kernel's persistent grid, the gather's block size, the team size and lanes per
// matrix group, and how many of the largest models leave the team launch.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <exception>
#include <thread>
#include <cmath>
#include <cstring>
#include <functional>
#include <numeric>
#include <vector>

#include "../../include/nla3d_gpu.h"
#include "batch.h"
#include "ctx.h"
#include "newton.h"
#include "thermal.h"

namespace nla3d_amd {

// a system larger than this never goes to the team launch (one workgroup per system)
constexpr long long kTeamMaxRows = 65536;

// host threads driving whole-device solves at once, each on its own stream (6 measured
// best at the default 4 hardware queues per process: 3 gave 5.5 s, 6 and 10 gave 4.2 s on
// the 256-model batch)
constexpr int kBigWorkers = 6;

struct Batch {
  int device = 0;
  hipStream_t stream = nullptr;
  hipStream_t workers[kBigWorkers] = {};  // the whole-device solves run on these
  std::string err;
  std::vector<Ctx*> m;
  std::vector<int> n2;
  DevBuf d_s81, d_blockBegin, d_gather, d_targetBegin;
  DevBuf d_team, d_queue, d_counter, d_out, d_relres, d_norms;
  int nBlocks = 0, nTargets = 0;
  nla3d_gpu_batch_config cfg{};
  StreamGraph pre;  // the pre-solve chain at time preT
  double preT = std::nan("");
  std::vector<double> rhs1;
  std::vector<int> done;
  ~Batch() {
    pre.reset();
    if (stream) (void)hipStreamDestroy(stream);
    for (hipStream_t w : workers)
      if (w) (void)hipStreamDestroy(w);
  }
  void dropGraph() {
    pre.reset();
    preT = std::nan("");
  }
};

static double nowMsB() {
  return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// the per-model chain after the batched assembly: R = 0, BCs at t, rhs, ||rhs||^2, preconditioner
static void preSolveChain(Batch& b, double t) {
  hipStream_t s = b.stream;
  const int grid = b.cfg.elem_grid > 0 ? std::min(b.cfg.elem_grid, b.nBlocks) : b.nBlocks;
  launch_solid81_batched(s, grid, (int)b.m.size(), b.d_s81.as<Solid81Batch>(), b.d_blockBegin.as<int>(), b.nBlocks);
  launch_gather_batched(s, b.cfg.gather_threads, (int)b.m.size(), b.d_gather.as<GatherBatch>(),
                        b.d_targetBegin.as<int>(), b.nTargets);
  for (size_t k = 0; k < b.m.size(); k++) {
    Ctx& c = *b.m[k];
    FEStorage& md = c.model;
    const int nC = (int)md.nConstrainedDofs;
    double* U = c.d_U.as<double>();
    double* R = c.d_R.as<double>();
    launch_fill(s, (int)md.nAll(), 0.0, R);
    newton_apply_bc(s, (int)md.fixs.size(), c.d_fixEq.as<int>(), c.d_fixVal.as<double>(), (int)md.loads.size(),
                    c.d_loadEq.as<int>(), c.d_loadVal.as<double>(), t, U, R);
    newton_rhs(s, c.n2, nC, c.d_vals.as<double>() + md.nnzAll(), R, U, c.d_Ucprev.as<double>(), c.d_k12tPtr.as<int>(),
               c.d_k12tPos.as<int>(), c.d_k12tRow.as<int>(), c.d_vals.as<double>() + md.offB12(), c.d_rhs.as<double>());
    vec_sumsq(s, c.n2, c.d_rhs.as<double>(), c.d_partial.as<double>(), b.d_norms.as<double>() + k);
    c.refreshSpmvCopy(s);
    krylov_precond_setup(s, c.krylovMatrix(), c.precond());
  }
}

static void runPreSolve(Batch& b, double t) {
  if (!(b.pre && b.preT == t)) {
    b.dropGraph();
    b.pre.begin(b.stream);
    preSolveChain(b, t);
    b.pre.end();
    b.preT = t;
  }
  b.pre.launch(b.stream);
}

static std::vector<double> readNorms(Batch& b) {
  std::vector<double> v(b.m.size());
  hipCheck(hipMemcpyAsync(v.data(), b.d_norms.p, v.size() * 8, hipMemcpyDeviceToHost, b.stream), "norms");
  hipCheck(hipStreamSynchronize(b.stream), "sync");
  return v;
}

// solve block(2) of the models in `ids` (their rhs and preconditioners are current);
// iterations and flags per model into its/flags
// pc >= 2: the whole-device solves add the Schur-row multigrid (a model where it is
// unavailable or whose solve fails with it is solved again with block Jacobi, counted in
// fallbacks[k]); the team launch always runs block Jacobi
static void solveModels(Batch& b, const std::vector<int>& ids, double tol, int maxit, std::vector<int>& its,
                        std::vector<int>& flags, std::vector<double>& relres, int pc = 0,
                        std::vector<int>* fallbacks = nullptr) {
  std::vector<int> team, big;
  for (int k : ids) ((long long)b.n2[k] > std::min(b.cfg.big_rows, kTeamMaxRows) ? big : team).push_back(k);
  std::sort(team.begin(), team.end(), [&](int a, int c) { return b.n2[a] > b.n2[c]; });  // largest first
  // the right-hand sides and preconditioners were written on the batch stream
  hipEvent_t ready = nullptr;
  if (!big.empty()) {
    hipCheck(hipEventCreateWithFlags(&ready, hipEventDisableTiming), "event");
    hipCheck(hipEventRecord(ready, b.stream), "record");
  }
  if (!team.empty()) {
    std::vector<TeamSolve> ts(b.m.size());
    for (size_t k = 0; k < b.m.size(); k++) {
      Ctx& c = *b.m[k];
      TeamSolve& t = ts[k];
      t.M = c.krylovMatrix();
      t.P = c.precond();
      t.b = c.d_rhs.as<double>();
      t.x = c.d_dU.as<double>();
      for (int v = 0; v < 8; v++) t.v[v] = c.d_mw[v].as<double>();
      t.tol = tol;
      t.maxit = maxit;
      t.out = b.d_out.as<int>() + 2 * k;
      t.relres = b.d_relres.as<double>() + k;
    }
    hipCheck(hipMemcpyAsync(b.d_team.p, ts.data(), ts.size() * sizeof(TeamSolve), hipMemcpyHostToDevice, b.stream),
             "team");
    hipCheck(hipMemcpyAsync(b.d_queue.p, team.data(), team.size() * 4, hipMemcpyHostToDevice, b.stream), "queue");
    const int grid = std::min((int)team.size(), b.cfg.team_grid > 0 ? b.cfg.team_grid : 1024);
    launch_minres_team(b.stream, b.cfg.team_threads, b.cfg.team_lanes, grid, (int)team.size(),
                       b.d_team.as<TeamSolve>(), b.d_queue.as<int>(), b.d_counter.as<int>());
    hipCheck(hipGetLastError(), "team launch");
  }
  // the largest systems: whole-device MINRES, each on its own model's stream, while the team
  // launch runs; kBigWorkers host threads take them largest first, so up to that many solves
  // (each too small to fill the device) overlap.  Every model is touched by one thread only.
  // worker w drives its models on b.workers[w] (the model's own stream is put back after
  // its solve, which ends synchronised), so concurrent solves use distinct streams
  auto solveBig = [&](int k, int w) {
    Ctx& c = *b.m[k];
    hipCheck(hipSetDevice(c.device), "hipSetDevice");
    struct StreamSwap {
      Ctx& c;
      hipStream_t own;
      ~StreamSwap() { c.stream = own; }
    } swap{c, c.stream};
    c.stream = b.workers[w];
    hipCheck(hipStreamWaitEvent(c.stream, ready, 0), "wait");
    double rr = 0;
    bool brk = false;
    PrecondHook hook;
    FusedStage fused;
    bool mg = false;
    int fb = 0;
    if (pc >= 2) {
      std::string why;
      mg = schurUsable(c, &why) && schurRefresh(c, &why);
      if (mg) {
        hook = schurHook(c);
        fused = schurFusedStage(c);
      } else {
        fb++;
      }
    }
    its[k] = c.minresSolve(c.d_rhs.as<double>(), c.d_dU.as<double>(), tol, maxit, &rr, &brk, mg ? &hook : nullptr,
                           0.0, nullptr, 0.0, mg ? &fused : nullptr);
    if (mg && (brk || rr > tol)) {  // block Jacobi from a fresh start
      fb++;
      brk = false;
      its[k] += c.minresSolve(c.d_rhs.as<double>(), c.d_dU.as<double>(), tol, maxit, &rr, &brk);
    }
    flags[k] = brk ? 2 : 0;
    relres[k] = rr;
    if (fallbacks) (*fallbacks)[k] += fb;
  };
  std::sort(big.begin(), big.end(), [&](int a, int c) { return b.n2[a] > b.n2[c]; });
  const int nThreads = std::min<int>(kBigWorkers, (int)big.size());
  if (nThreads <= 1) {
    for (int k : big) solveBig(k, 0);
  } else {
    std::atomic<int> next{0};
    std::vector<std::exception_ptr> errs(nThreads);
    std::vector<std::thread> pool;
    for (int w = 0; w < nThreads; w++)
      pool.emplace_back([&, w] {
        try {
          for (int i = next++; i < (int)big.size(); i = next++) solveBig(big[i], w);
        } catch (...) {
          errs[w] = std::current_exception();
          next = (int)big.size();  // the other workers stop after their current model
        }
      });
    for (auto& t : pool) t.join();
    for (auto& e : errs)
      if (e) {
        if (ready) (void)hipEventDestroy(ready);
        std::rethrow_exception(e);
      }
  }
  if (ready) (void)hipEventDestroy(ready);
  if (!team.empty()) {
    std::vector<int> out(2 * b.m.size());
    std::vector<double> rr(b.m.size());
    hipCheck(hipMemcpyAsync(out.data(), b.d_out.p, out.size() * 4, hipMemcpyDeviceToHost, b.stream), "out");
    hipCheck(hipMemcpyAsync(rr.data(), b.d_relres.p, rr.size() * 8, hipMemcpyDeviceToHost, b.stream), "relres");
    hipCheck(hipStreamSynchronize(b.stream), "team sync");
    for (int k : team) {
      its[k] = out[2 * k];
      flags[k] = out[2 * k + 1];
      relres[k] = rr[k];
    }
  }
}

// one Newton iteration of every model not yet converged in this load step
// (FESolver.cpp:443-490 per model); returns the number of models still iterating
static int batchNewtonIteration(Batch& b, double t, int it, const nla3d_gpu_newton_opts& o,
                                nla3d_gpu_newton_stats* st, int* conv) {
  const size_t n = b.m.size();
  double t0 = nowMsB();
  // criterion 1 may end a model's load step without a solve here: keep its last reactions
  const bool mayStop = o.criterion == 1 && it > 1;
  auto saveR = [&](size_t k, bool restore) {
    Ctx& c = *b.m[k];
    const size_t bytes = (size_t)c.model.nConstrainedDofs * 8;
    if (bytes)
      hipCheck(hipMemcpyAsync(restore ? c.d_R.p : c.d_Rsave.p, restore ? c.d_Rsave.p : c.d_R.p, bytes,
                              hipMemcpyDeviceToDevice, b.stream),
               "reactions");
  };
  if (mayStop)
    for (size_t k = 0; k < n; k++)
      if (!b.done[k]) saveR(k, false);
  runPreSolve(b, t);
  std::vector<double> nrm = readNorms(b);
  std::vector<int> active;
  for (size_t k = 0; k < n; k++) {
    if (b.done[k]) continue;
    const double rn = std::sqrt(nrm[k]);
    if (it == 1) b.rhs1[k] = rn;
    if (!(rn == rn)) {
      b.done[k] = 1;
      conv[k] = -6;
      continue;
    }
    if (o.criterion == 1) {
      st[k].last_criterion = b.rhs1[k] > 0 ? rn / b.rhs1[k] : 0.0;
      if ((it > 1 && rn <= o.convergence * b.rhs1[k]) || b.rhs1[k] == 0.0) {
        if (mayStop) saveR(k, true);
        b.done[k] = 1;
        conv[k] = 1;
        continue;
      }
    }
    active.push_back((int)k);
  }
  const double tAsm = nowMsB() - t0;
  for (int k : active) st[k].ms_assemble += tAsm / std::max<size_t>(1, active.size());
  if (active.empty()) return 0;
  t0 = nowMsB();
  std::vector<int> its(n, 0), flags(n, 0), fb(n, 0);
  std::vector<double> relres(n, 0.0);
  solveModels(b, active, o.krylov_tol, o.krylov_max_iter, its, flags, relres, o.preconditioner, &fb);
  for (int k : active) {
    Ctx& c = *b.m[k];
    FEStorage& md = c.model;
    st[k].newton_iterations++;
    st[k].linear_solves++;
    st[k].krylov_iterations += its[k];
    st[k].last_krylov_iterations = its[k];
    st[k].precond_fallbacks += fb[k];
    if (flags[k] == 2) st[k].krylov_breakdown = 1;
    computeReactions(c, b.stream);
    newton_update(b.stream, (int)md.nConstrainedDofs, (int)md.nUnknownDofs, (int)md.nMpc(), c.d_dU.as<double>(),
                  c.d_U.as<double>(), c.d_Ucprev.as<double>());
    if (o.criterion == 0)
      vec_abssum(b.stream, (int)md.nUnknownDofs, c.d_dU.as<double>(), c.d_partial.as<double>(),
                 b.d_norms.as<double>() + k);
  }
  hipCheck(hipGetLastError(), "update");
  if (o.criterion == 0) {
    nrm = readNorms(b);
    for (int k : active) {
      const double crit = nrm[k] / b.m[k]->model.nUnknownDofs / o.convergence;
      st[k].last_criterion = crit;
      if (crit < o.convergence) {
        b.done[k] = 1;
        conv[k] = 1;
      } else if (crit > 1.0e6 || crit != crit) {
        b.done[k] = 1;
        conv[k] = -6;
      }
    }
  } else {
    hipCheck(hipStreamSynchronize(b.stream), "sync");
  }
  const double tSol = nowMsB() - t0;
  for (int k : active) st[k].ms_solve += tSol / active.size();
  int left = 0;
  for (size_t k = 0; k < n; k++) left += b.done[k] ? 0 : 1;
  return left;
}

// one load step of every model; conv[k] = 1 converged, 0 not converged, < 0 error code
static void batchLoadStep(Batch& b, double t, const nla3d_gpu_newton_opts& o, nla3d_gpu_newton_stats* st,
                          int* conv) {
  const size_t n = b.m.size();
  b.rhs1.assign(n, 0.0);
  b.done.assign(n, 0);
  std::vector<int> before(n);
  for (size_t k = 0; k < n; k++) {
    conv[k] = 0;
    before[k] = st[k].newton_iterations;
  }
  const int maxIt = o.max_iterations + (o.criterion == 1 ? 1 : 0);
  for (int it = 1; it <= maxIt; it++)
    if (batchNewtonIteration(b, t, it, o, st, conv) == 0) break;
  for (size_t k = 0; k < n; k++) {
    st[k].last_step_iterations = st[k].newton_iterations - before[k];
    if (conv[k] == 1) {
      st[k].load_steps++;
      st[k].last_time = t;
    }
  }
}

// ---- autotuner ------------------------------------------------------------
template <class F>
static float timeIt(hipStream_t s, int reps, F&& fn) {
  hipEvent_t a, e;
  hipCheck(hipEventCreate(&a), "event");
  hipCheck(hipEventCreate(&e), "event");
  fn();  // warm-up
  hipCheck(hipEventRecord(a, s), "rec");
  for (int r = 0; r < reps; r++) fn();
  hipCheck(hipEventRecord(e, s), "rec");
  hipCheck(hipEventSynchronize(e), "sync");
  float ms = 0;
  hipCheck(hipEventElapsedTime(&ms, a, e), "elapsed");
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(e);
  return ms / reps;
}

static void autotune(Batch& b, double t, int probeIters, int pc, std::vector<double>& report) {
  hipStream_t s = b.stream;
  b.dropGraph();
  // the probes apply the prescribed values at t to U (fix_kernel); every model's U is
  // put back afterwards, so a Newton iteration after tuning starts from the caller's state
  std::vector<DevBuf> savedU(b.m.size());
  for (size_t k = 0; k < b.m.size(); k++) {
    Ctx& c = *b.m[k];
    savedU[k].alloc(c.d_U.bytes);
    hipCheck(hipMemcpyAsync(savedU[k].p, c.d_U.p, c.d_U.bytes, hipMemcpyDeviceToDevice, s), "save U");
  }
  preSolveChain(b, t);  // current matrices, right-hand sides and preconditioners
  hipCheck(hipStreamSynchronize(s), "sync");
  // element kernel: persistent grid
  const int gridCands[] = {512, 1024, 2048, 4096, 8192, 16384, 0};
  float best = 1e30f;
  for (int g : gridCands) {
    const int grid = g > 0 ? std::min(g, b.nBlocks) : b.nBlocks;
    const float ms = timeIt(s, 3, [&] {
      launch_solid81_batched(s, grid, (int)b.m.size(), b.d_s81.as<Solid81Batch>(), b.d_blockBegin.as<int>(),
                             b.nBlocks);
    });
    report.push_back(ms);
    if (ms < best) {
      best = ms;
      b.cfg.elem_grid = g;
    }
  }
  best = 1e30f;
  for (int th : {128, 256, 512, 1024}) {
    const float ms = timeIt(s, 3, [&] {
      launch_gather_batched(s, th, (int)b.m.size(), b.d_gather.as<GatherBatch>(), b.d_targetBegin.as<int>(),
                            b.nTargets);
    });
    report.push_back(ms);
    if (ms < best) {
      best = ms;
      b.cfg.gather_threads = th;
    }
  }
  // team shape: every model in the team launch, a fixed number of iterations (tol 0)
  std::vector<int> all(b.m.size());
  std::iota(all.begin(), all.end(), 0);
  std::vector<int> its(b.m.size()), flags(b.m.size());
  std::vector<double> rr(b.m.size());
  const long long keepBig = b.cfg.big_rows;
  b.cfg.big_rows = kTeamMaxRows;
  best = 1e30f;
  int bt = 256, bl = 8;
  for (int th : {256, 512, 1024})
    for (int ln : {4, 8, 16}) {
      b.cfg.team_threads = th;
      b.cfg.team_lanes = ln;
      const double t0 = nowMsB();
      solveModels(b, all, 0.0, probeIters, its, flags, rr);
      const float ms = (float)(nowMsB() - t0);
      report.push_back(ms);
      if (ms < best) {
        best = ms;
        bt = th;
        bl = ln;
      }
    }
  b.cfg.team_threads = bt;
  b.cfg.team_lanes = bl;
  // how many of the largest models leave the team launch for the whole-device solver.
  // Each side's probe time per iteration is scaled by an expected iteration count that
  // grows with the mesh edge e = rows^(1/3): block Jacobi ~ 100 e, block Jacobi + Schur
  // multigrid ~ 150 e^(1/3) (measured on SOLID81 boxes from 32^3 to 120x120x60).  The team
  // launch's makespan is set by its largest member; the whole-device solves run
  // kBigWorkers at a time, each preceded (multigrid) by its refresh.
  std::vector<int> order(all);
  std::sort(order.begin(), order.end(), [&](int a, int c) { return b.n2[a] > b.n2[c]; });
  auto edge = [&](int k) { return std::cbrt((double)b.n2[k]); };
  auto itersBJ = [&](int k) { return 100.0 * edge(k); };
  auto itersMG = [&](int k) { return 150.0 * std::cbrt(edge(k)); };
  std::vector<double> devMs;  // expected whole-device solve time of order[i]
  const int maxBig = std::min<int>(128, (int)order.size() - 1);
  for (int i = 0; i < maxBig; i++) {
    Ctx& c = *b.m[order[i]];
    double r = 0, tRefresh = 0;
    bool brk = false;
    PrecondHook hook;
    FusedStage fused;
    bool mg = false;
    if (pc >= 2) {
      std::string why;
      const double t0 = nowMsB();
      mg = schurUsable(c, &why) && schurRefresh(c, &why);
      tRefresh = nowMsB() - t0;
      if (mg) {
        hook = schurHook(c);
        fused = schurFusedStage(c);
      }
    }
    const double t0 = nowMsB();
    c.minresSolve(c.d_rhs.as<double>(), c.d_dU.as<double>(), 0.0, probeIters, &r, &brk, mg ? &hook : nullptr, 0.0,
                  nullptr, 0.0, mg ? &fused : nullptr);
    const double perIt = (nowMsB() - t0) / probeIters;
    devMs.push_back(tRefresh + perIt * (mg ? itersMG(order[i]) : itersBJ(order[i])));
  }
  double bestCost = 1e300;
  long long bestRows = (1ll << 62);
  for (int k : {0, 1, 2, 4, 8, 16, 24, 32, 48, 64, 96, 128}) {
    if (k > maxBig) break;
    if (k < maxBig && b.n2[order[k]] > kTeamMaxRows) continue;  // too large for one workgroup
    std::vector<int> rest(order.begin() + k, order.end());
    b.cfg.big_rows = kTeamMaxRows;
    const double t0 = nowMsB();
    solveModels(b, rest, 0.0, probeIters, its, flags, rr);
    double cost = (nowMsB() - t0) / probeIters * itersBJ(order[k]);
    double bigSum = 0.0, bigMax = 0.0;
    for (int i = 0; i < k; i++) {
      bigSum += devMs[i];
      bigMax = std::max(bigMax, devMs[i]);
    }
    cost = std::max(cost, std::max(bigMax, bigSum / kBigWorkers));  // the two run concurrently
    report.push_back(cost);
    if (cost < bestCost) {
      bestCost = cost;
      bestRows = k == 0 ? kTeamMaxRows : (long long)b.n2[order[k]];
    }
  }
  b.cfg.big_rows = bestRows;
  (void)keepBig;
  for (size_t k = 0; k < b.m.size(); k++) {
    Ctx& c = *b.m[k];
    hipCheck(hipMemcpyAsync(c.d_U.p, savedU[k].p, c.d_U.bytes, hipMemcpyDeviceToDevice, s), "restore U");
  }
  hipCheck(hipStreamSynchronize(s), "restore sync");
  b.dropGraph();
}

}  // namespace nla3d_amd

using namespace nla3d_amd;

static int batchGuard(nla3d_gpu_batch_t h, const std::function<void(Batch&)>& fn) {
  if (!h) return NLA3D_GPU_EINVAL;
  Batch& b = *reinterpret_cast<Batch*>(h);
  try {
    b.err.clear();
    hipCheck(hipSetDevice(b.device), "hipSetDevice");
    fn(b);
    return NLA3D_GPU_OK;
  } catch (const ApiError& e) {
    b.err = e.what();
    return e.code;
  } catch (const std::exception& e) {
    b.err = e.what();
    return NLA3D_GPU_EINVAL;
  }
}

extern "C" {

void nla3d_gpu_batch_default_config(nla3d_gpu_batch_config* c) {
  if (!c) return;
  c->elem_grid = 4096;
  c->gather_threads = 256;
  c->team_threads = 256;
  c->team_lanes = 8;
  c->team_grid = 1024;
  c->big_rows = (1ll << 62);
}

int nla3d_gpu_batch_create(nla3d_gpu_batch_t* out, const nla3d_gpu_t* models, int n) {
  if (!out || !models || n < 1) return NLA3D_GPU_EINVAL;
  *out = nullptr;
  Batch* b = new (std::nothrow) Batch();
  if (!b) return NLA3D_GPU_EINVAL;
  try {
    b->device = -2;
    for (int k = 0; k < n; k++) {
      if (!models[k]) throw ApiError(-1, "null model handle");
      Ctx* c = reinterpret_cast<Ctx*>(models[k]);
      c->requireSolid81();
      if (b->device == -2) b->device = c->device;
      if (c->device != b->device) throw ApiError(-1, "all models of a batch must be on one device");
      if (!c->krylovReady) throw ApiError(-3, "model has no Krylov data");
      if (!c->model.mpcs.empty() || !c->rbs.empty()) throw ApiError(-1, "batched models may not carry MPCs");
      if (c->model.solidCount != c->model.nElements())
        throw ApiError(-1, "batched models hold SOLID81 elements only");
      b->m.push_back(c);
      b->n2.push_back(c->n2);
    }
    hipCheck(hipSetDevice(b->device), "hipSetDevice");
    hipCheck(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking), "stream");
    for (hipStream_t& w : b->workers) hipCheck(hipStreamCreateWithFlags(&w, hipStreamNonBlocking), "stream");
    std::vector<Solid81Batch> s81(n);
    std::vector<GatherBatch> ga(n);
    std::vector<int> bb(n), tb(n);
    long long nb = 0, nt = 0;
    for (int k = 0; k < n; k++) {
      Ctx& c = *b->m[k];
      s81[k] = Solid81Batch{(int)c.model.solidCount, c.d_xyz.as<double>(), c.d_conn.as<int>(), c.d_nodeEq3.as<int>(),
                            c.d_elemEqP.as<int>(), c.d_U.as<double>(), c.mat, c.d_Ke.as<double>()};
      const int off2 = (int)c.model.offB2();
      ga[k] = GatherBatch{c.nTargets, c.d_seg.as<int>(), c.d_src.as<int>(), c.d_Ke.as<double>(), c.d_vals.as<double>(),
                          off2, off2 + (int)c.model.block2.nnz(), c.d_mir2.as<int>(), c.d_full2.as<double>()};
      bb[k] = (int)nb;
      tb[k] = (int)nt;
      nb += solid81_blocks((int)c.model.solidCount);
      nt += c.nTargets;
    }
    if (nb >= (1ll << 31) || nt >= (1ll << 31)) throw ApiError(-1, "batch too large for 32-bit indices");
    b->nBlocks = (int)nb;
    b->nTargets = (int)nt;
    b->d_s81.alloc(n * sizeof(Solid81Batch));
    hipCheck(hipMemcpy(b->d_s81.p, s81.data(), n * sizeof(Solid81Batch), hipMemcpyHostToDevice), "s81");
    b->d_gather.alloc(n * sizeof(GatherBatch));
    hipCheck(hipMemcpy(b->d_gather.p, ga.data(), n * sizeof(GatherBatch), hipMemcpyHostToDevice), "gather");
    b->d_blockBegin.upload(bb);
    b->d_targetBegin.upload(tb);
    b->d_team.alloc(n * sizeof(TeamSolve));
    b->d_queue.alloc(n * 4);
    b->d_counter.alloc(256);
    b->d_out.alloc(2 * n * 4);
    b->d_relres.alloc(n * 8);
    b->d_norms.alloc(n * 8);
    nla3d_gpu_batch_default_config(&b->cfg);
  } catch (const ApiError& e) {
    delete b;
    return e.code;
  } catch (const std::exception&) {
    delete b;
    return NLA3D_GPU_EHIP;
  }
  *out = reinterpret_cast<nla3d_gpu_batch_t>(b);
  return NLA3D_GPU_OK;
}

int nla3d_gpu_batch_destroy(nla3d_gpu_batch_t h) {
  if (!h) return NLA3D_GPU_EINVAL;
  Batch* b = reinterpret_cast<Batch*>(h);
  (void)hipSetDevice(b->device);
  delete b;
  return NLA3D_GPU_OK;
}

const char* nla3d_gpu_batch_last_error(nla3d_gpu_batch_t h) {
  if (!h) return "null batch handle";
  return reinterpret_cast<Batch*>(h)->err.c_str();
}

int nla3d_gpu_batch_get_config(nla3d_gpu_batch_t h, nla3d_gpu_batch_config* c) {
  return batchGuard(h, [&](Batch& b) {
    if (!c) throw ApiError(-1, "null config");
    *c = b.cfg;
  });
}

int nla3d_gpu_batch_set_config(nla3d_gpu_batch_t h, const nla3d_gpu_batch_config* c) {
  return batchGuard(h, [&](Batch& b) {
    if (!c) throw ApiError(-1, "null config");
    if (!team_config_valid(c->team_threads, c->team_lanes)) throw ApiError(-1, "unsupported team shape");
    if (c->gather_threads < 64 || c->gather_threads > 1024 || c->gather_threads % 64)
      throw ApiError(-1, "gather_threads must be a multiple of 64 in [64, 1024]");
    if (c->elem_grid < 0 || c->team_grid < 0) throw ApiError(-1, "negative grid");
    b.cfg = *c;
    b.dropGraph();
  });
}

int nla3d_gpu_batch_autotune(nla3d_gpu_batch_t h, double t, int probe_iterations, int preconditioner,
                             nla3d_gpu_batch_config* chosen, double* report, int report_cap) {
  return batchGuard(h, [&](Batch& b) {
    if (probe_iterations < 1) throw ApiError(-1, "probe_iterations must be positive");
    std::vector<double> rep;
    autotune(b, t, probe_iterations, preconditioner, rep);
    if (chosen) *chosen = b.cfg;
    if (report)
      for (int i = 0; i < (int)rep.size() && i < report_cap; i++) report[i] = rep[i];
  });
}

int nla3d_gpu_batch_assemble(nla3d_gpu_batch_t h, double t) {
  return batchGuard(h, [&](Batch& b) {
    runPreSolve(b, t);
    hipCheck(hipStreamSynchronize(b.stream), "sync");
  });
}

int nla3d_gpu_batch_newton_iteration(nla3d_gpu_batch_t h, double t, int iteration, const nla3d_gpu_newton_opts* o,
                                     nla3d_gpu_newton_stats* stats, int* converged) {
  return batchGuard(h, [&](Batch& b) {
    if (!o || !stats || !converged || iteration < 1) throw ApiError(-1, "bad arguments");
    if (iteration == 1) {
      b.rhs1.assign(b.m.size(), 0.0);
      b.done.assign(b.m.size(), 0);
      for (size_t k = 0; k < b.m.size(); k++) converged[k] = 0;
    }
    batchNewtonIteration(b, t, iteration, *o, stats, converged);
  });
}

int nla3d_gpu_batch_solve_load_step(nla3d_gpu_batch_t h, double t, const nla3d_gpu_newton_opts* o,
                                    nla3d_gpu_newton_stats* stats, int* converged) {
  return batchGuard(h, [&](Batch& b) {
    if (!o || !stats || !converged) throw ApiError(-1, "bad arguments");
    batchLoadStep(b, t, *o, stats, converged);
  });
}

int nla3d_gpu_batch_solve_nonlinear(nla3d_gpu_batch_t h, const nla3d_gpu_newton_opts* o,
                                    nla3d_gpu_newton_stats* stats, int* converged) {
  return batchGuard(h, [&](Batch& b) {
    if (!o || !stats || !converged || o->n_load_steps < 1) throw ApiError(-1, "bad arguments");
    const double t0 = nowMsB();
    const size_t n = b.m.size();
    std::vector<int> ok(n, 1), cv(n);
    const double endT = 1.0, delta = 1.0 / o->n_load_steps;
    double cur = 0.0;
    while (cur < endT) {  // TimeControl::nextStep (FESolver.cpp:34-56)
      const double d = (cur + delta >= endT) ? endT - cur : ((cur + 2 * delta >= endT) ? (endT - cur) / 2.0 : delta);
      cur += d;
      batchLoadStep(b, cur, *o, stats, cv.data());
      for (size_t k = 0; k < n; k++)
        if (cv[k] != 1 && ok[k] == 1) ok[k] = cv[k] < 0 ? cv[k] : 0;
    }
    for (size_t k = 0; k < n; k++) {
      converged[k] = ok[k];
      stats[k].ms_total += nowMsB() - t0;
    }
  });
}

}  // extern "C"
