// Learned block-Jacobi correction: the network object, the per-Newton-iteration
// set-up (features -> BF16 MFMA forward -> corrected blocks), and training by
// reverse-mode differentiation through k steps of preconditioned MINRES
// (the Lanczos recurrence of Ctx::minresSolve), the loss being the mean log
// residual-reduction factor per step.  The corrections touch only the
// preconditioner, so a converged solve is the same solve; block Jacobi (all
// outputs zero, or a node whose diagonal block is not positive definite)
// is the fallback.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <functional>
#include <vector>

#include "../../include/nla3d_gpu.h"
#include "ctx.h"
#include "learned.h"

namespace nla3d_amd {

struct Mlp {
  int device = 0;
  hipStream_t stream = nullptr;
  std::string err;
  DevBuf w, m, v, g;  // FP32 parameters, Adam moments, gradient
  DevBuf wp, bp;      // packed BF16 weights and biases for the MFMA kernel
  DevBuf mean, istd;  // input normalisation
  int step = 0;
  bool packed = false;
  // FP32 training activations
  DevBuf X, X0, Z[5], A[4], dZa, dZb, partial, Y;
  size_t cap = 0;
  ~Mlp() {
    if (stream) (void)hipStreamDestroy(stream);
  }
  void ensure(size_t n) {
    if (n <= cap) return;
    X.alloc(n * kMlpIn * 4);
    X0.alloc(n * kMlpIn * 4);
    for (int l = 0; l < 4; l++) {
      Z[l].alloc(n * kMlpWidth * 4);
      A[l].alloc(n * kMlpWidth * 4);
    }
    Z[4].alloc(n * kMlpOut * 4);
    Y.alloc(n * kMlpOut * 4);
    dZa.alloc(n * kMlpWidth * 4);
    dZb.alloc(n * kMlpWidth * 4);
    partial.alloc(((n + kGradChunk - 1) / kGradChunk) * kMlpWidth * (kMlpWidth + 1) * 4);
    cap = n;
  }
  void pack(hipStream_t s) {
    if (packed) return;
    launch_mlp_pack(s, w.as<float>(), wp.as<unsigned short>(), bp.as<float>());
    packed = true;
  }
};

struct LearnedWork {
  DevBuf X, Y, dY, Rinv, ok, Pbar;  // per learned node
  DevBuf vecs;                      // unrolled MINRES vectors
  int kCap = -1;
};

static LearnedWork& work(Ctx& c) {
  if (!c.lw) {
    c.lw = std::make_shared<LearnedWork>();
    const size_t nL = std::max(c.nLearned, 1);
    c.lw->X.alloc(nL * kMlpIn * 4);
    c.lw->Y.alloc(nL * kMlpOut * 4);
    c.lw->dY.alloc(nL * kMlpOut * 4);
    c.lw->Rinv.alloc(nL * 6 * 8);
    c.lw->ok.alloc(nL * 4);
    c.lw->Pbar.alloc(nL * 9 * 8);
  }
  return *c.lw;
}

static void requireLearnable(Ctx& c) {
  c.requireSolid81();
  if (!c.krylovReady) throw ApiError(-3, "Krylov data not set up");
}

// features of the current tangent (block-Jacobi blocks must be current)
static void nodeFeatures(Ctx& c, double t, float* X) {
  FeatureInputs in{};
  in.nL = c.nLearned;
  in.lRow0 = c.d_lRow0.as<int>();
  in.rowKind = c.d_rowKind2.as<signed char>();
  in.rowNode = c.d_rowNode2.as<int>();
  in.k12Ptr = c.d_k12tPtr.as<int>();
  in.rowPGroup = c.d_rowPG.as<int>();
  in.binv = c.d_binv.as<double>();
  in.U = c.d_U.as<double>();
  in.nC = (int)c.model.nConstrainedDofs;
  in.bulk = c.mat.k;
  in.t = t;
  launch_node_features(c.stream, c.krylovMatrix(), in, X);
}

// FP32 forward of n rows of features X (device) into mlp buffers; returns Y (device)
static float* forwardF32(Mlp& net, hipStream_t s, int n, const float* X) {
  net.ensure((size_t)n);
  const float* w = net.w.as<float>();
  launch_normalize(s, n, X, net.mean.as<float>(), net.istd.as<float>(), net.X0.as<float>());
  const float* in = net.X0.as<float>();
  int nin = kMlpIn;
  for (int l = 0; l < 4; l++) {
    launch_dense_fwd(s, n, nin, kMlpWidth, in, w + kMlpOffW[l], w + kMlpOffB[l], net.Z[l].as<float>(),
                     net.A[l].as<float>(), 1);
    in = net.A[l].as<float>();
    nin = kMlpWidth;
  }
  launch_dense_fwd(s, n, kMlpWidth, kMlpOut, in, w + kMlpOffW[4], w + kMlpOffB[4], net.Z[4].as<float>(), nullptr, 0);
  return net.Z[4].as<float>();
}

// FP32 backward from dY (device, n x 6) after forwardF32; writes the parameter gradient into net.g
static void backwardF32(Mlp& net, hipStream_t s, int n, const float* dY) {
  float* g = net.g.as<float>();
  const float* w = net.w.as<float>();
  const float* dZ = dY;
  float* bufs[2] = {net.dZa.as<float>(), net.dZb.as<float>()};
  int nout = kMlpOut;
  for (int l = 4; l >= 0; l--) {
    const int nin = l == 0 ? kMlpIn : kMlpWidth;
    const float* Xin = l == 0 ? net.X0.as<float>() : net.A[l - 1].as<float>();
    launch_dense_bwd_weight(s, n, nin, nout, dZ, Xin, net.partial.as<float>(), g + kMlpOffW[l], g + kMlpOffB[l]);
    if (l > 0) {
      float* dprev = bufs[l & 1];
      launch_dense_bwd_input(s, n, nin, nout, dZ, w + kMlpOffW[l], net.Z[l - 1].as<float>(), dprev);
      dZ = dprev;
    }
    nout = kMlpWidth;
  }
}

// the network's corrections over the learned nodes (block-Jacobi blocks must be current)
static void learnedPreconditioner(Ctx& c, Mlp& net, double t) {
  LearnedWork& lw = work(c);
  if (c.nLearned == 0) return;
  nodeFeatures(c, t, lw.X.as<float>());
  net.pack(c.stream);
  launch_mlp_bf16(c.stream, c.nLearned, lw.X.as<float>(), net.mean.as<float>(), net.istd.as<float>(),
                  net.wp.as<unsigned short>(), net.bp.as<float>(), lw.Y.as<float>());
  launch_learned_blocks(c.stream, c.nLearned, c.d_lGroup.as<int>(), c.precond(), c.krylovMatrix(), lw.Y.as<float>(),
                        nullptr, nullptr);
  hipCheck(hipGetLastError(), "learned preconditioner");
}

void learnedCorrect(Ctx& c, double t) {
  if (!c.mlp) throw ApiError(-3, "no network attached (nla3d_gpu_attach_mlp)");
  if (c.mlp->device != c.device) throw ApiError(-1, "network and model are on different devices");
  learnedPreconditioner(c, *c.mlp, t);
}

void learnedPrecondSetup(Ctx& c, double t) {
  krylov_precond_setup(c.stream, c.krylovMatrix(), c.precond());
  learnedCorrect(c, t);
}

// ---- unrolled MINRES --------------------------------------------------------
struct Dual {
  double v, d;
};
static inline Dual operator+(Dual a, Dual b) { return {a.v + b.v, a.d + b.d}; }
static inline Dual operator-(Dual a, Dual b) { return {a.v - b.v, a.d - b.d}; }
static inline Dual operator*(Dual a, Dual b) { return {a.v * b.v, a.d * b.v + a.v * b.d}; }
static inline Dual operator/(Dual a, Dual b) { return {a.v / b.v, (a.d * b.v - a.v * b.d) / (b.v * b.v)}; }
static inline Dual operator-(Dual a) { return {-a.v, -a.d}; }
static inline Dual dsqrt(Dual a) {
  const double r = std::sqrt(a.v);
  return {r, a.d / (2.0 * r)};
}
static inline Dual dlog(Dual a) { return {std::log(a.v), a.d / a.v}; }

// mean log of the MINRES residual-reduction factor over k steps, from the Lanczos
// coefficients alpha[0..k-1], beta[0..k] (the plane-rotation part of the solver)
static Dual minresLogRate(int k, const std::vector<Dual>& alpha, const std::vector<Dual>& beta) {
  Dual dbar{0, 0}, cs{-1, 0}, sn{0, 0}, L{0, 0};
  for (int j = 0; j < k; j++) {
    const Dual gbar = sn * dbar - cs * alpha[j];
    const Dual b = beta[j + 1];
    dbar = -(cs * b);
    const Dual gamma = dsqrt(gbar * gbar + b * b);
    cs = gbar / gamma;
    sn = b / gamma;
    L = L + dlog(sn);
  }
  return {L.v / k, L.d / k};
}

static double hostDot(Ctx& c, int n, const double* a, const double* b) {
  krylov_dot(c.stream, n, a, b, c.d_partial.as<double>(), c.d_scal.as<double>() + 8);
  double v;
  hipCheck(hipMemcpyAsync(&v, c.d_scal.as<double>() + 8, 8, hipMemcpyDeviceToHost, c.stream), "dot");
  hipCheck(hipStreamSynchronize(c.stream), "dot sync");
  return v;
}

// loss and dL/dY for rhs b (device, n2) with the corrected preconditioner of Y;
// returns false when the recurrence broke down (sample unusable)
static bool unrolledLossGrad(Ctx& c, int k, const double* b, double& loss, const PrecondHook* hook) {
  LearnedWork& lw = work(c);
  const int n = c.n2;
  const size_t nv = (size_t)n;
  if (lw.kCap < k) {
    lw.vecs.alloc((4 * (size_t)(k + 1) + 6) * nv * 8);
    lw.kCap = k;
  }
  double* base = lw.vecs.as<double>();
  auto R = [&](int j) { return base + (size_t)j * nv; };
  auto Yv = [&](int j) { return base + (size_t)(k + 1 + j) * nv; };
  auto Rb = [&](int j) { return base + (size_t)(2 * (k + 1) + j) * nv; };
  auto Yb = [&](int j) { return base + (size_t)(3 * (k + 1) + j) * nv; };
  double* V = base + 4 * (size_t)(k + 1) * nv;
  double* T = V + nv;
  double* Tb = T + nv;
  double* Vb = Tb + nv;
  double* W = Vb + nv;
  double* T3 = W + nv;
  const KrylovMatrix M = c.krylovMatrix();
  const Precond P = c.precond();
  hipStream_t s = c.stream;
  // the preconditioner: (corrected) block Jacobi, then the optional stage on the Schur rows
  // (a fixed symmetric operator, so its adjoint is itself)
  auto applyP = [&](const double* r, double* z) {
    launch_bj_apply(s, P, r, z);
    if (hook) (*hook)(s, r, z, c.d_partial.as<double>(), nullptr);
  };

  std::vector<double> alpha(k), beta(k + 1);
  hipCheck(hipMemcpyAsync(R(0), b, nv * 8, hipMemcpyDeviceToDevice, s), "rhs");
  applyP(R(0), Yv(0));
  beta[0] = std::sqrt(hostDot(c, n, R(0), Yv(0)));
  if (!(beta[0] > 0.0)) return false;
  for (int j = 0; j < k; j++) {
    launch_axpby(s, n, 1.0 / beta[j], Yv(j), 0.0, Yv(j), V);
    krylov_spmv(s, M, V, T);
    if (j >= 1) launch_axpby(s, n, 1.0, T, -beta[j] / beta[j - 1], R(j - 1), T);
    alpha[j] = hostDot(c, n, V, T);
    launch_axpby(s, n, 1.0, T, -alpha[j] / beta[j], R(j), R(j + 1));
    applyP(R(j + 1), Yv(j + 1));
    const double bb = hostDot(c, n, R(j + 1), Yv(j + 1));
    if (!(bb > 0.0)) return false;
    beta[j + 1] = std::sqrt(bb);
  }
  // scalar part: dL/dalpha, dL/dbeta by forward-mode duals (k is small)
  std::vector<Dual> ad(k), bd(k + 1);
  for (int j = 0; j < k; j++) ad[j] = {alpha[j], 0.0};
  for (int j = 0; j <= k; j++) bd[j] = {beta[j], 0.0};
  std::vector<double> abar(k, 0.0), bbar(k + 1, 0.0);
  loss = minresLogRate(k, ad, bd).v;
  if (!(loss == loss)) return false;
  for (int j = 0; j < k; j++) {
    ad[j].d = 1.0;
    abar[j] = minresLogRate(k, ad, bd).d;
    ad[j].d = 0.0;
  }
  for (int j = 1; j <= k; j++) {
    bd[j].d = 1.0;
    bbar[j] = minresLogRate(k, ad, bd).d;
    bd[j].d = 0.0;
  }
  // reverse sweep over the vector recurrence
  launch_zero_d(s, 2 * (size_t)(k + 1) * nv, Rb(0));  // Rb and Yb are contiguous
  launch_zero_d(s, (size_t)c.nLearned * 9, lw.Pbar.as<double>());
  for (int j = k - 1; j >= 0; j--) {
    // beta_{j+1} = sqrt(R_{j+1} . Y_{j+1})
    const double cb = bbar[j + 1] / (2.0 * beta[j + 1]);
    launch_axpby(s, n, cb, Yv(j + 1), 1.0, Rb(j + 1), Rb(j + 1));
    launch_axpby(s, n, cb, R(j + 1), 1.0, Yb(j + 1), Yb(j + 1));
    // Y_{j+1} = P R_{j+1}
    launch_outer_acc(s, c.nLearned, c.d_lRow0.as<int>(), Yb(j + 1), R(j + 1), lw.Pbar.as<double>());
    applyP(Yb(j + 1), W);
    launch_axpby(s, n, 1.0, W, 1.0, Rb(j + 1), Rb(j + 1));
    // R_{j+1} = T3 - d R_j, d = alpha_j / beta_j
    const double d = alpha[j] / beta[j];
    hipCheck(hipMemcpyAsync(Tb, Rb(j + 1), nv * 8, hipMemcpyDeviceToDevice, s), "Tb");
    const double dbarv = -hostDot(c, n, R(j), Rb(j + 1));
    launch_axpby(s, n, -d, Rb(j + 1), 1.0, Rb(j), Rb(j));
    abar[j] += dbarv / beta[j];
    bbar[j] += -dbarv * alpha[j] / (beta[j] * beta[j]);
    // alpha_j = V . T3, T3 = R_{j+1} + d R_j, V = Y_j / beta_j
    launch_axpby(s, n, 1.0, R(j + 1), d, R(j), T3);
    launch_axpby(s, n, abar[j], T3, 0.0, T3, Vb);
    launch_axpby(s, n, 1.0 / beta[j], Yv(j), 0.0, Yv(j), V);
    launch_axpby(s, n, abar[j], V, 1.0, Tb, Tb);
    // T3 = T2 - cj R_{j-1}, cj = beta_j / beta_{j-1}
    if (j >= 1) {
      const double cj = beta[j] / beta[j - 1];
      const double cbar = -hostDot(c, n, R(j - 1), Tb);
      launch_axpby(s, n, -cj, Tb, 1.0, Rb(j - 1), Rb(j - 1));
      bbar[j] += cbar / beta[j - 1];
      bbar[j - 1] += -cbar * beta[j] / (beta[j - 1] * beta[j - 1]);
    }
    // T2 = K V
    krylov_spmv(s, M, Tb, W);
    launch_axpby(s, n, 1.0, W, 1.0, Vb, Vb);
    // V = Y_j / beta_j
    launch_axpby(s, n, 1.0 / beta[j], Vb, 1.0, Yb(j), Yb(j));
    bbar[j] += -hostDot(c, n, Vb, Yv(j)) / (beta[j] * beta[j]);
  }
  // beta_0 = sqrt(R_0 . Y_0), Y_0 = P R_0 (R_0 = b is data)
  const double c0 = bbar[0] / (2.0 * beta[0]);
  launch_axpby(s, n, c0, R(0), 1.0, Yb(0), Yb(0));
  launch_outer_acc(s, c.nLearned, c.d_lRow0.as<int>(), Yb(0), R(0), lw.Pbar.as<double>());
  hipCheck(hipGetLastError(), "unrolled minres");
  return true;
}

// loss and parameter gradient (net.g) for one (tangent, rhs) sample
static bool lossGrad(Mlp& net, Ctx& c, int k, double t, const double* rhsDev, double& loss, bool multigrid) {
  requireLearnable(c);
  if (c.nLearned == 0) throw ApiError(-3, "model has no free UX UY UZ node triples");
  if (k < 1 || k > 1024) throw ApiError(-1, "k must be in [1, 1024]");
  LearnedWork& lw = work(c);
  krylov_precond_setup(c.stream, c.krylovMatrix(), c.precond());
  PrecondHook hook;
  if (multigrid) {  // S from the plain block-Jacobi blocks, as preconditioner 3 builds it
    std::string why;
    if (!schurUsable(c, &why) || !schurRefresh(c, &why)) throw ApiError(-3, "Schur multigrid: " + why);
    hook = schurHook(c);
  }
  nodeFeatures(c, t, lw.X.as<float>());
  const float* Y = forwardF32(net, c.stream, c.nLearned, lw.X.as<float>());
  launch_learned_blocks(c.stream, c.nLearned, c.d_lGroup.as<int>(), c.precond(), c.krylovMatrix(), Y,
                        lw.Rinv.as<double>(), lw.ok.as<int>());
  if (!unrolledLossGrad(c, k, rhsDev, loss, multigrid ? &hook : nullptr)) return false;
  launch_learned_grad(c.stream, c.nLearned, lw.Rinv.as<double>(), lw.ok.as<int>(), Y, lw.Pbar.as<double>(),
                      lw.dY.as<float>());
  backwardF32(net, c.stream, c.nLearned, lw.dY.as<float>());
  hipCheck(hipGetLastError(), "loss gradient");
  return true;
}

}  // namespace nla3d_amd

using namespace nla3d_amd;

static int mlpGuard(nla3d_gpu_mlp_t h, const std::function<void(Mlp&)>& fn) {
  if (!h) return NLA3D_GPU_EINVAL;
  Mlp& m = *reinterpret_cast<Mlp*>(h);
  try {
    m.err.clear();
    hipCheck(hipSetDevice(m.device), "hipSetDevice");
    fn(m);
    return NLA3D_GPU_OK;
  } catch (const ApiError& e) {
    m.err = e.what();
    return e.code;
  } catch (const std::exception& e) {
    m.err = e.what();
    return NLA3D_GPU_EINVAL;
  }
}

// errors of calls that take both objects are reported on the model handle
static int pairGuard(nla3d_gpu_mlp_t mh, nla3d_gpu_t h, const std::function<void(Mlp&, Ctx&)>& fn) {
  if (!mh || !h) return NLA3D_GPU_EINVAL;
  Mlp& m = *reinterpret_cast<Mlp*>(mh);
  Ctx& c = *reinterpret_cast<Ctx*>(h);
  try {
    c.err.clear();
    m.err.clear();
    c.requireDevice();
    if (m.device != c.device) throw ApiError(-1, "network and model are on different devices");
    hipCheck(hipSetDevice(c.device), "hipSetDevice");
    fn(m, c);
    return NLA3D_GPU_OK;
  } catch (const ApiError& e) {
    c.err = m.err = e.what();
    return e.code;
  } catch (const std::exception& e) {
    c.err = m.err = e.what();
    return NLA3D_GPU_EINVAL;
  }
}

static uint64_t splitmix64(uint64_t x) {
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}

extern "C" {

int nla3d_gpu_mlp_num_params(void) { return kMlpParams; }

int nla3d_gpu_mlp_create(nla3d_gpu_mlp_t* out, int device, unsigned long long seed) {
  if (!out) return NLA3D_GPU_EINVAL;
  *out = nullptr;
  Mlp* m = new (std::nothrow) Mlp();
  if (!m) return NLA3D_GPU_EINVAL;
  m->device = device;
  try {
    hipCheck(hipSetDevice(device), "hipSetDevice");
    hipCheck(hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking), "stream");
    // Glorot-uniform hidden weights from splitmix64(seed, index); zero biases; zero output
    // layer, so a fresh network is exactly block Jacobi
    std::vector<float> w(kMlpParams, 0.0f);
    const int fanIn[5] = {kMlpIn, kMlpWidth, kMlpWidth, kMlpWidth, kMlpWidth};
    const int fanOut[5] = {kMlpWidth, kMlpWidth, kMlpWidth, kMlpWidth, kMlpOut};
    for (int l = 0; l < 4; l++) {
      const double a = std::sqrt(6.0 / (fanIn[l] + fanOut[l]));
      for (int q = 0; q < fanIn[l] * fanOut[l]; q++) {
        const uint64_t r = splitmix64(seed * 0x100000001B3ull + (uint64_t)(kMlpOffW[l] + q));
        const double u = (double)(r >> 11) * (1.0 / 9007199254740992.0);
        w[kMlpOffW[l] + q] = (float)((2.0 * u - 1.0) * a);
      }
    }
    m->w.upload(w);
    std::vector<float> zero(kMlpParams, 0.0f);
    m->m.upload(zero);
    m->v.upload(zero);
    m->g.upload(zero);
    m->wp.alloc((size_t)kMlpPackedW * 2);
    m->bp.alloc((size_t)kMlpPackedB * 4);
    m->mean.upload(std::vector<float>(kMlpIn, 0.0f));
    m->istd.upload(std::vector<float>(kMlpIn, 1.0f));
  } catch (const std::exception& e) {
    delete m;
    return NLA3D_GPU_EHIP;
  }
  *out = reinterpret_cast<nla3d_gpu_mlp_t>(m);
  return NLA3D_GPU_OK;
}

int nla3d_gpu_mlp_destroy(nla3d_gpu_mlp_t h) {
  if (!h) return NLA3D_GPU_EINVAL;
  Mlp* m = reinterpret_cast<Mlp*>(h);
  (void)hipSetDevice(m->device);
  delete m;
  return NLA3D_GPU_OK;
}

const char* nla3d_gpu_mlp_last_error(nla3d_gpu_mlp_t h) {
  if (!h) return "null network handle";
  return reinterpret_cast<Mlp*>(h)->err.c_str();
}

int nla3d_gpu_mlp_get_params(nla3d_gpu_mlp_t h, float* w, int n) {
  return mlpGuard(h, [&](Mlp& m) {
    if (!w || n != kMlpParams) throw ApiError(-1, "parameter array must hold nla3d_gpu_mlp_num_params() floats");
    hipCheck(hipMemcpy(w, m.w.p, (size_t)n * 4, hipMemcpyDeviceToHost), "get params");
  });
}

int nla3d_gpu_mlp_set_params(nla3d_gpu_mlp_t h, const float* w, int n) {
  return mlpGuard(h, [&](Mlp& m) {
    if (!w || n != kMlpParams) throw ApiError(-1, "parameter array must hold nla3d_gpu_mlp_num_params() floats");
    for (int i = 0; i < n; i++)
      if (!std::isfinite(w[i])) throw ApiError(-1, "non-finite parameter");
    hipCheck(hipMemcpy(m.w.p, w, (size_t)n * 4, hipMemcpyHostToDevice), "set params");
    m.packed = false;
  });
}

int nla3d_gpu_mlp_set_normalization(nla3d_gpu_mlp_t h, const float* mean, const float* inv_std) {
  return mlpGuard(h, [&](Mlp& m) {
    if (!mean || !inv_std) throw ApiError(-1, "null normalisation arrays");
    hipCheck(hipMemcpy(m.mean.p, mean, kMlpIn * 4, hipMemcpyHostToDevice), "mean");
    hipCheck(hipMemcpy(m.istd.p, inv_std, kMlpIn * 4, hipMemcpyHostToDevice), "istd");
  });
}

int nla3d_gpu_mlp_get_normalization(nla3d_gpu_mlp_t h, float* mean, float* inv_std) {
  return mlpGuard(h, [&](Mlp& m) {
    if (!mean || !inv_std) throw ApiError(-1, "null normalisation arrays");
    hipCheck(hipMemcpy(mean, m.mean.p, kMlpIn * 4, hipMemcpyDeviceToHost), "mean");
    hipCheck(hipMemcpy(inv_std, m.istd.p, kMlpIn * 4, hipMemcpyDeviceToHost), "istd");
  });
}

int nla3d_gpu_mlp_forward(nla3d_gpu_mlp_t h, const float* X, int n, float* Y, int bf16_mfma) {
  return mlpGuard(h, [&](Mlp& m) {
    if (!X || !Y || n < 1) throw ApiError(-1, "bad forward arguments");
    m.ensure((size_t)n);
    hipCheck(hipMemcpyAsync(m.X.p, X, (size_t)n * kMlpIn * 4, hipMemcpyHostToDevice, m.stream), "X");
    const float* out;
    if (bf16_mfma) {
      m.pack(m.stream);
      launch_mlp_bf16(m.stream, n, m.X.as<float>(), m.mean.as<float>(), m.istd.as<float>(),
                      m.wp.as<unsigned short>(), m.bp.as<float>(), m.Y.as<float>());
      out = m.Y.as<float>();
    } else {
      out = forwardF32(m, m.stream, n, m.X.as<float>());
    }
    hipCheck(hipGetLastError(), "forward");
    hipCheck(hipMemcpyAsync(Y, out, (size_t)n * kMlpOut * 4, hipMemcpyDeviceToHost, m.stream), "Y");
    hipCheck(hipStreamSynchronize(m.stream), "sync");
  });
}

int nla3d_gpu_mlp_backward(nla3d_gpu_mlp_t h, const float* X, const float* dY, int n, float* dW) {
  return mlpGuard(h, [&](Mlp& m) {
    if (!X || !dY || !dW || n < 1) throw ApiError(-1, "bad backward arguments");
    m.ensure((size_t)n);
    hipCheck(hipMemcpyAsync(m.X.p, X, (size_t)n * kMlpIn * 4, hipMemcpyHostToDevice, m.stream), "X");
    forwardF32(m, m.stream, n, m.X.as<float>());
    hipCheck(hipMemcpyAsync(m.Y.p, dY, (size_t)n * kMlpOut * 4, hipMemcpyHostToDevice, m.stream), "dY");
    backwardF32(m, m.stream, n, m.Y.as<float>());
    hipCheck(hipGetLastError(), "backward");
    hipCheck(hipMemcpyAsync(dW, m.g.p, (size_t)kMlpParams * 4, hipMemcpyDeviceToHost, m.stream), "dW");
    hipCheck(hipStreamSynchronize(m.stream), "sync");
  });
}

int nla3d_gpu_attach_mlp(nla3d_gpu_t h, nla3d_gpu_mlp_t mh) {
  if (!h) return NLA3D_GPU_EINVAL;
  Ctx& c = *reinterpret_cast<Ctx*>(h);
  Mlp* m = reinterpret_cast<Mlp*>(mh);
  if (m && m->device != c.device) {
    c.err = "network and model are on different devices";
    return NLA3D_GPU_EINVAL;
  }
  c.mlp = m;
  return NLA3D_GPU_OK;
}

int nla3d_gpu_node_features(nla3d_gpu_t h, double t, float* X, int* node_rows, int cap, int* n_out) {
  if (!h || !n_out) return NLA3D_GPU_EINVAL;
  Ctx& c = *reinterpret_cast<Ctx*>(h);
  try {
    c.err.clear();
    hipCheck(hipSetDevice(c.device), "hipSetDevice");
    requireLearnable(c);
    *n_out = c.nLearned;
    if (cap < c.nLearned || !X) return NLA3D_GPU_OK;  // size query
    LearnedWork& lw = work(c);
    krylov_precond_setup(c.stream, c.krylovMatrix(), c.precond());
    nodeFeatures(c, t, lw.X.as<float>());
    hipCheck(hipGetLastError(), "features");
    hipCheck(hipMemcpyAsync(X, lw.X.p, (size_t)c.nLearned * kMlpIn * 4, hipMemcpyDeviceToHost, c.stream), "X");
    if (node_rows)
      hipCheck(hipMemcpyAsync(node_rows, c.d_lRow0.p, (size_t)c.nLearned * 4, hipMemcpyDeviceToHost, c.stream), "rows");
    hipCheck(hipStreamSynchronize(c.stream), "sync");
    return NLA3D_GPU_OK;
  } catch (const ApiError& e) {
    c.err = e.what();
    return e.code;
  } catch (const std::exception& e) {
    c.err = e.what();
    return NLA3D_GPU_EINVAL;
  }
}

int nla3d_gpu_learned_blocks(nla3d_gpu_t h, const float* Y, int n, double* blocks) {
  if (!h) return NLA3D_GPU_EINVAL;
  Ctx& c = *reinterpret_cast<Ctx*>(h);
  try {
    c.err.clear();
    hipCheck(hipSetDevice(c.device), "hipSetDevice");
    requireLearnable(c);
    if (!Y || !blocks || n != c.nLearned) throw ApiError(-1, "Y and blocks must cover every learned node");
    LearnedWork& lw = work(c);
    krylov_precond_setup(c.stream, c.krylovMatrix(), c.precond());
    hipCheck(hipMemcpyAsync(lw.Y.p, Y, (size_t)n * kMlpOut * 4, hipMemcpyHostToDevice, c.stream), "Y");
    launch_learned_blocks(c.stream, n, c.d_lGroup.as<int>(), c.precond(), c.krylovMatrix(), lw.Y.as<float>(), nullptr,
                          nullptr);
    hipCheck(hipGetLastError(), "learned blocks");
    std::vector<double> binv((size_t)c.nGroups * 9);
    std::vector<int> lg(n);
    hipCheck(hipMemcpyAsync(binv.data(), c.d_binv.p, binv.size() * 8, hipMemcpyDeviceToHost, c.stream), "binv");
    hipCheck(hipMemcpyAsync(lg.data(), c.d_lGroup.p, (size_t)n * 4, hipMemcpyDeviceToHost, c.stream), "groups");
    hipCheck(hipStreamSynchronize(c.stream), "sync");
    for (int i = 0; i < n; i++) std::memcpy(blocks + (size_t)i * 9, binv.data() + (size_t)lg[i] * 9, 9 * 8);
    return NLA3D_GPU_OK;
  } catch (const ApiError& e) {
    c.err = e.what();
    return e.code;
  } catch (const std::exception& e) {
    c.err = e.what();
    return NLA3D_GPU_EINVAL;
  }
}

static const double* rhsOnDevice(Ctx& c, const double* rhs) {
  if (!rhs) return c.d_rhs.as<double>();  // the last Newton right-hand side
  double* d = c.d_mw[7].as<double>();
  hipCheck(hipMemcpyAsync(d, rhs, (size_t)c.n2 * 8, hipMemcpyHostToDevice, c.stream), "rhs");
  return d;
}

int nla3d_gpu_mlp_loss_grad(nla3d_gpu_mlp_t mh, nla3d_gpu_t h, const double* rhs, int k, double t, int multigrid,
                            double* loss, float* grad) {
  return pairGuard(mh, h, [&](Mlp& m, Ctx& c) {
    if (!loss) throw ApiError(-1, "null loss");
    requireLearnable(c);
    double L = 0.0;
    if (!lossGrad(m, c, k, t, rhsOnDevice(c, rhs), L, multigrid != 0))
      throw ApiError(-6, "Lanczos recurrence broke down");
    *loss = L;
    if (grad) hipCheck(hipMemcpyAsync(grad, m.g.p, (size_t)kMlpParams * 4, hipMemcpyDeviceToHost, c.stream), "grad");
    hipCheck(hipStreamSynchronize(c.stream), "sync");
  });
}

void nla3d_gpu_adam_defaults(nla3d_gpu_adam_opts* o) {
  if (!o) return;
  o->lr = 1e-3f;
  o->beta1 = 0.9f;
  o->beta2 = 0.999f;
  o->eps = 1e-8f;
}

int nla3d_gpu_mlp_train_step(nla3d_gpu_mlp_t mh, nla3d_gpu_t h, const double* rhs, int k, double t, int multigrid,
                             const nla3d_gpu_adam_opts* o, double* loss) {
  return pairGuard(mh, h, [&](Mlp& m, Ctx& c) {
    if (!o || !loss) throw ApiError(-1, "null options or loss");
    requireLearnable(c);
    double L = 0.0;
    if (!lossGrad(m, c, k, t, rhsOnDevice(c, rhs), L, multigrid != 0))
      throw ApiError(-6, "Lanczos recurrence broke down");
    m.step++;
    launch_adam(c.stream, kMlpParams, o->lr, o->beta1, o->beta2, o->eps, m.step, m.g.as<float>(), m.w.as<float>(),
                m.m.as<float>(), m.v.as<float>());
    hipCheck(hipGetLastError(), "adam");
    m.packed = false;
    hipCheck(hipStreamSynchronize(c.stream), "sync");
    *loss = L;
  });
}

int nla3d_gpu_solve_block2_pc(nla3d_gpu_t h, const double* rhs, double* x, double tol, int max_iter,
                              int preconditioner, double t, int* iterations, double* relres) {
  if (!h) return NLA3D_GPU_EINVAL;
  Ctx& c = *reinterpret_cast<Ctx*>(h);
  try {
    c.err.clear();
    hipCheck(hipSetDevice(c.device), "hipSetDevice");
    requireLearnable(c);
    if (!rhs || !x) throw ApiError(-1, "null vectors");
    PrecondHook hook;
    FusedStage fused;
    if (preconditioner < 0 || preconditioner > 3)
      throw ApiError(-1, "preconditioner must be 0 (block Jacobi), 1 (learned), 2 (Schur multigrid) or 3 (both)");
    krylov_precond_setup(c.stream, c.krylovMatrix(), c.precond());
    const bool mg = preconditioner >= 2;
    if (mg) {  // S from the plain block-Jacobi blocks
      std::string why;
      if (!schurUsable(c, &why) || !schurRefresh(c, &why)) throw ApiError(-3, "Schur multigrid: " + why);
      hook = schurHook(c);
      fused = schurFusedStage(c);
    }
    if (preconditioner == 1 || preconditioner == 3) learnedCorrect(c, t);
    double* b = c.d_mw[7].as<double>();
    hipCheck(hipMemcpyAsync(b, rhs, (size_t)c.n2 * 8, hipMemcpyHostToDevice, c.stream), "rhs");
    double rr = 0;
    bool brk = false;
    const int it = c.minresSolve(b, c.d_dU.as<double>(), tol, max_iter, &rr, &brk, mg ? &hook : nullptr, 0.0,
                                 nullptr, 0.0, mg ? &fused : nullptr);
    hipCheck(hipMemcpy(x, c.d_dU.p, (size_t)c.n2 * 8, hipMemcpyDeviceToHost), "x");
    if (iterations) *iterations = it;
    if (relres) *relres = rr;
    if (brk) throw ApiError(-6, "MINRES breakdown (preconditioner not positive definite)");
    return NLA3D_GPU_OK;
  } catch (const ApiError& e) {
    c.err = e.what();
    return e.code;
  } catch (const std::exception& e) {
    c.err = e.what();
    return NLA3D_GPU_EINVAL;
  }
}

}  // extern "C"
