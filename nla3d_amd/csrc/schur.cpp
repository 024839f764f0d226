// Host side of the Schur-row multigrid preconditioner (see schur.h): structure set-up
// (once per model), the refresh of values with every tangent, the V-cycle launch
// sequence, its MINRES hook and the ABI entries that expose S and one V-cycle.
#include <algorithm>
#include <unordered_map>
#include <climits>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <memory>
#include <vector>

#include "../../include/nla3d_gpu.h"
#include "ctx.h"
#include "learned.h"
#include "schur.h"

namespace nla3d_amd {

constexpr int kCoarsest = 256;   // at most this many rows are inverted densely
constexpr int kMaxCoarse = 2048; // coarsening that stalls above this disables the preconditioner
constexpr int kMaxQPerNode = 64;
// coarse levels whose smoother copy holds at most this many slots (and coarsest levels) run
// inside one workgroup: below it a pass is cheaper than a launch, above it (C3's 1,372-row
// level with ~80 entries per row) one compute unit streams the level slower than the device
constexpr int kFuseSlots = 32768;
// Jacobi-Chebyshev smoothing on [lmax / div, 1.05 lmax] with `steps` terms.  The finest level
// gets a long polynomial over a wide interval: S's spectrum holds many small eigenvalues
// (checkerboard pressure modes, whose smooth envelopes no piecewise-constant aggregate
// represents), so a short smoother leaves them all to the coarse grid.  How long pays off
// grows with the matrix: finestSteps() gives 8 terms on [lmax / 100, .] up to 10^4 rows,
// rising with log10(rows) to kSteps0 = 16 terms on [lmax / 400, .] at 10^6 rows (the interval
// widening with the square of the length).  Coarser levels keep a short one.
constexpr int kSteps0 = 16;  // longest finest-level smoother
constexpr int kSteps0Min = 8;
constexpr int kStepsC = 3;
constexpr double kLoDivC = 10.0;
// S = alpha B^T D^-1 B - C (see schur.h): alpha = 1/2 under the V-cycle (120x120x60, MINRES to
// 1e-9: 1/4, 0.35, 1/2, 1 -> 840, 804, 795, 843 iterations); the one-sweep stage wants less
// (Newton iterations of the driver's bench: 0.2, 0.25, 0.3, 0.35, 0.5, 0.75 -> 361.0, 358.0,
// 359.4, 363.9, 369.6, 400.2 MINRES iterations each)
constexpr double kSchurAlpha = 0.5;
constexpr double kSchurAlphaSweep = 0.3;
constexpr int kPowerSteps = 40;     // power iteration for lmax of D^-1 A

// finest-level smoother length and interval for a Schur matrix of n rows
static int finestSteps(int n) {
  const double s = kSteps0Min + 4.0 * std::log10(std::max(n, 1) / 1.0e4);
  return std::max(kSteps0Min, std::min(kSteps0, (int)std::lround(s)));
}
static double finestDiv(int steps) { return 100.0 * (steps / 8.0) * (steps / 8.0); }
// The default Schur stage is no V-cycle but one Jacobi-Chebyshev sweep on the finest level,
// z = p(D^-1 S) D^-1 r, without coarse levels: at 120x120x60 (907,923 Schur rows) the
// V-cycle's coarse correction (35,721 / 1,372 / 75 rows) changed the MINRES count by nothing
// measurable (337.8 against 338.3 per Newton iteration with two 16-term sweeps) while its
// launches cost 10 % of an iteration, and one longer sweep beats two short ones.  The length
// grows with the matrix, 8 terms up to 10^4 rows and 8 more per decade up to 24 (the 907,923-row
// matrix: 24 terms on [lmax / 400, .]; 20, 28 and 32 terms, or the interval's lower end at
// lmax / 200 or / 900, all measured slower); the interval widens with the length to lmax / 400.
constexpr int kPolyMin = 8, kPolyMax = 24;
static int polySteps(int n) {
  const double s = kPolyMin + 8.0 * std::log10(std::max(n, 1) / 1.0e4);
  return std::max(kPolyMin, std::min(kPolyMax, (int)std::lround(s)));
}
static double polyDiv(int steps) { return std::min(400.0, 25.0 * steps); }
// the stage's output weight against the block-Jacobi rows: z = beta p(D^-1 S) D^-1 r (any
// weight keeps the preconditioner SPD), rising with the sweep's length from 1 at 8 terms to 2.5
// at 24.  At 120x120x60 (24 terms, the driver's bench, alpha = 0.3): beta 0.6, 1, 1.6, 2, 2.5,
// 3, 3.5, 5, 8 -> 394.3, 360.9, 340.8, 330.4, 332.3, 329.4, 338.0, 384.2, 453.6 MINRES iterations
// per Newton iteration; 2.5 on every size slowed the batched cubes (10^4 - 3 10^4 Schur rows)
// by 4 %.
static double polyBeta(int steps) { return 1.0 + 1.5 * (steps - kPolyMin) / double(kPolyMax - kPolyMin); }
constexpr int kCoefMax = kPolyMax > kSteps0 ? kPolyMax : kSteps0;  // longest sweep (coefficient arrays)
static_assert(kStepsC <= kCoefMax, "coefficient arrays sized by the longest sweep");
constexpr double kLmaxSafety = 1.05;  // the estimate approaches lmax from below
constexpr int kHalfRows = 1 << 17;    // finest levels at least this large smooth in half precision

struct AmgLevel {
  int n = 0, nnz = 0, nc = 0, nSlots = 0, nColSlots = 0;
  std::vector<int> h_rowPtr, h_col;
  DevBuf rowPtr, col, val, dinv, agg, memPtr, mem, gSeg, gSrc;
  DevBuf sOff, sCol, sCol16, sBase, sPos, sVal;  // the smoother's sliced-ELLPACK copy
  DevBuf sCOff, sDOff, sDPos, sDia;              // and the diagonals of its pattern slices
  int nd = 0, ldd = 0;
  bool narrow = false, patterned = false;
  // the finest level's half-precision smoother operator (AmgHalf): FP16 slots in the layout
  // above, the compensated diagonal and s = D^-1/2
  bool half = false;
  DevBuf hVal, hDia, hDiag, hS, hBlk;
  DevBuf r, x, res, da, db, t, start;
  double lmax = 1.0;
  AmgCsr csr() const { return AmgCsr{n, rowPtr.as<int>(), col.as<int>(), val.as<double>()}; }
  AmgSell sell() const {
    return AmgSell{n, sOff.as<int>(), narrow ? nullptr : sCol.as<int>(),
                   narrow ? sCol16.as<unsigned short>() : nullptr, narrow ? sBase.as<int>() : nullptr,
                   sVal.as<float>(), patterned ? sCOff.as<int>() : nullptr, sDOff.as<int>(), sDia.as<float>(), nd,
                   ldd};
  }
  AmgHalf halfView() const {
    return AmgHalf{n, sOff.as<int>(), narrow ? nullptr : sCol.as<int>(),
                   narrow ? sCol16.as<unsigned short>() : nullptr, narrow ? sBase.as<int>() : nullptr,
                   hVal.as<unsigned short>(), patterned ? sCOff.as<int>() : nullptr, sDOff.as<int>(),
                   patterned ? hDia.as<unsigned short>() : nullptr, hDiag.as<float>(), hS.as<double>(), nd, ldd,
                   hBlk.as<int>()};
  }
};

struct SchurAmg {
  bool available = false;
  std::string why;
  int nq = 0;
  std::vector<int> h_qrow;
  DevBuf qrow;
  int nNodes = 0, nPairs = 0, nContrib = 0;
  DevBuf pairNode, nodeGroup, nodeSize, qPtr, cPtr, posB, contrib;
  DevBuf sSeg, sSrc, sDiag, cPos;
  std::vector<std::unique_ptr<AmgLevel>> L;
  DevBuf cinv;
  bool fresh = false;
  // levels fuseFrom.. (each at most kFuseSlots smoother slots) run as one workgroup launch; -1: none
  int fuseFrom = -1;
  DevBuf fused;  // FusedLevel[L.size() - fuseFrom], refreshed with the values
  // the Schur stage: one Chebyshev sweep on level 0 (true) or the V-cycle over every level; a
  // hierarchy of one level (at most kCoarsest rows) is inverted exactly either way
  bool poly = true;
  bool polynomial() const { return poly && L.size() > 1; }
  DevBuf powerNorm;  // ||x||^2 of the power iteration's current vector
};

// stable counting sort of (target, source) pairs by target: seg[t]..seg[t+1] lists the sources
static void countingMap(int nTargets, const std::vector<int>& target, std::vector<int>& seg, std::vector<int>& src) {
  seg.assign(nTargets + 1, 0);
  for (int t : target) seg[t + 1]++;
  for (int t = 0; t < nTargets; t++) seg[t + 1] += seg[t];
  src.assign(target.size(), 0);
  std::vector<int> fill(seg.begin(), seg.end() - 1);
  for (size_t k = 0; k < target.size(); k++) src[fill[target[k]]++] = (int)k;
}

static int findIn(const std::vector<int>& a, int b, int e, int v) {
  auto it = std::lower_bound(a.begin() + b, a.begin() + e, v);
  return (it != a.begin() + e && *it == v) ? (int)(it - a.begin()) : -1;
}

// greedy distance-one aggregation of a CSR graph
static int aggregate(const std::vector<int>& rp, const std::vector<int>& cl, std::vector<int>& agg) {
  const int n = (int)rp.size() - 1;
  agg.assign(n, -1);
  int nc = 0;
  for (int i = 0; i < n; i++) {
    if (agg[i] >= 0) continue;
    bool free = true;
    for (int k = rp[i]; k < rp[i + 1] && free; k++)
      if (agg[cl[k]] >= 0) free = false;
    if (!free) continue;
    for (int k = rp[i]; k < rp[i + 1]; k++) agg[cl[k]] = nc;
    agg[i] = nc++;
  }
  for (int i = 0; i < n; i++) {
    if (agg[i] >= 0) continue;
    for (int k = rp[i]; k < rp[i + 1]; k++)
      if (agg[cl[k]] >= 0) {
        agg[i] = agg[cl[k]];
        break;
      }
    if (agg[i] < 0) agg[i] = nc++;
  }
  return nc;
}

static uint64_t mix64(uint64_t x) {
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}

// the most common list of column offsets (col - row) over the level's rows (empty when it has
// more than 64 entries or covers fewer than a quarter of the rows)
static std::vector<int> commonOffsets(const AmgLevel& A) {
  std::unordered_map<uint64_t, std::pair<int, int>> count;  // hash -> (rows, first row)
  for (int i = 0; i < A.n; i++) {
    uint64_t hsh = mix64((uint64_t)(A.h_rowPtr[i + 1] - A.h_rowPtr[i]));
    for (int k = A.h_rowPtr[i]; k < A.h_rowPtr[i + 1]; k++) hsh = mix64(hsh ^ (uint64_t)(uint32_t)(A.h_col[k] - i));
    auto it = count.find(hsh);
    if (it == count.end()) count.emplace(hsh, std::make_pair(1, i));
    else it->second.first++;
  }
  int best = -1, rows = 0;
  for (auto& kv : count)
    if (kv.second.first > rows || (kv.second.first == rows && kv.second.second < best)) {
      rows = kv.second.first;
      best = kv.second.second;
    }
  std::vector<int> pat;
  if (best < 0 || 4ll * rows < A.n) return pat;
  for (int k = A.h_rowPtr[best]; k < A.h_rowPtr[best + 1]; k++) pat.push_back(A.h_col[k] - best);
  if (pat.size() > 64) pat.clear();
  return pat;
}

// row i's column offsets are all in the (sorted) pattern
static bool rowFits(const AmgLevel& A, int i, const std::vector<int>& pat) {
  size_t p = 0;
  for (int k = A.h_rowPtr[i]; k < A.h_rowPtr[i + 1]; k++) {
    const int o = A.h_col[k] - i;
    while (p < pat.size() && pat[p] < o) p++;
    if (p == pat.size() || pat[p] != o) return false;
  }
  return true;
}

static void allocLevel(AmgLevel& A, bool coarsest) {
  A.rowPtr.upload(A.h_rowPtr);
  A.col.upload(A.h_col);
  A.val.alloc((size_t)std::max(A.nnz, 1) * 8);
  const size_t nb = (size_t)std::max(A.n, 1) * 8;
  A.r.alloc(nb);
  A.x.alloc(nb);
  if (coarsest) return;
  // sliced ELLPACK: 64-row slices, each as wide as its longest row, or as wide as the pattern
  // for a pattern slice (rows whose column offsets col - row are all in the level's most common
  // offset list, with every row + offset inside the matrix: the columns are not stored)
  const int nSl = (A.n + 63) / 64;
  std::vector<int> pat = commonOffsets(A);
  // the diagonal form needs a symmetric offset list holding 0
  for (size_t k = 0; k < pat.size(); k++)
    if (pat[k] != -pat[pat.size() - 1 - k]) pat.clear();
  if (!pat.empty() && pat.size() % 2 == 0) pat.clear();
  std::vector<char> patSlice(nSl, 0);
  if (!pat.empty()) {
    int nPat = 0;
    for (int sl = 0; sl < nSl; sl++) {
      bool fits = true;
      for (int i = sl * 64; i < std::min(A.n, sl * 64 + 64) && fits; i++)
        fits = i + pat.front() >= 0 && i + pat.back() < A.n && rowFits(A, i, pat);
      patSlice[sl] = fits ? 1 : 0;
      nPat += fits ? 1 : 0;
    }
    if (2 * nPat < nSl) std::fill(patSlice.begin(), patSlice.end(), 0);  // not worth a second path
  }
  std::vector<int> off(nSl + 1, 0), cOff(nSl, -1);
  int nCol = 0;
  for (int sl = 0; sl < nSl; sl++) {
    int w = 0;
    for (int i = sl * 64; i < std::min(A.n, sl * 64 + 64); i++) w = std::max(w, A.h_rowPtr[i + 1] - A.h_rowPtr[i]);
    if (patSlice[sl]) w = 0;  // held by the diagonals
    if ((int64_t)off[sl] + 64ll * w >= (1ll << 31)) throw ApiError(-1, "Schur level too large for its smoother copy");
    off[sl + 1] = off[sl] + 64 * w;
    if (!patSlice[sl]) {
      cOff[sl] = nCol;
      nCol += 64 * w;
    }
  }
  A.nSlots = off[nSl];
  A.nColSlots = nCol;
  A.patterned = !pat.empty() && std::find(patSlice.begin(), patSlice.end(), 1) != patSlice.end();
  // base column of each stored-column slice; 16-bit offsets when every one spans < 65,536 columns
  std::vector<int> base(nSl, 0);
  A.narrow = true;
  for (int sl = 0; sl < nSl; sl++) {
    if (patSlice[sl]) continue;
    int lo = INT_MAX, hi = -1;
    for (int i = sl * 64; i < std::min(A.n, sl * 64 + 64); i++)
      for (int k = A.h_rowPtr[i]; k < A.h_rowPtr[i + 1]; k++) {
        lo = std::min(lo, A.h_col[k]);
        hi = std::max(hi, A.h_col[k]);
      }
    if (hi < 0) lo = hi = sl * 64;
    base[sl] = lo;
    if (hi - lo > 65535) A.narrow = false;
  }
  // padding: value 0 at the slice's base column (a valid index)
  std::vector<int> scol(std::max(nCol, 1)), spos(std::max(A.nSlots, 1));
  for (int sl = 0; sl < nSl; sl++) {
    const int w = (off[sl + 1] - off[sl]) / 64;
    if (patSlice[sl]) continue;
    for (int l = 0; l < 64; l++) {
      const int i = sl * 64 + l;
      for (int k = 0; k < w; k++) {
        const bool real = i < A.n && k < A.h_rowPtr[i + 1] - A.h_rowPtr[i];
        scol[cOff[sl] + k * 64 + l] = real ? A.h_col[A.h_rowPtr[i] + k] : base[sl];
        spos[off[sl] + k * 64 + l] = real ? A.h_rowPtr[i] + k : -1;
      }
    }
  }
  A.sOff.upload(off);
  if (A.patterned) {
    // diagonals 0, d_1 .. d_m of every row j: the CSR position of (j, j + d_k) or -1
    const int m = (int)pat.size() / 2;
    std::vector<int> dOff(m + 1);
    for (int k = 0; k <= m; k++) dOff[k] = pat[m + k];
    A.nd = m + 1;
    A.ldd = (A.n + 63) / 64 * 64;
    std::vector<int> dpos((size_t)A.nd * A.ldd, -1);
    for (int j = 0; j < A.n; j++) {
      int k = A.h_rowPtr[j];
      for (int q = 0; q < A.nd; q++) {
        const int c = j + dOff[q];
        while (k < A.h_rowPtr[j + 1] && A.h_col[k] < c) k++;
        if (k < A.h_rowPtr[j + 1] && A.h_col[k] == c) dpos[(size_t)q * A.ldd + j] = k;
      }
    }
    A.sCOff.upload(cOff);
    A.sDOff.upload(dOff);
    A.sDPos.upload(dpos);
    A.sDia.alloc(dpos.size() * 4);
  }
  if (A.narrow) {
    std::vector<unsigned short> c16(std::max(nCol, 1));
    for (int sl = 0; sl < nSl; sl++) {
      if (patSlice[sl]) continue;
      for (int j = cOff[sl]; j < cOff[sl] + (off[sl + 1] - off[sl]); j++) c16[j] = (unsigned short)(scol[j] - base[sl]);
    }
    A.sCol16.upload(c16);
    A.sBase.upload(base);
  } else {
    A.sCol.upload(scol);
  }
  A.sPos.upload(spos);
  if (A.half) {
    // dispatch order of the 512-row blocks (AmgHalf::blk): blocks with an ELLPACK slice first,
    // round-robin over the XCDs, then contiguous ranges of the others
    const int nB = (A.n + 511) / 512;
    std::vector<int> ell, plain;
    for (int b = 0; b < nB; b++) {
      bool any = false;
      for (int sl = 8 * b; sl < std::min(nSl, 8 * b + 8); sl++) any = any || !patSlice[sl];
      (any ? ell : plain).push_back(b);
    }
    std::vector<std::vector<int>> seq(8);
    std::vector<int> cap(8);
    for (int x = 0; x < 8; x++) cap[x] = nB / 8 + (x < nB % 8 ? 1 : 0);
    int x = 0;
    for (int b : ell) {
      while ((int)seq[x].size() >= cap[x]) x = (x + 1) % 8;
      seq[x].push_back(b);
      x = (x + 1) % 8;
    }
    size_t p = 0;
    for (int y = 0; y < 8; y++)
      while ((int)seq[y].size() < cap[y]) seq[y].push_back(plain[p++]);
    // dispatch word: the block, and in bit 24 + v whether wave v's two slices are both in
    // diagonal form (the kernels' paired-row form)
    if (nB >= (1 << 24)) throw ApiError(-1, "Schur level too large for its dispatch order");
    std::vector<int> blk(std::max(nB, 1));
    for (int b = 0; b < nB; b++) {
      const int blkIdx = seq[b % 8][b / 8];
      int flags = 0;
      for (int v = 0; v < 4; v++) {
        const long long w = 4ll * blkIdx + v;
        if (A.patterned && w * 128 + 128 <= A.n && patSlice[2 * w] && patSlice[2 * w + 1]) flags |= 1 << v;
      }
      blk[b] = blkIdx | (flags << 24);
    }
    A.hBlk.upload(blk);
    A.hVal.alloc((size_t)std::max(A.nSlots, 1) * 2);
    if (A.patterned) A.hDia.alloc((size_t)A.nd * A.ldd * 2);
    A.hDiag.alloc((size_t)std::max(A.n, 1) * 4);
    A.hS.alloc((size_t)std::max(A.n, 1) * 8);
  }
  A.sVal.alloc((size_t)std::max(A.nSlots, 1) * 4);
  A.dinv.alloc(nb);
  A.res.alloc(nb);
  A.da.alloc(nb);
  A.db.alloc(nb);
  A.t.alloc(nb);
  std::vector<double> st(A.n);
  for (int i = 0; i < A.n; i++) st[i] = (double)(mix64((uint64_t)i) >> 11) * (1.0 / 9007199254740992.0) - 0.5;
  A.start.upload(st);
}

void setupSchur(Ctx& c, const std::vector<int64_t>& rp, const std::vector<int>& col,
                const std::vector<int64_t>& fullToVal, const std::vector<int>& gStart, const std::vector<int>& gInfo) {
  auto S = std::make_shared<SchurAmg>();
  c.schur = S;
  const int n2 = c.n2;
  std::vector<int> qOf(n2, -1);
  for (size_t g = 0; g < gStart.size(); g++)
    if ((gInfo[g] >> 2) == 1) {
      qOf[gStart[g]] = (int)S->h_qrow.size();
      S->h_qrow.push_back(gStart[g]);
    }
  S->nq = (int)S->h_qrow.size();
  if (S->nq == 0) {
    S->why = "no Schur rows";
    return;
  }
  for (int r : S->h_qrow)
    for (int64_t k = rp[r]; k < rp[r + 1]; k++)
      if (col[k] != r && qOf[col[k]] >= 0) {
        S->why = "Schur rows coupled to each other (element-less node DoFs)";
        return;
      }
  auto fullPos = [&](int r, int cc) -> int {
    auto b = col.begin() + rp[r], e = col.begin() + rp[r + 1];
    auto it = std::lower_bound(b, e, cc);
    return (it != e && *it == cc) ? (int)fullToVal[it - col.begin()] : -1;
  };
  // node groups with their Schur columns
  std::vector<int> nodeGroup, nodeSize, qPtr{0}, qlist, posB, pairNode;
  for (size_t g = 0; g < gStart.size(); g++) {
    if ((gInfo[g] >> 2) != 0) continue;
    const int r0 = gStart[g], sz = gInfo[g] & 3;
    std::vector<int> qs;
    for (int a = 0; a < sz; a++)
      for (int64_t k = rp[r0 + a]; k < rp[r0 + a + 1]; k++)
        if (qOf[col[k]] >= 0) qs.push_back(qOf[col[k]]);
    if (qs.empty()) continue;
    std::sort(qs.begin(), qs.end());
    qs.erase(std::unique(qs.begin(), qs.end()), qs.end());
    if ((int)qs.size() > kMaxQPerNode) {
      S->why = "a node couples to too many Schur rows";
      return;
    }
    const int node = (int)nodeGroup.size();
    nodeGroup.push_back((int)g);
    nodeSize.push_back(sz);
    for (int q : qs) {
      qlist.push_back(q);
      pairNode.push_back(node);
      for (int r = 0; r < 3; r++) posB.push_back(r < sz ? fullPos(r0 + r, S->h_qrow[q]) : -1);
    }
    qPtr.push_back((int)qlist.size());
  }
  const int nNodes = (int)nodeGroup.size();
  // pattern of S: all Schur-column pairs of a node, and the diagonal
  std::vector<std::vector<int>> srow(S->nq);
  for (int q = 0; q < S->nq; q++) srow[q].push_back(q);
  for (int nd = 0; nd < nNodes; nd++)
    for (int a = qPtr[nd]; a < qPtr[nd + 1]; a++)
      for (int b = qPtr[nd]; b < qPtr[nd + 1]; b++) srow[qlist[a]].push_back(qlist[b]);
  auto L0 = std::make_unique<AmgLevel>();
  L0->n = S->nq;
  L0->h_rowPtr.assign(S->nq + 1, 0);
  for (int q = 0; q < S->nq; q++) {
    auto& v = srow[q];
    std::sort(v.begin(), v.end());
    v.erase(std::unique(v.begin(), v.end()), v.end());
    L0->h_rowPtr[q + 1] = L0->h_rowPtr[q] + (int)v.size();
  }
  if ((int64_t)L0->h_rowPtr[S->nq] >= (1ll << 31)) throw ApiError(-1, "Schur matrix too large");
  L0->nnz = L0->h_rowPtr[S->nq];
  L0->h_col.reserve(L0->nnz);
  for (int q = 0; q < S->nq; q++) L0->h_col.insert(L0->h_col.end(), srow[q].begin(), srow[q].end());
  std::vector<std::vector<int>>().swap(srow);
  // contributions -> S entries
  std::vector<int> cPtr(nNodes + 1, 0), target;
  for (int nd = 0; nd < nNodes; nd++) {
    const int m = qPtr[nd + 1] - qPtr[nd];
    cPtr[nd + 1] = cPtr[nd] + m * m;
  }
  target.reserve(cPtr[nNodes]);
  for (int nd = 0; nd < nNodes; nd++)
    for (int a = qPtr[nd]; a < qPtr[nd + 1]; a++) {
      const int qa = qlist[a];
      for (int b = qPtr[nd]; b < qPtr[nd + 1]; b++)
        target.push_back(findIn(L0->h_col, L0->h_rowPtr[qa], L0->h_rowPtr[qa + 1], qlist[b]));
    }
  std::vector<int> seg, src;
  countingMap(L0->nnz, target, seg, src);
  std::vector<int>().swap(target);
  std::vector<int> sDiag(S->nq), cPos(S->nq);
  for (int q = 0; q < S->nq; q++) {
    sDiag[q] = findIn(L0->h_col, L0->h_rowPtr[q], L0->h_rowPtr[q + 1], q);
    cPos[q] = fullPos(S->h_qrow[q], S->h_qrow[q]);
  }
  // aggregation levels
  S->L.push_back(std::move(L0));
  for (;;) {
    AmgLevel& F = *S->L.back();
    if (F.n <= kCoarsest) break;
    std::vector<int> agg;
    const int nc = aggregate(F.h_rowPtr, F.h_col, agg);
    if (nc > 0.85 * F.n) {
      if (F.n > kMaxCoarse) {
        S->why = "aggregation stalled";
        S->L.clear();
        return;
      }
      break;
    }
    F.nc = nc;
    std::vector<std::vector<int>> crow(nc), members(nc);
    for (int i = 0; i < F.n; i++) {
      members[agg[i]].push_back(i);
      for (int k = F.h_rowPtr[i]; k < F.h_rowPtr[i + 1]; k++) crow[agg[i]].push_back(agg[F.h_col[k]]);
    }
    auto C = std::make_unique<AmgLevel>();
    C->n = nc;
    C->h_rowPtr.assign(nc + 1, 0);
    for (int I = 0; I < nc; I++) {
      auto& v = crow[I];
      std::sort(v.begin(), v.end());
      v.erase(std::unique(v.begin(), v.end()), v.end());
      C->h_rowPtr[I + 1] = C->h_rowPtr[I] + (int)v.size();
    }
    C->nnz = C->h_rowPtr[nc];
    for (int I = 0; I < nc; I++) C->h_col.insert(C->h_col.end(), crow[I].begin(), crow[I].end());
    std::vector<int> tgt(F.nnz);
    for (int i = 0; i < F.n; i++)
      for (int k = F.h_rowPtr[i]; k < F.h_rowPtr[i + 1]; k++) {
        const int I = agg[i], J = agg[F.h_col[k]];
        tgt[k] = findIn(C->h_col, C->h_rowPtr[I], C->h_rowPtr[I + 1], J);
      }
    std::vector<int> gseg, gsrc, mptr(nc + 1, 0), mem;
    countingMap(C->nnz, tgt, gseg, gsrc);
    for (int I = 0; I < nc; I++) {
      mptr[I + 1] = mptr[I] + (int)members[I].size();
      mem.insert(mem.end(), members[I].begin(), members[I].end());
    }
    F.agg.upload(agg);
    F.memPtr.upload(mptr);
    F.mem.upload(mem);
    F.gSeg.upload(gseg);
    F.gSrc.upload(gsrc);
    S->L.push_back(std::move(C));
  }
  if (S->L.back()->n > kMaxCoarse) {
    S->why = "coarsest level too large";
    S->L.clear();
    return;
  }
  // the finest level smooths in half precision when it has at least kHalfRows rows (below that
  // its launches are bound by latency, not bytes, and the two-rows-per-thread kernels only add
  // latency: the reference's 3d_damper, 7,513 Schur rows, took 15.9 s with the FP32 copy and
  // 21.6 s with the half-precision one); NLA3D_GPU_SCHUR_SMOOTHER=half or =fp32 forces either
  {
    const char* e = std::getenv("NLA3D_GPU_SCHUR_SMOOTHER");
    bool half = S->L[0]->n >= kHalfRows;
    if (e && std::strcmp(e, "half") == 0) half = true;
    if (e && std::strcmp(e, "fp32") == 0) half = false;
    S->L[0]->half = S->L.size() > 1 && half;
  }
  {
    const char* e = std::getenv("NLA3D_GPU_SCHUR_CYCLE");
    S->poly = !(e && std::strcmp(e, "vcycle") == 0);
  }
  for (size_t l = 0; l < S->L.size(); l++) allocLevel(*S->L[l], l + 1 == S->L.size());
  const int ncst = S->L.back()->n;
  S->cinv.alloc((size_t)ncst * ncst * 8);
  // (a range holding only the coarsest level stays a multi-workgroup dense product)
  for (size_t l = 1; l + 1 < S->L.size(); l++)
    // (the diagonal form reads 2 nd - 1 entries a row: count those, as the slots they replace)
    if ((int64_t)S->L[l]->nSlots + (2ll * S->L[l]->nd - 1) * S->L[l]->ldd <= kFuseSlots) {
      S->fuseFrom = (int)l;
      S->fused.alloc((S->L.size() - l) * sizeof(FusedLevel));
      break;
    }
  S->qrow.upload(S->h_qrow);
  S->powerNorm.alloc(8);
  S->nNodes = nNodes;
  S->nPairs = (int)qlist.size();
  S->nContrib = cPtr[nNodes];
  S->pairNode.upload(pairNode);
  S->nodeGroup.upload(nodeGroup);
  S->nodeSize.upload(nodeSize);
  S->qPtr.upload(qPtr);
  S->cPtr.upload(cPtr);
  S->posB.upload(posB);
  S->contrib.alloc((size_t)std::max(S->nContrib, 1) * 8);
  S->sSeg.upload(seg);
  S->sSrc.upload(src);
  S->sDiag.upload(sDiag);
  S->cPos.upload(cPos);
  S->available = true;
}

static double hostDot(Ctx& c, int n, const double* a, const double* b) {
  krylov_dot(c.stream, n, a, b, c.d_partial.as<double>(), c.d_scal.as<double>() + 9);
  double v;
  hipCheck(hipMemcpyAsync(&v, c.d_scal.as<double>() + 9, 8, hipMemcpyDeviceToHost, c.stream), "dot");
  hipCheck(hipStreamSynchronize(c.stream), "dot sync");
  return v;
}

// Chebyshev coefficients on [lmax/div, 1.05 lmax] (Jacobi-preconditioned, steps terms)
struct ChebCoef {
  double c0;
  double c1[kCoefMax], c2[kCoefMax];
};
static ChebCoef chebCoef(double lmax, int steps, double div) {
  const double hi = 1.05 * lmax, lo = lmax / div;
  const double theta = 0.5 * (hi + lo), delta = 0.5 * (hi - lo), sigma = theta / delta;
  ChebCoef k{};
  k.c0 = 1.0 / theta;
  double rho = 1.0 / sigma;
  for (int i = 1; i < steps && i < kCoefMax; i++) {
    const double rn = 1.0 / (2.0 * sigma - rho);
    k.c1[i] = rn * rho;
    k.c2[i] = 2.0 * rn / delta;
    rho = rn;
  }
  return k;
}

// values of every level for the current tangent (block-Jacobi blocks must be current);
// returns false (with the reason) when the coarsest matrix is not positive definite
bool schurRefresh(Ctx& c, std::string* why) {
  SchurAmg& S = *c.schur;
  hipStream_t s = c.stream;
  const double* kval = c.d_full2.as<double>();
  AmgLevel& L0 = *S.L[0];
  launch_schur_contrib(s, S.nPairs, S.pairNode.as<int>(), S.nodeGroup.as<int>(), S.nodeSize.as<int>(),
                       S.qPtr.as<int>(), S.cPtr.as<int>(), S.posB.as<int>(), kval, c.d_binv.as<double>(),
                       S.polynomial() ? kSchurAlphaSweep : kSchurAlpha, S.contrib.as<double>());
  launch_gather_sum(s, L0.nnz, S.sSeg.as<int>(), S.sSrc.as<int>(), S.contrib.as<double>(), L0.val.as<double>());
  launch_schur_sub_c(s, S.nq, S.sDiag.as<int>(), S.cPos.as<int>(), kval, L0.val.as<double>());
  // the polynomial stage smooths level 0 only and needs no coarse matrices
  const size_t nSmoothed = S.polynomial() ? 1 : S.L.size() - 1;
  for (size_t l = 0; l < nSmoothed; l++) {
    AmgLevel& A = *S.L[l];
    if (!A.half) {
      launch_sell_fill(s, A.nSlots, A.sPos.as<int>(), A.val.as<double>(), A.sVal.as<float>());
      if (A.patterned)
        launch_sell_fill(s, A.nd * A.ldd, A.sDPos.as<int>(), A.val.as<double>(), A.sDia.as<float>());
    }
    launch_inv_diag(s, A.csr(), A.dinv.as<double>());
    if (A.half)
      launch_half_fill(s, A.halfView(), A.csr(), A.sPos.as<int>(), A.patterned ? A.sDPos.as<int>() : nullptr,
                       A.dinv.as<double>(), A.hS.as<double>(), A.hVal.as<unsigned short>(),
                       A.patterned ? A.hDia.as<unsigned short>() : nullptr, A.hDiag.as<float>());
    // largest eigenvalue of D^-1 A (of the stored operator S^' on a half-precision level, which
    // the sweep applies) by power iteration from a fixed pseudo-random start, each step scaled
    // by the norm of its input on the device, raised by a safety factor (the Chebyshev
    // polynomial grows quickly above its interval)
    double* x = A.da.as<double>();
    double* y = A.db.as<double>();
    double* nsq = S.powerNorm.as<double>();
    hipCheck(hipMemcpyAsync(x, A.start.p, (size_t)A.n * 8, hipMemcpyDeviceToDevice, s), "start");
    for (int it = 0; it < kPowerSteps; it++) {
      krylov_dot(s, A.n, x, x, c.d_partial.as<double>(), nsq);
      if (A.half)
        launch_half_power(s, A.halfView(), x, nsq, y);
      else
        launch_dinv_spmv(s, A.csr(), A.dinv.as<double>(), x, nsq, y);
      std::swap(x, y);
    }
    double lam = std::sqrt(hostDot(c, A.n, x, x));
    if (!(lam > 0.0) || !std::isfinite(lam)) lam = 1.0;  // (a zero product: any interval serves)
    A.lmax = kLmaxSafety * lam;
    if (S.polynomial()) break;
    AmgLevel& C = *S.L[l + 1];
    launch_gather_sum(s, C.nnz, A.gSeg.as<int>(), A.gSrc.as<int>(), A.val.as<double>(), C.val.as<double>());
  }
  hipCheck(hipGetLastError(), "schur refresh");
  if (S.polynomial()) {
    hipCheck(hipStreamSynchronize(s), "sync");
    S.fresh = true;
    return true;
  }
  // coarsest: explicit inverse by Cholesky on the host
  AmgLevel& Z = *S.L.back();
  const int n = Z.n;
  std::vector<double> v(Z.nnz), A((size_t)n * n, 0.0);
  hipCheck(hipMemcpyAsync(v.data(), Z.val.p, v.size() * 8, hipMemcpyDeviceToHost, s), "coarse");
  hipCheck(hipStreamSynchronize(s), "sync");
  for (int i = 0; i < n; i++)
    for (int k = Z.h_rowPtr[i]; k < Z.h_rowPtr[i + 1]; k++) A[(size_t)i * n + Z.h_col[k]] = v[k];
  for (int j = 0; j < n; j++) {
    double d = A[(size_t)j * n + j];
    for (int k = 0; k < j; k++) d -= A[(size_t)j * n + k] * A[(size_t)j * n + k];
    if (!(d > 0.0)) {
      if (why) *why = "coarsest Schur matrix not positive definite";
      S.fresh = false;
      return false;
    }
    const double ljj = std::sqrt(d);
    A[(size_t)j * n + j] = ljj;
    for (int i = j + 1; i < n; i++) {
      double t = A[(size_t)i * n + j];
      for (int k = 0; k < j; k++) t -= A[(size_t)i * n + k] * A[(size_t)j * n + k];
      A[(size_t)i * n + j] = t / ljj;
    }
  }
  // inverse = L^-T L^-1, column by column
  std::vector<double> inv((size_t)n * n), e(n);
  for (int j = 0; j < n; j++) {
    std::fill(e.begin(), e.end(), 0.0);
    e[j] = 1.0;
    for (int i = 0; i < n; i++) {  // forward: L y = e
      double t = e[i];
      for (int k = 0; k < i; k++) t -= A[(size_t)i * n + k] * e[k];
      e[i] = t / A[(size_t)i * n + i];
    }
    for (int i = n - 1; i >= 0; i--) {  // backward: L^T x = y
      double t = e[i];
      for (int k = i + 1; k < n; k++) t -= A[(size_t)k * n + i] * e[k];
      e[i] = t / A[(size_t)i * n + i];
    }
    for (int i = 0; i < n; i++) inv[(size_t)i * n + j] = e[i];
  }
  hipCheck(hipMemcpyAsync(S.cinv.p, inv.data(), inv.size() * 8, hipMemcpyHostToDevice, s), "cinv");
  if (S.fuseFrom > 0) {  // descriptors of the fused coarse levels with this tangent's coefficients
    std::vector<FusedLevel> fl(S.L.size() - S.fuseFrom);
    for (size_t l = S.fuseFrom; l < S.L.size(); l++) {
      AmgLevel& A = *S.L[l];
      FusedLevel& f = fl[l - S.fuseFrom];
      std::memset(&f, 0, sizeof f);
      f.n = A.n;
      f.nc = A.nc;
      f.r = A.r.as<double>();
      f.x = A.x.as<double>();
      if (l + 1 == S.L.size()) {
        f.cinv = S.cinv.as<double>();
        continue;
      }
      const ChebCoef k = chebCoef(A.lmax, kStepsC, kLoDivC);
      f.steps = kStepsC;
      f.A = A.sell();
      f.dinv = A.dinv.as<double>();
      f.c0 = k.c0;
      for (int i = 1; i < kStepsC; i++) {
        f.c1[i] = k.c1[i];
        f.c2[i] = k.c2[i];
      }
      f.res = A.res.as<double>();
      f.da = A.da.as<double>();
      f.db = A.db.as<double>();
      f.t = A.t.as<double>();
      f.memPtr = A.memPtr.as<int>();
      f.mem = A.mem.as<int>();
      f.agg = A.agg.as<int>();
    }
    hipCheck(hipMemcpyAsync(S.fused.p, fl.data(), fl.size() * sizeof(FusedLevel), hipMemcpyHostToDevice, s),
             "fused levels");
  }
  hipCheck(hipStreamSynchronize(s), "sync");
  S.fresh = true;
  return true;
}

// the coarse level l + 1's cheb_start for its restriction to do, when that level is smoothed by
// the one-level kernels (not the coarsest, not the one-workgroup range)
static ChebStartNext startNext(SchurAmg& S, size_t l) {
  ChebStartNext nx;
  const size_t c = l + 1;
  if (c + 1 >= S.L.size() || (int)c == S.fuseFrom) return nx;
  AmgLevel& C = *S.L[c];
  nx.dinv = C.dinv.as<double>();
  nx.c0 = chebCoef(C.lmax, kStepsC, kLoDivC).c0;
  nx.res = C.res.as<double>();
  nx.d = C.da.as<double>();
  nx.x = C.x.as<double>();
  return nx;
}

// the polynomial stage: one Chebyshev sweep of polySteps() terms on level 0, from zero, its
// result weighted by polyBeta; the MINRES
// stage (rext != null) gathers the right-hand side from rext and scatters the result into zext
// (st gates the write), otherwise A.r -> A.x
static void polySweep(SchurAmg& S, hipStream_t s, const double* rext, double* zext, const MinresState* st) {
  AmgLevel& A = *S.L[0];
  const int steps = polySteps(A.n);
  const ChebCoef k = chebCoef(A.lmax, steps, polyDiv(steps));
  const int* qrow = rext ? S.qrow.as<int>() : nullptr;
  if (A.half) {
    // the scaled three-term sweep (schur.h, AmgHalf); the last step writes z = s x^
    const AmgHalf H = A.halfView();
    double* rhat = A.res.as<double>();
    double* cur = A.da.as<double>();
    double* prev = nullptr;
    launch_half_start(s, A.n, qrow, rext ? rext : A.r.as<double>(), A.hS.as<double>(), k.c0, rhat, cur, st);
    for (int i = 1; i < steps; i++) {
      const bool last = i + 1 == steps;
      double* out = prev ? prev : A.db.as<double>();
      launch_half_step(s, H, cur, prev, rhat, k.c1[i], k.c2[i], out, last ? qrow : nullptr,
                       last ? (zext ? zext : A.x.as<double>()) : nullptr, st, nullptr, polyBeta(steps));
      prev = cur;
      cur = out;
    }
    return;
  }
  const AmgSell M = A.sell();
  double* dinv = A.dinv.as<double>();
  double* din = A.da.as<double>();
  double* dout = A.db.as<double>();
  if (rext)
    launch_cheb_start_gather(s, A.n, qrow, rext, dinv, k.c0, A.r.as<double>(), A.res.as<double>(), din,
                             A.x.as<double>());
  else
    launch_cheb_start(s, A.n, A.r.as<double>(), dinv, k.c0, A.res.as<double>(), din, A.x.as<double>());
  for (int i = 1; i < steps; i++) {
    launch_cheb_step(s, M, din, dinv, k.c1[i], k.c2[i], A.res.as<double>(), dout, A.x.as<double>(),
                     chebXMode(i, steps));
    std::swap(din, dout);
  }
  if (zext)
    launch_finish_scatter(s, A.n, qrow, din, A.x.as<double>(), st, zext, polyBeta(steps));
  else
    launch_axpby(s, A.n, polyBeta(steps), din, polyBeta(steps), A.x.as<double>(), A.x.as<double>());
}

// the same sweep fused with MINRES's preconditioner pass (FusedStage): the start also does the
// pass's T -= c_r2 R2 on the Schur rows, the last step writes Z there and the partials of T . Z
static void polySweepFused(SchurAmg& S, hipStream_t s, const double* R2, double* T, double* Z, double* partial,
                           const MinresState* st) {
  AmgLevel& A = *S.L[0];
  const int steps = polySteps(A.n);
  const ChebCoef k = chebCoef(A.lmax, steps, polyDiv(steps));
  const int* qrow = S.qrow.as<int>();
  if (!A.half) {  // the FP32 copy: polySweep's two-term form
    const AmgSell M = A.sell();
    double* dinv = A.dinv.as<double>();
    double* din = A.da.as<double>();
    double* dout = A.db.as<double>();
    launch_cheb_start_fused(s, A.n, qrow, R2, T, dinv, k.c0, A.r.as<double>(), A.res.as<double>(), din,
                            A.x.as<double>(), st);
    for (int i = 1; i < steps; i++) {
      launch_cheb_step(s, M, din, dinv, k.c1[i], k.c2[i], A.res.as<double>(), dout, A.x.as<double>(),
                       chebXMode(i, steps));
      std::swap(din, dout);
    }
    launch_finish_fused(s, A.n, qrow, din, A.x.as<double>(), A.r.as<double>(), st, Z, partial, polyBeta(steps));
    return;
  }
  const AmgHalf H = A.halfView();
  double* rhat = A.res.as<double>();
  double* cur = A.da.as<double>();
  double* prev = nullptr;
  launch_half_start_fused(s, A.n, qrow, R2, T, A.hS.as<double>(), k.c0, rhat, cur, st);
  for (int i = 1; i < steps; i++) {
    const bool last = i + 1 == steps;
    double* out = prev ? prev : A.db.as<double>();
    launch_half_step(s, H, cur, prev, rhat, k.c1[i], k.c2[i], out, last ? qrow : nullptr, last ? Z : nullptr, st,
                     last ? partial : nullptr, polyBeta(steps));
    prev = cur;
    cur = out;
  }
}

// one V-cycle from level l.  At level 0 of the MINRES stage (rext != null) the right-hand side
// is gathered from the Krylov vector rext and the result scattered into zext (the Schur rows,
// S.qrow; st gates the write) inside the first and last passes; otherwise it is A.r -> A.x.
// started: the restriction that produced A.r already did this level's cheb_start.  Returns
// the last smoothing increment still to be added to A.x (the prolongation above adds it), or
// null when A.x is complete.
static const double* vcycle(SchurAmg& S, hipStream_t s, size_t l, const double* rext = nullptr,
                            double* zext = nullptr, const MinresState* st = nullptr, bool started = false) {
  if (l == 0 && S.polynomial()) {
    polySweep(S, s, rext, zext, st);
    return nullptr;
  }
  if (l > 0 && (int)l == S.fuseFrom) {
    launch_coarse_cycle(s, S.fused.as<FusedLevel>(), (int)(S.L.size() - l));
    return nullptr;
  }
  AmgLevel& A = *S.L[l];
  if (l + 1 == S.L.size()) {
    if (rext) launch_q_gather(s, S.nq, S.qrow.as<int>(), rext, st, A.r.as<double>());
    launch_dense_matvec(s, A.n, S.cinv.as<double>(), A.r.as<double>(), A.x.as<double>());
    if (zext) launch_q_scatter(s, S.nq, S.qrow.as<int>(), A.x.as<double>(), st, zext);
    return nullptr;
  }
  const int steps = l == 0 ? finestSteps(A.n) : kStepsC;
  const ChebCoef k = chebCoef(A.lmax, steps, l == 0 ? finestDiv(steps) : kLoDivC);
  AmgLevel& C = *S.L[l + 1];
  const ChebStartNext nx = startNext(S, l);
  if (A.half) {
    // the scaled three-term sweep (schur.h, AmgHalf): r^ in res, the iterates in da / db,
    // x^_(k+1) written over x^_(k-1); the post-smoother's last step writes z = s x^
    const AmgHalf H = A.halfView();
    const double* sc = A.hS.as<double>();
    double* rhat = A.res.as<double>();
    double* xa = A.da.as<double>();
    double* xb = A.db.as<double>();
    launch_half_start(s, A.n, rext ? S.qrow.as<int>() : nullptr, rext ? rext : A.r.as<double>(), sc, k.c0, rhat, xa,
                      st);
    double* cur = xa;
    double* prev = nullptr;
    for (int i = 1; i < steps; i++) {
      double* out = prev ? prev : xb;
      launch_half_step(s, H, cur, prev, rhat, k.c1[i], k.c2[i], out, nullptr, nullptr, st);
      prev = cur;
      cur = out;
    }
    launch_half_residual_restrict(s, H, cur, rhat, A.t.as<double>(), A.nc, A.memPtr.as<int>(), A.mem.as<int>(),
                                  C.r.as<double>(), st, nx);
    const double* pend = vcycle(S, s, l + 1, nullptr, nullptr, nullptr, nx.res != nullptr);
    launch_half_prolong(s, A.n, A.agg.as<int>(), C.x.as<double>(), pend, sc, cur, st);
    prev = cur;
    cur = cur == xa ? xb : xa;
    launch_half_start_x(s, H, prev, rhat, k.c0, cur, st);
    for (int i = 1; i < steps; i++) {
      const bool last = i + 1 == steps;
      launch_half_step(s, H, cur, prev, rhat, k.c1[i], k.c2[i], prev, last && zext ? S.qrow.as<int>() : nullptr,
                       last ? (zext ? zext : A.x.as<double>()) : nullptr, st);
      std::swap(prev, cur);
    }
    return nullptr;
  }
  const AmgSell M = A.sell();
  double* dinv = A.dinv.as<double>();
  double* din = A.da.as<double>();
  double* dout = A.db.as<double>();
  if (rext)
    launch_cheb_start_gather(s, A.n, S.qrow.as<int>(), rext, dinv, k.c0, A.r.as<double>(), A.res.as<double>(), din,
                             A.x.as<double>());
  else if (!started)
    launch_cheb_start(s, A.n, A.r.as<double>(), dinv, k.c0, A.res.as<double>(), din, A.x.as<double>());
  for (int i = 1; i < steps; i++) {
    launch_cheb_step(s, M, din, dinv, k.c1[i], k.c2[i], A.res.as<double>(), dout, A.x.as<double>(),
                     chebXMode(i, steps));
    std::swap(din, dout);
  }
  launch_residual_restrict(s, M, A.res.as<double>(), din, A.t.as<double>(), A.nc, A.memPtr.as<int>(),
                           A.mem.as<int>(), C.r.as<double>(), A.x.as<double>(), nx);
  const double* pend = vcycle(S, s, l + 1, nullptr, nullptr, nullptr, nx.res != nullptr);
  launch_prolong(s, A.n, A.agg.as<int>(), C.x.as<double>(), pend, A.x.as<double>());
  din = A.da.as<double>();
  dout = A.db.as<double>();
  launch_cheb_start_x(s, M, A.r.as<double>(), dinv, k.c0, A.res.as<double>(), din, A.x.as<double>());
  for (int i = 1; i < steps; i++) {
    launch_cheb_step(s, M, din, dinv, k.c1[i], k.c2[i], A.res.as<double>(), dout, A.x.as<double>(),
                     chebXMode(i, steps));
    std::swap(din, dout);
  }
  if (zext) {
    launch_finish_scatter(s, A.n, S.qrow.as<int>(), din, A.x.as<double>(), st, zext);
    return nullptr;
  }
  if (l == 0) {  // the cycle's result in A.x
    launch_axpby(s, A.n, 1.0, din, 1.0, A.x.as<double>(), A.x.as<double>());
    return nullptr;
  }
  return din;  // x + din: the prolongation above adds both
}

// the MINRES stage: z[Schur rows] = V-cycle(r[Schur rows]); partials of r . z
PrecondHook schurHook(Ctx& c) {
  std::shared_ptr<SchurAmg> S = c.schur;
  const int n2 = c.n2;
  return [S, n2](hipStream_t s, const double* r, double* z, double* partial, const MinresState* st) {
    vcycle(*S, s, 0, r, z, st);
    launch_dot_partials(s, n2, r, z, st, partial);
  };
}

FusedStage schurFusedStage(Ctx& c) {
  FusedStage f;
  std::shared_ptr<SchurAmg> S = c.schur;
  if (!S || !S->available || !S->polynomial()) return f;  // partials = 0: not fused
  f.partials = (S->L[0]->n + 511) / 512;  // one per 512 rows (the last step's / the exit's workgroups)
  f.run = [S](hipStream_t s, const double* R2, double* T, double* Z, double* partial, const MinresState* st) {
    polySweepFused(*S, s, R2, T, Z, partial, st);
  };
  return f;
}

bool schurUsable(Ctx& c, std::string* why) {
  if (!c.schur || !c.schur->available) {
    if (why) *why = c.schur ? c.schur->why : "Krylov data not set up";
    return false;
  }
  return true;
}

}  // namespace nla3d_amd

using namespace nla3d_amd;

static int schurGuard(nla3d_gpu_t h, const std::function<void(Ctx&)>& fn) {
  if (!h) return NLA3D_GPU_EINVAL;
  Ctx& c = *reinterpret_cast<Ctx*>(h);
  try {
    c.err.clear();
    c.requireSolid();
    hipCheck(hipSetDevice(c.device), "hipSetDevice");
    if (!c.krylovReady) c.setupKrylov();
    fn(c);
    return NLA3D_GPU_OK;
  } catch (const ApiError& e) {
    c.err = e.what();
    return e.code;
  } catch (const std::exception& e) {
    c.err = e.what();
    return NLA3D_GPU_EINVAL;
  }
}

static void refreshOrThrow(Ctx& c) {
  std::string why;
  if (!schurUsable(c, &why)) throw ApiError(-3, "Schur multigrid unavailable: " + why);
  krylov_precond_setup(c.stream, c.krylovMatrix(), c.precond());
  if (!schurRefresh(c, &why)) throw ApiError(-6, "Schur multigrid: " + why);
}

extern "C" {

int nla3d_gpu_schur_info(nla3d_gpu_t h, int* available, int* nq, int* nnz0, int* levels, int* sizes, int cap) {
  return schurGuard(h, [&](Ctx& c) {
    SchurAmg* S = c.schur.get();
    if (available) *available = S && S->available ? 1 : 0;
    if (nq) *nq = S ? S->nq : 0;
    if (nnz0) *nnz0 = S && S->available ? S->L[0]->nnz : 0;
    if (levels) *levels = S && S->available ? (int)S->L.size() : 0;
    if (sizes && S && S->available)
      for (int l = 0; l < (int)S->L.size() && l < cap; l++) sizes[l] = S->L[l]->n;
  });
}

int nla3d_gpu_schur_cycle_info(nla3d_gpu_t h, int* finest_steps, int* fuse_from) {
  return schurGuard(h, [&](Ctx& c) {
    SchurAmg* S = c.schur.get();
    if (!S || !S->available) throw ApiError(-3, "the Schur multigrid is unavailable for this model");
    if (finest_steps)
      *finest_steps = S->L.size() < 2 ? 0 : S->polynomial() ? polySteps(S->L[0]->n) : finestSteps(S->L[0]->n);
    if (fuse_from) *fuse_from = S->fuseFrom;
  });
}

int nla3d_gpu_schur_set_cycle(nla3d_gpu_t h, int vcycle) {
  return schurGuard(h, [&](Ctx& c) {
    SchurAmg* S = c.schur.get();
    if (!S || !S->available) throw ApiError(-3, "the Schur multigrid is unavailable for this model");
    if (vcycle != 0 && vcycle != 1) throw ApiError(-1, "vcycle must be 0 or 1");
    if (S->poly != (vcycle == 0)) {
      S->poly = vcycle == 0;
      S->fresh = false;  // the V-cycle needs the coarse matrices of the current values
    }
  });
}

int nla3d_gpu_schur_get_cycle(nla3d_gpu_t h, int* vcycle) {
  return schurGuard(h, [&](Ctx& c) {
    SchurAmg* S = c.schur.get();
    if (!S || !S->available) throw ApiError(-3, "the Schur multigrid is unavailable for this model");
    if (vcycle) *vcycle = S->poly ? 0 : 1;
  });
}

int nla3d_gpu_schur_finest_precision(nla3d_gpu_t h, int* bits) {
  return schurGuard(h, [&](Ctx& c) {
    SchurAmg* S = c.schur.get();
    if (!S || !S->available) throw ApiError(-3, "the Schur multigrid is unavailable for this model");
    if (bits) *bits = S->L[0]->half ? 16 : (S->L.size() > 1 ? 32 : 64);
  });
}

int nla3d_gpu_schur_matrix(nla3d_gpu_t h, int* qrow, int* rowPtr, int* col, double* val) {
  return schurGuard(h, [&](Ctx& c) {
    refreshOrThrow(c);
    SchurAmg& S = *c.schur;
    AmgLevel& L0 = *S.L[0];
    if (qrow) std::memcpy(qrow, S.h_qrow.data(), S.h_qrow.size() * 4);
    if (rowPtr) std::memcpy(rowPtr, L0.h_rowPtr.data(), L0.h_rowPtr.size() * 4);
    if (col) std::memcpy(col, L0.h_col.data(), L0.h_col.size() * 4);
    if (val) hipCheck(hipMemcpy(val, L0.val.p, (size_t)L0.nnz * 8, hipMemcpyDeviceToHost), "S");
  });
}

int nla3d_gpu_schur_level(nla3d_gpu_t h, int level, int* agg, double* lmax) {
  return schurGuard(h, [&](Ctx& c) {
    SchurAmg* S = c.schur.get();
    if (!S || !S->available || level < 0 || level >= (int)S->L.size()) throw ApiError(-1, "no such level");
    AmgLevel& A = *S->L[level];
    if (lmax) *lmax = A.lmax;
    if (agg && level + 1 < (int)S->L.size())
      hipCheck(hipMemcpy(agg, A.agg.p, (size_t)A.n * 4, hipMemcpyDeviceToHost), "agg");
  });
}

int nla3d_gpu_schur_smoother_operator(nla3d_gpu_t h, double* sc, double* val) {
  return schurGuard(h, [&](Ctx& c) {
    if (!sc || !val) throw ApiError(-1, "null arrays");
    if (!c.schur || !c.schur->fresh) refreshOrThrow(c);
    SchurAmg& S = *c.schur;
    AmgLevel& A = *S.L[0];
    if (!A.half) throw ApiError(-3, "the finest level has no half-precision smoother operator");
    auto down = [&](const DevBuf& b, auto& v) {
      v.resize(b.bytes / sizeof(v[0]));
      if (b.bytes) hipCheck(hipMemcpy(v.data(), b.p, b.bytes, hipMemcpyDeviceToHost), "operator");
    };
    std::vector<int> off, cOff, dOff, pos;
    std::vector<unsigned short> hv, hd;
    std::vector<float> dg;
    down(A.sOff, off);
    down(A.sPos, pos);
    down(A.hVal, hv);
    down(A.hDiag, dg);
    if (A.patterned) {
      down(A.sCOff, cOff);
      down(A.sDOff, dOff);
      down(A.hDia, hd);
    }
    hipCheck(hipMemcpy(sc, A.hS.p, (size_t)A.n * 8, hipMemcpyDeviceToHost), "s");
    auto half2d = [](unsigned short b) {
      _Float16 f;
      std::memcpy(&f, &b, 2);
      return (double)f;
    };
    for (int i = 0; i < A.n; i++) {
      const int sl = i >> 6;
      const bool pat = A.patterned && cOff[sl] < 0;
      for (int k = A.h_rowPtr[i]; k < A.h_rowPtr[i + 1]; k++) {
        const int col = A.h_col[k], o = col - i;
        if (o == 0) {
          val[k] = dg[i];
        } else if (pat) {
          const int d = std::abs(o);
          const int q = (int)(std::find(dOff.begin(), dOff.end(), d) - dOff.begin());
          val[k] = half2d(hd[(size_t)q * A.ldd + (o > 0 ? i : col)]);
        } else {
          val[k] = 0.0;
        }
      }
      if (!pat) {
        const int w = (off[sl + 1] - off[sl]) / 64;
        for (int k = 0; k < w; k++) {
          const int slot = off[sl] + k * 64 + (i & 63);
          const int p = pos[slot];
          if (p >= 0 && A.h_col[p] != i) val[p] = half2d(hv[slot]);
        }
      }
    }
  });
}

int nla3d_gpu_time_smoother(nla3d_gpu_t h, int iters, double* ms, long long* bytes) {
  return schurGuard(h, [&](Ctx& c) {
    if (iters < 1 || !ms) throw ApiError(-1, "bad arguments");
    if (!c.schur || !c.schur->fresh) refreshOrThrow(c);
    SchurAmg& S = *c.schur;
    if (S.L.size() < 2) throw ApiError(-3, "the hierarchy has no smoothed level");
    AmgLevel& A = *S.L[0];
    // the coefficients of the active stage's finest-level sweep
    const int steps0 = S.polynomial() ? polySteps(A.n) : finestSteps(A.n);
    const ChebCoef k = chebCoef(A.lmax, steps0, S.polynomial() ? polyDiv(steps0) : finestDiv(steps0));
    hipStream_t s = c.stream;
    if (A.half) {
      const AmgHalf H = A.halfView();
      double* rhat = A.res.as<double>();
      double* cur = A.da.as<double>();
      double* prev = A.db.as<double>();
      // a finite start (the power-iteration start vector), one warm-up step, then the steps
      launch_half_start(s, A.n, nullptr, A.start.as<double>(), A.hS.as<double>(), k.c0, rhat, cur, nullptr);
      launch_half_step(s, H, cur, nullptr, rhat, k.c1[1], k.c2[1], prev, nullptr, nullptr, nullptr);
      std::swap(prev, cur);
      hipEvent_t e0, e1;
      hipCheck(hipEventCreate(&e0), "event");
      hipCheck(hipEventCreate(&e1), "event");
      hipCheck(hipEventRecord(e0, s), "rec");
      for (int i = 0; i < iters; i++) {
        const int j = 1 + i % (steps0 - 1);
        launch_half_step(s, H, cur, prev, rhat, k.c1[j], k.c2[j], prev, nullptr, nullptr, nullptr);
        std::swap(prev, cur);
      }
      hipCheck(hipEventRecord(e1, s), "rec");
      hipCheck(hipEventSynchronize(e1), "sync");
      float t = 0;
      hipCheck(hipEventElapsedTime(&t, e0, e1), "elapsed");
      *ms = t / iters;
      (void)hipEventDestroy(e0);
      (void)hipEventDestroy(e1);
      if (bytes) {
        // stored operator (FP16 slots, column offsets or indices, slice offsets and bases, the
        // FP16 diagonals k >= 1 of the pattern slices, the FP32 diagonal) and four vectors per
        // step (x^_k, x^_(k-1), r^ read; x^_(k+1) written)
        const long long nSl = (A.n + 63) / 64;
        *bytes = (long long)A.nSlots * 2 + (long long)A.nColSlots * (A.narrow ? 2 : 4) + (nSl + 1) * 4 +
                 (A.narrow ? nSl * 4 : 0) + (A.patterned ? nSl * 4 + (long long)(A.nd - 1) * A.ldd * 2 : 0) +
                 (long long)A.n * 4 + 4ll * A.n * 8;
      }
      return;
    }
    const AmgSell M = A.sell();
    double* din = A.da.as<double>();
    double* dout = A.db.as<double>();
    // a finite start (the power-iteration start vector), then the finest level's steps
    hipCheck(hipMemcpyAsync(A.r.p, A.start.p, (size_t)A.n * 8, hipMemcpyDeviceToDevice, s), "start");
    launch_cheb_start(s, A.n, A.r.as<double>(), A.dinv.as<double>(), k.c0, A.res.as<double>(), din,
                      A.x.as<double>());
    launch_cheb_step(s, M, din, A.dinv.as<double>(), k.c1[1], k.c2[1], A.res.as<double>(), dout,
                     A.x.as<double>(), kChebPlain);  // warm-up
    std::swap(din, dout);
    hipEvent_t e0, e1;
    hipCheck(hipEventCreate(&e0), "event");
    hipCheck(hipEventCreate(&e1), "event");
    hipCheck(hipEventRecord(e0, s), "rec");
    long long xPasses = 0;  // vector passes of x over the timed steps (2 when added to, else 0)
    for (int i = 0; i < iters; i++) {
      const int j = 1 + i % (steps0 - 1);
      const int mode = chebXMode(j, steps0);
      launch_cheb_step(s, M, din, A.dinv.as<double>(), k.c1[j], k.c2[j], A.res.as<double>(), dout,
                       A.x.as<double>(), mode);
      xPasses += mode == kChebSkip ? 0 : 2;
      std::swap(din, dout);
    }
    hipCheck(hipEventRecord(e1, s), "rec");
    hipCheck(hipEventSynchronize(e1), "sync");
    float t = 0;
    hipCheck(hipEventElapsedTime(&t, e0, e1), "elapsed");
    *ms = t / iters;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (bytes) {
      // stored matrix (FP32 values, column offsets or indices, slice offsets and bases), the
      // five vectors of every step (din, res, D^-1 read; res, dout written) and x, read and
      // written by the steps that add to it: the mean over the timed steps, which follow the
      // sweep's order of modes
      // (pattern slices store values only)
      const long long nSl = (A.n + 63) / 64;
      *bytes = (long long)A.nSlots * 4 + (long long)A.nColSlots * (A.narrow ? 2 : 4) + (nSl + 1) * 4 +
               (A.narrow ? nSl * 4 : 0) + (A.patterned ? nSl * 4 + (long long)A.nd * A.ldd * 4 : 0) +
               5ll * A.n * 8 + (xPasses * A.n * 8 + iters / 2) / iters;
    }
  });
}

int nla3d_gpu_schur_apply(nla3d_gpu_t h, const double* r, double* z, int refresh) {
  return schurGuard(h, [&](Ctx& c) {
    if (!r || !z) throw ApiError(-1, "null vectors");
    if (refresh || !c.schur || !c.schur->fresh) refreshOrThrow(c);
    SchurAmg& S = *c.schur;
    AmgLevel& L0 = *S.L[0];
    hipCheck(hipMemcpyAsync(L0.r.p, r, (size_t)S.nq * 8, hipMemcpyHostToDevice, c.stream), "r");
    vcycle(S, c.stream, 0);
    hipCheck(hipGetLastError(), "vcycle");
    hipCheck(hipMemcpyAsync(z, L0.x.p, (size_t)S.nq * 8, hipMemcpyDeviceToHost, c.stream), "z");
    hipCheck(hipStreamSynchronize(c.stream), "sync");
  });
}

}  // extern "C"
