// Host model store: DoF registration, reference equation numbering and the
// upper-triangle CSR sparsity of the three K blocks.  See model.h.
#include "model.h"

#include <algorithm>
#include <atomic>
#include <cstring>
#include <functional>
#include <thread>

namespace nla3d_amd {

unsigned hostThreads() {
  unsigned n = std::thread::hardware_concurrency();
  if (n == 0) n = 1;
  return std::min(n, 16u);
}

void parallelFor(size_t n, const std::function<void(size_t, size_t)>& fn) {
  unsigned nt = hostThreads();
  if (n < 4096 || nt == 1) {
    fn(0, n);
    return;
  }
  std::vector<std::thread> th;
  size_t chunk = (n + nt - 1) / nt;
  for (unsigned t = 0; t < nt; t++) {
    size_t b = t * chunk, e = std::min(n, b + chunk);
    if (b >= e) break;
    th.emplace_back([&fn, b, e] { fn(b, e); });
  }
  for (auto& x : th) x.join();
}

std::string FEStorage::addElements(int kind, uint32_t n, const uint32_t* conn1, const double* params) {
  int nn = (kind == ELEM_SOLID81) ? 8
           : (kind == ELEM_QUADTH || kind == ELEM_PLANE41) ? 4
           : (kind == ELEM_SURFACE_LINETH ? 2 : 0);
  if (nn == 0) return "unknown element kind " + std::to_string(kind);
  if (kind == ELEM_SOLID81 || kind == ELEM_PLANE41) {
    if (solidCount > 0 && upKind != kind) return "SOLID81 and PLANE41 elements cannot be mixed in one model";
    if (solidCount > 0 && solidFirst + solidCount != nElements())
      return "u/p solid elements must be added as one contiguous range";
    if (solidCount == 0) solidFirst = nElements();
    solidCount += n;
    upKind = kind;
  }
  for (uint32_t e = 0; e < n; e++) {
    elemKind.push_back((uint8_t)kind);
    for (int i = 0; i < nn; i++) {
      uint32_t node = conn1[(size_t)e * nn + i];
      if (node < 1 || node > nNodes()) return "element node number out of range";
      elemNodes.push_back(node);
    }
    elemNodePtr.push_back(elemNodes.size());
    for (int i = 0; i < 4; i++) elemParams.push_back(params ? params[(size_t)e * 4 + i] : 0.0);
  }
  return "";
}

std::string FEStorage::initDofs() {
  const uint32_t nN = nNodes(), nE = nElements();
  if (nN < 1 && nE < 1) return "No any nodes or elements";
  nodeDofMask.assign(nN, 0);
  elemDofMask.assign(nE, 0);
  // Element::pre registration, in element order (FEStorage.cpp:347-349)
  for (uint32_t e = 0; e < nE; e++) {
    uint8_t nodeMask = 0, elMask = 0;
    switch (elemKind[e]) {
      case ELEM_SOLID81:  // SOLID81.cpp:21-25
        nodeMask = (1u << UX) | (1u << UY) | (1u << UZ);
        elMask = 1u << HYDRO_PRESSURE;
        break;
      case ELEM_PLANE41:  // PLANE41.cpp:24-28
        nodeMask = (1u << UX) | (1u << UY);
        elMask = 1u << HYDRO_PRESSURE;
        break;
      case ELEM_QUADTH:  // QUADTH.cpp:16-19
      case ELEM_SURFACE_LINETH:  // QUADTH.cpp:83-85
        nodeMask = 1u << TEMP;
        break;
    }
    for (uint64_t k = elemNodePtr[e]; k < elemNodePtr[e + 1]; k++) nodeDofMask[elemNodes[k] - 1] |= nodeMask;
    elemDofMask[e] |= elMask;
  }
  // MpcCollection::pre + registerMpcsInStorage (FEStorage.cpp:351-354; RigidBodyMpc::pre, Mpc.cpp:52-68)
  // the MPCs added directly come first (main.cpp:207-212 adds them before the collections
  // register theirs), then each rigid body's rows
  mpcs = ceMpcs;
  for (auto& rb : rigidBodies) {
    if (rb.masterNode < 1 || rb.masterNode > nN) return "rigid body master node out of range";
    rb.firstMpc = (uint32_t)mpcs.size();
    for (uint32_t s : rb.slaveNodes) {
      if (s < 1 || s > nN) return "rigid body slave node out of range";
      for (int d : rb.dofs) {
        Mpc m;
        m.b = 0.0;
        m.eq.push_back({s, d, 1.0});
        m.eq.push_back({rb.masterNode, d, -1.0});
        m.eq.push_back({rb.masterNode, ROTX, 0.0});
        m.eq.push_back({rb.masterNode, ROTY, 0.0});
        m.eq.push_back({rb.masterNode, ROTZ, 0.0});
        mpcs.push_back(m);
      }
    }
    nodeDofMask[rb.masterNode - 1] |= 0x3F;  // UX..ROTZ
  }
  uint64_t cnt = 0;
  for (auto m : nodeDofMask) cnt += __builtin_popcount(m);
  for (auto m : elemDofMask) cnt += __builtin_popcount(m);
  if (cnt == 0) return "no DoFs registered";
  if (cnt >= 0xFFFFFFFFull) return "too many DoFs for 32-bit equation numbers";
  nDofs = (uint32_t)cnt;
  return "";
}

std::string FEStorage::setConstrainedDofs() {
  nodeConstr.assign(nNodes(), 0);
  elemConstr.assign(nElements(), 0);
  nConstrainedDofs = 0;
  for (auto& f : fixs) {  // FEStorage::setConstrainedNodeDof, FEStorage.cpp:81-91
    if (f.node < 1 || f.node > nNodes()) return "fix node out of range";
    uint8_t bit = (uint8_t)(1u << f.dof);
    if (!(nodeDofMask[f.node - 1] & bit)) return "fix on an unregistered DoF";
    if (nodeConstr[f.node - 1] & bit) continue;
    nodeConstr[f.node - 1] |= bit;
    nConstrainedDofs++;
  }
  return "";
}

std::string FEStorage::assignEquationNumbers() {
  nUnknownDofs = nDofs - nConstrainedDofs;
  if (nUnknownDofs == 0) return "no unknown DoFs";
  uint32_t next_s = nConstrainedDofs + 1, next_c = 1;
  elemEq.assign((size_t)nElements() * 8, 0);
  nodeEq.assign((size_t)nNodes() * 8, 0);
  for (uint32_t e = 0; e < nElements(); e++)
    for (int t = 0; t < NUM_DOF_TYPES; t++)
      if (elemDofMask[e] & (1u << t))
        elemEq[(size_t)e * 8 + t] = (elemConstr[e] & (1u << t)) ? next_c++ : next_s++;
  for (uint32_t n = 0; n < nNodes(); n++)
    for (int t = 0; t < NUM_DOF_TYPES; t++)
      if (nodeDofMask[n] & (1u << t))
        nodeEq[(size_t)n * 8 + t] = (nodeConstr[n] & (1u << t)) ? next_c++ : next_s++;
  for (auto& m : mpcs) m.eqNum = next_s++;
  return "";
}

// Sparsity (FEStorage.cpp:488-541): the union over elements of all pairs of
// the DoFs of the element's nodes and of the element itself, plus
// (MPC equation, term DoF) pairs; stored upper (i <= j) and split into
// blocks by the constrained count.
std::string FEStorage::buildSparsity(uint32_t maxInRow) {
  const uint32_t nN = nNodes(), nE = nElements();
  const uint32_t nC = nConstrainedDofs, nTot = nDofs + nMpc();
  // node -> elements (learnTopology, FEStorage.cpp:580-589)
  std::vector<uint64_t> n2ePtr(nN + 1, 0);
  for (uint32_t e = 0; e < nE; e++)
    for (uint64_t k = elemNodePtr[e]; k < elemNodePtr[e + 1]; k++) n2ePtr[elemNodes[k]]++;
  for (uint32_t n = 0; n < nN; n++) n2ePtr[n + 1] += n2ePtr[n];
  std::vector<uint32_t> n2e(n2ePtr[nN]);
  {
    std::vector<uint64_t> fill(n2ePtr.begin(), n2ePtr.end() - 1);
    for (uint32_t e = 0; e < nE; e++)
      for (uint64_t k = elemNodePtr[e]; k < elemNodePtr[e + 1]; k++) n2e[fill[elemNodes[k] - 1]++] = e;
  }
  // MPC columns keyed by the term DoF equation
  std::vector<std::pair<uint32_t, uint32_t>> mpcEntries;  // (row eq, mpc eq)
  for (auto& m : mpcs)
    for (auto& t : m.eq) {
      if (t.node < 1 || t.node > nN || !(nodeDofMask[t.node - 1] & (1u << t.dof))) return "MPC term on an unregistered DoF";
      mpcEntries.push_back({nodeDofEq(t.node, t.dof), m.eqNum});
    }
  std::sort(mpcEntries.begin(), mpcEntries.end());
  mpcEntries.erase(std::unique(mpcEntries.begin(), mpcEntries.end()), mpcEntries.end());

  // rows per global equation: count pass then fill pass
  std::vector<uint32_t> rowCount(nTot + 1, 0);  // index eq (1-based)
  std::vector<uint64_t> rowPtr;
  std::vector<uint32_t> cols;

  auto entityRows = [&](bool fill, size_t b, size_t e, std::atomic<bool>* bad) {
    std::vector<uint32_t> cand;
    cand.reserve(512);
    // node rows
    for (size_t n = b; n < e; n++) {
      if (!nodeDofMask[n]) continue;
      cand.clear();
      for (uint64_t k = n2ePtr[n]; k < n2ePtr[n + 1]; k++) {
        uint32_t el = n2e[k];
        for (int t = 0; t < NUM_DOF_TYPES; t++)
          if (elemDofMask[el] & (1u << t)) cand.push_back(elemEq[(size_t)el * 8 + t]);
        for (uint64_t q = elemNodePtr[el]; q < elemNodePtr[el + 1]; q++) {
          uint32_t m = elemNodes[q] - 1;
          for (int t = 0; t < NUM_DOF_TYPES; t++)
            if (nodeDofMask[m] & (1u << t)) cand.push_back(nodeEq[(size_t)m * 8 + t]);
        }
      }
      std::sort(cand.begin(), cand.end());
      cand.erase(std::unique(cand.begin(), cand.end()), cand.end());
      for (int t = 0; t < NUM_DOF_TYPES; t++) {
        if (!(nodeDofMask[n] & (1u << t))) continue;
        uint32_t row = nodeEq[n * 8 + t];
        auto it = std::lower_bound(cand.begin(), cand.end(), row);
        auto mp = std::lower_bound(mpcEntries.begin(), mpcEntries.end(), std::make_pair(row, 0u));
        size_t cntRow = (size_t)(cand.end() - it);
        auto mq = mp;
        while (mq != mpcEntries.end() && mq->first == row) mq++;
        cntRow += (size_t)(mq - mp);
        if (!fill) {
          rowCount[row] = (uint32_t)cntRow;
        } else {
          uint64_t o = rowPtr[row - 1];
          for (auto x = it; x != cand.end(); ++x) cols[o++] = *x;
          for (auto x = mp; x != mq; ++x) cols[o++] = x->second;  // MPC eqs exceed every DoF eq
        }
      }
    }
    (void)bad;
  };
  auto elemRows = [&](bool fill, size_t b, size_t e) {
    std::vector<uint32_t> cand;
    for (size_t el = b; el < e; el++) {
      if (!elemDofMask[el]) continue;
      cand.clear();
      for (int t = 0; t < NUM_DOF_TYPES; t++)
        if (elemDofMask[el] & (1u << t)) cand.push_back(elemEq[el * 8 + t]);
      for (uint64_t q = elemNodePtr[el]; q < elemNodePtr[el + 1]; q++) {
        uint32_t m = elemNodes[q] - 1;
        for (int t = 0; t < NUM_DOF_TYPES; t++)
          if (nodeDofMask[m] & (1u << t)) cand.push_back(nodeEq[(size_t)m * 8 + t]);
      }
      std::sort(cand.begin(), cand.end());
      cand.erase(std::unique(cand.begin(), cand.end()), cand.end());
      for (int t = 0; t < NUM_DOF_TYPES; t++) {
        if (!(elemDofMask[el] & (1u << t))) continue;
        uint32_t row = elemEq[el * 8 + t];
        auto it = std::lower_bound(cand.begin(), cand.end(), row);
        if (!fill) {
          rowCount[row] = (uint32_t)(cand.end() - it);
        } else {
          uint64_t o = rowPtr[row - 1];
          for (auto x = it; x != cand.end(); ++x) cols[o++] = *x;
        }
      }
    }
  };
  std::atomic<bool> bad{false};
  parallelFor(nN, [&](size_t b, size_t e) { entityRows(false, b, e, &bad); });
  parallelFor(nE, [&](size_t b, size_t e) { elemRows(false, b, e); });
  rowPtr.assign(nTot + 1, 0);
  for (uint32_t r = 1; r <= nTot; r++) rowPtr[r] = rowPtr[r - 1] + rowCount[r];
  cols.resize(rowPtr[nTot]);
  parallelFor(nN, [&](size_t b, size_t e) { entityRows(true, b, e, &bad); });
  parallelFor(nE, [&](size_t b, size_t e) { elemRows(true, b, e); });

  // split into block 1 (c x c), block 12 (c x s+l), block 2 (s+l x s+l)
  // (BlockSparseSymMatrix::addEntry, BlockSparseMatrix.h:139-152)
  block1 = CsrPattern();
  block12 = CsrPattern();
  block2 = CsrPattern();
  block1.nRows = block1.nColumns = nC;
  block12.nRows = nC;
  block12.nColumns = nTot - nC;
  block2.nRows = block2.nColumns = nTot - nC;
  block1.iofeir.assign(nC + 1, 1);
  block12.iofeir.assign(nC + 1, 1);
  block2.iofeir.assign(nTot - nC + 1, 1);
  const uint32_t mir1 = std::min(maxInRow, block1.nColumns), mir12 = std::min(maxInRow, block12.nColumns),
                 mir2 = std::min(maxInRow, block2.nColumns);
  for (uint32_t r = 1; r <= nTot; r++) {
    uint64_t b = rowPtr[r - 1], e = rowPtr[r];
    if (r <= nC) {
      uint64_t split = b;
      while (split < e && cols[split] <= nC) split++;
      uint32_t c1 = (uint32_t)(split - b), c12 = (uint32_t)(e - split);
      if (c1 > mir1 || c12 > mir12) return "maxInRow overflow!";  // SparseMatrix.cpp:84
      block1.iofeir[r] = block1.iofeir[r - 1] + c1;
      block12.iofeir[r] = block12.iofeir[r - 1] + c12;
    } else {
      uint32_t c2 = (uint32_t)(e - b);
      if (c2 > mir2) return "maxInRow overflow!";
      block2.iofeir[r - nC] = block2.iofeir[r - nC - 1] + c2;
    }
  }
  block1.columns.resize(block1.iofeir[nC] - 1);
  block12.columns.resize(block12.iofeir[nC] - 1);
  block2.columns.resize(block2.iofeir[nTot - nC] - 1);
  parallelFor(nTot, [&](size_t rb, size_t re) {
    for (size_t r0 = rb; r0 < re; r0++) {
      uint32_t r = (uint32_t)r0 + 1;
      uint64_t b = rowPtr[r - 1], e = rowPtr[r];
      if (r <= nC) {
        uint64_t o1 = block1.iofeir[r - 1] - 1, o12 = block12.iofeir[r - 1] - 1;
        for (uint64_t k = b; k < e; k++) {
          if (cols[k] <= nC) block1.columns[o1++] = cols[k];
          else block12.columns[o12++] = cols[k] - nC;
        }
      } else {
        uint64_t o2 = block2.iofeir[r - nC - 1] - 1;
        for (uint64_t k = b; k < e; k++) block2.columns[o2++] = cols[k] - nC;
      }
    }
  });
  return "";
}

static int64_t findIn(const CsrPattern& p, uint32_t i, uint32_t j) {
  if (i < 1 || i > p.nRows) return -1;
  const uint32_t* b = p.columns.data() + (p.iofeir[i - 1] - 1);
  const uint32_t* e = p.columns.data() + (p.iofeir[i] - 1);
  const uint32_t* it = std::lower_bound(b, e, j);
  if (it == e || *it != j) return -1;
  return (int64_t)(it - p.columns.data());
}

int64_t FEStorage::findEntry(uint32_t i, uint32_t j) const {
  if (i > j) std::swap(i, j);
  const uint32_t nC = nConstrainedDofs;
  if (j <= nC) return findIn(block1, i, j);
  if (i <= nC) {
    int64_t k = findIn(block12, i, j - nC);
    return k < 0 ? -1 : (int64_t)offB12() + k;
  }
  int64_t k = findIn(block2, i - nC, j - nC);
  return k < 0 ? -1 : (int64_t)offB2() + k;
}

}  // namespace nla3d_amd
