// Host-side model store: the library's counterpart of nla3d's FEStorage
// (FEStorage.h/.cpp), DofCollection (Dof.h/.cpp), the sparsity of
// BlockSparseSymMatrix<2> (BlockSparseMatrix.h, SparseMatrix.cpp) and the
// MPC records (Mpc.h).  It owns the mesh, the DoF tables and equation
// numbers, and the 1-based upper-triangle CSR patterns of the three K blocks;
// the values live on the device (see Device in capi.cpp).
//
// Names follow the reference so that a reader can map each method to the
// FEStorage call it replaces.  Errors are returned as strings instead of the
// reference's LOG(FATAL) aborts.
#pragma once

#include <cstdint>
#include <functional>
#include <string>
#include <vector>

namespace nla3d_amd {

// Dof::dofType, Dof.h:15-25 (enum order defines the numbering order inside an entity)
enum DofType { UX = 0, UY, UZ, ROTX, ROTY, ROTZ, HYDRO_PRESSURE, TEMP, NUM_DOF_TYPES };

// element types handled by the library (ElementType, sys.h:151-163)
enum ElemKind { ELEM_SOLID81 = 1, ELEM_QUADTH = 2, ELEM_SURFACE_LINETH = 3, ELEM_PLANE41 = 4 };

struct MpcTerm {  // Mpc.h:15-27
  uint32_t node;
  int dof;
  double coef;
};

struct Mpc {  // Mpc.h:30-36
  std::vector<MpcTerm> eq;
  double b = 0.0;
  uint32_t eqNum = 0;
};

struct RigidBodySpec {  // RigidBodyMpc, Mpc.h:49-60
  uint32_t masterNode = 0;
  std::vector<uint32_t> slaveNodes;
  std::vector<int> dofs;
  uint32_t firstMpc = 0;  // index of its first Mpc in FEStorage::mpcs
};

struct BC {  // fixBC / loadBC, Mpc.h:64-84
  uint32_t node;
  int dof;
  double value;
};

// One 1-based CSR block in the reference layout (SparsityInfo, SparseMatrix.h:24-60):
// iofeir[nRows+1] starts at 1, columns sorted per row.
struct CsrPattern {
  uint32_t nRows = 0, nColumns = 0;
  std::vector<uint32_t> iofeir;
  std::vector<uint32_t> columns;
  uint64_t nnz() const { return columns.size(); }
};

class FEStorage {
 public:
  // ---- mesh (FEStorage::createNodes / createElements) ----
  std::vector<double> nodePos;          // 3 per node
  uint32_t nNodes() const { return (uint32_t)(nodePos.size() / 3); }

  // all elements in global (1-based) order; each element has a kind and nodes
  std::vector<uint8_t> elemKind;
  std::vector<uint64_t> elemNodePtr{0};  // CSR offsets into elemNodes
  std::vector<uint32_t> elemNodes;       // 1-based node numbers
  uint32_t nElements() const { return (uint32_t)elemKind.size(); }
  // contiguous range of the u/p solid elements (the GPU kernel needs them dense): all
  // SOLID81 or all PLANE41 (upKind)
  uint32_t solidFirst = 0, solidCount = 0;
  int upKind = 0;
  // thermal element data (QUADTH k, rho, c, volFlux ; SurfaceLINETH flux, htc, etemp0, etemp1)
  std::vector<double> elemParams;        // 4 per element

  // ---- BCs and MPCs ----
  std::vector<BC> fixs, loads;
  std::vector<Mpc> mpcs;
  std::vector<Mpc> ceMpcs;  // constant linear MPCs (FEStorage::addMpc; CE records): first in mpcs
  std::vector<RigidBodySpec> rigidBodies;

  // ---- DoF tables (DofCollection) ----
  std::vector<uint8_t> nodeDofMask, elemDofMask;      // registered types per entity
  std::vector<uint8_t> nodeConstr, elemConstr;        // constrained types per entity
  std::vector<uint32_t> nodeEq, elemEq;               // [entity*8 + type], 1-based, 0 = unused
  uint32_t nDofs = 0, nConstrainedDofs = 0, nUnknownDofs = 0;
  uint32_t nMpc() const { return (uint32_t)mpcs.size(); }
  uint32_t nAll() const { return nDofs + nMpc(); }

  // ---- K sparsity (BlockSparseSymMatrix<2>({nC, nU + nMpc})) ----
  CsrPattern block1, block12, block2;

  // adds elements; returns "" or an error
  std::string addElements(int kind, uint32_t n, const uint32_t* conn1, const double* params);

  // FEStorage::initDofs (FEStorage.cpp:338-379) + Element::pre DoF registration
  std::string initDofs();
  // FESolver::setConstrainedDofs (FESolver.cpp:221-226)
  std::string setConstrainedDofs();
  // FEStorage::assignEquationNumbers (FEStorage.cpp:382-432)
  std::string assignEquationNumbers();
  // FEStorage::initSolutionData sparsity part (FEStorage.cpp:434-550) + compress
  std::string buildSparsity(uint32_t maxInRow = 100);

  uint32_t nodeDofEq(uint32_t node1, int t) const { return nodeEq[(size_t)(node1 - 1) * 8 + t]; }
  uint32_t elemDofEq(uint32_t el1, int t) const { return elemEq[(size_t)(el1 - 1) * 8 + t]; }

  // locate global (i,j), 1-based, in the block patterns: returns the index into
  // the concatenated value array [block1 | block12 | block2], or -1 if absent
  // (BlockSparseSymMatrix::addValue -> SparsityInfo::getIndex).
  int64_t findEntry(uint32_t i, uint32_t j) const;
  uint64_t offB12() const { return block1.nnz(); }
  uint64_t offB2() const { return block1.nnz() + block12.nnz(); }
  uint64_t nnzAll() const { return block1.nnz() + block12.nnz() + block2.nnz(); }
};

// number of host threads used by the setup loops
unsigned hostThreads();
// fn(begin, end) over [0, n) split into contiguous chunks, one per host thread (inline below 4096)
void parallelFor(size_t n, const std::function<void(size_t, size_t)>& fn);

}  // namespace nla3d_amd
