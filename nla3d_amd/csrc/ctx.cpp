// Device set-up and the assembly pipeline of a handle.
#include "ctx.h"

#include <algorithm>

namespace nla3d_amd {

// defined in capi.cpp
void buildUpScatterMap(struct Ctx& c);

void Ctx::setupDevice() {
  FEStorage& m = model;
  const uint32_t nN = m.nNodes();
  d_xyz.upload(m.nodePos);
  std::vector<int> eq3((size_t)nN * 3);
  for (uint32_t n = 0; n < nN; n++)
    for (int d = 0; d < 3; d++) eq3[(size_t)n * 3 + d] = (int)m.nodeEq[(size_t)n * 8 + d] - 1;
  d_nodeEq3.upload(eq3);
  if (m.solidCount) {
    const int nn = m.upKind == ELEM_PLANE41 ? 4 : 8;
    const int slots = m.upKind == ELEM_PLANE41 ? kPlane41Slots : kSolid81Slots;
    std::vector<int> conn((size_t)m.solidCount * nn), peq(m.solidCount);
    for (uint32_t s = 0; s < m.solidCount; s++) {
      const uint32_t el = m.solidFirst + s;
      for (int i = 0; i < nn; i++) conn[(size_t)s * nn + i] = (int)m.elemNodes[m.elemNodePtr[el] + i] - 1;
      peq[s] = (int)m.elemEq[(size_t)el * 8 + HYDRO_PRESSURE] - 1;
    }
    d_conn.upload(conn);
    d_elemEqP.upload(peq);
    d_Ke.alloc((size_t)m.solidCount * slots * 8);
    buildUpScatterMap(*this);
  }
  d_U.alloc((size_t)m.nAll() * 8);
  hipCheck(hipMemset(d_U.p, 0, d_U.bytes), "memset U");
  d_vals.alloc((size_t)(m.nnzAll() + m.nAll()) * 8);
  hipCheck(hipMemset(d_vals.p, 0, d_vals.bytes), "memset vals");

  // constant MPC rows (the first m.ceMpcs.size() entries of m.mpcs)
  {
    std::vector<int> ptr{0}, tgt, ft;
    std::vector<double> coef, bb;
    for (size_t r = 0; r < m.ceMpcs.size(); r++) {
      const Mpc& mp = m.mpcs[r];
      for (const MpcTerm& t : mp.eq) {
        const int64_t k = m.findEntry(m.nodeDofEq(t.node, t.dof), mp.eqNum);
        if (k < 0) throw ApiError(-4, "MPC entry absent from the pattern");
        tgt.push_back((int)k);
        coef.push_back(t.coef);
      }
      ptr.push_back((int)tgt.size());
      ft.push_back((int)(m.nnzAll() + mp.eqNum - 1));
      bb.push_back(mp.b);
    }
    nCeRows = (int)m.ceMpcs.size();
    ceTermPtr.upload(ptr);
    ceTgt.upload(tgt);
    ceCoef.upload(coef);
    ceFTgt.upload(ft);
    ceB.upload(bb);
  }
  // rigid-body MPC rows: targets of (term DoF, MPC eq) and of F[MPC eq]
  rbs.clear();
  rbs.resize(m.rigidBodies.size());
  for (size_t r = 0; r < m.rigidBodies.size(); r++) {
    const RigidBodySpec& spec = m.rigidBodies[r];
    RB& rb = rbs[r];
    for (int k = 0; k < 6; k++) rb.dev.masterEq[k] = (int)m.nodeDofEq(spec.masterNode, k) - 1;
    for (int k = 0; k < 3; k++) rb.dev.masterPos[k] = m.nodePos[(size_t)(spec.masterNode - 1) * 3 + k];
    rb.dev.nDofs = (int)spec.dofs.size();
    for (int k = 0; k < rb.dev.nDofs; k++) rb.dev.dofs[k] = spec.dofs[k];
    rb.dev.nRows = (int)(spec.slaveNodes.size() * spec.dofs.size());
    std::vector<int> sn(spec.slaveNodes.size()), se(rb.dev.nRows), tg((size_t)rb.dev.nRows * 5), ft(rb.dev.nRows);
    for (size_t s = 0; s < spec.slaveNodes.size(); s++) sn[s] = (int)spec.slaveNodes[s] - 1;
    for (int row = 0; row < rb.dev.nRows; row++) {
      const Mpc& mp = m.mpcs[spec.firstMpc + row];
      se[row] = (int)m.nodeDofEq(mp.eq[0].node, mp.eq[0].dof) - 1;
      for (int t = 0; t < 5; t++) {
        int64_t k = m.findEntry(m.nodeDofEq(mp.eq[t].node, mp.eq[t].dof), mp.eqNum);
        if (k < 0) throw ApiError(-4, "MPC entry absent from the pattern");
        tg[(size_t)row * 5 + t] = (int)k;
      }
      ft[row] = (int)(m.nnzAll() + mp.eqNum - 1);
    }
    rb.slaveNode.upload(sn);
    rb.slaveEq.upload(se);
    rb.tgt.upload(tg);
    rb.fTgt.upload(ft);
  }
}

void Ctx::runElements(double* gp_state) {
  if (model.upKind == ELEM_PLANE41) {
    if (gp_state) throw ApiError(-3, "Gauss-point state is available for SOLID81 only");
    launch_plane41(stream, (int)model.solidCount, d_xyz.as<double>(), d_conn.as<int>(), d_nodeEq3.as<int>(),
                   d_elemEqP.as<int>(), d_U.as<double>(), mat, d_Ke.as<double>());
    hipCheck(hipGetLastError(), "plane41 kernel launch");
    return;
  }
  launch_solid81(stream, (int)model.solidCount, d_xyz.as<double>(), d_conn.as<int>(), d_nodeEq3.as<int>(),
                 d_elemEqP.as<int>(), d_U.as<double>(), mat, d_Ke.as<double>(), gp_state);
  hipCheck(hipGetLastError(), "solid81 kernel launch");
}

void Ctx::scatter() {
  const int off2 = (int)model.offB2();
  launch_segmented_gather(stream, nTargets, d_seg.as<int>(), d_src.as<int>(), d_Ke.as<double>(), d_vals.as<double>(),
                          off2, off2 + (int)model.block2.nnz(), d_mir2.as<int>(),
                          krylovReady ? d_full2.as<double>() : nullptr);
  hipCheck(hipGetLastError(), "scatter kernel launch");
}

void Ctx::applyMpc() {
  launch_ce_mpc(stream, nCeRows, ceTermPtr.as<int>(), ceTgt.as<int>(), ceCoef.as<double>(), ceFTgt.as<int>(),
                ceB.as<double>(), d_vals.as<double>());
  for (auto& rb : rbs) {
    launch_rigid_body_mpc(stream, rb.dev, d_xyz.as<double>(), rb.slaveNode.as<int>(), rb.slaveEq.as<int>(),
                          d_U.as<double>(), rb.tgt.as<int>(), rb.fTgt.as<int>(), d_vals.as<double>());
    hipCheck(hipGetLastError(), "mpc kernel launch");
  }
  if (krylovReady && nMpcMir > 0) {
    launch_mirror(stream, nMpcMir, d_vals.as<double>() + model.offB2(), d_mpcMirIdx.as<int>(), d_mir2.as<int>(),
                  d_full2.as<double>());
    hipCheck(hipGetLastError(), "mirror launch");
  }
  refreshSpmvCopy(stream);
}

// FEStorage::assembleGlobalEqMatrices (FEStorage.cpp:26-78): the gather writes
// every K and F entry (entries without element contributions become +0.0,
// which is zeroK/zeroF), then the MPC rows are set.
void Ctx::assemble() {
  runElements(nullptr);
  scatter();
  applyMpc();
}

}  // namespace nla3d_amd
