// C ABI of libnla3d_gpu.so (include/nla3d_gpu.h): handle, model set-up,
// scatter-map construction and the device assembly path.
#include "../../include/nla3d_gpu.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <functional>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

#include "ctx.h"

using namespace nla3d_amd;

namespace nla3d_amd {

void hipCheck(hipError_t e, const char* what) {
  if (e != hipSuccess) throw ApiError(NLA3D_GPU_EHIP, std::string(what) + ": " + hipGetErrorString(e));
}

void DevBuf::free_() {
  if (p) (void)hipFree(p);
  p = nullptr;
  bytes = 0;
}

void DevBuf::alloc(size_t nbytes) {
  if (nbytes == bytes && p) return;
  free_();
  if (nbytes == 0) return;
  hipCheck(hipMalloc(&p, nbytes), "hipMalloc");
  bytes = nbytes;
}

void setupThermal(Ctx& c);  // thermal.cpp

// Build the u/p element (element, slot) -> target map and sort it by target (stable in
// element order): the precomputed element-to-CSR index map of the deterministic scatter.
// SOLID81 outputs are element-major ([e][350], library slot order), PLANE41 outputs
// slot-major ([54][e], assembleK order); src holds the position in that layout.
void buildUpScatterMap(Ctx& c) {
  FEStorage& m = c.model;
  const uint32_t nS = m.solidCount;
  const bool plane = m.upKind == ELEM_PLANE41;
  const int slots = plane ? kPlane41Slots : kSolid81Slots;
  const int nn = plane ? 4 : 8, nd = plane ? 2 : 3, nLoc = nn * nd;  // node DoFs; nLoc = the pressure
  const uint64_t nK = m.nnzAll();
  const uint64_t nT = nK + m.nAll();
  if (nT >= (1ull << 31) || (uint64_t)nS * slots >= (1ull << 31))
    throw ApiError(NLA3D_GPU_EINVAL, "model too large for 32-bit scatter indices");
  // local DoF pair of each slot: node DoFs nd*i+d, then the pressure; F slots have b = -1
  int la[kSolid81Slots], lb[kSolid81Slots];
  if (plane) {
    plane41_slot_pairs(la, lb);
  } else {
    int slot2a3[kSolid81Slots];
    solid81_slot_to_assemble3(slot2a3);
    int k = 0;
    // recover (a,b) of every packed Kuu index
    int pa[300], pb[300];
    for (int a = 0; a < 24; a++)
      for (int b = a; b < 24; b++) {
        pa[k] = a;
        pb[k] = b;
        k++;
      }
    for (int s = 0; s < 300; s++) {
      la[s] = pa[slot2a3[s]];
      lb[s] = pb[slot2a3[s]];
    }
    for (int s = 300; s < 324; s++) { la[s] = s - 300; lb[s] = 24; }
    la[324] = 24; lb[324] = 24;
    for (int s = 325; s < 349; s++) { la[s] = s - 325; lb[s] = -1; }
    la[349] = 24; lb[349] = -1;
  }
  std::vector<int> tgt((size_t)nS * slots);
  std::string err;
  std::mutex mu;
  parallelFor(nS, [&](size_t b, size_t e) {
    uint32_t eq[25];
    for (size_t s = b; s < e; s++) {
      const uint32_t el = m.solidFirst + (uint32_t)s;  // 0-based global element
      for (int i = 0; i < nn; i++) {
        uint32_t node = m.elemNodes[m.elemNodePtr[el] + i];
        for (int d = 0; d < nd; d++) eq[nd * i + d] = m.nodeDofEq(node, d);
      }
      eq[nLoc] = m.elemDofEq(el + 1, HYDRO_PRESSURE);
      for (int k = 0; k < slots; k++) {
        int64_t t;
        if (lb[k] < 0) {
          t = (int64_t)nK + eq[la[k]] - 1;
        } else {
          t = m.findEntry(eq[la[k]], eq[lb[k]]);
          if (t < 0) {
            std::lock_guard<std::mutex> g(mu);
            err = "The position(" + std::to_string(std::min(eq[la[k]], eq[lb[k]])) + ", " +
                  std::to_string(std::max(eq[la[k]], eq[lb[k]])) + ") is absent in the matrix";
            return;
          }
        }
        tgt[s * slots + k] = (int)t;
      }
    }
  });
  if (!err.empty()) throw ApiError(NLA3D_GPU_EPATTERN, err);
  // counting sort by target, stable in (element, slot) order
  std::vector<int> seg(nT + 1, 0);
  for (size_t i = 0; i < tgt.size(); i++) seg[(size_t)tgt[i] + 1]++;
  for (uint64_t t = 0; t < nT; t++) seg[t + 1] += seg[t];
  std::vector<int> src(tgt.size());
  {
    std::vector<int> pos(seg.begin(), seg.end() - 1);
    for (size_t i = 0; i < tgt.size(); i++)
      src[pos[tgt[i]]++] = plane ? (int)((i % slots) * nS + i / slots) : (int)i;
  }
  c.nTargets = (int)nT;
  c.nContrib = (int64_t)src.size();
  c.d_seg.upload(seg);
  c.d_src.upload(src);
}

}  // namespace nla3d_amd

// ---------------------------------------------------------------------------
// ABI entry points
// ---------------------------------------------------------------------------

static int guard(nla3d_gpu_t h, const std::function<void(Ctx&)>& fn) {
  if (!h) return NLA3D_GPU_EINVAL;
  Ctx& c = *reinterpret_cast<Ctx*>(h);
  try {
    c.err.clear();
    if (c.device >= 0) hipCheck(hipSetDevice(c.device), "hipSetDevice");
    fn(c);
    return NLA3D_GPU_OK;
  } catch (const ApiError& e) {
    c.err = e.what();
    return e.code;
  } catch (const std::bad_alloc&) {
    c.err = "host out of memory";
    return NLA3D_GPU_EINVAL;
  } catch (const std::exception& e) {
    c.err = e.what();
    return NLA3D_GPU_EINVAL;
  }
}

extern "C" {

int nla3d_gpu_create(nla3d_gpu_t* h, int device) {
  if (!h) return NLA3D_GPU_EINVAL;
  *h = nullptr;
  if (device == -1) {  // host-only handle: model set-up, numbering and sparsity, no device work
    Ctx* c = new (std::nothrow) Ctx();
    if (!c) return NLA3D_GPU_EINVAL;
    c->device = -1;
    *h = reinterpret_cast<nla3d_gpu_t>(c);
    return NLA3D_GPU_OK;
  }
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || device < 0 || device >= n) return NLA3D_GPU_EHIP;
  if (hipSetDevice(device) != hipSuccess) return NLA3D_GPU_EHIP;
  Ctx* c = new (std::nothrow) Ctx();
  if (!c) return NLA3D_GPU_EINVAL;
  c->device = device;
  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
    delete c;
    return NLA3D_GPU_EHIP;
  }
  *h = reinterpret_cast<nla3d_gpu_t>(c);
  return NLA3D_GPU_OK;
}

int nla3d_gpu_destroy(nla3d_gpu_t h) {
  if (!h) return NLA3D_GPU_EINVAL;
  Ctx* c = reinterpret_cast<Ctx*>(h);
  if (c->device >= 0) {
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->stream);
  }
  delete c;  // DevBuf destructors free device memory; Ctx dtor destroys the stream
  return NLA3D_GPU_OK;
}

const char* nla3d_gpu_last_error(nla3d_gpu_t h) {
  if (!h) return "null handle";
  return reinterpret_cast<Ctx*>(h)->err.c_str();
}

int nla3d_gpu_set_nodes(nla3d_gpu_t h, uint32_t n_nodes, const double* xyz) {
  return guard(h, [&](Ctx& c) {
    if (c.initialized) throw ApiError(NLA3D_GPU_ESTATE, "model already initialised");
    if (n_nodes && !xyz) throw ApiError(NLA3D_GPU_EINVAL, "xyz is NULL");
    c.model.nodePos.assign(xyz, xyz + (size_t)n_nodes * 3);
  });
}

int nla3d_gpu_add_elements(nla3d_gpu_t h, int kind, uint32_t n, const uint32_t* conn, const double* params) {
  return guard(h, [&](Ctx& c) {
    if (c.initialized) throw ApiError(NLA3D_GPU_ESTATE, "model already initialised");
    if (n && !conn) throw ApiError(NLA3D_GPU_EINVAL, "conn is NULL");
    std::string e = c.model.addElements(kind, n, conn, params);
    if (!e.empty()) throw ApiError(NLA3D_GPU_EINVAL, e);
  });
}

int nla3d_gpu_set_material(nla3d_gpu_t h, int type, const double* MC, int n) {
  return guard(h, [&](Ctx& c) {
    int need = (type == 1) ? 2 : (type == 2 ? 5 : (type == 3 ? 3 : -1));
    if (need < 0) throw ApiError(NLA3D_GPU_EINVAL, "unknown material type");
    if (n != need || !MC) throw ApiError(NLA3D_GPU_EINVAL, "wrong number of material constants");
    c.mat.type = type;
    for (int i = 0; i < 5; i++) c.mat.mc[i] = (i < n) ? MC[i] : 0.0;
    c.mat.k = MC[n - 1];  // getK(): last constant for all three materials
    c.hasMaterial = true;
  });
}

int nla3d_gpu_add_fix(nla3d_gpu_t h, uint32_t node, int dof, double value) {
  return guard(h, [&](Ctx& c) {
    if (c.initialized) throw ApiError(NLA3D_GPU_ESTATE, "model already initialised");
    if (dof < 0 || dof >= NUM_DOF_TYPES) throw ApiError(NLA3D_GPU_EINVAL, "bad dof type");
    c.model.fixs.push_back({node, dof, value});
  });
}

int nla3d_gpu_add_load(nla3d_gpu_t h, uint32_t node, int dof, double value) {
  return guard(h, [&](Ctx& c) {
    if (dof < 0 || dof >= NUM_DOF_TYPES) throw ApiError(NLA3D_GPU_EINVAL, "bad dof type");
    c.model.loads.push_back({node, dof, value});
  });
}

int nla3d_gpu_add_rigid_body_mpc(nla3d_gpu_t h, uint32_t master, uint32_t n_slaves, const uint32_t* slaves,
                                 int n_dofs, const int* dofs) {
  return guard(h, [&](Ctx& c) {
    if (c.initialized) throw ApiError(NLA3D_GPU_ESTATE, "model already initialised");
    if (n_dofs < 1 || n_dofs > 3 || !dofs || (n_slaves && !slaves))
      throw ApiError(NLA3D_GPU_EINVAL, "bad rigid body arguments");
    RigidBodySpec rb;
    rb.masterNode = master;
    rb.slaveNodes.assign(slaves, slaves + n_slaves);
    for (int i = 0; i < n_dofs; i++) {
      if (dofs[i] < 0 || dofs[i] > 2) throw ApiError(NLA3D_GPU_EINVAL, "which degree of freedom?");  // Mpc.cpp:118
      rb.dofs.push_back(dofs[i]);
    }
    c.model.rigidBodies.push_back(rb);
  });
}

int nla3d_gpu_add_mpc(nla3d_gpu_t h, int n_terms, const uint32_t* nodes, const int* dofs, const double* coefs,
                      double b) {
  return guard(h, [&](Ctx& c) {
    if (c.initialized) throw ApiError(NLA3D_GPU_ESTATE, "model already initialised");
    if (n_terms < 1 || !nodes || !dofs || !coefs) throw ApiError(NLA3D_GPU_EINVAL, "an MPC needs terms");  // FEStorage.cpp:176-177
    Mpc mp;
    mp.b = b;
    for (int i = 0; i < n_terms; i++) {
      if (dofs[i] < 0 || dofs[i] >= NUM_DOF_TYPES) throw ApiError(NLA3D_GPU_EINVAL, "bad MPC degree of freedom");
      mp.eq.push_back({nodes[i], dofs[i], coefs[i]});
    }
    c.model.ceMpcs.push_back(mp);
  });
}

int nla3d_gpu_set_max_in_row(nla3d_gpu_t h, uint32_t max_in_row) {
  return guard(h, [&](Ctx& c) {
    if (c.initialized) throw ApiError(NLA3D_GPU_ESTATE, "model already initialised");
    c.maxInRow = max_in_row ? max_in_row : 0xFFFFFFFFu;
  });
}

int nla3d_gpu_init_solution(nla3d_gpu_t h) {
  return guard(h, [&](Ctx& c) {
    if (c.initialized) throw ApiError(NLA3D_GPU_ESTATE, "model already initialised");
    FEStorage& m = c.model;
    std::string e;
    if (!(e = m.initDofs()).empty()) throw ApiError(NLA3D_GPU_EINVAL, e);
    if (!(e = m.setConstrainedDofs()).empty()) throw ApiError(NLA3D_GPU_EINVAL, e);
    if (!(e = m.assignEquationNumbers()).empty()) throw ApiError(NLA3D_GPU_EINVAL, e);
    if (!(e = m.buildSparsity(c.maxInRow)).empty()) throw ApiError(NLA3D_GPU_EPATTERN, e);
    if (c.device >= 0) {
      c.setupDevice();
      c.setupKrylov();
      setupThermal(c);
    }
    c.initialized = true;
  });
}

int nla3d_gpu_get_sizes(nla3d_gpu_t h, uint64_t* s) {
  return guard(h, [&](Ctx& c) {
    if (!s) throw ApiError(NLA3D_GPU_EINVAL, "NULL");
    FEStorage& m = c.model;
    s[0] = m.nNodes();
    s[1] = m.nElements();
    s[2] = m.nDofs;
    s[3] = m.nConstrainedDofs;
    s[4] = m.nUnknownDofs;
    s[5] = m.nMpc();
    s[6] = m.block1.nnz();
    s[7] = m.block12.nnz();
    s[8] = m.block2.nnz();
  });
}

int nla3d_gpu_get_equation_numbers(nla3d_gpu_t h, uint32_t* node_eq, uint32_t* elem_eq) {
  return guard(h, [&](Ctx& c) {
    if (!c.initialized) throw ApiError(NLA3D_GPU_ESTATE, "call nla3d_gpu_init_solution first");
    if (node_eq) std::memcpy(node_eq, c.model.nodeEq.data(), c.model.nodeEq.size() * 4);
    if (elem_eq) std::memcpy(elem_eq, c.model.elemEq.data(), c.model.elemEq.size() * 4);
  });
}

int nla3d_gpu_get_pattern(nla3d_gpu_t h, int block, uint32_t* iofeir, uint32_t* columns) {
  return guard(h, [&](Ctx& c) {
    if (!c.initialized) throw ApiError(NLA3D_GPU_ESTATE, "call nla3d_gpu_init_solution first");
    const CsrPattern* p = block == 1 ? &c.model.block1 : block == 12 ? &c.model.block12 : block == 2 ? &c.model.block2 : nullptr;
    if (!p) throw ApiError(NLA3D_GPU_EINVAL, "block must be 1, 12 or 2");
    if (iofeir) std::memcpy(iofeir, p->iofeir.data(), p->iofeir.size() * 4);
    if (columns) std::memcpy(columns, p->columns.data(), p->columns.size() * 4);
  });
}

int nla3d_gpu_set_solution(nla3d_gpu_t h, const double* U) {
  return guard(h, [&](Ctx& c) {
    c.requireDevice();
    if (!c.initialized) throw ApiError(NLA3D_GPU_ESTATE, "call nla3d_gpu_init_solution first");
    if (!U) throw ApiError(NLA3D_GPU_EINVAL, "NULL");
    hipCheck(hipMemcpyAsync(c.d_U.p, U, (size_t)c.model.nAll() * 8, hipMemcpyHostToDevice, c.stream), "copy U");
    // the next load step's prescribed increments are measured from this state's constrained values
    if (c.model.nConstrainedDofs > 0)
      hipCheck(hipMemcpyAsync(c.d_Ucprev.p, U, (size_t)c.model.nConstrainedDofs * 8, hipMemcpyHostToDevice, c.stream),
               "copy Uc");
    hipCheck(hipStreamSynchronize(c.stream), "sync");
    c.lsValid = false;  // the predictors' history no longer describes this state
    c.predN[0] = c.predN[1] = 0;
    c.lsTargetT = 0.0;
  });
}

int nla3d_gpu_get_solution(nla3d_gpu_t h, double* U) {
  return guard(h, [&](Ctx& c) {
    c.requireDevice();
    if (!c.initialized) throw ApiError(NLA3D_GPU_ESTATE, "call nla3d_gpu_init_solution first");
    if (!U) throw ApiError(NLA3D_GPU_EINVAL, "NULL");
    hipCheck(hipMemcpyAsync(U, c.d_U.p, (size_t)c.model.nAll() * 8, hipMemcpyDeviceToHost, c.stream), "copy U");
    hipCheck(hipStreamSynchronize(c.stream), "sync");
  });
}

int nla3d_gpu_assemble(nla3d_gpu_t h) {
  return guard(h, [&](Ctx& c) {
    c.requireSolid();
    c.assemble();
    hipCheck(hipStreamSynchronize(c.stream), "assemble");
  });
}

int nla3d_gpu_get_K(nla3d_gpu_t h, int block, double* values) {
  return guard(h, [&](Ctx& c) {
    c.requireDevice();
    if (!c.initialized) throw ApiError(NLA3D_GPU_ESTATE, "call nla3d_gpu_init_solution first");
    if (!values) throw ApiError(NLA3D_GPU_EINVAL, "NULL");
    uint64_t off, n;
    if (block == 1) { off = 0; n = c.model.block1.nnz(); }
    else if (block == 12) { off = c.model.offB12(); n = c.model.block12.nnz(); }
    else if (block == 2) { off = c.model.offB2(); n = c.model.block2.nnz(); }
    else throw ApiError(NLA3D_GPU_EINVAL, "block must be 1, 12 or 2");
    hipCheck(hipMemcpyAsync(values, c.d_vals.as<double>() + off, n * 8, hipMemcpyDeviceToHost, c.stream), "copy K");
    hipCheck(hipStreamSynchronize(c.stream), "sync");
  });
}

int nla3d_gpu_get_F(nla3d_gpu_t h, double* F) {
  return guard(h, [&](Ctx& c) {
    c.requireDevice();
    if (!c.initialized) throw ApiError(NLA3D_GPU_ESTATE, "call nla3d_gpu_init_solution first");
    if (!F) throw ApiError(NLA3D_GPU_EINVAL, "NULL");
    hipCheck(hipMemcpyAsync(F, c.d_vals.as<double>() + c.model.nnzAll(), (size_t)c.model.nAll() * 8,
                            hipMemcpyDeviceToHost, c.stream), "copy F");
    hipCheck(hipStreamSynchronize(c.stream), "sync");
  });
}

int nla3d_gpu_solid81_element_matrices(nla3d_gpu_t h, double* out) {
  return guard(h, [&](Ctx& c) {
    c.requireSolid81();
    if (!out) throw ApiError(NLA3D_GPU_EINVAL, "NULL");
    c.runElements(nullptr);
    const uint32_t nS = c.model.solidCount;
    std::vector<double> ke((size_t)nS * kSolid81Slots);
    hipCheck(hipMemcpyAsync(ke.data(), c.d_Ke.p, ke.size() * 8, hipMemcpyDeviceToHost, c.stream), "copy Ke");
    hipCheck(hipStreamSynchronize(c.stream), "sync");
    int map[kSolid81Slots];
    solid81_slot_to_assemble3(map);
    for (size_t e = 0; e < nS; e++)
      for (int k = 0; k < kSolid81Slots; k++) out[e * kSolid81Slots + map[k]] = ke[e * kSolid81Slots + k];
  });
}

int nla3d_gpu_solid81_gp_state(nla3d_gpu_t h, double* out) {
  return guard(h, [&](Ctx& c) {
    c.requireSolid81();
    if (!out) throw ApiError(NLA3D_GPU_EINVAL, "NULL");
    DevBuf st;
    st.alloc((size_t)c.model.solidCount * 8 * kSolid81GpState * 8);
    c.runElements(st.as<double>());
    hipCheck(hipMemcpyAsync(out, st.p, st.bytes, hipMemcpyDeviceToHost, c.stream), "copy state");
    hipCheck(hipStreamSynchronize(c.stream), "sync");
  });
}

int nla3d_gpu_plane41_element_matrices(nla3d_gpu_t h, double* out) {
  return guard(h, [&](Ctx& c) {
    c.requireSolid();
    if (c.model.upKind != ELEM_PLANE41) throw ApiError(NLA3D_GPU_ESTATE, "this entry point needs PLANE41 elements");
    if (!out) throw ApiError(NLA3D_GPU_EINVAL, "NULL");
    c.runElements(nullptr);
    const size_t nS = c.model.solidCount;
    std::vector<double> ke(nS * kPlane41Slots);
    hipCheck(hipMemcpyAsync(ke.data(), c.d_Ke.p, ke.size() * 8, hipMemcpyDeviceToHost, c.stream), "copy Ke");
    hipCheck(hipStreamSynchronize(c.stream), "sync");
    for (size_t e = 0; e < nS; e++)
      for (int k = 0; k < kPlane41Slots; k++) out[e * kPlane41Slots + k] = ke[(size_t)k * nS + e];
  });
}

int nla3d_gpu_element_query(nla3d_gpu_t h, int kind, int query, int gp, double* out) {
  return guard(h, [&](Ctx& c) {
    c.requireSolid();
    if (!out) throw ApiError(NLA3D_GPU_EINVAL, "NULL");
    const bool hex = c.model.upKind == ELEM_SOLID81;
    const int ngp = hex ? 8 : 4;
    if (gp < -1 || gp >= ngp) throw ApiError(NLA3D_GPU_EINVAL, "Gauss point out of range");
    // the queries each element answers (SOLID81.cpp:191-345, PLANE41.cpp:153-300); PLANE41's
    // IC indexes past its argument (PLANE41.cpp:205-207) and is not offered
    bool ok;
    if (kind == 0) ok = hex ? (query == 1 || query == 3 || query == 4 || query == 5) : query == 1;
    else if (kind == 1) ok = hex && query == 1;
    else if (kind == 2) ok = query >= 1 && query <= 4;
    else ok = false;
    if (!ok) throw ApiError(NLA3D_GPU_EINVAL, "query not answered by this element type");
    const int width = kind == 0 ? 1 : (kind == 1 ? 3 : 6);
    const size_t nS = c.model.solidCount;
    DevBuf buf;
    buf.alloc(nS * width * 8);
    launch_element_query(c.stream, hex, (int)nS, c.d_xyz.as<double>(), c.d_conn.as<int>(), c.d_nodeEq3.as<int>(),
                         c.d_elemEqP.as<int>(), c.d_U.as<double>(), c.mat, kind, query, gp, buf.as<double>());
    hipCheck(hipGetLastError(), "query kernel launch");
    hipCheck(hipMemcpyAsync(out, buf.p, buf.bytes, hipMemcpyDeviceToHost, c.stream), "copy query");
    hipCheck(hipStreamSynchronize(c.stream), "sync");
  });
}

int nla3d_gpu_time_assembly(nla3d_gpu_t h, int iters, double* ms_element, double* ms_scatter) {
  return guard(h, [&](Ctx& c) {
    c.requireSolid();
    if (iters < 1) throw ApiError(NLA3D_GPU_EINVAL, "iters < 1");
    hipEvent_t ev[3];
    for (auto& e : ev) hipCheck(hipEventCreate(&e), "event");
    double te = 0, ts = 0;
    for (int it = 0; it < iters; it++) {
      hipCheck(hipEventRecord(ev[0], c.stream), "rec");
      c.runElements(nullptr);
      hipCheck(hipEventRecord(ev[1], c.stream), "rec");
      c.scatter();
      hipCheck(hipEventRecord(ev[2], c.stream), "rec");
      hipCheck(hipEventSynchronize(ev[2]), "sync");
      float a = 0, b = 0;
      hipCheck(hipEventElapsedTime(&a, ev[0], ev[1]), "elapsed");
      hipCheck(hipEventElapsedTime(&b, ev[1], ev[2]), "elapsed");
      te += a;
      ts += b;
    }
    for (auto& e : ev) (void)hipEventDestroy(e);
    c.applyMpc();  // the timed scatters left the MPC rows and the SpMV's copy to be redone
    if (ms_element) *ms_element = te / iters;
    if (ms_scatter) *ms_scatter = ts / iters;
  });
}

int nla3d_gpu_copy_calibration(nla3d_gpu_t h, uint64_t bytes, double* ms) {
  return guard(h, [&](Ctx& c) {
    c.requireDevice();
    if (bytes < 8 || !ms) throw ApiError(NLA3D_GPU_EINVAL, "bad arguments");
    const size_t n = bytes / 8;
    DevBuf a, b;
    a.alloc(n * 8);
    b.alloc(n * 8);
    hipCheck(hipMemset(a.p, 0, n * 8), "memset");
    hipEvent_t e0, e1;
    hipCheck(hipEventCreate(&e0), "event");
    hipCheck(hipEventCreate(&e1), "event");
    hipCheck(hipEventRecord(e0, c.stream), "rec");
    launch_calibration_copy(c.stream, n, a.as<double>(), b.as<double>());
    hipCheck(hipEventRecord(e1, c.stream), "rec");
    hipCheck(hipEventSynchronize(e1), "sync");
    float t = 0;
    hipCheck(hipEventElapsedTime(&t, e0, e1), "elapsed");
    *ms = t;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
  });
}

}  // extern "C"
