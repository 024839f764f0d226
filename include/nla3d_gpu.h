/*
 * nla3d_gpu.h -- C ABI of libnla3d_gpu.so, the MI355X (gfx950) implementation
 * of nla3d's Newton-Raphson equilibrium iteration.
 *
 * Conventions
 *   - Every function returns int: 0 = OK, negative = error.  The reference
 *     aborts through LOG(FATAL)/CHECK instead; here the message of the last
 *     error is kept per handle (nla3d_gpu_last_error) and nothing aborts.
 *   - Host pointers are owned by the caller and only read/written during the
 *     call.  Device buffers are owned by the library.
 *   - Node, element and equation numbers are 1-based, as in nla3d; CSR
 *     arrays use the reference's 1-based MKL layout (iofeir, columns).
 *   - One HIP stream per handle.  A handle is not thread-safe; distinct
 *     handles may be used concurrently.
 *   - DoF type codes are nla3d's Dof::dofType (Dof.h:15-25):
 *     UX=0 UY=1 UZ=2 ROTX=3 ROTY=4 ROTZ=5 HYDRO_PRESSURE=6 TEMP=7.
 *   - Material type codes are MaterialFactory's ids (MaterialFactory.cpp:9-13):
 *     1 Neo-Hookean {G, K}, 2 Biderman {C10, C20, C30, C01, K},
 *     3 Mooney-Rivlin {C10, C01, K}.
 *
 * Citations are path:line in dmitryikh/nla3d.
 */
#ifndef NLA3D_GPU_H
#define NLA3D_GPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define NLA3D_GPU_OK 0
#define NLA3D_GPU_EINVAL (-1)     /* bad argument */
#define NLA3D_GPU_EHIP (-2)       /* HIP runtime error */
#define NLA3D_GPU_ESTATE (-3)     /* call out of order (e.g. assemble before init) */
#define NLA3D_GPU_EPATTERN (-4)   /* CSR position absent / maxInRow overflow (SparseMatrix.cpp:84,401,476) */
#define NLA3D_GPU_ENOTCONV (-5)   /* Newton/Krylov did not converge (FESolver.cpp:490) */
#define NLA3D_GPU_EDIVERGED (-6)  /* Newton diverged (FESolver.cpp:484) */

/* element kinds (ElementType, sys.h:151-163) */
#define NLA3D_GPU_SOLID81 1
#define NLA3D_GPU_QUADTH 2
#define NLA3D_GPU_SURFACE_LINETH 3
#define NLA3D_GPU_PLANE41 4 /* plane-strain u/p quadrilateral (PLANE41.h:21-47) */

typedef struct nla3d_gpu_ctx* nla3d_gpu_t;

/* ---- lifetime ---------------------------------------------------------- */
int nla3d_gpu_create(nla3d_gpu_t* h, int device);
int nla3d_gpu_destroy(nla3d_gpu_t h);
const char* nla3d_gpu_last_error(nla3d_gpu_t h);

/* ---- model (FEStorage, FEStorage.h:38-270) ------------------------------ */
/* FEStorage::createNodes + Node::pos (FEStorage.cpp:199-214): xyz[3*n] */
int nla3d_gpu_set_nodes(nla3d_gpu_t h, uint32_t n_nodes, const double* xyz);
/* FEStorage::createElements (FEStorage.h:675-696).  conn: nodes per element
 * (8 SOLID81, 4 QUADTH, 2 SurfaceLINETH, 4 PLANE41), 1-based.  The u/p solids
 * (SOLID81 or PLANE41, not both) are added as one contiguous range.  params: 4 doubles per
 * element or NULL: QUADTH {k, rho, c, volFlux} (QUADTH.h:27-34),
 * SurfaceLINETH {flux, htc, etemp0, etemp1} (QUADTH.h:49-51). */
int nla3d_gpu_add_elements(nla3d_gpu_t h, int kind, uint32_t n, const uint32_t* conn, const double* params);
/* FEStorage::material (FEStorage.h:46) through MaterialFactory + Material::Ci */
int nla3d_gpu_set_material(nla3d_gpu_t h, int type, const double* MC, int n);
/* FESolver::addFix / addLoad (FESolver.cpp:250-259) */
int nla3d_gpu_add_fix(nla3d_gpu_t h, uint32_t node, int dof, double value);
int nla3d_gpu_add_load(nla3d_gpu_t h, uint32_t node, int dof, double value);
/* FEStorage::addMpc (FEStorage.cpp:175-179) with an Mpc (Mpc.h:30-36): the linear
 * constraint sum_i coefs[i] * U(nodes[i], dofs[i]) = b, e.g. an ANSYS CE record.  Its
 * multiplier equation is numbered after the DoFs, before the rigid bodies' rows, in the
 * order of the calls; every term's DoF must belong to an element. */
int nla3d_gpu_add_mpc(nla3d_gpu_t h, int n_terms, const uint32_t* nodes, const int* dofs, const double* coefs,
                      double b);
/* RigidBodyMpc (Mpc.h:49-60): masterNode, slaveNodes, dofs */
int nla3d_gpu_add_rigid_body_mpc(nla3d_gpu_t h, uint32_t master, uint32_t n_slaves, const uint32_t* slaves,
                                 int n_dofs, const int* dofs);
/* Per-row capacity of the sparsity build, the max_in_row argument of
 * BlockSparseSymMatrix<2> (BlockSparseMatrix.h:42; FEStorage.cpp:465 uses the
 * default 100).  A row with more entries fails init_solution with
 * NLA3D_GPU_EPATTERN, as the reference's "maxInRow overflow!" (SparseMatrix.cpp:84).
 * 0 = no limit (needed when a rigid-body master DoF couples to more than 100
 * MPC rows).  Must be called before nla3d_gpu_init_solution. */
int nla3d_gpu_set_max_in_row(nla3d_gpu_t h, uint32_t max_in_row);
/* initDofs + setConstrainedDofs + assignEquationNumbers + initSolutionData
 * (FESolver.cpp:418-421; FEStorage.cpp:338-550), then builds the scatter map
 * and allocates device storage. */
int nla3d_gpu_init_solution(nla3d_gpu_t h);
/* sizes[9] = {nNodes, nElements, nDofs, nConstrained, nUnknown, nMpc,
 *             nnz(block1), nnz(block12), nnz(block2)} */
int nla3d_gpu_get_sizes(nla3d_gpu_t h, uint64_t* sizes);
/* equation numbers: node_eq[nNodes*8], elem_eq[nElements*8] indexed by DoF type,
 * 0 = DoF not registered (FEStorage::getNodeDofEqNumber, FEStorage.h:606-609) */
int nla3d_gpu_get_equation_numbers(nla3d_gpu_t h, uint32_t* node_eq, uint32_t* elem_eq);
/* BlockSparseSymMatrix<2>::block(1), block(1,2), block(2): block = 1, 12, 2.
 * iofeir[nRows+1], columns[nnz], 1-based (SparseMatrix.h:114-116). */
int nla3d_gpu_get_pattern(nla3d_gpu_t h, int block, uint32_t* iofeir, uint32_t* columns);

/* ---- solution vector and assembly ---------------------------------------- */
/* vecU (FEStorage.h:250), length nDofs + nMpc, equation order.  set_solution makes U the state
 * the next load step starts from: its constrained values are the ones that step's prescribed
 * increments are measured from, and the first-solve predictor's history restarts. */
int nla3d_gpu_set_solution(nla3d_gpu_t h, const double* U);
int nla3d_gpu_get_solution(nla3d_gpu_t h, double* U);
/* FEStorage::assembleGlobalEqMatrices (FEStorage.cpp:26-78) on the device for
 * SOLID81 or PLANE41 at the current solution: per element update() + buildK()
 * (SOLID81.cpp:29-115, PLANE41.cpp:31-149), deterministic scatter into K and F,
 * then the MPC rows. */
int nla3d_gpu_assemble(nla3d_gpu_t h);
/* copy K block values (block = 1, 12, 2) or F (length nDofs + nMpc) to the host */
int nla3d_gpu_get_K(nla3d_gpu_t h, int block, double* values);
int nla3d_gpu_get_F(nla3d_gpu_t h, double* F);
/* Per-element SOLID81 buildK output at the current solution, in the reference's
 * assemble3 order (SOLID81.h:54-94): Kuu packed upper 24x24 (300), Kup (24),
 * Kpp, Fu (24), Fp -> out[nSolid81 * 350]. */
int nla3d_gpu_solid81_element_matrices(nla3d_gpu_t h, double* out);
/* Gauss-point state after update() (SOLID81.h:38-45): per element and Gauss
 * point O[9], C[6], S[6] -> out[nSolid81 * 8 * 21]. */
int nla3d_gpu_solid81_gp_state(nla3d_gpu_t h, double* out);
/* Per-element PLANE41 buildK output at the current solution, in the reference's
 * assembleK order (PLANE41.h:55-89): 36 Kuu entries (node pairs i <= j, DoF
 * pairs UX UY, the lower one only on a diagonal node block), Kup (8), Kpp,
 * Fe = [-Qe (8), -Fp] -> out[nPlane41 * 54]. */
int nla3d_gpu_plane41_element_matrices(nla3d_gpu_t h, double* out);

/* Element results at the current solution for every u/p element: Element::getScalar /
 * getVector / getTensor (element.h:122-131; SOLID81.cpp:191-345, PLANE41.cpp:153-300) with
 * the reference's query numbers (query.h): kind 0 scalarQuery (1 SP, 3 WU, 4 WP, 5 VOL;
 * PLANE41: SP only) -> out[nE]; kind 1 vectorQuery (1 IC; SOLID81 only) -> out[nE * 3];
 * kind 2 tensorQuery (1 COUCHY, 2 PK2, 3 E, 4 C) -> out[nE * 6] in the order XX XY XZ YY YZ
 * ZZ.  gp: Gauss point (0..7 SOLID81, 0..3 PLANE41) or -1 for GP_MEAN.  The reference's
 * arithmetic is kept, quirks included (SOLID81 COUCHY/PK2 count p J C^-1 twice; PLANE41
 * COUCHY is F^T S F / J without the Gauss weight). */
int nla3d_gpu_element_query(nla3d_gpu_t h, int kind, int query, int gp, double* out);

/* ---- solution procedures (FESolver, FESolver.h:60-242) ------------------- */
/* Options of the Newton-Raphson loop (NonlinearFESolver, FESolver.h:205-218)
 * and of the Krylov solve that replaces EquationSolver::solveEquations. */
typedef struct {
  int n_load_steps;      /* numberOfLoadsteps; TimeControl delta = 1/N (FESolver.cpp:439) */
  int max_iterations;    /* numberOfIterations (FESolver.h:210) */
  int criterion;         /* 0: reference, mean|dU_s| / c < c (FESolver.cpp:474-483, 511-521)
                            1: relative residual, ||rhs_k|| <= c ||rhs_1|| in each load step */
  double convergence;    /* c: convergenceCriteria (reference default 1e-3) or the residual ratio */
  double krylov_tol;     /* MINRES: preconditioned residual <= krylov_tol * initial */
  int krylov_max_iter;
  int preconditioner;    /* 0: block Jacobi; 1: block Jacobi with the learned node-block corrections of the
                            network attached by nla3d_gpu_attach_mlp; 2: block Jacobi on the node DoFs and the
                            Schur-row stage (one Jacobi-Chebyshev sweep, or a V-cycle) on the sparse
                            Schur approximation S = B^T D^-1 B - C for the pressure and multiplier
                            rows; 3: 1 and 2 together
                            (S is formed from the plain block-Jacobi blocks).  A solve of 1, 2 or 3 that
                            breaks down or does not reach its tolerance is repeated with plain block Jacobi
                            (stats.precond_fallbacks) */
  /* inexact Newton (criterion 1 only; the single-model solvers).  krylov_tol_loose > 0: the first
     solve of a load step runs to max(krylov_tol, krylov_tol_loose) (its update leaves a residual
     set by the nonlinearity whatever its accuracy).  krylov_forcing > 0: every later solve also
     stops, at the end of a 12-iteration group, once its true linear residual ||rhs - K2 dU||_2 is
     at most krylov_forcing * convergence * ||rhs_1||, i.e. that fraction of the residual the
     criterion accepts (one more SpMV per group); the next Newton residual is then that plus
     Newton's quadratic term, in the norm the criterion measures.  0 (the defaults) leaves every
     solve at krylov_tol.  The converged state is set by the residual criterion either way. */
  double krylov_tol_loose;
  double krylov_forcing;
  /* 1: the first solve of a load step starts from a prediction instead of zero and solves for
     the correction to the same absolute accuracy in MINRES's norm (its tolerance is taken
     relative to sqrt(rhs^T M rhs), M the preconditioner, instead of the correction's right-hand
     side): x0 = sum_i a_i [p_i,u ; 0] + [0 ; U_l] over the increments p_i of
     the last three load steps (node and pressure rows; the multipliers from their current
     values), the a_i minimising ||rhs - K2 x0||_2.  2: also the second solve of a load step,
     the same way from the last three load steps' second Newton increments.  The history is the
     handle's: it restarts after nla3d_gpu_set_solution and when a load step does not follow the
     last one.  0 (the default): start from zero. */
  int first_solve_predictor;
} nla3d_gpu_newton_opts;

typedef struct {
  int load_steps;            /* load steps completed */
  int newton_iterations;     /* total equilibrium iterations (TimeControl::getTotalNumberOfEquilibriumSteps) */
  int last_step_iterations;  /* iterations of the last load step */
  int linear_solves;
  long long krylov_iterations;  /* summed over all linear solves */
  int last_krylov_iterations;   /* iterations of the last linear solve */
  int krylov_breakdown;      /* a MINRES solve hit an indefinite preconditioner */
  double last_criterion;     /* value of the convergence measure at exit */
  double last_time;          /* normalised time of the last completed step */
  double ms_assemble, ms_solve, ms_total;  /* wall time (host clock around device work) */
  int precond_fallbacks;     /* solves with preconditioner 1-3 repeated with block Jacobi, or run with it
                                because the multigrid was unavailable */
} nla3d_gpu_newton_stats;

/* Fills opts with the reference defaults (20 iterations, 10 load steps,
 * criterion 0 with 1e-3) and krylov_tol 1e-12, krylov_max_iter 20000. */
int nla3d_gpu_newton_defaults(nla3d_gpu_newton_opts* opts);
/* One load step of NonlinearFESolver::solve (FESolver.cpp:441-490): Newton
 * iterations at normalised time t until the criterion holds.  The model state
 * (vecU, previous constrained values) carries over between calls. */
int nla3d_gpu_solve_load_step(nla3d_gpu_t h, double t, const nla3d_gpu_newton_opts* opts,
                              nla3d_gpu_newton_stats* stats);
/* One equilibrium iteration (FESolver.cpp:443-490) at normalised time t:
 * assemble, boundary conditions, rhs, Krylov solve, update, criterion.
 * iteration is the 1-based count within the load step and *rhs1 carries the
 * first iteration's ||rhs|| between calls.  *converged = 1 when the criterion
 * holds (criterion 1: the state was already in equilibrium, nothing solved). */
int nla3d_gpu_newton_iteration(nla3d_gpu_t h, double t, int iteration, double* rhs1,
                               const nla3d_gpu_newton_opts* opts, nla3d_gpu_newton_stats* stats, int* converged);
/* NonlinearFESolver::solve: TimeControl over [0, 1] with n_load_steps steps,
 * each solved by nla3d_gpu_solve_load_step. */
int nla3d_gpu_solve_nonlinear(nla3d_gpu_t h, const nla3d_gpu_newton_opts* opts, nla3d_gpu_newton_stats* stats);
/* EquationSolver::solveEquations(K.block(2), rhs, x) (EquationSolver.h:22) with the
 * currently assembled K: MINRES + block Jacobi.  rhs and x are host arrays of
 * length nUnknown + nMpc.  iters / relres (preconditioned estimate) may be NULL. */
int nla3d_gpu_solve_block2(nla3d_gpu_t h, const double* rhs, double* x, double tol, int max_iter, int* iters,
                           double* relres);
/* No counterpart in the reference: nla3d_gpu_solve_block2 for callers that cannot allocate
 * device memory (code under stream capture, tight serving loops, harnesses that forbid it).
 * rhs and x are DEVICE arrays of length nUnknown + nMpc on the handle's device (x is
 * overwritten); the handle must be initialised (nla3d_gpu_init_solution) and assembled.  The
 * solve calls no hipMalloc / hipFree and instantiates no graph: its MINRES work space is a
 * static per-device array of NLA3D_GPU_NOALLOC_MAX_ROWS rows (56 MiB on every GPU the library
 * runs on, held until the process ends); larger systems return NLA3D_GPU_EINVAL.  Calls on one
 * GPU share that array: calls that overlap in time on one GPU race.  Synchronous (reads the
 * state back every 12 iterations). */
#define NLA3D_GPU_NOALLOC_MAX_ROWS (1 << 20)
int nla3d_gpu_solve_block2_noalloc(nla3d_gpu_t h, const double* rhs, double* x, double tol, int max_iter,
                                   int* iters, double* relres);
/* vecR (length nDofs + nMpc, FEStorage::getR): the constrained-DoF reactions of the last
 * Newton iteration, vecRc = K11 dUc + K12 dUsl - Fc (FESolver.cpp:464-467), and the
 * applied loads at the other equations; FEStorage::getReaction(node, dof) is
 * R[eq - 1] for the DoF's equation number. */
int nla3d_gpu_get_R(nla3d_gpu_t h, double* R);
/* The right-hand side of block(2) of the last Newton iteration (FESolver.cpp:454-457:
 * F_sl + R_sl - K12^T dU_c), host array of length nUnknown + nMpc; with the tangent
 * still assembled it is the system that iteration solved. */
int nla3d_gpu_get_rhs(nla3d_gpu_t h, double* rhs);

/* ---- equation solver on a caller's matrix ---------------------------------------
 * EquationSolver (math/EquationSolver.h:19-78) for K.block(2) as nla3d's FESolver hands it over
 * (FESolver.cpp:377, 457, 601, 613): n rows of a SparseSymMatrix as its raw 1-based upper CSR,
 * values[nnz], iofeir[n+1], columns[nnz] (getValuesArray / getIofeirArray / getColumnsArray,
 * SparseMatrix.h:114-116), host arrays read during the call.  factorize = PARDISO's phases 11
 * (first call, or a changed pattern: the pattern is analysed and device storage allocated) and
 * 22 (every call: values copied to the device, block-Jacobi blocks inverted, Schur-row multigrid
 * refreshed); substitute = phase 33 for one right-hand side (MINRES); solve = both
 * (EquationSolver.cpp:141-251).  Errors are return codes instead of the reference's CHECK:
 * NLA3D_GPU_EPATTERN for arrays that break SparsityInfo's invariants (1-based, sorted, upper),
 * NLA3D_GPU_ESTATE for substitute before factorize, NLA3D_GPU_ENOTCONV when MINRES does not
 * reach tol (the unknowns then hold its last iterate).  One HIP stream per solver handle. */
typedef struct nla3d_gpu_eqsolver* nla3d_gpu_eqsolver_t;
typedef struct {
  double tol;          /* MINRES: preconditioned residual <= tol * initial (default 1e-12) */
  int max_iter;        /* default 20000 */
  int preconditioner;  /* 0 (default): block Jacobi (3x3 node blocks, the diagonal Schur estimate on the
                          other rows); 2: plus the Schur-row stage (one Jacobi-Chebyshev sweep, or a V-cycle)
                          on the rows that are not node DoFs (a solve it fails is repeated with block
                          Jacobi) */
} nla3d_gpu_eqsolver_opts;
typedef struct {
  int iterations;         /* MINRES iterations of the last substitute (both runs after a fallback) */
  double relres;          /* its preconditioned residual estimate relative to the initial one */
  int precond_fallbacks;  /* 1 or 2 when preconditioner 2 was unavailable or failed */
  int breakdown;          /* the preconditioner was not positive definite */
  double ms_factorize, ms_substitute;  /* host wall time of the last calls */
  int spmv_dia_nodes;     /* nodes the SpMV reads through its node-diagonal copy (nla3d_gpu_spmv_info) */
} nla3d_gpu_eqsolver_stats;

void nla3d_gpu_eqsolver_defaults(nla3d_gpu_eqsolver_opts* opts);
/* EquationSolver's constructor: a solver on `device` (no host-only form) */
int nla3d_gpu_eqsolver_create(nla3d_gpu_eqsolver_t* solver, int device);
int nla3d_gpu_eqsolver_destroy(nla3d_gpu_eqsolver_t solver);
const char* nla3d_gpu_eqsolver_last_error(nla3d_gpu_eqsolver_t solver);
/* setSymmetric / setPositive have no counterpart: MINRES takes symmetric indefinite matrices */
int nla3d_gpu_eqsolver_set_options(nla3d_gpu_eqsolver_t solver, const nla3d_gpu_eqsolver_opts* opts);
/* Optional description of the rows for the preconditioner, FEStorage's numbering read backwards
 * (FEStorage::getNodeDofEqNumber / getElementDofEqNumber, FEStorage.h:148-149): row i of block(2)
 * (equation nC + 1 + i) is the DoF row_dof[i] (Dof::dofType, Dof.h:15-25) of node row_node[i]
 * (1-based), or with row_node[i] = 0 an element DoF (row_dof[i] = 6, HYDRO_PRESSURE) or an MPC
 * multiplier (row_dof[i] = -1).  n = 0 clears the map: rows are then classified from the matrix
 * (up to three consecutive rows with identical column lists are one node's DoFs; a row of its own
 * with a positive diagonal is a one-DoF node; the others are pressure/multiplier rows).  Takes
 * effect at the next factorize. */
int nla3d_gpu_eqsolver_set_dof_map(nla3d_gpu_eqsolver_t solver, uint32_t n, const uint32_t* row_node,
                                   const int* row_dof);
/* EquationSolver::factorizeEquations(matrix) (EquationSolver.h:23) */
int nla3d_gpu_eqsolver_factorize(nla3d_gpu_eqsolver_t solver, uint32_t n, const double* values, const uint32_t* iofeir,
                                 const uint32_t* columns);
/* EquationSolver::substituteEquations(matrix, rhs, unknowns) (EquationSolver.h:24) with the matrix of the
 * last factorize: rhs[n] -> unknowns[n] */
int nla3d_gpu_eqsolver_substitute(nla3d_gpu_eqsolver_t solver, const double* rhs, double* unknowns);
/* EquationSolver::solveEquations(matrix, rhs, unknowns) (EquationSolver.h:22): factorize + substitute */
int nla3d_gpu_eqsolver_solve(nla3d_gpu_eqsolver_t solver, uint32_t n, const double* values, const uint32_t* iofeir,
                             const uint32_t* columns, const double* rhs, double* unknowns);
int nla3d_gpu_eqsolver_get_stats(nla3d_gpu_eqsolver_t solver, nla3d_gpu_eqsolver_stats* stats);

/* ---- batched models ---------------------------------------------------------
 * Many independent SOLID81 models (no MPCs) on one device, each handle set up as
 * usual through nla3d_gpu_init_solution; the batch does not own them.  One
 * Newton iteration of the batch = one persistent element-kernel launch over all
 * models' elements and one gather launch over all their CSR targets, the
 * per-model boundary conditions / rhs / norms / block-Jacobi set-up replayed
 * from a hipGraph, then one persistent team-MINRES launch (each workgroup solves
 * whole systems taken from a queue) plus whole-device solves for the largest
 * systems.  Every model keeps its own Newton and Krylov convergence.  No
 * counterpart in the reference, which solves one FEStorage at a time
 * (FESolver.cpp:407-508). */
typedef struct nla3d_gpu_batch* nla3d_gpu_batch_t;
typedef struct {
  int elem_grid;       /* workgroups of the persistent element kernel (0: one per element block) */
  int gather_threads;  /* block size of the batched gather (multiple of 64, <= 1024) */
  int team_threads;    /* team-MINRES workgroup size: 256, 512 or 1024 */
  int team_lanes;      /* lanes per matrix row group in the team SpMV: 4, 8 or 16 */
  int team_grid;       /* teams launched (0: 1024); at most one per queued system */
  long long big_rows;  /* systems with more rows use the whole-device solver (always above 65,536 rows) */
} nla3d_gpu_batch_config;

void nla3d_gpu_batch_default_config(nla3d_gpu_batch_config* cfg);
int nla3d_gpu_batch_create(nla3d_gpu_batch_t* batch, const nla3d_gpu_t* models, int n);
int nla3d_gpu_batch_destroy(nla3d_gpu_batch_t batch);
const char* nla3d_gpu_batch_last_error(nla3d_gpu_batch_t batch);
int nla3d_gpu_batch_get_config(nla3d_gpu_batch_t batch, nla3d_gpu_batch_config* cfg);
int nla3d_gpu_batch_set_config(nla3d_gpu_batch_t batch, const nla3d_gpu_batch_config* cfg);
/* Times candidate launch shapes on the batch itself at load time t (the assembly and
 * the element/gather grids; team size and lanes, and the large-system split, on
 * probe_iterations MINRES steps of every system) and keeps the fastest.  preconditioner:
 * the value the Newton options will carry (2 or 3: the whole-device probes include the
 * Schur multigrid and its refresh).  report (may be NULL) receives up to report_cap
 * measured times or costs in ms, in the order tried.  Every model's solution vector U is
 * left as it was. */
int nla3d_gpu_batch_autotune(nla3d_gpu_batch_t batch, double t, int probe_iterations, int preconditioner,
                             nla3d_gpu_batch_config* chosen,
                             double* report, int report_cap);
/* assembly + rhs + preconditioner of every model at time t (the pre-solve graph); the
 * prescribed values at t are written into each model's U (as the first Newton iteration
 * of a load step does) */
int nla3d_gpu_batch_assemble(nla3d_gpu_batch_t batch, double t);
/* One Newton iteration (FESolver.cpp:443-490) of every model not yet converged in
 * the load step; stats and converged hold one entry per model (converged: 1 done,
 * 0 iterating, < 0 error code).  iteration 1 starts the load step.  Preconditioner:
 * the team launch runs block Jacobi; the models solved on the whole device (above
 * the configuration's big_rows) add the Schur-row multigrid when opts->preconditioner
 * is 2 or 3 (the learned corrections are not used in a batch). */
int nla3d_gpu_batch_newton_iteration(nla3d_gpu_batch_t batch, double t, int iteration,
                                     const nla3d_gpu_newton_opts* opts, nla3d_gpu_newton_stats* stats,
                                     int* converged);
int nla3d_gpu_batch_solve_load_step(nla3d_gpu_batch_t batch, double t, const nla3d_gpu_newton_opts* opts,
                                    nla3d_gpu_newton_stats* stats, int* converged);
/* TimeControl over [0, 1] with opts->n_load_steps steps for every model. */
int nla3d_gpu_batch_solve_nonlinear(nla3d_gpu_batch_t batch, const nla3d_gpu_newton_opts* opts,
                                    nla3d_gpu_newton_stats* stats, int* converged);
/* As nla3d_gpu_solve_block2 with a choice of preconditioner (the values of
 * nla3d_gpu_newton_opts.preconditioner; t is the load-step time fed to an attached
 * network).  No fallback here: this is the entry that compares them. */
int nla3d_gpu_solve_block2_pc(nla3d_gpu_t h, const double* rhs, double* x, double tol, int max_iter,
                              int preconditioner, double t, int* iters, double* relres);

/* ---- Schur-row stage of preconditioner 2 -----------------------------------
 * No counterpart in the reference.  The rows that block Jacobi treats by a scalar Schur
 * estimate (element pressures, multipliers) get z = B r with S = alpha B^T D^-1 B - C
 * (schur.h; alpha = 0.3 for the sweep, 1/2 for the V-cycle): by default one Jacobi-Chebyshev
 * sweep on S (8 terms up to 10^4 rows, 8 more per decade, at most 24), its result weighted
 * against the block-Jacobi rows by 1 at 8 terms rising to 2.5 at 24, or a V-cycle over an
 * aggregation hierarchy of S
 * (nla3d_gpu_schur_set_cycle; NLA3D_GPU_SCHUR_CYCLE=vcycle in the environment at set-up
 * makes the V-cycle the default).  The sweep runs inside MINRES's block-Jacobi pass
 * (NLA3D_GPU_SCHUR_FUSED=0 in the environment, read at each solve: after it, the same
 * arithmetic).  A level-0 matrix of at most 256 rows is inverted
 * exactly either way.  Inspection entries: sizes of the hierarchy (sizes[l] = rows of level l,
 * up to cap); S of level 0 for the current tangent (qrow[nq] = block(2) row of each Schur row;
 * CSR rowPtr[nq+1], col[nnz0], val[nnz0]); the aggregate of each row of a level and its
 * power-iteration estimate of lambda_max(D^-1 A); one application z = B r of the active
 * stage on host vectors of length nq (refresh = 1 recomputes its values from the current
 * tangent first). */
int nla3d_gpu_schur_info(nla3d_gpu_t h, int* available, int* nq, int* nnz0, int* levels, int* sizes, int cap);
int nla3d_gpu_schur_matrix(nla3d_gpu_t h, int* qrow, int* rowPtr, int* col, double* val);
/* the stage's shape: the finest level's Chebyshev length in the active stage (terms; 0 with one
 * level) and the first level of the V-cycle's coarse range run as one workgroup launch (-1: none) */
int nla3d_gpu_schur_cycle_info(nla3d_gpu_t h, int* finest_steps, int* fuse_from);
/* the active stage: vcycle = 0 the Chebyshev sweep on level 0, 1 the V-cycle; -3 when the model
 * has no Schur stage */
int nla3d_gpu_schur_set_cycle(nla3d_gpu_t h, int vcycle);
int nla3d_gpu_schur_get_cycle(nla3d_gpu_t h, int* vcycle);
int nla3d_gpu_schur_level(nla3d_gpu_t h, int level, int* agg, double* lmax);
int nla3d_gpu_schur_apply(nla3d_gpu_t h, const double* r, double* z, int refresh);
/* the precision of the finest level's smoother operator: 16 (the half-precision operator, levels
 * of at least 131,072 rows), 32 (an FP32 copy) or 64 (one level: its dense FP64 inverse);
 * NLA3D_GPU_SCHUR_SMOOTHER=half / fp32 in the environment at set-up forces 16 / 32 */
int nla3d_gpu_schur_finest_precision(nla3d_gpu_t h, int* bits);
/* the finest level's stored smoother operator for the current tangent: s[nq] = D^-1/2 and, in
 * S's CSR pattern (val[nnz0]), the scaled entries as the smoother reads them (FP16 off the
 * diagonal, 1 + Delta on it); -3 when the finest level smooths with its FP32 copy */
int nla3d_gpu_schur_smoother_operator(nla3d_gpu_t h, double* s, double* val);

/* ---- learned block-Jacobi correction --------------------------------------
 * No counterpart in the reference, whose only solver is PARDISO
 * (EquationSolver.cpp:158-249): a per-node MLP 24-128-128-128-128-6 (GELU)
 * maps features of a node's rows of the tangent to the lower-triangular
 * factor C (exp on its diagonal) of a correction of that node's block-Jacobi
 * block: P_i = L^-T C C^T L^-1 with K_ii = L L^T.  C = I (all outputs zero,
 * the state of a new network) is block Jacobi.  Inference runs in BF16 on
 * MFMA with FP32 accumulation; training runs FP32 with Adam on the loss
 * mean_j log(phi_j / phi_{j-1}) of k steps of the preconditioned MINRES
 * recurrence, differentiated in reverse through the recurrence. */
typedef struct nla3d_gpu_mlp* nla3d_gpu_mlp_t;
typedef struct {
  float lr, beta1, beta2, eps;
} nla3d_gpu_adam_opts;

int nla3d_gpu_mlp_num_params(void); /* 53,510 FP32 values: W1 b1 W2 b2 W3 b3 W4 b4 W5 b5, W row-major [out][in] */
/* A new network on `device`: Glorot-uniform hidden weights from splitmix64(seed), zero biases and
 * a zero output layer. */
int nla3d_gpu_mlp_create(nla3d_gpu_mlp_t* mlp, int device, unsigned long long seed);
int nla3d_gpu_mlp_destroy(nla3d_gpu_mlp_t mlp);
const char* nla3d_gpu_mlp_last_error(nla3d_gpu_mlp_t mlp);
int nla3d_gpu_mlp_get_params(nla3d_gpu_mlp_t mlp, float* w, int n);
int nla3d_gpu_mlp_set_params(nla3d_gpu_mlp_t mlp, const float* w, int n);
/* input normalisation x' = (x - mean) * inv_std, 24 values each (default 0 and 1) */
int nla3d_gpu_mlp_set_normalization(nla3d_gpu_mlp_t mlp, const float* mean, const float* inv_std);
int nla3d_gpu_mlp_get_normalization(nla3d_gpu_mlp_t mlp, float* mean, float* inv_std);
/* Y[n x 6] = MLP(X[n x 24]) on host arrays; bf16_mfma = 1: the BF16 MFMA inference kernel,
 * 0: the FP32 training forward */
int nla3d_gpu_mlp_forward(nla3d_gpu_mlp_t mlp, const float* X, int n, float* Y, int bf16_mfma);
/* dW = d(sum(dY . Y))/dparams through the FP32 forward of X (host arrays; dW holds 53,510) */
int nla3d_gpu_mlp_backward(nla3d_gpu_mlp_t mlp, const float* X, const float* dY, int n, float* dW);
/* Attach (NULL: detach) the network used by preconditioner 1; not owned by the model. */
int nla3d_gpu_attach_mlp(nla3d_gpu_t h, nla3d_gpu_mlp_t mlp);
/* Features of every learned node (free UX UY UZ triple) of the current tangent:
 * X[n x 24] and the node's first block(2) row (0-based).  *n_out = node count;
 * with cap < that count only the count is returned. */
int nla3d_gpu_node_features(nla3d_gpu_t h, double t, float* X, int* node_rows, int cap, int* n_out);
/* The corrected 3x3 blocks (row-major, 9 per node) for outputs Y[n x 6] given by the caller. */
int nla3d_gpu_learned_blocks(nla3d_gpu_t h, const float* Y, int n, double* blocks);
/* Loss of k unrolled MINRES steps for rhs (host, length nUnknown + nMpc; NULL: the last
 * Newton right-hand side) with the current tangent, and its gradient (53,510, may be NULL). */
int nla3d_gpu_mlp_loss_grad(nla3d_gpu_mlp_t mlp, nla3d_gpu_t h, const double* rhs, int k, double t, int multigrid,
                            double* loss, float* grad);
void nla3d_gpu_adam_defaults(nla3d_gpu_adam_opts* opts); /* lr 1e-3, 0.9, 0.999, 1e-8 */
/* One Adam step on that gradient.  multigrid = 1 (both entries): the unrolled MINRES uses
 * preconditioner 3's operator, the Schur-row multigrid (built from the plain block-Jacobi
 * blocks) beside the corrected node blocks; 0: preconditioner 1's. */
int nla3d_gpu_mlp_train_step(nla3d_gpu_mlp_t mlp, nla3d_gpu_t h, const double* rhs, int k, double t, int multigrid,
                             const nla3d_gpu_adam_opts* opts, double* loss);

/* ---- thermal path (QUADTH / SurfaceLINETH) -------------------------------- */
/* Options of LinearTransientFESolver (FESolver.h:223-242) and of the CG solve. */
typedef struct {
  double alpha, delta;      /* Newmark parameters (0.25, 0.5: trapezoidal rule) */
  double time0, time1;      /* time interval */
  int n_timesteps;          /* numberOfTimesteps; the loop runs while curTime <= time1 (FESolver.cpp:604) */
  double init_value;        /* initValue: every DoF of vecU at t = time0 (FESolver.cpp:577-579) */
  int dirichlet_coupling;   /* 0: as the reference, the rhs has no -K12^T Uc term (FESolver.cpp:606-608);
                               1: the coupling of prescribed temperatures is included */
  double krylov_tol;        /* CG: preconditioned residual <= tol * ||b||_M */
  int krylov_max_iter;
  int predictor;            /* 1 (the default): each step's CG starts from the current unknowns
                               plus the least-squares combination of the last three steps'
                               increments (minimising ||b - A x0||_2); 0: from the current unknowns.
                               The tolerance is relative to ||b||_M either way, so the accuracy is
                               the same */
} nla3d_gpu_transient_opts;

/* FEStorage::assembleGlobalEqMatrices with transient matrices for QUADTH and
 * SurfaceLINETH (QUADTH.cpp:22-125): K, C and F on the device. */
int nla3d_gpu_assemble_thermal(nla3d_gpu_t h);
/* C block values (block = 1, 12, 2), as nla3d_gpu_get_K */
int nla3d_gpu_get_C(nla3d_gpu_t h, int block, double* values);
/* LinearFESolver::solve (FESolver.cpp:337-397) with CG + Jacobi on K.block(2). */
int nla3d_gpu_solve_linear(nla3d_gpu_t h, double krylov_tol, int krylov_max_iter, nla3d_gpu_newton_stats* stats);
/* reference defaults: alpha 0.25, delta 0.5, [0, 1], 100 steps, init 0, no coupling; tol 1e-12 */
int nla3d_gpu_transient_defaults(nla3d_gpu_transient_opts* opts);
/* LinearTransientFESolver::solve, split for callers that read results per step:
 * begin (initial field, assembly, BCs, Kmod = a1 C + K), then one call per time
 * step (*done = 1 once curTime > time1), or the whole loop at once. */
int nla3d_gpu_transient_begin(nla3d_gpu_t h, const nla3d_gpu_transient_opts* opts);
int nla3d_gpu_transient_step(nla3d_gpu_t h, nla3d_gpu_newton_stats* stats, int* done);
int nla3d_gpu_solve_transient(nla3d_gpu_t h, const nla3d_gpu_transient_opts* opts, nla3d_gpu_newton_stats* stats);

/* ---- timing -------------------------------------------------------------- */
/* Runs `iters` device assemblies on the resident solution and returns the mean
 * duration in ms of the element kernel and of the scatter kernel, measured with
 * HIP events on the handle's stream. */
int nla3d_gpu_time_assembly(nla3d_gpu_t h, int iters, double* ms_element, double* ms_scatter);
/* Mean duration in ms of the block(2) SpMV kernel of the Krylov solve (full-symmetric
 * CSR, y = K2 x) over `iters` launches, HIP events on the handle's stream. */
int nla3d_gpu_time_spmv(nla3d_gpu_t h, int iters, double* ms);
/* No counterpart in the reference (a check of the Krylov solve's operator): y = K.block(2) x
 * with the currently assembled K through the solve's own SpMV kernel; x and y are host arrays
 * of length nUnknown + nMpc. */
int nla3d_gpu_apply_block2(nla3d_gpu_t h, const double* x, double* y);
/* How the SpMV reads block(2): dia_nodes nodes through the symmetric node-diagonal copy with
 * dia_offsets upper node offsets each (0, 0 when there is no copy: small or unstructured
 * models, or NLA3D_GPU_NODE_DIA=0 in the environment at nla3d_gpu_init_solution), csr_groups
 * groups of rows through the grouped full-symmetric CSR.  Any pointer may be NULL. */
int nla3d_gpu_spmv_info(nla3d_gpu_t h, int* dia_nodes, int* dia_offsets, int* csr_groups);
/* Mean duration in ms of one finest-level step of the Schur multigrid's Chebyshev
 * smoother (cheb_step_kernel: res -= A d, d' = c1 d + c2 D^-1 res over the FP32
 * sliced-ELLPACK copy; x += d + d' on odd steps and nothing on the even steps after them,
 * in a sweep's order) over `iters` launches, HIP events on the handle's stream; bytes (may
 * be NULL) receives the mean bytes one launch must move (stored matrix + 5 vectors + x
 * read and written by the steps that add to it). */
int nla3d_gpu_time_smoother(nla3d_gpu_t h, int iters, double* ms, long long* bytes);
/* One launch of a plain device copy of `bytes` bytes (read + write), timed with
 * HIP events: the reference point for scaling hardware byte counters. */
int nla3d_gpu_copy_calibration(nla3d_gpu_t h, uint64_t bytes, double* ms);

#ifdef __cplusplus
}
#endif
#endif
