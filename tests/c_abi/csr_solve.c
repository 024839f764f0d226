/* A caller of the solver-level entry points written in plain C against include/nla3d_gpu.h only:
 * what a GpuKrylovEquationSolver : EquationSolver would do with the SparseSymMatrix nla3d's
 * FESolver hands it (EquationSolver.h:21-23).
 *
 *   csr_solve <device> <matrix file> <preconditioner 0|2> <use DoF map 0|1>
 *
 * The file holds, whitespace-separated: n nnz, iofeir[n+1], columns[nnz] (1-based upper CSR),
 * values[nnz], then two right-hand sides rhs1[n] rhs2[n], then (optional) the DoF map as n pairs
 * (node, dof).  The program factorizes once, substitutes both right-hand sides, then calls
 * solve (factorize + substitute) on the first one again with every value doubled, and prints
 * "x<k> i value" lines for the three solutions and "iterations <k> <it>".  Exit status 0 on
 * success, the library's code (negated) otherwise. */
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>

#include "nla3d_gpu.h"

#define CHECK(call)                                                                          \
  do {                                                                                       \
    int rc_ = (call);                                                                        \
    if (rc_ != NLA3D_GPU_OK) {                                                               \
      fprintf(stderr, "%s failed (%d): %s\n", #call, rc_, nla3d_gpu_eqsolver_last_error(s)); \
      return rc_ < 0 ? -rc_ : rc_;                                                           \
    }                                                                                        \
  } while (0)

static void print_x(int k, uint32_t n, const double* x) {
  for (uint32_t i = 0; i < n; i++) printf("x%d %u %.17g\n", k, i, x[i]);
}

int main(int argc, char** argv) {
  if (argc != 5) {
    fprintf(stderr, "usage: %s device file preconditioner use_map\n", argv[0]);
    return 2;
  }
  const int device = atoi(argv[1]), pc = atoi(argv[3]), use_map = atoi(argv[4]);
  FILE* f = fopen(argv[2], "r");
  if (!f) {
    fprintf(stderr, "cannot open %s\n", argv[2]);
    return 2;
  }
  unsigned n = 0, nnz = 0;
  if (fscanf(f, "%u %u", &n, &nnz) != 2) return 2;
  uint32_t* iofeir = malloc(sizeof(uint32_t) * (n + 1));
  uint32_t* columns = malloc(sizeof(uint32_t) * nnz);
  double* values = malloc(sizeof(double) * nnz);
  double* rhs = malloc(sizeof(double) * 2 * n);
  double* x = malloc(sizeof(double) * n);
  uint32_t* row_node = malloc(sizeof(uint32_t) * n);
  int* row_dof = malloc(sizeof(int) * n);
  for (unsigned i = 0; i <= n; i++)
    if (fscanf(f, "%u", &iofeir[i]) != 1) return 2;
  for (unsigned k = 0; k < nnz; k++)
    if (fscanf(f, "%u", &columns[k]) != 1) return 2;
  for (unsigned k = 0; k < nnz; k++)
    if (fscanf(f, "%lf", &values[k]) != 1) return 2;
  for (unsigned i = 0; i < 2 * n; i++)
    if (fscanf(f, "%lf", &rhs[i]) != 1) return 2;
  int have_map = 1;
  for (unsigned i = 0; i < n && have_map; i++)
    if (fscanf(f, "%u %d", &row_node[i], &row_dof[i]) != 2) have_map = 0;
  fclose(f);

  nla3d_gpu_eqsolver_t s = NULL;
  int rc = nla3d_gpu_eqsolver_create(&s, device);
  if (rc != NLA3D_GPU_OK) {
    fprintf(stderr, "nla3d_gpu_eqsolver_create failed (%d)\n", rc);
    return 3;
  }
  nla3d_gpu_eqsolver_opts opts;
  nla3d_gpu_eqsolver_defaults(&opts);
  opts.tol = 1e-13;
  opts.max_iter = 50000;
  opts.preconditioner = pc;
  CHECK(nla3d_gpu_eqsolver_set_options(s, &opts));
  if (use_map && have_map) CHECK(nla3d_gpu_eqsolver_set_dof_map(s, n, row_node, row_dof));
  nla3d_gpu_eqsolver_stats st;
  /* factorizeEquations once, substituteEquations twice (FESolver.cpp:601, 613) */
  CHECK(nla3d_gpu_eqsolver_factorize(s, n, values, iofeir, columns));
  for (int k = 0; k < 2; k++) {
    CHECK(nla3d_gpu_eqsolver_substitute(s, rhs + (size_t)k * n, x));
    CHECK(nla3d_gpu_eqsolver_get_stats(s, &st));
    print_x(k, n, x);
    printf("iterations %d %d %d\n", k, st.iterations, st.precond_fallbacks);
  }
  /* solveEquations with new values on the same pattern (FESolver.cpp:457) */
  for (unsigned k = 0; k < nnz; k++) values[k] *= 2.0;
  CHECK(nla3d_gpu_eqsolver_solve(s, n, values, iofeir, columns, rhs, x));
  CHECK(nla3d_gpu_eqsolver_get_stats(s, &st));
  print_x(2, n, x);
  printf("iterations %d %d %d\n", 2, st.iterations, st.precond_fallbacks);
  CHECK(nla3d_gpu_eqsolver_destroy(s));
  free(iofeir);
  free(columns);
  free(values);
  free(rhs);
  free(x);
  free(row_node);
  free(row_dof);
  return 0;
}
