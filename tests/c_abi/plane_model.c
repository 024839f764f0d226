/* A caller of libnla3d_gpu.so written in plain C against include/nla3d_gpu.h only: a PLANE41
 * plate of nx x ny unit-square quadrilaterals (node i + (nx+1) j at (i/nx, j/ny), counter-
 * clockwise connectivity), bottom edge fixed, top edge moved by UY = -0.1 with UX = 0,
 * Neo-Hookean G = 1, K = 500.
 *
 *   plane_model <device> <nx> <ny>
 *
 * device -1 (host-only handle): prints the sizes and every node's UX / UY equation numbers.
 * device >= 0: also runs two load steps (t = 0.5, 1) with the residual criterion 1e-10 and
 * prints the solution vector.  Exit status 0 on success, the library's code otherwise. */
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>

#include "nla3d_gpu.h"

#define CHECK(call)                                                                   \
  do {                                                                                \
    int rc_ = (call);                                                                 \
    if (rc_ != NLA3D_GPU_OK) {                                                        \
      fprintf(stderr, "%s failed (%d): %s\n", #call, rc_, nla3d_gpu_last_error(h));   \
      return rc_ < 0 ? -rc_ : rc_;                                                    \
    }                                                                                 \
  } while (0)

int main(int argc, char** argv) {
  if (argc != 4) {
    fprintf(stderr, "usage: %s device nx ny\n", argv[0]);
    return 2;
  }
  const int device = atoi(argv[1]), nx = atoi(argv[2]), ny = atoi(argv[3]);
  const uint32_t nn = (uint32_t)((nx + 1) * (ny + 1)), ne = (uint32_t)(nx * ny);
  nla3d_gpu_t h = NULL;
  int rc = nla3d_gpu_create(&h, device);
  if (rc != NLA3D_GPU_OK) {
    fprintf(stderr, "nla3d_gpu_create failed (%d)\n", rc);
    return 3;
  }
  double* xyz = malloc(sizeof(double) * 3 * nn);
  uint32_t* conn = malloc(sizeof(uint32_t) * 4 * ne);
  for (int j = 0; j <= ny; j++)
    for (int i = 0; i <= nx; i++) {
      const int n = i + (nx + 1) * j;
      xyz[3 * n + 0] = (double)i / nx;
      xyz[3 * n + 1] = (double)j / ny;
      xyz[3 * n + 2] = 0.0;
    }
  for (int j = 0; j < ny; j++)
    for (int i = 0; i < nx; i++) {
      const int e = i + nx * j;
      const uint32_t a = (uint32_t)(1 + i + (nx + 1) * j);
      conn[4 * e + 0] = a;
      conn[4 * e + 1] = a + 1;
      conn[4 * e + 2] = a + 1 + (uint32_t)(nx + 1);
      conn[4 * e + 3] = a + (uint32_t)(nx + 1);
    }
  CHECK(nla3d_gpu_set_nodes(h, nn, xyz));
  CHECK(nla3d_gpu_add_elements(h, NLA3D_GPU_PLANE41, ne, conn, NULL));
  const double mc[2] = {1.0, 500.0};
  CHECK(nla3d_gpu_set_material(h, 1, mc, 2));
  for (int i = 0; i <= nx; i++) {
    CHECK(nla3d_gpu_add_fix(h, (uint32_t)(1 + i), 0, 0.0));
    CHECK(nla3d_gpu_add_fix(h, (uint32_t)(1 + i), 1, 0.0));
    CHECK(nla3d_gpu_add_fix(h, (uint32_t)(1 + i + (nx + 1) * ny), 0, 0.0));
    CHECK(nla3d_gpu_add_fix(h, (uint32_t)(1 + i + (nx + 1) * ny), 1, -0.1));
  }
  CHECK(nla3d_gpu_init_solution(h));
  uint64_t sizes[9];
  CHECK(nla3d_gpu_get_sizes(h, sizes));
  printf("sizes");
  for (int k = 0; k < 9; k++) printf(" %llu", (unsigned long long)sizes[k]);
  printf("\n");
  uint32_t* neq = malloc(sizeof(uint32_t) * 8 * nn);
  uint32_t* eeq = malloc(sizeof(uint32_t) * 8 * ne);
  CHECK(nla3d_gpu_get_equation_numbers(h, neq, eeq));
  for (uint32_t n = 0; n < nn; n++) printf("eq %u %u %u\n", n + 1, neq[8 * n + 0], neq[8 * n + 1]);
  if (device >= 0) {
    nla3d_gpu_newton_opts o;
    nla3d_gpu_newton_stats st;
    CHECK(nla3d_gpu_newton_defaults(&o));
    o.criterion = 1;
    o.convergence = 1e-10;
    o.krylov_tol = 1e-13;
    o.max_iterations = 20;
    o.preconditioner = 2;
    CHECK(nla3d_gpu_solve_load_step(h, 0.5, &o, &st));
    CHECK(nla3d_gpu_solve_load_step(h, 1.0, &o, &st));
    const uint64_t nall = sizes[2] + sizes[5];
    double* U = malloc(sizeof(double) * nall);
    CHECK(nla3d_gpu_get_solution(h, U));
    for (uint64_t k = 0; k < nall; k++) printf("U %llu %.17g\n", (unsigned long long)k, U[k]);
    printf("newton %d krylov %lld fallbacks %d\n", st.newton_iterations, st.krylov_iterations, st.precond_fallbacks);
    free(U);
  }
  free(neq);
  free(eeq);
  free(xyz);
  free(conn);
  nla3d_gpu_destroy(h);
  return 0;
}
