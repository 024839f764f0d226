// Batched solve of many independent models (batch.hip, batch.cpp).
#pragma once

#include <hip/hip_runtime.h>

#include "krylov.h"

namespace nla3d_amd {

// one block(2) system of a team solve: matrix, preconditioner, rhs, solution and
// eight work vectors of length M.n (all device pointers of that model)
struct TeamSolve {
  KrylovMatrix M;
  Precond P;
  const double* b;
  double* x;
  double* v[8];
  double tol;
  int maxit;
  int* out;        // [iterations, flag: 0 ok, 1 zero rhs, 2 indefinite preconditioner]
  double* relres;  // preconditioned residual estimate / initial
};

bool team_config_valid(int threads, int lanes);
// persistent launch: `grid` teams of `threads` threads pull queue[0..nQ) (indices into
// solves) until the queue is empty; counter is reset by the launch
void launch_minres_team(hipStream_t s, int threads, int lanes, int grid, int nQ, const TeamSolve* solves,
                        const int* queue, int* counter);

}  // namespace nla3d_amd
