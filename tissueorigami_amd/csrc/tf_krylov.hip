// Entry of the linear solves of the NR step: the replacement of the Pardiso
// call of NewtonRaphsonSolver::solveWithPardiso (NewtonRaphsonSolver.cpp:498-763),
// and the harvesting of the per-launch SpMV timings.  The solvers themselves
// (tile-based BiCGStab and CG over SELL-64 3x3 blocks, one or many tissues)
// are in tf_tiled.hip; a bound system arrives here as its explicit operator
// N^T J N + I_slave, so every solve takes the same path.
#include "tf_internal.h"

namespace tf {

void prof_harvest(Prof* p) {
  if (!p || !p->on) return;
  for (int i = 0; i < p->used; ++i) {
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, p->ev[2 * i], p->ev[2 * i + 1]) == hipSuccess) {
      p->ms[p->kind[i]] += ms;
      p->count[p->kind[i]]++;
    }
  }
  p->used = 0;
}

hipError_t krylov_bicgstab(const KrylovArgs& k, const double* b, double* x, double rtol,
                           int maxit, int* iters, int* state, int sync_every) {
  static thread_local TiledResult res;
  hipError_t e = tiled_bicgstab(k, b, x, rtol, maxit, sync_every, &res);
  *iters = res.iters[0];
  *state = res.state[0];
  return e;
}

}  // namespace tf
