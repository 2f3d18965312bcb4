/*
 * oracle_batch.c -- TEST INFRASTRUCTURE ONLY (CPU baseline leg).
 * One replica per OpenMP thread, as the reference's main does
 * (src/main.cpp:21-30); each replica's loop time is recorded.
 */
#include <omp.h>
#include "mdn_oracle.h"

int orc_verlet_batch(orc_ctx **ctx, orc_state **st, int n, int nsteps, int eflag_every, int64_t step0,
                     int64_t *atom_steps, int nthreads, double *seconds) {
  int rc = 0;
  if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 1) reduction(| : rc)
  for (int r = 0; r < n; r++) {
    int64_t a = 0;
    double t0 = omp_get_wtime();
    rc |= orc_verlet(ctx[r], st[r], nsteps, eflag_every, step0, &a) != 0;
    if (seconds) seconds[r] = omp_get_wtime() - t0;
    atom_steps[r] = a;
  }
  return rc ? -1 : 0;
}
