/*
 * lepic_baseline_omp.c -- CPU baseline timing over the oracle's restatement
 * (test and measurement infrastructure; never the product).
 *
 * The reference runs its particle phases inside one OpenMP region per step
 * (main.f90:1024-1089): the particles of every species are split into nproc
 * partitions, and thread iproc calls part_mover and charge_deposition on its
 * own partition, depositing into its own density array np(:,:,ptype,iproc)
 * (summed later by calc_rhs, calc_rho.f90:56-69).  collisions opens its own
 * region and runs collision_OMP on each thread's share of the candidates
 * (collisions.f90:160-181).  These functions time the restated routines
 * (lo_part_mover, lo_charge_deposition, lo_collisions) split the same way.
 */
#include <omp.h>
#include <stdlib.h>
#include <string.h>

#include "lepic_oracle.h"

/* Seconds of `reps` steps of push + deposit of one species, nthreads
 * partitions.  vxp (np particles) is copied first and left unchanged. */
double lo_time_push_deposit_omp(const lo_domain *d, const lo_species *sp, const double *Ei,
                                const double *kq, const double *vxp, long long np, int nthreads,
                                int reps) {
  const size_t nodes = (size_t)(d->n1 + 3) * (size_t)(d->n2 + 3);
  double *work = (double *)malloc(sizeof(double) * 6 * (size_t)(np > 0 ? np : 1));
  int8_t *fd = (int8_t *)calloc((size_t)(np > 0 ? np : 1), 1);
  double *rho = (double *)malloc(sizeof(double) * nodes * (size_t)nthreads);
  double *sink = (double *)malloc(sizeof(double) * nodes * (size_t)nthreads);
  double *pmac = (double *)malloc(sizeof(double) * 2 * (size_t)(d->ngrid + 1) * (size_t)nthreads);
  long long *cnt = (long long *)malloc(sizeof(long long) * (size_t)nthreads);
  double t0 = 0.0, t1 = 0.0;
#pragma omp parallel num_threads(nthreads)
  {
    const int t = omp_get_thread_num();
    const long long a = np * t / nthreads, b = np * (t + 1) / nthreads;
    memcpy(work + 6 * a, vxp + 6 * a, sizeof(double) * 6 * (size_t)(b - a));   /* first touch */
    memset(rho + nodes * t, 0, sizeof(double) * nodes);
    memset(sink + nodes * t, 0, sizeof(double) * nodes);
    cnt[t] = b - a;
#pragma omp barrier
#pragma omp master
    t0 = omp_get_wtime();
#pragma omp barrier
    for (int r = 0; r < reps; ++r) {
      lo_mover_tally tal;
      memset(&tal, 0, sizeof(tal));
      memset(pmac + 2 * (d->ngrid + 1) * t, 0, sizeof(double) * 2 * (size_t)(d->ngrid + 1));
      lo_part_mover(d, sp, Ei, work + 6 * a, fd + a, &cnt[t], &tal, pmac + 2 * (d->ngrid + 1) * t,
                    sink + nodes * t);
      memset(rho + nodes * t, 0, sizeof(double) * nodes);
      lo_charge_deposition(d, sp, work + 6 * a, cnt[t], kq, rho + nodes * t);
#pragma omp barrier
    }
#pragma omp master
    t1 = omp_get_wtime();
  }
  free(work); free(fd); free(rho); free(sink); free(pmac); free(cnt);
  return t1 - t0;
}

/* Seconds of one collisions call (all charged species) with the particles
 * split into nthreads partitions, each thread running lo_collisions on its
 * own partition with its own ran2 seed (123456 * thread, main.f90:529-531)
 * and tallies.  The partitions are copied first; p is left unchanged. */
double lo_time_collisions_omp(const lo_domain *d, const lo_mcc_setup *s, const lo_reaction_table *t,
                              const lo_particles *p, const double *const *np_red, int nthreads) {
  lo_particles *part = (lo_particles *)calloc((size_t)nthreads, sizeof(lo_particles));
  lo_mcc_out *out = (lo_mcc_out *)calloc((size_t)nthreads, sizeof(lo_mcc_out));
  const size_t ncc = 3 * (size_t)(t->ncol > 0 ? t->ncol : 1) * (size_t)(s->n_rg > 0 ? s->n_rg : 1);
  double t0 = 0.0, t1 = 0.0;
#pragma omp parallel num_threads(nthreads)
  {
    const int th = omp_get_thread_num();
    lo_particles *q = &part[th];
    for (int pt = 1; pt <= s->ntype; ++pt) {
      const long long n = p->np_tot[pt];
      const long long a = n * th / nthreads, b = n * (th + 1) / nthreads;
      const long long cap = (b - a) + (b - a) / 4 + 64;
      q->vxp[pt] = (double *)malloc(sizeof(double) * 6 * (size_t)cap);
      q->flag_dead[pt] = (int8_t *)calloc((size_t)cap, 1);
      memcpy(q->vxp[pt], p->vxp[pt] + 6 * a, sizeof(double) * 6 * (size_t)(b - a));
      q->np_tot[pt] = b - a;
      q->cap[pt] = cap;
    }
    out[th].cnt_col = (double *)calloc(ncc, sizeof(double));
    int32_t seed = 123456 * (th + 1);
#pragma omp barrier
#pragma omp master
    t0 = omp_get_wtime();
#pragma omp barrier
    lo_collisions(d, s, t, q, np_red, &seed, &out[th]);
#pragma omp barrier
#pragma omp master
    t1 = omp_get_wtime();
    for (int pt = 1; pt <= s->ntype; ++pt) { free(q->vxp[pt]); free(q->flag_dead[pt]); }
    free(out[th].cnt_col);
  }
  free(part);
  free(out);
  return t1 - t0;
}
