/*
 * lepic_oracle.c -- CPU FP64 restatement of the LePIC hot path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker, never the product).  See
 * lepic_oracle.h for the pinning status of each routine.  Build with strict
 * IEEE arithmetic: -O2 -ffp-contract=off, no -ffast-math.
 *
 * Layout conventions shared with the reference's Fortran arrays:
 *   node arrays A(0:n1+2,0:n2+2) -> A[iy*(n1+3)+ix]
 *   rhs(n1+1,n2+1) (1-based)     -> rhs[(iy-1)*(n1+1)+(ix-1)]
 *   Ei(2,0:n1+2,0:n2+2)          -> Ei[2*node + c]
 *   Bi(4,0:nB1+2,0:nB2+2)        -> Bi[4*bnode + c]
 *   vxp(6,nmax)                  -> vxp[6*i + c]   (i 0-based)
 */
#include "lepic_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define NODE(ix, iy, n1) ((long)(iy) * ((n1) + 3) + (ix))
#define RHS(ix, iy, n1) ((long)((iy)-1) * ((n1) + 1) + ((ix)-1))

/* ran2.f90:1-15 -- Park & Miller minimal standard generator via Schrage. */
double lo_ran2(int32_t *irand) {
  const int32_t ia = 16807, im = 2147483647, iq = 127773, ir = 2836;
  const double am = 1.0 / (double)im;
  int32_t k = *irand / iq;
  *irand = ia * (*irand - k * iq) - ir * k;
  if (*irand < 0) *irand += im;
  return am * (double)(*irand);
}

/* load_gauss.f90:1-24 */
void lo_load_gauss(double *vx, double *vy, double vt, const double rnd[2], double pi) {
  double vp = vt * sqrt(-log(1.0 - rnd[0]));
  double theta = 2.0 * pi * rnd[1];
  *vx = vp * cos(theta);
  *vy = vp * sin(theta);
}

/* part_expmover.f90:612-627 */
static void check_outofbounds(int *ix, int *iy, int n1, int n2) {
  if (*ix < 0) *ix = 0;
  if (*ix > n1 + 1) *ix = n1 + 1;
  if (*iy < 0) *iy = 0;
  if (*iy > n2 + 1) *iy = n2 + 1;
}

/* part_expmover.f90:1-517, for the options of the hot path: 2D, no periodic
 * walls, no Neumann wall, no z losses (hz<0), no dielectric, no secondary
 * emission, no phase-space dump.  Those options return -1 from the callers'
 * validation (the domain struct cannot express them), so this routine
 * follows lines 69-258, 314-335, 397-446 (flag_lost==1 branch) and 495-513. */
int lo_part_mover(const lo_domain *d, const lo_species *sp, const double *Ei,
                  double *vxp, int8_t *flag_dead, long long *np_tot,
                  lo_mover_tally *t, double *p_mac, double *ss2D_sink) {
  const int n1 = d->n1, n2 = d->n2;
  const double h1 = d->h1, h2 = d->h2;
  const double dt = d->dt;
  const int ng1 = d->ngrid + 1;
  long long np_lost = 0;
  /* Boris coefficient, part_expmover.f90:61 */
  const double k1 = dt * sp->charge / (2.0 * sp->mass);
  const long long N = *np_tot;
  const int nB1 = d->nB1;
  const long nbs = d->nB1 + 3;

  for (long long i = 0; i < N; ++i) {
    double xp = vxp[6 * i + 0], yp = vxp[6 * i + 1], zp = vxp[6 * i + 2];
    double vx = vxp[6 * i + 3], vy = vxp[6 * i + 4], vz = vxp[6 * i + 5];
    double Eki;
    if (flag_dead[i] == 1) { /* lines 84-98 */
      np_lost += 1;
      flag_dead[i] = 0;
      if (sp->tag_neg) t->cnt_dead = t->cnt_dead + sp->Nm * sp->charge;
      Eki = 0.5 * sp->Nm * sp->mass * (vx * vx + vy * vy + vz * vz);
      t->P_loss[2] = t->P_loss[2] - Eki;
      continue;
    }
    int ix = (int)(xp / h1) + 1; /* INT(): truncation, lines 103-104 */
    int iy = (int)(yp / h2) + 1;
    double px = (ix * h1 - xp) / h1;
    double py = (iy * h2 - yp) / h2;
    double ki[4];
    ki[0] = px * py;
    ki[1] = (1.0 - px) * py;
    ki[2] = (1.0 - px) * (1.0 - py);
    ki[3] = px * (1.0 - py);
    const double *E00 = Ei + 2 * NODE(ix, iy, n1), *E10 = Ei + 2 * NODE(ix + 1, iy, n1);
    const double *E11 = Ei + 2 * NODE(ix + 1, iy + 1, n1), *E01 = Ei + 2 * NODE(ix, iy + 1, n1);
    double Exp = ki[0] * E00[0] + ki[1] * E10[0] + ki[2] * E11[0] + ki[3] * E01[0];
    double Eyp = ki[0] * E00[1] + ki[1] * E10[1] + ki[2] * E11[1] + ki[3] * E01[1];

    int flag_B_tmp = 0;
    double Bpx = 0, Bpy = 0, Bpz = 0;
    if (d->flag_B == 1) { /* lines 131-182 */
      if (nB1 == 1) {
        const double *B = d->Bi + 4 * (1 * nbs + 1);
        Bpx = B[0]; Bpy = B[1]; Bpz = B[2];
        flag_B_tmp = 1;
      } else {
        int bx = ix, by = iy;
        if (d->flag_gridB == 1) {
          bx = (int)(xp / d->hB1) + 1;
          by = (int)(yp / d->hB2) + 1;
        }
        if (d->Bi[4 * (by * nbs + bx) + 3] > 1.0e-5) {
          if (d->flag_gridB == 1) {
            px = (bx * d->hB1 - xp) / d->hB1;
            py = (by * d->hB2 - yp) / d->hB2;
            ki[0] = px * py;
            ki[1] = (1.0 - px) * py;
            ki[2] = (1.0 - px) * (1.0 - py);
            ki[3] = px * (1.0 - py);
          }
          const double *B00 = d->Bi + 4 * (by * nbs + bx);
          const double *B10 = d->Bi + 4 * (by * nbs + bx + 1);
          const double *B11 = d->Bi + 4 * ((by + 1) * nbs + bx + 1);
          const double *B01 = d->Bi + 4 * ((by + 1) * nbs + bx);
          Bpx = ki[0] * B00[0] + ki[1] * B10[0] + ki[2] * B11[0] + ki[3] * B01[0];
          Bpy = ki[0] * B00[1] + ki[1] * B10[1] + ki[2] * B11[1] + ki[3] * B01[1];
          Bpz = ki[0] * B00[2] + ki[1] * B10[2] + ki[2] * B11[2] + ki[3] * B01[2];
          flag_B_tmp = 1;
        }
        /* the map branch reuses ix,iy for the new-cell step below only
         * after they are recomputed (line 244), so nothing leaks. */
      }
    }
    if (d->flag_B_pos == 1 && sp->charge > 0.0) flag_B_tmp = 0; /* line 185 */

    if (flag_B_tmp == 1) { /* Boris, lines 187-206 */
      double B_tot = sqrt(Bpx * Bpx + Bpy * Bpy + Bpz * Bpz);
      double k2 = k1 * (2.0 / (1.0 + (k1 * B_tot) * (k1 * B_tot)));
      double v1x = vx + k1 * Exp;
      double v1y = vy + k1 * Eyp;
      double v1z = vz;
      double v3x = v1x + k1 * (v1y * Bpz - v1z * Bpy);
      double v3y = v1y + k1 * (v1z * Bpx - v1x * Bpz);
      double v3z = v1z + k1 * (v1x * Bpy - v1y * Bpx);
      double v2x = v1x + k2 * (v3y * Bpz - v3z * Bpy);
      double v2y = v1y + k2 * (v3z * Bpx - v3x * Bpz);
      double v2z = v1z + k2 * (v3x * Bpy - v3y * Bpx);
      vx = v2x + k1 * Exp;
      vy = v2y + k1 * Eyp;
      vz = v2z;
    } else { /* leap-frog, lines 210-211 */
      vx = vx + 2.0 * k1 * Exp;
      vy = vy + 2.0 * k1 * Eyp;
    }
    xp = xp + dt * vx; /* lines 216-219 */
    yp = yp + dt * vy;
    zp = zp + dt * vz;

    ix = (int)floor(xp / h1) + 1; /* FLOOR(), lines 244-245 */
    iy = (int)floor(yp / h2) + 1;
    check_outofbounds(&ix, &iy, n1, n2);
    const int32_t *bc = d->bcnd;
    int flag_lost = 0;
    if (bc[NODE(ix, iy, n1)] >= 1 && bc[NODE(ix + 1, iy, n1)] >= 1 &&
        bc[NODE(ix + 1, iy + 1, n1)] >= 1 && bc[NODE(ix, iy + 1, n1)] >= 1)
      flag_lost = 1; /* lines 254-258 */

    if (flag_lost >= 1) { /* lines 314-446 */
      np_lost += 1;
      Eki = 0.5 * sp->Nm * sp->mass * (vx * vx + vy * vy + vz * vz);
      t->P_loss[0] = t->P_loss[0] + Eki;
      if (ss2D_sink) { /* plt_src == 0 sink term, lines 327-332 */
        int ixw = (int)lround(xp / h1) + 1;
        int iyw = (int)lround(yp / h2) + 1;
        check_outofbounds(&ixw, &iyw, n1, n2);
        ss2D_sink[NODE(ixw, iyw, n1)] += 1.0;
      }
      int igrid = bc[NODE(ix, iy, n1)];
      if (p_mac) {
        p_mac[0 * ng1 + igrid] += 1.0;
        p_mac[1 * ng1 + igrid] += Eki;
      }
    } else { /* order-preserving compaction, lines 495-503 */
      long long is = i - np_lost;
      vxp[6 * is + 0] = xp; vxp[6 * is + 1] = yp; vxp[6 * is + 2] = zp;
      vxp[6 * is + 3] = vx; vxp[6 * is + 4] = vy; vxp[6 * is + 5] = vz;
    }
  }
  *np_tot = N - np_lost; /* line 513 */
  t->np_lost += np_lost;
  return 0;
}

/* main.f90:334-344 */
void lo_kq(const lo_domain *d, double *kq) {
  for (int iy = 0; iy <= d->n2 + 2; ++iy)
    for (int ix = 0; ix <= d->n1 + 2; ++ix)
      kq[NODE(ix, iy, d->n1)] = (d->bcnd[NODE(ix, iy, d->n1)] != -1) ? 2.0 : 1.0;
}

/* part_expmover.f90:520-610 (heating tallies off: Pabs(1)<=0) */
void lo_charge_deposition(const lo_domain *d, const lo_species *sp,
                          const double *vxp, long long np_tot, const double *kq,
                          double *np) {
  const int n1 = d->n1;
  const double h1 = d->h1, h2 = d->h2;
  for (long long i = 0; i < np_tot; ++i) {
    double xp = vxp[6 * i + 0], yp = vxp[6 * i + 1];
    int ix = (int)(xp / h1) + 1;
    int iy = (int)(yp / h2) + 1;
    double px = (ix * h1 - xp) / h1;
    double py = (iy * h2 - yp) / h2;
    double k1 = sp->Nm / (h1 * h2 * d->zmax);
    double ki0 = k1 * px * py;
    double ki1 = k1 * (1.0 - px) * py;
    double ki2 = k1 * (1.0 - px) * (1.0 - py);
    double ki3 = k1 * px * (1.0 - py);
    long a = NODE(ix, iy, n1), b = NODE(ix + 1, iy, n1);
    long c = NODE(ix + 1, iy + 1, n1), e = NODE(ix, iy + 1, n1);
    np[a] = np[a] + kq[a] * ki0;
    np[b] = np[b] + kq[b] * ki1;
    np[c] = np[c] + kq[c] * ki2;
    np[e] = np[e] + kq[e] * ki3;
  }
}

/* part_moments, utils.f90:1-136, one species and partition, no Neumann or
 * periodic walls (flag_nmn = flag_pbc = 0), 2D.  p_mts(4,na1+1,na2+1) with
 * index 0 = sum k v^2, 1..3 = sum k v; the coarse node (ix,iy) is
 * p_mts[4*((iy-1)*(na1+1)+(ix-1)) + q].  Writes that would land outside
 * p_mts (a particle exactly on xmax or ymax) are dropped. */
void lo_part_moments(const lo_domain *d, const lo_species *sp, const double *vxp,
                     long long np_tot, const double *kq, int size_na, double *p_mts) {
  const int n1 = d->n1;
  const int na1 = d->n1 / size_na, na2 = d->n2 / size_na;
  const double ha1 = size_na * d->h1, ha2 = size_na * d->h2;
  const double k1 = sp->Nm / (ha1 * ha2 * d->zmax);
  for (long i = 0; i < (long)(na1 + 1) * (na2 + 1) * 4; ++i) p_mts[i] = 0.0;
  for (long long i = 0; i < np_tot; ++i) {
    const double vx = vxp[6 * i + 3], vy = vxp[6 * i + 4], vz = vxp[6 * i + 5];
    const double xp = vxp[6 * i + 0] - vx * d->dt / 2.0;
    const double yp = vxp[6 * i + 1] - vy * d->dt / 2.0;
    if (xp > d->xmax || xp < 0.0 || yp > d->ymax || yp < 0.0) continue;
    const double Eki = vx * vx + vy * vy + vz * vz;
    const int ix = (int)(xp / ha1) + 1;
    const double px = (ix * ha1 - xp) / ha1;
    const int iy = (int)(yp / ha2) + 1;
    const double py = (iy * ha2 - yp) / ha2;
    const double ki[4] = {k1 * px * py, k1 * (1.0 - px) * py, k1 * (1.0 - px) * (1.0 - py), k1 * px * (1.0 - py)};
    const int ix_f = size_na * ix - (size_na - 1), iy_f = size_na * iy - (size_na - 1);
    const int cx[4] = {ix, ix + 1, ix + 1, ix}, cy[4] = {iy, iy, iy + 1, iy + 1};
    const int fx[4] = {ix_f, ix_f + 1, ix_f + 1, ix_f}, fy[4] = {iy_f, iy_f, iy_f + 1, iy_f + 1};
    for (int c = 0; c < 4; ++c) {
      if (cx[c] > na1 + 1 || cy[c] > na2 + 1) continue;
      const double k_tmp = kq[NODE(fx[c], fy[c], n1)] * ki[c];
      double *m = p_mts + 4 * ((long)(cy[c] - 1) * (na1 + 1) + (cx[c] - 1));
      m[1] = m[1] + k_tmp * vx;
      m[2] = m[2] + k_tmp * vy;
      m[3] = m[3] + k_tmp * vz;
      m[0] = m[0] + k_tmp * Eki;
    }
  }
}

/* calc_avg, utils.f90:138-265, one partition and rank, plt_src = 0.
 * np: node arrays per species; src, sink: ss2D(1,...), ss2D(2,...) node
 * arrays per species (either may be NULL: zero);
 * data_pavg(7,na1+1,na2+1,ntype) accumulates:
 * [(p*(na2+1) + iy-1)*(na1+1) + ix-1]*7 + {0 n, 1 Tp, 2..4 j, 5 source, 6 sink}. */
void lo_calc_avg(const lo_domain *d, int ntype, const lo_species *sp, const double *const *np,
                 const double *const *p_mts, const double *const *src, const double *const *sink,
                 int size_na, double *data_pavg) {
  const int n1 = d->n1;
  const int na1 = d->n1 / size_na, na2 = d->n2 / size_na;
  const double Vc = d->h1 * d->h2 * d->zmax;
  for (int p = 0; p < ntype; ++p) {
    const double k = sp[p].Nm / Vc;
    for (int iy = 1; iy <= na2 + 1; ++iy)
      for (int ix = 1; ix <= na1 + 1; ++ix) {
        const int ix_f = size_na * ix - (size_na - 1), iy_f = size_na * iy - (size_na - 1);
        const double *m = p_mts[p] + 4 * ((long)(iy - 1) * (na1 + 1) + (ix - 1));
        const double np_tmp = np[p][NODE(ix_f, iy_f, n1)];
        const double jx = m[1], jy = m[2], jz = m[3];
        double *o = data_pavg + 7 * (((long)p * (na2 + 1) + (iy - 1)) * (na1 + 1) + (ix - 1));
        o[0] = o[0] + np_tmp;
        o[2] = o[2] + jx;
        o[3] = o[3] + jy;
        o[4] = o[4] + jz;
        double Tp = 0.0;
        if (np_tmp > 0)
          Tp = sp[p].mass / d->qe * (1.0 / 3.0) * (m[0] / np_tmp - (jx * jx + jy * jy + jz * jz) / (np_tmp * np_tmp));
        if (Tp > 0) o[1] = o[1] + Tp;
        o[5] = o[5] + (src ? src[p][NODE(ix_f, iy_f, n1)] * k : 0.0);
        o[6] = o[6] + (sink ? sink[p][NODE(ix_f, iy_f, n1)] * k : 0.0);
      }
  }
}

/* eheating, e_heating.f90:1-112, one partition: electrons (species 1) in
 * the heating region (slit flag_c=0 or disk flag_c=1) are heated with
 * probability nu*dt per call, taking (eheat_type=1) or adding a Maxwellian
 * of thermal speed vt(i_icp); returns the change of P_loss(2). */
double lo_eheating(const lo_eheat *e, double *vxp, long long np_tot, int32_t *iseed) {
  const double dt = e->dt, dr = e->dr;
  double vz_sav = 0.0;
  double nu_tmp = e->nudt / dt;
  double dP = 0.0;
  for (long long i = 0; i < np_tot; ++i) {
    double *v = vxp + 6 * i;
    const double xp = v[0], yp = v[1];
    int flag_eh = 0;
    double rp2;
    if (e->flag_c == 0) {
      rp2 = (xp - e->xa) * (xp - e->xa);
      if (rp2 <= dr * dr && yp >= e->yl_pow && yp <= e->yr_pow) flag_eh = 1;
    } else {
      rp2 = (xp - e->xa) * (xp - e->xa) + (yp - e->ymax / 2.0) * (yp - e->ymax / 2.0);
      if (rp2 <= dr * dr) flag_eh = 1;
    }
    if (!flag_eh) continue;
    if (e->flag_gpp == 1) nu_tmp = (e->nudt / dt) * 9.0 * e->Pabs_cor * exp(-9.0 * rp2 / (dr * dr));
    double rnd[2];
    rnd[0] = lo_ran2(iseed);
    if (rnd[0] <= nu_tmp * dt) {
      const double v2old = v[3] * v[3] + v[4] * v[4] + v[5] * v[5];
      const int i_icp = (int)(e->n_icp * yp / e->ymax) + 1;
      const double vt = e->vt[i_icp - 1];
      double dvx, dvy, dvz;
      rnd[0] = lo_ran2(iseed);
      rnd[1] = lo_ran2(iseed);
      lo_load_gauss(&dvx, &dvy, vt, rnd, e->pi);
      if (vz_sav == 0.0) {
        rnd[0] = lo_ran2(iseed);
        rnd[1] = lo_ran2(iseed);
        lo_load_gauss(&dvz, &vz_sav, vt, rnd, e->pi);
      } else {
        dvz = vz_sav;
        vz_sav = 0.0;
      }
      if (e->eheat_type == 1) {
        v[3] = dvx; v[4] = dvy; v[5] = dvz;
      } else {
        v[3] = v[3] + dvx; v[4] = v[4] + dvy; v[5] = v[5] + dvz;
      }
      dP = dP + 0.5 * e->Nm * e->mass * (v[3] * v[3] + v[4] * v[4] + v[5] * v[5] - v2old);
    }
  }
  return dP;
}

/* calc_rho.f90:37-77, one partition (nproc=1, nproc_mpi=1). */
void lo_calc_rhs(int n1, int n2, int ntype, const double *charge,
                 const double *const *np, double *rhs) {
  for (int iy = 1; iy <= n2 + 1; ++iy)
    for (int ix = 1; ix <= n1 + 1; ++ix) {
      double r = 0.0;
      for (int p = 0; p < ntype; ++p) r = r - charge[p] * np[p][NODE(ix, iy, n1)];
      rhs[RHS(ix, iy, n1)] = r;
    }
}

/* calc_rho.f90:134-199, one partition. */
void lo_dens_red(int n1, int n2, int ntype, const int32_t *bcnd,
                 const double *const *np, double *const *np_red, double *np_mx) {
  for (int p = 0; p < ntype; ++p) {
    np_mx[p] = 0.0;
    memset(np_red[p], 0, sizeof(double) * (size_t)(n1 + 3) * (n2 + 3));
  }
  for (int iy = 1; iy <= n2 + 1; ++iy)
    for (int ix = 1; ix <= n1 + 1; ++ix)
      for (int p = 0; p < ntype; ++p)
        np_red[p][NODE(ix, iy, n1)] = np_red[p][NODE(ix, iy, n1)] + np[p][NODE(ix, iy, n1)];
  for (int p = 0; p < ntype; ++p)
    for (int iy = 1; iy <= n2 + 1; ++iy)
      for (int ix = 1; ix <= n1 + 1; ++ix)
        if (bcnd[NODE(ix, iy, n1)] == -1 && np_red[p][NODE(ix, iy, n1)] > np_mx[p])
          np_mx[p] = np_red[p][NODE(ix, iy, n1)];
}

/* Efield.f90:1-128, without the Neumann (-2) and periodic (0) branches. */
void lo_calc_efield(int n1, int n2, const double h[2], const double *phi,
                    const int32_t *bcnd, double *Ei) {
  for (int iy = 1; iy <= n2 + 1; ++iy)
    for (int ix = 1; ix <= n1 + 1; ++ix)
      if (bcnd[NODE(ix, iy, n1)] <= 0) {
        Ei[2 * NODE(ix, iy, n1) + 0] = -(phi[NODE(ix + 1, iy, n1)] - phi[NODE(ix - 1, iy, n1)]) / (2.0 * h[0]);
        Ei[2 * NODE(ix, iy, n1) + 1] = -(phi[NODE(ix, iy + 1, n1)] - phi[NODE(ix, iy - 1, n1)]) / (2.0 * h[1]);
      }
  for (int iy = 1; iy <= n2 + 1; ++iy)
    for (int ix = 1; ix <= n1 + 1; ++ix) {
      long nd = NODE(ix, iy, n1);
      if (bcnd[nd] == -1) continue;
      if (bcnd[nd] >= 1) {
        Ei[2 * nd + 0] = 0.0;
        Ei[2 * nd + 1] = 0.0;
        if (bcnd[NODE(ix + 1, iy, n1)] <= 0)
          Ei[2 * nd + 0] = 2.0 * Ei[2 * NODE(ix + 1, iy, n1) + 0] - Ei[2 * NODE(ix + 2, iy, n1) + 0];
        if (bcnd[NODE(ix - 1, iy, n1)] <= 0)
          Ei[2 * nd + 0] = 2.0 * Ei[2 * NODE(ix - 1, iy, n1) + 0] - Ei[2 * NODE(ix - 2, iy, n1) + 0];
        if (bcnd[NODE(ix, iy + 1, n1)] <= 0)
          Ei[2 * nd + 1] = 2.0 * Ei[2 * NODE(ix, iy + 1, n1) + 1] - Ei[2 * NODE(ix, iy + 2, n1) + 1];
        if (bcnd[NODE(ix, iy - 1, n1)] <= 0)
          Ei[2 * nd + 1] = 2.0 * Ei[2 * NODE(ix, iy - 1, n1) + 1] - Ei[2 * NODE(ix, iy - 2, n1) + 1];
      }
    }
}

/* sorting.f90:92-251 with one thread: a stable counting sort by the y-major
 * cell index ic=(ix-1)+n1*(iy-1)+1; Plist(ic) ends as the 1-based index of
 * the last particle of cell ic (0 for leading empty cells). */
void lo_part_sorting(int n1, int n2, const double h[2], double *vxp,
                     long long np_tot, int32_t *Plist) {
  const long pl_max = (long)n1 + (long)n1 * n2 + 1;
  memset(Plist, 0, sizeof(int32_t) * (size_t)(pl_max + 1));
  double *tmp = (double *)malloc(sizeof(double) * 6 * (size_t)(np_tot > 0 ? np_tot : 1));
  for (long long i = 0; i < np_tot; ++i) {
    int ix = (int)(vxp[6 * i] / h[0]) + 1;
    int iy = (int)(vxp[6 * i + 1] / h[1]) + 1;
    long ic = (ix - 1) + (long)n1 * (iy - 1) + 1;
    Plist[ic] += 1;
    memcpy(tmp + 6 * i, vxp + 6 * i, 6 * sizeof(double));
  }
  int32_t k = 0;
  for (long i = 1; i <= pl_max; ++i) {
    int32_t j = Plist[i];
    Plist[i] = k;
    k += j;
  }
  for (long long i = 0; i < np_tot; ++i) {
    int ix = (int)(tmp[6 * i] / h[0]) + 1;
    int iy = (int)(tmp[6 * i + 1] / h[1]) + 1;
    long ic = (ix - 1) + (long)n1 * (iy - 1) + 1;
    Plist[ic] += 1;
    long long j = Plist[ic];
    memcpy(vxp + 6 * (j - 1), tmp + 6 * i, 6 * sizeof(double));
  }
  /* Plist_thd for one thread: Plist(i) where positive, 0 otherwise (:226-235) */
  free(tmp);
}

/* mg.h: stencil coefficients for cell size h */
typedef struct { double aw, ae, an, as, ac; } coef_t;
static coef_t mg_coef(const double h[2], double eps0) {
  coef_t c;
  c.aw = eps0 / (h[0] * h[0]);
  c.ae = c.aw;
  c.an = eps0 / (h[1] * h[1]);
  c.as = c.an;
  c.ac = -(c.an + c.as + c.aw + c.ae);
  return c;
}

#define U(i, j) u[(long)(j) * (n1 + 3) + (i)]
#define B(i, j) b[(long)((j)-1) * (n1 + 1) + ((i)-1)]
#define BC(i, j) bcnd[(long)(j) * (n1 + 3) + (i)]

/* sors.f90:1-435 with one MPI rank and no periodic/Neumann rows. */
void lo_sor_rb(double *u, const double *b, const double h[2], double eps0,
               const int32_t *bcnd, int n1, int n2, double omega, double eps,
               int kmg, int dig, double *res_out, int *ksor) {
  coef_t c = mg_coef(h, eps0);
  const double eta = 0.6;
  long nr = 10;
  if (kmg == 1 && dig == 0) nr = 1000000;
  double res_s = 0.0, res = 0.0;
  *ksor = 0;
  for (long k = 1; k <= nr; ++k) {
    for (int j = 1; j <= n2 + 1; ++j) { /* red: i+j odd, lines 82-116 */
      int is = (j % 2 == 0) ? 1 : 2;
      for (int i = is; i <= n1 + 1; i += 2)
        if (BC(i, j) <= 0)
          U(i, j) = (1.0 - omega) * U(i, j) +
                    omega * (B(i, j) - c.aw * U(i - 1, j) - c.an * U(i, j + 1) -
                             c.as * U(i, j - 1) - c.ae * U(i + 1, j)) / c.ac;
    }
    for (int j = 1; j <= n2 + 1; ++j) { /* black: i+j even, lines 212-246 */
      int is = (j % 2 == 0) ? 2 : 1;
      for (int i = is; i <= n1 + 1; i += 2)
        if (BC(i, j) <= 0)
          U(i, j) = (1.0 - omega) * U(i, j) +
                    omega * (B(i, j) - c.aw * U(i - 1, j) - c.an * U(i, j + 1) -
                             c.as * U(i, j - 1) - c.ae * U(i + 1, j)) / c.ac;
    }
    res = 0.0; /* L1 residual over interior nodes, lines 338-357 */
    for (int j = 1; j <= n2 + 1; ++j)
      for (int i = 1; i <= n1 + 1; ++i) {
        double rij = 0.0;
        if (BC(i, j) == -1)
          rij = B(i, j) - (c.aw * U(i - 1, j) + c.an * U(i, j + 1) + c.as * U(i, j - 1) +
                           c.ae * U(i + 1, j) + c.ac * U(i, j));
        res = res + fabs(rij);
      }
    *ksor += 1;
    if (dig == 0) {
      if (res <= eps) break;
    } else {
      if (res_s > 0.0 && (res / res_s) <= eta) break;
    }
    res_s = res;
  }
  *res_out = res;
}

/* mg.f90:512-547 (flag_nmn = 0) */
static double getres(const double *u, const double *b, coef_t c, int ir, int jr, int n1) {
  double uij = U(ir, jr), un = U(ir, jr + 1), us = U(ir, jr - 1), ue = U(ir + 1, jr);
  double uw = U(ir - 1, jr);
  double urhs = B(ir, jr);
  return urhs - (c.aw * uw + c.ae * ue + c.an * un + c.as * us + c.ac * uij);
}

/* mg.f90:235-509 with one MPI rank and no periodic rows. */
void lo_restriction(const double *u, const double *b, const double h[2],
                    double eps0, const int32_t *bcnd, int32_t *bcnd2h,
                    double *r2h, double *e2h, int n1, int n2) {
  coef_t c = mg_coef(h, eps0);
  const int m1 = n1 / 2, m2 = n2 / 2;
#define BC2(i, j) bcnd2h[(long)(j) * (m1 + 3) + (i)]
#define R2(i, j) r2h[(long)((j)-1) * (m1 + 1) + ((i)-1)]
#define E2(i, j) e2h[(long)(j) * (m1 + 3) + (i)]
  /* corners are never written by the reference; copy the fine corners */
  BC2(0, 0) = BC(0, 0);
  BC2(m1 + 2, 0) = BC(n1 + 2, 0);
  BC2(0, m2 + 2) = BC(0, n2 + 2);
  BC2(m1 + 2, m2 + 2) = BC(n1 + 2, n2 + 2);
  for (int j = 1; j <= m2 + 1; ++j)
    for (int i = 1; i <= m1 + 1; ++i) {
      int ir = 2 * i - 1, jr = 2 * j - 1;
      BC2(i, j) = BC(ir, jr);
      if (ir == 1) BC2(0, j) = BC(0, jr);
      if (ir == n1 + 1) BC2(m1 + 2, j) = BC(n1 + 2, jr);
      if (jr == 1) BC2(i, 0) = BC(ir, 0);
      if (jr == n2 + 1) BC2(i, m2 + 2) = BC(ir, n2 + 2);
      if (BC2(i, j) <= 0) {
        double r1 = 0, r2 = 0, r3 = 0, r4 = 0, r5 = 0, r6 = 0, r7 = 0, r8 = 0, r9;
        r9 = getres(u, b, c, ir, jr, n1);
        if (ir != n1 + 1 && jr != n2 + 1 && BC(ir + 1, jr + 1) <= 0) r1 = getres(u, b, c, ir + 1, jr + 1, n1);
        if (ir != n1 + 1 && BC(ir + 1, jr) <= 0) r2 = getres(u, b, c, ir + 1, jr, n1);
        if (BC(ir + 1, jr - 1) <= 0 && ir != n1 + 1 && jr != 1) r3 = getres(u, b, c, ir + 1, jr - 1, n1);
        if (BC(ir, jr - 1) <= 0 && jr != 1) r4 = getres(u, b, c, ir, jr - 1, n1);
        if (BC(ir - 1, jr - 1) <= 0 && ir != 1 && jr != 1) r5 = getres(u, b, c, ir - 1, jr - 1, n1);
        if (BC(ir - 1, jr) <= 0 && ir != 1) r6 = getres(u, b, c, ir - 1, jr, n1);
        if (BC(ir - 1, jr + 1) <= 0 && ir != 1 && jr != n2 + 1) r7 = getres(u, b, c, ir - 1, jr + 1, n1);
        if (BC(ir, jr + 1) <= 0 && jr != n2 + 1) r8 = getres(u, b, c, ir, jr + 1, n1);
        R2(i, j) = (4.0 * r9 + 2.0 * (r2 + r4 + r6 + r8) + r1 + r3 + r5 + r7) / 16.0;
      } else {
        R2(i, j) = 0.0;
      }
    }
  for (int j = 0; j <= m2 + 2; ++j)
    for (int i = 0; i <= m1 + 2; ++i) E2(i, j) = 0.0;
#undef BC2
#undef R2
#undef E2
}

/* mg.f90:549-616 */
void lo_prolongation(double *u, const double *e2h, const int32_t *bcnd, int n1, int n2) {
  const int m1 = n1 / 2, m2 = n2 / 2;
#define E2(i, j) e2h[(long)(j) * (m1 + 3) + (i)]
  for (int j = 1; j <= m2 + 1; ++j)
    for (int i = 1; i <= m1 + 1; ++i) {
      double r1 = E2(i, j), r2 = E2(i, j - 1), r3 = E2(i - 1, j - 1), r4 = E2(i - 1, j);
      int ir = 2 * i - 1, jr = 2 * j - 1;
      if (BC(ir, jr) <= 0) U(ir, jr) = U(ir, jr) + r1;
      if (BC(ir, jr - 1) <= 0) U(ir, jr - 1) = U(ir, jr - 1) + (r1 + r2) / 2.0;
      if (BC(ir - 1, jr - 1) <= 0) U(ir - 1, jr - 1) = U(ir - 1, jr - 1) + (r1 + r2 + r3 + r4) / 4.0;
      if (BC(ir - 1, jr) <= 0) U(ir - 1, jr) = U(ir - 1, jr) + (r1 + r4) / 2.0;
    }
#undef E2
}
#undef U
#undef B
#undef BC

/* mg.f90:1-232: one V-cycle over ng levels. */
int lo_mg(const lo_poisson *P, double *u, const double *b, double omega,
          double eps, int k, double *ktot, double *res, lo_sweep_log *log) {
  const int ng = P->ng;
  if (ng < 1 || ng > LO_MAX_LEVELS) return -1;
  int n1[LO_MAX_LEVELS + 1], n2[LO_MAX_LEVELS + 1];
  double hh[LO_MAX_LEVELS + 1][2];
  double *e[LO_MAX_LEVELS + 1], *r[LO_MAX_LEVELS + 1];
  int32_t *bc[LO_MAX_LEVELS + 1];
  int iter[LO_MAX_LEVELS + 1];
  e[1] = u; r[1] = (double *)b; bc[1] = (int32_t *)P->bcnd;
  n1[1] = P->n1; n2[1] = P->n2; hh[1][0] = P->h1; hh[1][1] = P->h2;
  for (int l = 2; l <= ng; ++l) {
    double f = (double)(1 << (l - 1));
    n1[l] = P->n1 >> (l - 1);
    n2[l] = P->n2 >> (l - 1);
    hh[l][0] = f * P->h1;
    hh[l][1] = f * P->h2;
    e[l] = (double *)calloc((size_t)(n1[l] + 3) * (n2[l] + 3), sizeof(double));
    r[l] = (double *)calloc((size_t)(n1[l] + 1) * (n2[l] + 1), sizeof(double));
    bc[l] = (int32_t *)calloc((size_t)(n1[l] + 3) * (n2[l] + 3), sizeof(int32_t));
  }
  for (int l = 1; l <= ng; ++l) iter[l] = 0;
  int ksor = 0;
  for (int ig = 1; ig <= 2 * ng - 1; ++ig) {
    /* level il is smoothed at ig=il on the down leg and at ig=2ng-il on the
     * up leg; the coarsest level ng at ig=ng (mg.f90:85-225) */
    int il = (ig < ng) ? ig : (ig > ng) ? 2 * ng - ig : ng;
    lo_sor_rb(e[il], r[il], hh[il], P->eps0, bc[il], n1[il], n2[il], omega, eps, k,
              ig - ng, res, &ksor);
    iter[il] += ksor;
    if (log && log->ksor && log->count < log->cap) log->ksor[log->count] = ksor;
    if (log) log->count++;
    if (ig < ng)
      lo_restriction(e[il], r[il], hh[il], P->eps0, bc[il], bc[il + 1], r[il + 1], e[il + 1],
                     n1[il], n2[il]);
    if (ig >= ng && il >= 2)
      lo_prolongation(e[il - 1], e[il], bc[il - 1], n1[il - 1], n2[il - 1]);
  }
  for (int l = 1; l <= ng; ++l) *ktot = *ktot + (double)iter[l] / pow(2.0, 2 * (l - 1));
  for (int l = 2; l <= ng; ++l) { free(e[l]); free(r[l]); free(bc[l]); }
  return 0;
}

/* iterative_solver.f90:1-69.  Returns 0, or -2 on divergence (the
 * reference's stop_calculation at lines 42-49). */
int lo_iterative_solver(const lo_poisson *P, double *u, const double *b,
                        int ncycl, double eps, int *k_out, double *ktot,
                        double *res, lo_sweep_log *log) {
  double h[2] = {P->h1, P->h2};
  coef_t c = mg_coef(h, P->eps0);
  const double u0 = 1.0;
  double alpha = fabs((double)((P->n1 + 1) * (P->n2 + 1)) * c.ac) * u0;
  double res_s = 0.0;
  const double omega = 1.4;
  *ktot = 0.0;
  int k = 1;
  int converged = 0;
  while (!converged) {
    int rc = lo_mg(P, u, b, omega, alpha * eps, k, ktot, res, log);
    if (rc) return rc;
    *res = *res / alpha;
    if (k == 1) res_s = *res;
    else if (fabs(*res / res_s) > 10.0) { *k_out = k; return -2; }
    if (*res <= eps || k == ncycl) converged = 1;
    k = k + 1;
  }
  *k_out = k - 1; /* number of V-cycles done (the reference's k is one past) */
  return 0;
}

/* ---- cross-section union grid: read_reactions.f90:511-697 ---- */
static int cmp_double(const void *a, const void *b) {
  double x = *(const double *)a, y = *(const double *)b;
  return (x < y) ? -1 : (x > y) ? 1 : 0;
}

int lo_ordering(const lo_reaction_input *in, const double *mass, double qe,
                int ntype_all, double *sig_Er, double *sig, lo_reaction_table *t) {
  int total = 0;
  memset(t, 0, sizeof(*t));
  t->ncol = in->ncol;
  for (int c = 0; c < in->ncol; ++c) total += in->npt[c];
  double *all = (double *)malloc(sizeof(double) * (size_t)(total > 0 ? total : 1));
  int q = 0;
  for (int c = 0; c < in->ncol; ++c)
    for (int i = 0; i < in->npt[c]; ++i) all[q++] = in->E[c][i];
  qsort(all, (size_t)total, sizeof(double), cmp_double);
  int nu = 0;
  double sav = 1.0e10;
  for (int i = 0; i < total; ++i)
    if (all[i] != sav) { sig_Er[nu++] = all[i]; sav = sig_Er[nu - 1]; }
  free(all);
  t->npt_mx = nu;
  t->sig_Er = sig_Er;
  t->sig = sig;
  for (int c = 0; c < in->ncol; ++c) {
    memcpy(t->col_info[c], in->col_info[c], sizeof(int) * 10);
    t->Eth[c] = in->Eth[c];
    int npt = in->npt[c];
    double sig_p = 0.0;
    for (int ip = 0; ip < nu; ++ip) {
      double Ekp = sig_Er[ip];
      double EkL = 0, EkR = 0, sL = 0, sR = 0;
      for (int j = 0; j < npt; ++j) {
        EkL = in->E[c][j];
        EkR = (j + 1 < npt) ? in->E[c][j + 1] : 0.0;
        sL = in->S[c][j];
        sR = (j + 1 < npt) ? in->S[c][j + 1] : 0.0;
        if (Ekp < EkL) { sig_p = sL; break; }
        if (Ekp >= EkL && Ekp <= EkR) { sig_p = sL + (Ekp - EkL) * (sR - sL) / (EkR - EkL); break; }
      }
      if (sL == 0.0 && sR > 0.0 && in->Eth[c] > EkR) return -3;
      sig[(long)c * nu + ip] = sig_p;
    }
  }
  int ind_col[LO_MAX_TYPES + 1] = {0};
  for (int c = 0; c < in->ncol; ++c) {
    int r1 = in->col_info[c][2], r2 = in->col_info[c][3]; /* col_info(icol,3:4) */
    ind_col[r1] += 1;
    t->sig_list[r1][ind_col[r1]] = c + 1;
    for (int ip = 0; ip < nu; ++ip) {
      double mu = fabs(mass[r1]) * fabs(mass[r2]) / (fabs(mass[r1]) + fabs(mass[r2]));
      double vr = sqrt(2.0 * sig_Er[ip] * qe / mu);
      if (sig_Er[ip] >= 0.1) {
        double v = vr * sig[(long)c * nu + ip];
        if (v > t->sigv_mx[r1][c]) t->sigv_mx[r1][c] = v;
      }
    }
  }
  for (int p = 1; p <= ntype_all; ++p) t->p_ncol[p] = ind_col[p];
  return 0;
}

/* indexx.f90:20-34: straight insertion for n < 7 (ascending, stable). */
static void indexx_small(int n, const double *arr, int *indx) {
  for (int j = 1; j <= n; ++j) indx[j] = j;
  for (int j = 2; j <= n; ++j) {
    int indxt = indx[j];
    double a = arr[indxt];
    int i;
    for (i = j - 1; i >= 1; --i) {
      if (arr[indx[i]] <= a) break;
      indx[i + 1] = indx[i];
    }
    indx[i + 1] = indxt;
  }
}

/* collisions.f90:1-207 and collision_OMP :209-1040, one thread, neutral
 * targets (MC branch), flag_ionrep = 0 (energy/momentum conserving
 * scattering).  Returns 0, or <0 for an unsupported option or an abort of
 * the reference (Pmax>1, negative energy, energy above the table, ...). */
int lo_collisions(const lo_domain *d, const lo_mcc_setup *s,
                  const lo_reaction_table *t, lo_particles *p,
                  const double *const *np_red, int32_t *iseed, lo_mcc_out *o) {
  const double pi = 4.0 * atan(1.0);
  const double qe = d->qe;
  const int n1 = d->n1;
  const double h1 = d->h1, h2 = d->h2;
  int sav_ttype[LO_MAX_REAC + 1];
  long long Nc[LO_MAX_TYPES + 1], np_add[LO_MAX_TYPES + 1];
  long long np_tot_rg[LO_MAX_TYPES + 1][LO_MAX_RG];
  double nu_max[LO_MAX_TYPES + 1];
  if (s->flag_ionrep != 0) return -10;
  for (int q = 0; q <= LO_MAX_TYPES; ++q) { Nc[q] = 0; np_add[q] = 0; nu_max[q] = 0.0; }
  memset(np_tot_rg, 0, sizeof(np_tot_rg));

  for (int pt = 1; pt <= s->ntype; ++pt) { /* lines 92-157 */
    if (t->p_ncol[pt] == 0 || s->mass[pt] <= 0) continue;
    for (int ic = 1; ic <= t->p_ncol[pt]; ++ic) {
      int c = t->sig_list[pt][ic];
      int n_re = t->col_info[c - 1][0];
      int tt = t->col_info[c - 1][1 + n_re]; /* col_info(c, ind_nby+n_re) */
      sav_ttype[c] = tt;
      nu_max[pt] = nu_max[pt] + s->np_mx[tt] * t->sigv_mx[pt][c - 1];
    }
    if (s->nu_uplim[pt] < nu_max[pt]) nu_max[pt] = s->nu_uplim[pt];
    long long sum_np_tot = p->np_tot[pt];
    double Pmax = nu_max[pt] * (double)(float)s->ns_coll * d->dt;
    if (Pmax > 1.0) return -11;
    double dNc = (double)sum_np_tot * Pmax / (double)(float)1;
    long long Nc_tmp = (long long)dNc;
    double rnd = lo_ran2(iseed);
    if (rnd <= (dNc - (double)Nc_tmp)) Nc_tmp = Nc_tmp + 1;
    Nc[pt] = Nc_tmp;
    if (s->n_cell > 0 && p->np_tot[1] / s->n_cell >= 5 && pt == 1 && Nc_tmp <= 1) return -12;
    double kk = h1 * h2 * d->zmax / s->Nm[pt];
    for (int rg = 0; rg < s->n_rg; ++rg) {
      double sum = 0.0;
      for (int iy = s->iyl_rg; iy <= s->iyr_rg; ++iy)
        for (int ix = s->ixl_rg[rg]; ix <= s->ixr_rg[rg]; ++ix) sum += np_red[pt - 1][NODE(ix, iy, n1)];
      np_tot_rg[pt][rg] = (long long)(kk * sum);
    }
  }

  for (int pt = 1; pt <= s->ntype; ++pt) { /* lines 167-179 */
    if (t->p_ncol[pt] == 0 || s->mass[pt] <= 0 || Nc[pt] == 0 || p->np_tot[pt] == 0) continue;
    float fnum = (float)Nc[pt] / (float)p->np_tot[pt];
    double nu_max_OMP = (double)fnum / ((double)(float)s->ns_coll * d->dt);
    double vz_sav = 0.0;
    double *vxp = p->vxp[pt];
    const long long Ncand = Nc[pt];
    for (long long icand = 1; icand <= Ncand; ++icand) {
      double rnd1 = lo_ran2(iseed);
      long long ip = (long long)((double)p->np_tot[pt] * rnd1) + 1;
      if (ip == p->np_tot[pt] + 1) return -13;
      double xp = vxp[6 * (ip - 1)], yp = vxp[6 * (ip - 1) + 1];
      int ix = (int)(xp / h1) + 1;
      int iy = (int)(yp / h2) + 1;
      double vx[3], vy[3], vz[3];
      vx[1] = vxp[6 * (ip - 1) + 3];
      vy[1] = vxp[6 * (ip - 1) + 4];
      vz[1] = vxp[6 * (ip - 1) + 5];
      int flag_ttype[LO_MAX_TYPES + 1] = {0};
      double Ekr[LO_MAX_TYPES + 1] = {0}, vr[LO_MAX_TYPES + 1] = {0}, mu[LO_MAX_TYPES + 1] = {0};
      double vxcm[LO_MAX_TYPES + 1] = {0}, vycm[LO_MAX_TYPES + 1] = {0}, vzcm[LO_MAX_TYPES + 1] = {0};
      double nu[LO_MAX_REAC + 1] = {0}, sort_arr[LO_MAX_REAC + 1] = {0};
      double sum_nu = 0.0;
      int c_ind = 0;
      for (int icol = 1; icol <= t->p_ncol[pt]; ++icol) {
        int c = t->sig_list[pt][icol];
        int tt = sav_ttype[c];
        if (flag_ttype[tt] == 0) {
          if (s->charge[tt] != 0.0) return -14; /* DSMC target: not in this restatement */
          double rr[2];
          rr[0] = lo_ran2(iseed);
          rr[1] = lo_ran2(iseed);
          lo_load_gauss(&vx[2], &vy[2], s->vt0[tt], rr, pi);
          if (vz_sav == 0.0) {
            rr[0] = lo_ran2(iseed);
            rr[1] = lo_ran2(iseed);
            lo_load_gauss(&vz[2], &vz_sav, s->vt0[tt], rr, pi);
          } else {
            vz[2] = vz_sav;
            vz_sav = 0.0;
          }
          double mp = fabs(s->mass[pt]), mt = fabs(s->mass[tt]);
          mu[tt] = mp * mt / (mp + mt);
          vr[tt] = sqrt((vx[1] - vx[2]) * (vx[1] - vx[2]) + (vy[1] - vy[2]) * (vy[1] - vy[2]) +
                        (vz[1] - vz[2]) * (vz[1] - vz[2]));
          Ekr[tt] = (0.5 * mu[tt] * (vr[tt] * vr[tt])) / qe;
          double sum_mass = mp + mt;
          vxcm[tt] = (mp * vx[1] + mt * vx[2]) / sum_mass;
          vycm[tt] = (mp * vy[1] + mt * vy[2]) / sum_mass;
          vzcm[tt] = (mp * vz[1] + mt * vz[2]) / sum_mass;
          if (Ekr[tt] > t->sig_Er[t->npt_mx - 1]) return -15;
          flag_ttype[tt] = 1;
        }
        /* dichotomy + scan, lines 485-611 (1-based indices into sig_Er) */
        int npt = t->npt_mx, L = 1, R = npt;
        for (int q = 1; q <= 5; ++q) {
          int M = (int)((double)(float)(R - L) / 2.0) + L;
          double EkL = t->sig_Er[L - 1], EkR = t->sig_Er[R - 1], EkM = t->sig_Er[M - 1];
          if (Ekr[tt] >= EkL && Ekr[tt] <= EkM) R = M;
          if (Ekr[tt] > EkM && Ekr[tt] <= EkR) L = M;
        }
        for (int ipt = L + 1; ipt <= R; ++ipt) {
          double EkL = t->sig_Er[ipt - 2], EkR = t->sig_Er[ipt - 1];
          if (Ekr[tt] > EkL && Ekr[tt] <= EkR) {
            double sL = t->sig[(long)(c - 1) * npt + ipt - 2], sR = t->sig[(long)(c - 1) * npt + ipt - 1];
            double sig_p = sL + (Ekr[tt] - EkL) * (sR - sL) / (EkR - EkL);
            double np_t = s->np_mx[tt];
            nu[icol] = np_t * sig_p * vr[tt] / nu_max_OMP;
            sum_nu = sum_nu + nu[icol];
            for (int rg = 0; rg < s->n_rg; ++rg)
              if (ix >= s->ixl_rg[rg] && ix <= s->ixr_rg[rg] && iy >= s->iyl_rg && iy <= s->iyr_rg) {
                if (o->cnt_col) {
                  o->cnt_col[(0 * (long)t->ncol + (c - 1)) * s->n_rg + rg] += nu[icol] * nu_max_OMP;
                  o->cnt_col[(1 * (long)t->ncol + (c - 1)) * s->n_rg + rg] += 1.0;
                }
                break;
              }
            sort_arr[icol] = nu[icol];
            break;
          }
        }
      }
      int indx[LO_MAX_REAC + 1];
      indexx_small(t->p_ncol[pt], sort_arr, indx);
      double rnd = lo_ran2(iseed);
      int flag_coll = 0;
      if (rnd <= sum_nu) {
        double cs = 0.0;
        for (int icol = 1; icol <= t->p_ncol[pt]; ++icol) {
          cs = cs + nu[indx[icol]];
          if (rnd <= cs) { flag_coll = 1; c_ind = t->sig_list[pt][indx[icol]]; break; }
        }
      }
      if (o->event_log && o->log_len < o->log_cap) {
        o->event_log[2 * o->log_len] = (int32_t)ip;
        o->event_log[2 * o->log_len + 1] = flag_coll ? c_ind : 0;
        o->log_len++;
      }
      if (!flag_coll) continue;
      const int *ci = t->col_info[c_ind - 1];
      int n_re = ci[0], n_by = ci[1];
      int rt = ci[2 + n_re + n_by]; /* col_info(c,2+n_re+n_by+1) */
      int tt = sav_ttype[c_ind];
      o->events[c_ind - 1] += 1;
      if (rt != 4) {
        double Eth = t->Eth[c_ind - 1];
        double Ee = 0.5 * mu[tt] * (vr[tt] * vr[tt]) - Eth * qe;
        double sum_mass = 0.0;
        for (int q = 1; q <= n_by; ++q) sum_mass = sum_mass + 1.0 / fabs(s->mass[ci[1 + n_re + q]]);
        double r1 = lo_ran2(iseed), r2 = lo_ran2(iseed);
        double th_add = 2.0 * pi / (double)(float)n_by;
        double cos_th = 1.0 - 2.0 * r1;
        double th = acos(cos_th);
        double phi = 2.0 * pi * r2;
        double cos_ph = cos(phi), sin_ph = sin(phi);
        (void)lo_ran2(iseed); /* cos_th_s, phi_s: drawn, used only by SCATTER, */
        (void)lo_ran2(iseed); /* whose result is discarded (lines 841-846) */
        if (Ee < 0.0) return -16;
        int flag_add[11];
        flag_add[1] = flag_add[2] = -1;
        for (int q = n_re + 1; q <= n_re + n_by; ++q) flag_add[q] = 0;
        for (int iby = 1; iby <= n_by; ++iby) {
          int bt = ci[1 + n_re + iby];
          if (s->mass[bt] <= 0) continue;
          int opt_add = 1;
          long long ib = 0;
          if (iby == 1 && bt == pt) { ib = ip; opt_add = 0; flag_add[1] = 0; }
          if (rt == 1 || rt == 3) {
            if (bt == pt) ib = ip;
            if (bt == tt) return -14;
            opt_add = 0;
            flag_add[1] = flag_add[2] = 0;
          }
          double *vb;
          double v2old;
          if (opt_add == 1) {
            np_add[bt] += 1;
            ib = p->np_tot[bt] + np_add[bt];
            if (ib > p->cap[bt]) return -17;
            double *dst = p->vxp[bt] + 6 * (ib - 1);
            const double *src = vxp + 6 * (ip - 1);
            dst[0] = src[0]; dst[1] = src[1]; dst[2] = src[2];
            flag_add[n_re + iby] = 1;
            if (s->plt_src == 0 && p->src[bt]) p->src[bt][NODE(ix, iy, n1)] += 1.0;   /* line 809 */
            vb = dst;
            v2old = 0.0;
          } else {
            vb = p->vxp[bt] + 6 * (ib - 1);
            v2old = vb[3] * vb[3] + vb[4] * vb[4] + vb[5] * vb[5];
          }
          double c_t = cos(th), s_t = sin(th);
          th = th + th_add;
          double ex = c_t, ey = s_t * sin_ph, ez = s_t * cos_ph;
          double mb = fabs(s->mass[bt]);
          double vp = sqrt(2.0 * (Ee / sum_mass) / (mb * mb));
          vb[3] = vxcm[tt] + vp * ex;
          vb[4] = vycm[tt] + vp * ey;
          vb[5] = vzcm[tt] + vp * ez;
          o->P_loss3[bt] = o->P_loss3[bt] + 0.5 * s->Nm[bt] * s->mass[bt] *
                                                (vb[3] * vb[3] + vb[4] * vb[4] + vb[5] * vb[5] - v2old);
        }
        for (int ire = 1; ire <= n_re; ++ire) {
          int ct = ci[1 + ire];
          if (s->mass[ct] <= 0) { flag_add[ire] = 0; continue; }
          if (flag_add[ire] == -1) {
            if (ct == tt) return -14;
            p->flag_dead[ct][ip - 1] = 1;
            if (s->plt_src == 0 && p->sink[ct]) p->sink[ct][NODE(ix, iy, n1)] += 1.0;   /* line 931 */
          }
        }
      } else { /* charge exchange with a neutral target, lines 942-1000 */
        for (int iby = 1; iby <= n_by; ++iby) {
          int bt = ci[1 + n_re + iby];
          if (s->mass[bt] <= 0) continue;
          if (bt != pt) return -14;
          double *vb = p->vxp[bt] + 6 * (ip - 1);
          double v2old = vb[3] * vb[3] + vb[4] * vb[4] + vb[5] * vb[5];
          vb[3] = vx[2]; vb[4] = vy[2]; vb[5] = vz[2];
          o->P_loss3[bt] = o->P_loss3[bt] + 0.5 * s->Nm[bt] * s->mass[bt] *
                                                (vb[3] * vb[3] + vb[4] * vb[4] + vb[5] * vb[5] - v2old);
        }
      }
      for (int rg = 0; rg < s->n_rg; ++rg)
        if (ix >= s->ixl_rg[rg] && ix <= s->ixr_rg[rg] && iy >= s->iyl_rg && iy <= s->iyr_rg) {
          if (o->cnt_col) o->cnt_col[(2 * (long)t->ncol + (c_ind - 1)) * s->n_rg + rg] += 1.0 / (double)np_tot_rg[pt][rg];
          break;
        }
    }
  }
  for (int pt = 1; pt <= s->ntype; ++pt) {
    p->np_tot[pt] += np_add[pt];
    o->Nc[pt] = Nc[pt];
    o->nu_max[pt] = nu_max[pt];
    if (p->np_tot[pt] > p->cap[pt]) return -17;
  }
  return 0;
}

/* part_flux_emitter, part_flux_emitter.f90:1-217 (part_flux_emitter with
 * nproc = nproc_mpi = 1, then part_flux_OMP on the one partition): Nh
 * particles of ptype_fx (tag_b on the chamfer) leave the electrode at xg1
 * with a flux-Maxwellian normal speed and a Maxwellian tangential pair of
 * temperature THm, spread over the ns_Hm*dt they travel.  A candidate
 * whose cell is outside the node range counts as off the electrode (the
 * reference would read bcnd out of bounds there); more than max_tries
 * candidates for one particle returns -1 where the reference loops forever.
 * Returns -2 when the chamfer is longer than ymax (line 39), -17 when a
 * species overflows its capacity. */
int lo_flux_emitter(const lo_domain *d, const lo_emitter *e, lo_particles *p, int32_t *iseed,
                    long long max_tries, lo_emitter_out *o) {
  const int n1 = d->n1, n2 = d->n2;
  const double h1 = d->h1, h2 = d->h2, ymax = d->ymax, zmax = d->zmax, dt = d->dt, qe = d->qe;
  const double pi = e->pi;
  const int chamfered = e->ptype_fx == e->tag_neg && e->flag_chmfrd == 1;
  const double js = e->jne * 10.0;
  double Lgrd = (double)e->n_icp * (e->Lg - e->Lh * (double)e->n_holes);
  double th_chrd = 0.0, y1 = 0.0;
  if (chamfered) {
    const double x1 = e->chmfrd_array[0], x2 = e->chmfrd_array[1];
    y1 = e->chmfrd_array[2];
    const double y2 = e->chmfrd_array[3];
    const double d_chrd = sqrt((x2 - x1) * (x2 - x1) + (y2 - y1) * (y2 - y1));
    th_chrd = acos((x2 - x1) / d_chrd);
    Lgrd = 2.0 * (y1 + d_chrd);
    if (Lgrd > ymax) return -2;
  }
  o->Lgrd = Lgrd;
  const double I_sec = js * zmax * Lgrd;
  const double dNh = I_sec * ((double)(float)e->ns_Hm * dt) / (e->Nm[e->ptype_fx] * qe) / 1.0 / 1.0;
  long long Nh = (long long)dNh;
  if (lo_ran2(iseed) <= dNh - (double)Nh) Nh = Nh + 1;
  o->dNh = dNh;
  o->Nh = Nh;
  const double x_injtmp = e->xg1;
  long long cnt_lhs = 0, cnt_rhs = 0;
  const int32_t *bc = d->bcnd;
  for (long long i = 1; i <= Nh; ++i) {
    double x_inj, y_inj, z_inj, dir = 0.0;
    int ptype, ix, iy;
    long long tries = 0;
    for (;;) { /* label 90 */
      if (++tries > max_tries) return -1;
      x_inj = x_injtmp;
      const double r = lo_ran2(iseed);
      if (e->n_icp == 1) y_inj = (ymax - e->Lg) / 2.0 + e->Lg * r;
      else y_inj = ymax * (1.0 - r);
      ptype = e->ptype_fx;
      if (chamfered && ((y_inj <= ymax / 2.0 && y_inj > y1) || (y_inj > ymax / 2.0 && fabs(ymax - y_inj) > y1))) {
        if (e->tag_b > 0) ptype = e->tag_b;
        double th = th_chrd, d0 = y_inj - y1, d1 = y1;
        if (y_inj > ymax / 2.0) {
          th = -th_chrd;
          d0 = fabs(y_inj - ymax) - y1;
          d1 = ymax - y1;
        }
        x_inj = x_injtmp + d0 * cos(th);
        y_inj = d1 + d0 * sin(th);
      }
      ix = (int)(x_inj / h1) + 1;
      iy = (int)(y_inj / h2) + 1;
      z_inj = zmax * (1.0 - lo_ran2(iseed));
      if (ix < 1 || ix > n1 + 1 || iy < 1 || iy > n2 + 1) continue;
      if ((bc[NODE(ix, iy, n1)] >= 1 && bc[NODE(ix, iy + 1, n1)] >= 1) ||
          (bc[NODE(ix + 1, iy, n1)] >= 1 && bc[NODE(ix + 1, iy + 1, n1)] >= 1)) {
        if (bc[NODE(ix - 1, iy, n1)] == -1) cnt_lhs = cnt_lhs + 1;
        if (bc[NODE(ix + 1, iy, n1)] == -1) cnt_rhs = cnt_rhs + 1;
        if (cnt_lhs != cnt_rhs) {
          dir = cnt_lhs > cnt_rhs ? 1.0 : -1.0;
          break;
        }
      }
    }
    if (ptype == e->tag_b) o->cnt_b[1] += 1;
    else o->cnt_b[0] += 1;
    const long long k = p->np_tot[ptype] + 1;
    if (k > p->cap[ptype]) return -17;
    p->np_tot[ptype] = k;
    if (p->src[ptype]) p->src[ptype][NODE(ix, iy, n1)] += 1.0;
    const double vt = sqrt(2.0 * qe * e->THm / fabs(e->mass[ptype]));
    const double vx = dir * vt * sqrt(-log(1.0 - lo_ran2(iseed)));
    double rnd[2], vy, vz;
    rnd[0] = lo_ran2(iseed);
    rnd[1] = lo_ran2(iseed);
    lo_load_gauss(&vy, &vz, vt, rnd, pi);
    double th = pi / 2.0;
    if (chamfered && ((y_inj <= ymax / 2.0 && y_inj > y1) || (y_inj > ymax / 2.0 && fabs(ymax - y_inj) > y1))) {
      th = th_chrd;
      if (y_inj > ymax / 2.0) th = pi / 2.0 + th_chrd;
    }
    double *v = p->vxp[ptype] + 6 * (k - 1);
    v[3] = -(vx * sin(th)) + vy * cos(th);
    v[4] = vx * cos(th) + vy * sin(th);
    v[5] = vz;
    const double dt_tmp = lo_ran2(iseed) * ((double)(float)e->ns_Hm * dt);
    v[0] = x_inj + v[3] * dt_tmp;
    v[1] = y_inj + v[4] * dt_tmp;
    v[2] = z_inj + v[5] * dt_tmp;
    o->P_loss4[ptype] = o->P_loss4[ptype] + 0.5 * e->Nm[ptype] * e->mass[ptype] *
                                                (v[3] * v[3] + v[4] * v[4] + v[5] * v[5]);
  }
  return 0;
}

/* shifted_maxwellian_flux, part_injection.f90:239-265: acceptance-rejection
 * of v*exp(-(v-vb)^2/vt^2) on [max(0, vb-4vt), vb+4vt] under fmax. */
static int lo_shifted_flux(double *v, double vb, double vt, double f_max, int32_t *iseed, long long max_tries) {
  const double vm = (vb - 4.0 * vt) > 0.0 ? vb - 4.0 * vt : 0.0, vp = vb + 4.0 * vt;
  for (long long t = 0; t < max_tries; ++t) {
    const double r1 = lo_ran2(iseed), r2 = lo_ran2(iseed);
    *v = vm + r1 * (vp - vm);
    const double f = *v * exp(-(*v - vb) * (*v - vb) / (vt * vt));
    if (!(r2 > f / f_max)) return 0;
  }
  return -1;
}

/* The main loop's injection (main.f90:1063-1071): part_injection,
 * part_injection.f90:1-237, for every species in turn with one ran2 stream,
 * then N_inj = 0.  Candidates whose cell lies outside the node range are
 * drawn again (the reference would index bcnd out of bounds); more than
 * max_tries candidates for one particle return -1 where the reference loops
 * forever; -17 when a species overflows its capacity. */
int lo_part_injection(const lo_domain *d, const lo_injection *in, const long long *N_inj_in,
                      const double *phi, lo_particles *p, int32_t *iseed, long long max_tries,
                      lo_injection_out *o) {
  const int n1 = d->n1, n2 = d->n2, ntype = in->ntype, opt = in->opt_inj, aopt = opt < 0 ? -opt : opt;
  const double h1 = d->h1, h2 = d->h2, xmax = d->xmax, ymax = d->ymax, zmax = d->zmax, dt = d->dt, qe = d->qe;
  const double dr = in->dr, xa = in->xa;
  const int32_t *bc = d->bcnd;
  long long N_inj[LO_MAX_TYPES + 1];
  for (int q = 0; q <= LO_MAX_TYPES; ++q) N_inj[q] = q <= ntype && N_inj_in ? N_inj_in[q] : 0;
  for (int ptype = 1; ptype <= ntype; ++ptype) {
    if (!(fabs(in->charge[ptype]) > 0.0)) continue;
    double vz_sav = 0.0, x_sav = 0.0, vb = 0.0, f_max = 0.0;
    int flag = 0;
    if (aopt == 3 || aopt == 4) {
      vb = sqrt(2.0 * qe * in->THm / fabs(in->mass[ptype]));
      if (aopt == 4) {
        if (ptype == 1 && in->flag_nmn == 1) vb = 0.0;
        const double vt0 = in->vt0[ptype];
        const double vmax = (vb + sqrt(vb * vb + 2.0 * vt0 * vt0)) / 2.0;   /* lines 268-286 */
        f_max = vmax * exp(-(vmax - vb) * (vmax - vb) / (vt0 * vt0));
      }
    }
    double dN_inj = 0.0;
    if (aopt == 2) {
      if (ptype == 1 || (ptype == in->tag_neg && in->tag_neg > 0)) {
        flag = 1;
        long long tmp = 0;
        for (int q = 2; q <= ntype; ++q) tmp += N_inj[q];
        if (in->tag_neg > 0) tmp = tmp - N_inj[in->tag_neg];
        dN_inj = in->ni0[ptype] * (double)tmp;
      }
    } else {
      flag = 1;
      dN_inj = in->I_inj * (double)in->ns_inj * dt * in->ni0[ptype] / (qe * in->Nm[1]) / 1.0 / 1.0;
    }
    if (flag == 1) {
      long long n = (long long)dN_inj;
      if (lo_ran2(iseed) <= dN_inj - (double)n) n = n + 1;
      N_inj[ptype] = n;
    }
    if (aopt == 3 || (aopt == 4 && in->flag_nmn == 0))
      for (int q = 2; q <= ntype; ++q) N_inj[q] = 0;
    for (long long i = 1; i <= N_inj[ptype]; ++i) {
      double x = 0.0, y = 0.0, z = 0.0;
      int ix = 0, iy = 0;
      long long tries = 0;
      for (;;) { /* label 70 */
        if (++tries > max_tries) return -1;
        if (opt > 0) {
          for (;;) { /* label 60 */
            x = lo_ran2(iseed) * (in->xl_pow - in->xr_pow) + in->xr_pow;
            if (in->flag_c != 1) break;
            y = lo_ran2(iseed) * (2.0 * dr) + ymax / 2.0 - dr;
            if (!(((x - xa) * (x - xa) + (y - ymax / 2.0) * (y - ymax / 2.0)) > dr * dr)) break;
            if (++tries > max_tries) return -1;
          }
        }
        if (opt < 0) {
          if (in->flag_c == 0) {
            for (;;) { /* label 50 */
              if (x_sav == 0.0) {
                double rnd[2];
                rnd[0] = lo_ran2(iseed);
                rnd[1] = lo_ran2(iseed);
                lo_load_gauss(&x, &x_sav, dr, rnd, in->pi);
                x = x + xa;
                x_sav = x_sav + xa;
              } else {
                x = x_sav;
                x_sav = 0.0;
              }
              if (in->flag_nmn == 1 && xa == 0.0) x = fabs(x);
              if (!(x < (xa - 4.0 * dr) || x > (xa + 4.0 * dr) || x < 0.0 || x > xmax)) break;
              if (++tries > max_tries) return -1;
            }
          } else {
            for (;;) { /* label 40 */
              double rnd[2];
              rnd[0] = lo_ran2(iseed);
              rnd[1] = lo_ran2(iseed);
              lo_load_gauss(&x, &y, dr, rnd, in->pi);
              x = x + xa;
              y = y + ymax / 2.0;
              if (!(((x - xa) * (x - xa) + (y - ymax / 2.0) * (y - ymax / 2.0)) > 4.0 * dr * dr || x < 0.0 ||
                    x > xmax || y < 0.0 || y > ymax))
                break;
              if (++tries > max_tries) return -1;
            }
          }
        }
        if (in->flag_c == 0) y = lo_ran2(iseed) * (in->yl_pow - in->yr_pow) + in->yr_pow;
        z = lo_ran2(iseed) * zmax;
        ix = (int)(x / h1) + 1;
        iy = (int)(y / h2) + 1;
        if (ix < 0 || ix > n1 + 1 || iy < 0 || iy > n2 + 1) continue;
        if (bc[NODE(ix, iy, n1)] >= 1 && bc[NODE(ix + 1, iy, n1)] >= 1 && bc[NODE(ix + 1, iy + 1, n1)] >= 1 &&
            bc[NODE(ix, iy + 1, n1)] >= 1)
          continue;
        break;
      }
      const long long k = p->np_tot[ptype] + 1;
      if (k > p->cap[ptype]) return -17;
      p->np_tot[ptype] = k;
      double vt = in->vt0[ptype];
      if (aopt <= 2 && in->flag_Tp == 1 && ptype == 1) vt = sqrt(2.0 * qe * in->THm / fabs(in->mass[ptype]));
      double vx, vy, vz;
      if (aopt == 4) {
        const double vb_tmp = vb * cos(in->th_B);
        long long t4 = 0;
        for (;;) { /* label 75 */
          if (++t4 > max_tries) return -1;
          if (lo_shifted_flux(&vx, vb_tmp, vt, f_max, iseed, max_tries)) return -1;
          double rnd[2];
          rnd[0] = lo_ran2(iseed);
          rnd[1] = lo_ran2(iseed);
          lo_load_gauss(&vy, &vz, vt, rnd, in->pi);
          vy = vy + vb * sin(in->th_B);
          const double dt_tmp = lo_ran2(iseed) * ((double)(float)in->ns_inj * dt);
          x = vx * dt_tmp;
          if (!(x < 0.0 || x > xmax)) break;
        }
        ix = (int)floor(x / h1) + 1;
      } else {
        double rnd[2];
        rnd[0] = lo_ran2(iseed);
        rnd[1] = lo_ran2(iseed);
        lo_load_gauss(&vx, &vy, vt, rnd, in->pi);
        if (vz_sav == 0.0) {
          rnd[0] = lo_ran2(iseed);
          rnd[1] = lo_ran2(iseed);
          lo_load_gauss(&vx, &vy, vt, rnd, in->pi);
          vz = vx;
          vz_sav = vy;
        } else {
          vz = vz_sav;
          vz_sav = 0.0;
        }
      }
      double *v = p->vxp[ptype] + 6 * (k - 1);
      v[0] = x; v[1] = y; v[2] = z; v[3] = vx; v[4] = vy;
      if (aopt == 3) vz = vz + vb;
      v[5] = vz;
      if (p->src[ptype]) p->src[ptype][NODE(ix, iy, n1)] += 1.0;
      o->injected[ptype] += 1;
      o->P_loss4[ptype] = o->P_loss4[ptype] + 0.5 * in->Nm[ptype] * in->mass[ptype] *
                                                  (v[3] * v[3] + v[4] * v[4] + v[5] * v[5]);
      if (in->flag_diffsrc == 1 && phi) {
        const double px = (ix * h1 - x) / h1, py = (iy * h2 - y) / h2;
        const double k1 = px * py, k2 = (1.0 - px) * py, k3 = (1.0 - px) * (1.0 - py), k4 = px * (1.0 - py);
        const double phip = k1 * phi[NODE(ix, iy, n1)] + k2 * phi[NODE(ix + 1, iy, n1)] +
                            k3 * phi[NODE(ix + 1, iy + 1, n1)] + k4 * phi[NODE(ix, iy + 1, n1)];
        o->P_loss1[ptype] = o->P_loss1[ptype] - in->Nm[ptype] * in->charge[ptype] * phip;
      }
    }
  }
  return 0;
}
