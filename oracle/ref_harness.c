/*
 * ref_harness.c -- drives the reference's own Fortran routines, compiled from
 * the reference checkout by oracle/Makefile, on inputs written by the tests.
 *
 * TEST INFRASTRUCTURE ONLY: this program is the checker of the checker (it
 * pins oracle/lepic_oracle.c); nothing in the product runs it.
 *
 * Usage: ref_harness <cmd> <in.bin> <out.bin>
 *   in.bin  = int64 ni, int64[ni], int64 nd, double[nd], then raw arrays
 *   out.bin = raw arrays (order per command, see tests/refharness.py)
 * Commands: ran2, gauss, mover, deposit, efield, sort, heat, emit, bench (a timed
 * push + deposition loop, one thread) and ompbench (the same with one particle
 * partition per OpenMP thread, as the reference's main loop runs it).
 *
 * The routines take every argument by reference (implicit interfaces) and
 * read their scalars from the COMMON blocks declared in particle_info.h:56-73
 * and constants.h:4-6; the structs below mirror those blocks member for
 * member (sizes checked against the symbol sizes flang emits: 588, 1336,
 * 480, 768 and 40 bytes).
 */
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <math.h>
#include <time.h>

#include <omp.h>

#define NPART 32

/* /part_info_int/ particle_info.h:56-66, 147 default integers */
extern struct {
  int32_t np_cell, ncol, sig_npt_mx, flag_pdf, n_cell, nbak, flag_restart, flag_B,
      ns_heat, ns_coll;
  int32_t p_ncol[NPART];
  int32_t nscol;
  int32_t p_nscol[NPART];
  int32_t gtype, n_rg, flag_pbc, n_holes, flag_grd[2], ind_g, nB, cnt_seg, tag_neg,
      cnt_neg, tag_neu, num_grd, ptype_pdf, ns_Hm, ixl_rg[5], ixr_rg[5], iyl_rg, iyr_rg,
      nx_PE, size_na, flag_heat, flag_inj, ns_inj, flag_nmn, opt_inj, flag_B_pos, flag_c,
      flag_convP, eheat_type, flag_gpp, flag_dielec, mpi_rank_max, omp_rank_max,
      flag_chmfrd, tag_b, flag_Tp, cnt_plt, flag_bak, flag_spec, flag_float, dtype[10],
      opt_solver, flag_ps, igrid_sec, dir_sec, flag_sec, flag_ionrep, ptype_fx, flag_1D,
      flag_diffsrc, flag_icp, flag_gridB, plt_src, flag_RFpot;
} part_info_int_;

/* /part_info_dp_1/ particle_info.h:67-69 */
extern struct {
  double charge[NPART], mass[NPART], Ti[NPART], vt0[NPART], k_eps0, B_scale[5], B0[5],
      dL[5], x0[5], y0[5], jne, RNeta, THm, phi0, xl_pow, xr_pow, chmfrd_array[4], yl_pow,
      yr_pow, th_B;
} part_info_dp_1_;

/* /part_info_dp_2/ particle_info.h:70-72 */
extern struct {
  double Pabs[4], n0, ngas, Nm[NPART], dt, xmax, ymax, zmax, x_load, nu_h, Lg, Lh, xg1,
      xg2, Bmax, nudt, hz, xa, dr, tseq, tseq_init, tseq_final, gam_sec, xg_sec, fEk, Lgrd;
} part_info_dp_2_;

/* /constants/ constants.h:6 */
extern struct { double qe, c, pi, amu, eps0; } constants_;

/* reference routines (Fortran, trailing underscore, all by reference) */
double ran2_(int32_t *);
void load_gauss_(double *vx, double *vy, double *vt, double *rnd);
void part_mover_(int32_t *istep, int32_t *n, int32_t *na, double *h, double *Ei, double *Bi,
                 double *p_mac, double *P_loss, double *vxp, int32_t *bcnd, int32_t *nmax,
                 int32_t *ntype, int32_t *ngrid, int8_t *flag_dead, int32_t *nproc,
                 int32_t *np_tot, double *phi, double *uB, int32_t *iproc, int32_t *ptype,
                 int32_t *iseed, double *cnt_dead, double *sum_q, double *sum_q_y,
                 double *phase_space, int32_t *ip_ps, int32_t *nmax_ps, int32_t *n_B,
                 double *h_B, double *ss2D);
void charge_deposition_(int32_t *n, double *h, double *vxp, int32_t *nmax, int32_t *ntype,
                        double *kq, double *np, int32_t *nproc, int32_t *np_tot,
                        double *sum_dEk, int32_t *Nh, int32_t *iproc, int32_t *ptype,
                        int32_t *n_icp);
void calc_efield_(int32_t *n, double *h, double *phi, double *Ei, int32_t *bcnd);
void eheating_(double *vxp, int32_t *nmax, int32_t *ntype, int32_t *nproc, int32_t *iseed,
               int32_t *np_tot, double *vt, int32_t *n_icp, int32_t *iproc, double *P_loss,
               double *Pabs_cor);
void part_flux_emitter_(int32_t *istep, double *vxp, int32_t *n, double *h, int32_t *ntype,
                        int32_t *nmax, int32_t *bcnd, double *ss2D, int32_t *nproc, int32_t *iseed,
                        int32_t *np_tot, int32_t *nproc_mpi, int32_t *mpi_rank, int32_t *cnt_b,
                        int32_t *n_icp, double *P_loss);
void omp_set_num_threads(int);
void part_sorting_(double *vxp, double *sorting, int32_t *nmax, int32_t *ntype, int32_t *n,
                   double *h, int32_t *Plist_thd, int32_t *pl_max, int32_t *nproc,
                   int32_t *np_tot, int32_t *ptype);

static FILE *fin, *fout;
static int64_t ni, nd, I[64];
static double D[64];

static void *rd(size_t bytes) {
  void *p = calloc(bytes ? bytes : 1, 1);
  if (bytes && fread(p, 1, bytes, fin) != bytes) { fprintf(stderr, "short input\n"); exit(2); }
  return p;
}
static void wr(const void *p, size_t bytes) { fwrite(p, 1, bytes, fout); }

static void set_constants(void) {
  constants_.qe = D[0];
  constants_.amu = D[1];
  constants_.pi = 4.0 * atan(1.0); /* main.f90:123 */
  constants_.c = 2.99792458e8;
  constants_.eps0 = 8.854187817e-12;
}

int main(int argc, char **argv) {
  if (argc != 4) { fprintf(stderr, "usage: %s cmd in out\n", argv[0]); return 2; }
  fin = fopen(argv[2], "rb");
  fout = fopen(argv[3], "wb");
  if (!fin || !fout) { fprintf(stderr, "cannot open files\n"); return 2; }
  if (fread(&ni, 8, 1, fin) != 1 || ni > 64 || fread(I, 8, (size_t)ni, fin) != (size_t)ni) return 3;
  if (fread(&nd, 8, 1, fin) != 1 || nd > 64 || fread(D, 8, (size_t)nd, fin) != (size_t)nd) return 3;
  memset(&part_info_int_, 0, sizeof(part_info_int_));
  memset(&part_info_dp_1_, 0, sizeof(part_info_dp_1_));
  memset(&part_info_dp_2_, 0, sizeof(part_info_dp_2_));
  const char *cmd = argv[1];

  if (!strcmp(cmd, "ran2")) { /* I: seed, count */
    int32_t s = (int32_t)I[0];
    for (int64_t k = 0; k < I[1]; ++k) { double r = ran2_(&s); wr(&r, 8); }
    wr(&s, 4);
  } else if (!strcmp(cmd, "gauss")) { /* I: count; D: qe, amu, vt; arrays: rnd[2*count] */
    set_constants();
    double vt = D[2];
    double *rnd = rd(16 * (size_t)I[0]);
    for (int64_t k = 0; k < I[0]; ++k) {
      double vx, vy;
      load_gauss_(&vx, &vy, &vt, rnd + 2 * k);
      wr(&vx, 8);
      wr(&vy, 8);
    }
  } else if (!strcmp(cmd, "mover")) {
    /* I: n1 n2 nB1 nB2 N ngrid flag_B flag_gridB flag_B_pos tag_neg
     * D: qe amu h1 h2 hB1 hB2 dt charge mass Nm zmax xmax ymax */
    set_constants();
    int32_t n[3] = {(int32_t)I[0], (int32_t)I[1], 1}, na[3] = {n[0], n[1], 1};
    int32_t nB[2] = {(int32_t)I[2], (int32_t)I[3]};
    int32_t N = (int32_t)I[4], nmax = N > 0 ? N : 1, ngrid = (int32_t)I[5];
    part_info_int_.flag_B = (int32_t)I[6];
    part_info_int_.flag_gridB = (int32_t)I[7];
    part_info_int_.flag_B_pos = (int32_t)I[8];
    part_info_int_.tag_neg = I[9] ? 1 : 0;
    part_info_int_.size_na = 1;
    part_info_int_.opt_inj = 0;  /* refluxing (needs part_injection.f90) off */
    part_info_int_.flag_nmn = 0;
    part_info_int_.plt_src = 0;
    double h[3] = {D[2], D[3], 1.0}, hB[2] = {D[4], D[5]};
    part_info_dp_2_.dt = D[6];
    part_info_dp_1_.charge[0] = D[7];
    part_info_dp_1_.mass[0] = D[8];
    part_info_dp_2_.Nm[0] = D[9];
    part_info_dp_2_.zmax = D[10];
    part_info_dp_2_.xmax = D[11];
    part_info_dp_2_.ymax = D[12];
    part_info_dp_2_.hz = -1.0; /* 2D: no z losses */
    size_t nodes = (size_t)(n[0] + 3) * (n[1] + 3), bnodes = (size_t)(nB[0] + 3) * (nB[1] + 3);
    double *Ei = rd(16 * nodes), *Bi = rd(32 * bnodes);
    int32_t *bcnd = rd(4 * nodes);
    double *vxp = rd(48 * (size_t)nmax);
    int8_t *fd = rd((size_t)nmax);
    double *phi = rd(8 * nodes);
    double *p_mac = calloc(2 * (size_t)(ngrid + 1), 8), P_loss[4] = {0, 0, 0, 0};
    double *uB = calloc((size_t)(na[0] + 1) * (na[1] + 1), 8), cnt_dead = 0, sum_q = 0;
    double *sum_q_y = calloc((size_t)(ngrid > 0 ? ngrid : 1) * (n[1] + 3), 8);
    double ps[4];
    double *ss2D = calloc(2 * nodes, 8);
    int32_t istep = 1, ntype = 1, nproc = 1, iproc = 1, ptype = 1, iseed = 123456, ip_ps = 0,
            nmax_ps = 1, np_tot = N;
    part_mover_(&istep, n, na, h, Ei, Bi, p_mac, P_loss, vxp, bcnd, &nmax, &ntype, &ngrid, fd,
                &nproc, &np_tot, phi, uB, &iproc, &ptype, &iseed, &cnt_dead, &sum_q, sum_q_y,
                ps, &ip_ps, &nmax_ps, nB, hB, ss2D);
    int64_t npt = np_tot;
    wr(&npt, 8);
    wr(vxp, 48 * (size_t)np_tot);
    wr(P_loss, 32);
    wr(p_mac, 16 * (size_t)(ngrid + 1));
    wr(&cnt_dead, 8);
    wr(ss2D, 16 * nodes);
    wr(fd, (size_t)nmax);
  } else if (!strcmp(cmd, "heat")) {
    /* I: N n_icp flag_c flag_gpp eheat_type seed
     * D: qe amu xa dr yl_pow yr_pow ymax nudt dt Pabs_cor Nm mass; arrays: vxp, vt */
    set_constants();
    int32_t N = (int32_t)I[0], nmax = N > 0 ? N : 1, n_icp = (int32_t)I[1];
    part_info_int_.flag_c = (int32_t)I[2];
    part_info_int_.flag_gpp = (int32_t)I[3];
    part_info_int_.eheat_type = (int32_t)I[4];
    int32_t iseed = (int32_t)I[5];
    part_info_dp_2_.xa = D[2];
    part_info_dp_2_.dr = D[3];
    part_info_dp_1_.yl_pow = D[4];
    part_info_dp_1_.yr_pow = D[5];
    part_info_dp_2_.ymax = D[6];
    part_info_dp_2_.nudt = D[7];
    part_info_dp_2_.dt = D[8];
    double Pabs_cor = D[9];
    part_info_dp_2_.Nm[0] = D[10];
    part_info_dp_1_.mass[0] = D[11];
    double *vxp = rd(48 * (size_t)nmax), *vt = rd(8 * (size_t)n_icp);
    double P_loss[4] = {0, 0, 0, 0};
    int32_t ntype = 1, nproc = 1, iproc = 1, np_tot = N;
    eheating_(vxp, &nmax, &ntype, &nproc, &iseed, &np_tot, vt, &n_icp, &iproc, P_loss, &Pabs_cor);
    wr(vxp, 48 * (size_t)N);
    wr(&P_loss[1], 8);
    wr(&iseed, 4);
  } else if (!strcmp(cmd, "emit")) {
    /* I: n1 n2 ntype nmax ptype_fx tag_neg tag_b flag_chmfrd n_holes n_icp ns_Hm seed
     * D: qe amu h1 h2 jne Lg Lh xg1 THm ymax zmax dt chmfrd_array(4) Nm(ntype) mass(ntype)
     * arrays: bcnd, np_tot(ntype), vxp(6,nmax,ntype).  One OpenMP thread: the
     * partition index of part_flux_OMP is the thread number. */
    set_constants();
    omp_set_num_threads(1);
    int32_t n[3] = {(int32_t)I[0], (int32_t)I[1], 1};
    int32_t ntype = (int32_t)I[2], nmax = (int32_t)I[3];
    part_info_int_.ptype_fx = (int32_t)I[4];
    part_info_int_.tag_neg = (int32_t)I[5];
    part_info_int_.tag_b = (int32_t)I[6];
    part_info_int_.flag_chmfrd = (int32_t)I[7];
    part_info_int_.n_holes = (int32_t)I[8];
    int32_t n_icp = (int32_t)I[9];
    part_info_int_.ns_Hm = (int32_t)I[10];
    int32_t iseed = (int32_t)I[11];
    double h[3] = {D[2], D[3], 0.0};
    part_info_dp_1_.jne = D[4];
    part_info_dp_2_.Lg = D[5];
    part_info_dp_2_.Lh = D[6];
    part_info_dp_2_.xg1 = D[7];
    part_info_dp_1_.THm = D[8];
    part_info_dp_2_.ymax = D[9];
    part_info_dp_2_.zmax = D[10];
    part_info_dp_2_.dt = D[11];
    for (int k = 0; k < 4; ++k) part_info_dp_1_.chmfrd_array[k] = D[12 + k];
    for (int k = 0; k < ntype; ++k) {
      part_info_dp_2_.Nm[k] = D[16 + k];
      part_info_dp_1_.mass[k] = D[16 + ntype + k];
    }
    const size_t nodes = (size_t)(n[0] + 3) * (n[1] + 3);
    int32_t *bcnd = rd(4 * nodes), *np_tot = rd(4 * (size_t)ntype);
    double *vxp = rd(48 * (size_t)nmax * ntype);
    double *ss2D = calloc(2 * nodes * ntype, 8), *P_loss = calloc(4 * (size_t)ntype, 8);
    int32_t cnt_b[2] = {0, 0}, nproc = 1, nproc_mpi = 1, mpi_rank = 0, istep = 2;
    part_flux_emitter_(&istep, vxp, n, h, &ntype, &nmax, bcnd, ss2D, &nproc, &iseed, np_tot, &nproc_mpi,
                       &mpi_rank, cnt_b, &n_icp, P_loss);
    wr(vxp, 48 * (size_t)nmax * ntype);
    wr(np_tot, 4 * (size_t)ntype);
    wr(ss2D, 16 * nodes * ntype);
    wr(cnt_b, 8);
    wr(P_loss, 32 * (size_t)ntype);
    wr(&iseed, 4);
    wr(&part_info_dp_2_.Lgrd, 8);
  } else if (!strcmp(cmd, "bench")) {
    /* I: n1 n2 nB1 nB2 N ngrid flag_B flag_gridB flag_B_pos tag_neg reps
     * D: as "mover"; arrays: Ei, Bi, bcnd, vxp, flag_dead, phi, kq.
     * reps x (part_mover + charge_deposition), main.f90:1041-1086 with one
     * thread; writes the seconds spent and the final particle count. */
    set_constants();
    int32_t n[3] = {(int32_t)I[0], (int32_t)I[1], 1}, na[3] = {n[0], n[1], 1};
    int32_t nB[2] = {(int32_t)I[2], (int32_t)I[3]};
    int32_t N = (int32_t)I[4], nmax = N > 0 ? N : 1, ngrid = (int32_t)I[5];
    const int64_t reps = I[10];
    part_info_int_.flag_B = (int32_t)I[6];
    part_info_int_.flag_gridB = (int32_t)I[7];
    part_info_int_.flag_B_pos = (int32_t)I[8];
    part_info_int_.tag_neg = I[9] ? 1 : 0;
    part_info_int_.size_na = 1;
    double h[3] = {D[2], D[3], 1.0}, hB[2] = {D[4], D[5]};
    part_info_dp_2_.dt = D[6];
    part_info_dp_1_.charge[0] = D[7];
    part_info_dp_1_.mass[0] = D[8];
    part_info_dp_2_.Nm[0] = D[9];
    part_info_dp_2_.zmax = D[10];
    part_info_dp_2_.xmax = D[11];
    part_info_dp_2_.ymax = D[12];
    part_info_dp_2_.hz = -1.0;
    size_t nodes = (size_t)(n[0] + 3) * (n[1] + 3), bnodes = (size_t)(nB[0] + 3) * (nB[1] + 3);
    double *Ei = rd(16 * nodes), *Bi = rd(32 * bnodes);
    int32_t *bcnd = rd(4 * nodes);
    double *vxp = rd(48 * (size_t)nmax);
    int8_t *fd = rd((size_t)nmax);
    double *phi = rd(8 * nodes), *kq = rd(8 * nodes);
    double *p_mac = calloc(2 * (size_t)(ngrid + 1), 8), P_loss[4] = {0, 0, 0, 0};
    double *uB = calloc((size_t)(na[0] + 1) * (na[1] + 1), 8), cnt_dead = 0, sum_q = 0;
    double *sum_q_y = calloc((size_t)(ngrid > 0 ? ngrid : 1) * (n[1] + 3), 8);
    double ps[4], sum_dEk = 0;
    double *ss2D = calloc(2 * nodes, 8), *np = calloc(nodes, 8);
    int32_t istep = 1, ntype = 1, nproc = 1, iproc = 1, ptype = 1, iseed = 123456, ip_ps = 0,
            nmax_ps = 1, np_tot = N, Nh = 0, n_icp = 1;
    struct timespec t0, t1;
    clock_gettime(CLOCK_MONOTONIC, &t0);
    for (int64_t k = 0; k < reps; ++k) {
      part_mover_(&istep, n, na, h, Ei, Bi, p_mac, P_loss, vxp, bcnd, &nmax, &ntype, &ngrid, fd,
                  &nproc, &np_tot, phi, uB, &iproc, &ptype, &iseed, &cnt_dead, &sum_q, sum_q_y,
                  ps, &ip_ps, &nmax_ps, nB, hB, ss2D);
      memset(np, 0, 8 * nodes);
      charge_deposition_(n, h, vxp, &nmax, &ntype, kq, np, &nproc, &np_tot, &sum_dEk, &Nh, &iproc,
                         &ptype, &n_icp);
    }
    clock_gettime(CLOCK_MONOTONIC, &t1);
    double sec = (double)(t1.tv_sec - t0.tv_sec) + 1e-9 * (double)(t1.tv_nsec - t0.tv_nsec);
    int64_t npt = np_tot;
    wr(&sec, 8);
    wr(&npt, 8);
  } else if (!strcmp(cmd, "ompbench")) {
    /* As "bench" with I[11] = nthreads: the particles split into nthreads
     * partitions vxp(6,nmax,1,nproc), and inside one OpenMP region per step
     * thread iproc calls part_mover and charge_deposition on partition iproc
     * with its own density np(:,:,1,iproc), the way main.f90:1024-1089 runs
     * them.  Writes the seconds spent and the final particle count. */
    set_constants();
    int32_t n[3] = {(int32_t)I[0], (int32_t)I[1], 1}, na[3] = {n[0], n[1], 1};
    int32_t nB[2] = {(int32_t)I[2], (int32_t)I[3]};
    const int64_t N = I[4], reps = I[10];
    const int32_t T = (int32_t)I[11], ngrid = (int32_t)I[5];
    part_info_int_.flag_B = (int32_t)I[6];
    part_info_int_.flag_gridB = (int32_t)I[7];
    part_info_int_.flag_B_pos = (int32_t)I[8];
    part_info_int_.tag_neg = I[9] ? 1 : 0;
    part_info_int_.size_na = 1;
    double h[3] = {D[2], D[3], 1.0}, hB[2] = {D[4], D[5]};
    part_info_dp_2_.dt = D[6];
    part_info_dp_1_.charge[0] = D[7];
    part_info_dp_1_.mass[0] = D[8];
    part_info_dp_2_.Nm[0] = D[9];
    part_info_dp_2_.zmax = D[10];
    part_info_dp_2_.xmax = D[11];
    part_info_dp_2_.ymax = D[12];
    part_info_dp_2_.hz = -1.0;
    size_t nodes = (size_t)(n[0] + 3) * (n[1] + 3), bnodes = (size_t)(nB[0] + 3) * (nB[1] + 3);
    double *Ei = rd(16 * nodes), *Bi = rd(32 * bnodes);
    int32_t *bcnd = rd(4 * nodes);
    double *src = rd(48 * (size_t)(N > 0 ? N : 1));
    int8_t *fd_in = rd((size_t)(N > 0 ? N : 1));
    double *phi = rd(8 * nodes), *kq = rd(8 * nodes);
    (void)fd_in;
    int32_t nmax = (int32_t)((N + T - 1) / T);
    if (nmax < 1) nmax = 1;
    double *vxp = calloc((size_t)6 * nmax * T, 8);
    int8_t *fd = calloc((size_t)nmax * T, 1);
    int32_t *np_tot = calloc((size_t)T, 4), *ip_ps = calloc((size_t)T, 4), *Nh = calloc((size_t)T, 4);
    double *p_mac = calloc(2 * (size_t)(ngrid + 1) * T, 8), *P_loss = calloc(4 * (size_t)T, 8);
    double *uB = calloc((size_t)(na[0] + 1) * (na[1] + 1), 8), *cnt_dead = calloc((size_t)T, 8);
    double *sum_q = calloc((size_t)T, 8), *sum_dEk = calloc((size_t)T, 8);
    double *sum_q_y = calloc((size_t)(ngrid > 0 ? ngrid : 1) * (n[1] + 3) * T, 8);
    double *ps = calloc(4 * (size_t)T, 8);
    double *ss2D = calloc(2 * nodes * T, 8), *np = calloc(nodes * T, 8);
    for (int32_t t = 0; t < T; ++t) {
      const int64_t a = N * t / T, b = N * (t + 1) / T;
      memcpy(vxp + (size_t)6 * nmax * t, src + 6 * a, 48 * (size_t)(b - a));
      np_tot[t] = (int32_t)(b - a);
    }
    struct timespec t0, t1;
    clock_gettime(CLOCK_MONOTONIC, &t0);
    for (int64_t k = 0; k < reps; ++k) {
#pragma omp parallel num_threads(T)
      {
        int32_t iproc = omp_get_thread_num() + 1, istep = 1, ntype = 1, ptype = 1, nproc = T,
                iseed = 123456 * iproc, nmax_ps = 1, ng = ngrid, nmx = nmax, n_icp = 1;
        part_mover_(&istep, n, na, h, Ei, Bi, p_mac, P_loss, vxp, bcnd, &nmx, &ntype, &ng, fd, &nproc, np_tot,
                    phi, uB, &iproc, &ptype, &iseed, cnt_dead, sum_q, sum_q_y, ps, ip_ps, &nmax_ps, nB, hB, ss2D);
        memset(np + nodes * (iproc - 1), 0, 8 * nodes);
        charge_deposition_(n, h, vxp, &nmx, &ntype, kq, np, &nproc, np_tot, sum_dEk, Nh, &iproc, &ptype,
                           &n_icp);
      }
    }
    clock_gettime(CLOCK_MONOTONIC, &t1);
    double sec = (double)(t1.tv_sec - t0.tv_sec) + 1e-9 * (double)(t1.tv_nsec - t0.tv_nsec);
    int64_t npt = 0;
    for (int32_t t = 0; t < T; ++t) npt += np_tot[t];
    wr(&sec, 8);
    wr(&npt, 8);
  } else if (!strcmp(cmd, "deposit")) {
    /* I: n1 n2 N; D: qe amu h1 h2 Nm zmax; arrays: vxp, kq */
    set_constants();
    int32_t n[3] = {(int32_t)I[0], (int32_t)I[1], 1};
    int32_t N = (int32_t)I[2], nmax = N > 0 ? N : 1;
    double h[3] = {D[2], D[3], 1.0};
    part_info_dp_2_.Nm[0] = D[4];
    part_info_dp_2_.zmax = D[5];
    part_info_dp_2_.Pabs[0] = 0.0;
    size_t nodes = (size_t)(n[0] + 3) * (n[1] + 3);
    double *vxp = rd(48 * (size_t)nmax), *kq = rd(8 * nodes);
    double *np = calloc(nodes, 8), sum_dEk = 0;
    int32_t Nh = 0, ntype = 1, nproc = 1, iproc = 1, ptype = 1, n_icp = 1, np_tot = N;
    charge_deposition_(n, h, vxp, &nmax, &ntype, kq, np, &nproc, &np_tot, &sum_dEk, &Nh, &iproc,
                       &ptype, &n_icp);
    wr(np, 8 * nodes);
  } else if (!strcmp(cmd, "efield")) {
    /* I: n1 n2; D: qe amu h1 h2; arrays: phi, bcnd, Ei (initial) */
    set_constants();
    int32_t n[3] = {(int32_t)I[0], (int32_t)I[1], 1};
    double h[3] = {D[2], D[3], 1.0};
    size_t nodes = (size_t)(n[0] + 3) * (n[1] + 3);
    double *phi = rd(8 * nodes);
    int32_t *bcnd = rd(4 * nodes);
    double *Ei = rd(16 * nodes);
    calc_efield_(n, h, phi, Ei, bcnd);
    wr(Ei, 16 * nodes);
  } else if (!strcmp(cmd, "sort")) {
    /* I: n1 n2 N; D: qe amu h1 h2; arrays: vxp */
    set_constants();
    int32_t n[3] = {(int32_t)I[0], (int32_t)I[1], 1};
    int32_t N = (int32_t)I[2], nmax = N > 0 ? N : 1;
    double h[3] = {D[2], D[3], 1.0};
    double *vxp = rd(48 * (size_t)nmax), *work = calloc(6 * (size_t)nmax, 8);
    int32_t pl_max = n[0] + n[0] * n[1] + 1;
    int32_t *Plist = calloc((size_t)pl_max + 1, 4);
    int32_t ntype = 1, nproc = 1, ptype = 1, np_tot = N;
    part_sorting_(vxp, work, &nmax, &ntype, n, h, Plist, &pl_max, &nproc, &np_tot, &ptype);
    int64_t npt = np_tot;
    wr(&npt, 8);
    wr(vxp, 48 * (size_t)np_tot);
    wr(Plist, 4 * ((size_t)pl_max + 1));
  } else {
    fprintf(stderr, "unknown command %s\n", cmd);
    return 2;
  }
  fclose(fout);
  return 0;
}
