/*
 * lepic.h -- C ABI of the MI355X (gfx950) implementation of LePIC's
 * per-timestep particle/field hot path.
 *
 * The reference calls its hot-path routines as implicit-interface Fortran
 * externals (all arguments by reference, arrays by base address, globals in
 * COMMON blocks).  This library keeps the same routine split, one entry point
 * per reference routine, but state lives on the GPU inside an opaque context:
 * the caller uploads particles and static fields once, then calls the phase
 * entry points each step in the reference's order (main.f90:689-1090), and
 * downloads only for diagnostics or checks.  Every function returns 0 or a
 * negative LEPIC_ERR_* code; lepic_last_error() gives the message (the
 * reference instead stops through stop_calculation, restart.f90:263-279).
 *
 * Array layouts are the reference's Fortran layouts (column-major):
 *   node arrays   A(0:n1+2,0:n2+2)      -> A[iy*(n1+3)+ix]
 *   rhs(n1+1,n2+1)                       -> rhs[(iy-1)*(n1+1)+(ix-1)]
 *   Ei(2,0:n1+2,0:n2+2)                  -> Ei[2*node+c]
 *   Bi(4,0:nB1+2,0:nB2+2)                -> Bi[4*bnode+c]
 *   vxp(6,nmax) for one species          -> vxp[6*i+c]  (x,y,z,vx,vy,vz)
 * Species indices are 1-based like ptype.  Not re-entrant per context; one
 * host thread per context.  No OpenMP inside.
 */
#ifndef LEPIC_H
#define LEPIC_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct lepic_ctx lepic_ctx;

enum {
  LEPIC_OK = 0,
  LEPIC_ERR_ARG = -1,          /* bad argument or size */
  LEPIC_ERR_UNSUPPORTED = -2,  /* option outside the library's supported hot path */
  LEPIC_ERR_HIP = -3,          /* HIP runtime error */
  LEPIC_ERR_CAPACITY = -4,     /* particle capacity exceeded (nmax, collisions.f90:187-191) */
  LEPIC_ERR_DIVERGED = -5,     /* MG divergence (iterative_solver.f90:39-51) */
  LEPIC_ERR_MCC = -6,          /* a collisions.f90 abort (Pmax>1, Ee<0, energy above table) */
  LEPIC_ERR_STATE = -7         /* call out of order (e.g. no reactions set) */
};

enum { LEPIC_FP64 = 0, LEPIC_MIXED = 1 };
enum { LEPIC_MCC_CHECK = 0, LEPIC_MCC_PROD = 1 };
enum { LEPIC_START_PREV = 0, LEPIC_START_COLD = 1, LEPIC_START_GIVEN = 2, LEPIC_START_NET = 3 };
enum {
  LEPIC_GRID_PHI = 0,   /* phi(0:n1+2,0:n2+2) */
  LEPIC_GRID_RHS = 1,   /* rhs(n1+1,n2+1) */
  LEPIC_GRID_E = 2,     /* Ei(2,0:n1+2,0:n2+2) */
  LEPIC_GRID_NP = 3,    /* np(0:n1+2,0:n2+2,ptype): deposited density */
  LEPIC_GRID_NPRED = 4, /* np_red(0:n1+2,0:n2+2,ptype) */
  LEPIC_GRID_SINK = 5,  /* ss2D(2,:,:,ptype): sink counts (wall losses, killed reactants) */
  LEPIC_GRID_SRC = 6    /* ss2D(1,:,:,ptype): source counts (created, injected, emitted) */
};

typedef struct {
  int32_t n1, n2;         /* cells (n(1), n(2)) */
  double h1, h2;          /* cell size h(1), h(2) in m */
  double zmax;            /* third-dimension length (1 m when hz<0) */
  double dt;              /* time step (s) */
  double eps0;            /* k_eps0 * eps0 (main.f90:161) */
  double qe;              /* elementary charge as the reference holds it */
  int32_t ng;             /* MG levels; 1 = red-black SOR only */
  int32_t ngrid;          /* number of wall labels (p_mac(:,:,0:ngrid)) */
  int32_t nspecies;       /* charged species carried as particles (ntype) */
  int32_t precision;      /* LEPIC_FP64 or LEPIC_MIXED */
  int32_t flag_B_pos;     /* positive ions not magnetised (part_expmover.f90:185) */
  int32_t tag_neg;        /* tagged negative-ion species (1-based) or 0 */
  int32_t track_z;        /* mixed precision only: keep the virtual z coordinate */
  int32_t mg_tail_nodes;  /* MG levels with at most this many nodes ((n1+1)(n2+1)) run in
                             one single-workgroup kernel; 0 = library default (5000) */
} lepic_grid_desc;

typedef struct {
  double charge;          /* C */
  double mass;            /* kg */
  double Nm;              /* macroparticle weight */
  int64_t capacity;       /* maximum particle count (nmax) */
} lepic_species_desc;

typedef struct {
  int32_t vcycles;        /* V-cycles done (the reference's niter-1) */
  int32_t status;         /* 0 converged, 1 hit ncycl */
  double res;             /* final residual / alpha (iterative_solver.f90:36) */
  double ktot;            /* SOR-equivalent work (mg.f90:227-229) */
  double res0;            /* residual / alpha of the starting guess (-1 if not computed) */
  int32_t start_used;     /* LEPIC_START_* actually used */
  int32_t nlog;           /* entries written to the sweep log */
  double res0_alt;        /* LEPIC_START_NET: residual / alpha of the start not used (-1 otherwise) */
} lepic_solve_info;

typedef struct {
  int64_t Nc[8];          /* candidates per species (index ptype-1) */
  double nu_max[8];       /* null collision frequency per species */
  int64_t events[64];     /* real events per reaction (index icol-1) */
  int64_t created[8];     /* particles appended per species */
  int64_t missed;         /* production mode: candidates dropped after claim retries */
  int32_t seed;           /* check mode: ran2 state after the call */
  int32_t reserved;
} lepic_mcc_info;

/* ---- lifecycle ---- */
/* replaces the allocations of main.f90:192-214 */
int lepic_create(const lepic_grid_desc *g, const lepic_species_desc *sp, lepic_ctx **out);
int lepic_destroy(lepic_ctx *ctx);
const char *lepic_last_error(const lepic_ctx *ctx);
int lepic_sync(lepic_ctx *ctx);
/* HIP device index of the context and its stream handle (hipStream_t) */
int lepic_stream(lepic_ctx *ctx, void **stream);
/* Measurement hook, no counterpart in the reference: when enabled, each
 * particle-mover kernel launch is bracketed by HIP events on the context's
 * stream (the density clearing before it is not); lepic_get_kernel_ms waits
 * for the last mover kernel of ptype and returns its duration. */
int lepic_kernel_timing(lepic_ctx *ctx, int32_t enable);
int lepic_get_kernel_ms(lepic_ctx *ctx, int32_t ptype, float *ms);

/* ---- static inputs ---- */
/* bcnd and the Dirichlet values of generate_boundary.f90 (phi at walls). */
int lepic_set_boundary(lepic_ctx *ctx, const int32_t *bcnd, const double *phi);
/* Bi(4,0:nB1+2,0:nB2+2) with |B| in component 4 (main.f90:361-422); nB1 == 1
 * means a constant field Bi(1:3,1,1). */
int lepic_set_bfield(lepic_ctx *ctx, int32_t nB1, int32_t nB2, double hB1, double hB2,
                     const double *Bi);

/* ---- particle transfer (reference vxp(6,N) layout) ---- */
int lepic_upload_particles(lepic_ctx *ctx, int32_t ptype, const double *vxp, int64_t np);
int lepic_download_particles(lepic_ctx *ctx, int32_t ptype, double *vxp, int64_t cap,
                             int64_t *np);
/* particles first .. first+count-1 (0-based, the species' order) only, e.g.
 * a sample of a large run; no counterpart routine in the reference, whose
 * diagnostics read vxp(:,i) in place */
int lepic_download_particles_range(lepic_ctx *ctx, int32_t ptype, int64_t first, int64_t count,
                                   double *vxp);
int lepic_get_count(lepic_ctx *ctx, int32_t ptype, int64_t *np);
/* flag_dead(nmax,ptype) as int8, for the first np particles */
int lepic_set_flag_dead(lepic_ctx *ctx, int32_t ptype, const int8_t *flags, int64_t np);
int lepic_get_flag_dead(lepic_ctx *ctx, int32_t ptype, int8_t *flags, int64_t np);

/* ---- grid transfer ---- */
int lepic_get_grid(lepic_ctx *ctx, int32_t which, int32_t ptype, double *out);
int lepic_set_grid(lepic_ctx *ctx, int32_t which, int32_t ptype, const double *in);

/* ---- phases ---- */
/* part_mover (part_expmover.f90:1-517): Boris/leap-frog push, wall loss,
 * order-preserving compaction, removal of particles flagged dead; with
 * deposit=1 the survivors' charge is also deposited (charge_deposition,
 * part_expmover.f90:520-610) into a zeroed np(:,:,ptype). */
int lepic_part_mover(lepic_ctx *ctx, int32_t ptype, int32_t deposit);
/* charge_deposition (part_expmover.f90:520-610): zero np(:,:,ptype), deposit */
int lepic_charge_deposition(lepic_ctx *ctx, int32_t ptype);
/* mover tallies since the last reset: P_loss(1:4), p_mac(1:2,0:ngrid), cnt_dead */
int lepic_get_mover_tally(lepic_ctx *ctx, int32_t ptype, double *P_loss, double *p_mac,
                          double *cnt_dead);
int lepic_reset_tallies(lepic_ctx *ctx);
/* calc_rhs (calc_rho.f90:37-77): rhs = -sum_ptype charge*np */
int lepic_calc_rhs(lepic_ctx *ctx);
/* calc_Efield (Efield.f90:1-128) */
int lepic_calc_efield(lepic_ctx *ctx);
/* iterative_solver (iterative_solver.f90:1-69) with mg/sor_rb; start selects
 * the initial phi (previous, zero, or the grid set by lepic_set_grid).
 * sweep_log (optional, cap entries) receives ksor per level visit. */
int lepic_iterative_solver(lepic_ctx *ctx, double eps, int32_t ncycl, int32_t start,
                           lepic_solve_info *info, int32_t *sweep_log, int32_t log_cap);
/* part_sorting (sorting.f90:92-251): stable counting sort by cell; Plist
 * (optional, n1+n1*n2+2 int32) receives the cell-end list, which is only
 * computed when Plist is not NULL (the device MCC does not need it). */
int lepic_part_sorting(lepic_ctx *ctx, int32_t ptype, int32_t *Plist);
/* lepic_part_sorting without any allocation: no counterpart in the
 * reference.  For callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * Same kernels and results as lepic_part_sorting, with the scratch in a
 * static device array of the library (about 21 MiB per device, held until
 * the process ends): species of at most 512 segments of 1024 slots (524288
 * particles) and, when Plist is requested, grids with n1 + n1*n2 + 2 <=
 * 1052672 (up to 1024 x 1024 cells); larger inputs return
 * LEPIC_ERR_CAPACITY.  Calls on one GPU share that array: calls that
 * overlap on one GPU (two host threads, two contexts) race. */
int lepic_part_sorting_noalloc(lepic_ctx *ctx, int32_t ptype, int32_t *Plist);
/* dens_red (calc_rho.f90:134-199): np_red and np_mx per charged species */
int lepic_dens_red(lepic_ctx *ctx, double *np_mx);

/* ---- collisions ---- */
/* Cross-section tables already built on the union grid (read_reactions.f90:
 * 511-697): sig_Er[npt], sig[ncol*npt] (reaction-major), col_info[ncol*10],
 * Eth[ncol], sigv_mx[(ntype_all+1)*ncol] (row ptype), p_ncol[ntype_all+1],
 * sig_list[(ntype_all+1)*ncol] (1-based reaction ids), and per species
 * (1..ntype_all, index 0 unused): charge, mass (neutrals negative), Nm, vt0,
 * np_mx (neutral densities; charged entries come from lepic_dens_red) and
 * nu_uplim.  Regions: n_rg boxes [ixl,ixr] x [iyl,iyr] for the rate tallies. */
typedef struct {
  int32_t ntype_all, ncol, npt, ns_coll, n_cell, n_rg, iyl_rg, iyr_rg;
  int32_t ixl_rg[5], ixr_rg[5];
  int32_t flag_ionrep, reserved;
  const double *sig_Er, *sig, *Eth, *sigv_mx;
  const int32_t *col_info, *p_ncol, *sig_list;
  const double *charge, *mass, *Nm, *vt0, *np_mx, *nu_uplim;
} lepic_reactions;
int lepic_set_reactions(lepic_ctx *ctx, const lepic_reactions *r);
/* collisions (collisions.f90:1-207, collision_OMP :209-1040).  mode
 * LEPIC_MCC_CHECK replays the reference's ran2 stream from *iseed in order
 * (one GPU lane; identical events); LEPIC_MCC_PROD draws from a counter-based
 * generator keyed by (seed, it) with one lane per candidate. */
int lepic_collisions(lepic_ctx *ctx, int32_t it, int32_t mode, uint64_t seed, int32_t *iseed,
                     lepic_mcc_info *info);
/* cnt_col(3,ncol,n_rg) tallies since the last reset (collisions.f90:535-543, 1026-1033) */
int lepic_get_cnt_col(lepic_ctx *ctx, double *cnt_col);
/* P_loss(3,ptype) energy change of collisions since the last reset */
int lepic_get_mcc_ploss(lepic_ctx *ctx, double *P_loss3);

/* ---- electron heating (eheating, e_heating.f90:1-112; called at main.f90:1033-1037
 *      every ns_heat steps before the push) ----
 * Electrons (species 1) inside the slit (flag_c = 0: |x-xa| <= dr,
 * yl_pow <= y <= yr_pow) or the disk (flag_c = 1: radius dr about
 * (xa, ymax/2)) are heated with probability nu*dt, nu = nudt/dt or, with
 * flag_gpp = 1, (nudt/dt)*9*Pabs_cor*exp(-9 r^2/dr^2): a Maxwellian kick of
 * thermal speed vt(i_icp), i_icp = int(n_icp*y/ymax)+1, added to
 * (eheat_type = 0) or replacing (1) the velocity.  *dP receives the energy
 * change 0.5*Nm*m*sum(v^2 - v_old^2), the reference's P_loss(2,1).
 * LEPIC_MCC_CHECK replays the reference's ran2 stream from *iseed in
 * particle order (FP64 contexts); LEPIC_MCC_PROD draws per particle slot
 * from a counter-based generator keyed by (seed, it). */
typedef struct {
  int32_t flag_c, flag_gpp, eheat_type, n_icp;   /* n_icp <= 64 */
  double xa, dr, yl_pow, yr_pow, nudt, Pabs_cor;
  const double *vt;       /* vt(n_icp), host memory */
} lepic_heating;
int lepic_eheating(lepic_ctx *ctx, const lepic_heating *h, int32_t mode, uint64_t seed, int32_t it,
                   int32_t *iseed, double *dP);

/* ---- surface emitter (part_flux_emitter.f90:1-217; called at main.f90:1092-1096
 *      after the deposition when jne > 0 and (ns_Hm == 1 or it%ns_Hm == 1)) ----
 * Emits round(js*zmax*Lgrd*ns_Hm*dt/(Nm*qe)) particles of ptype_fx (js =
 * 10*jne A/m^2, Lgrd = n_icp*(Lg - Lh*n_holes), or 2*(y1 + chamfer length)
 * on a chamfered grid) from the electrode face at x = xg1: y uniform over the
 * grid (n_icp = 1: [(ymax-Lg)/2, (ymax+Lg)/2], else [0, ymax)), on the chamfer
 * of a tagged negative-ion grid (ptype_fx == tag_neg, flag_chmfrd = 1) moved
 * onto the chamfer and typed tag_b (when > 0); a candidate is kept when its
 * cell touches the electrode, and leaves toward the plasma side with a
 * flux-Maxwellian normal speed and a Maxwellian tangential pair of
 * temperature THm, spread over its ns_Hm*dt flight.  Appends after each
 * species' particles, adds one count per particle to the source map
 * (LEPIC_GRID_SRC).  LEPIC_MCC_CHECK replays the reference's ran2 stream
 * from *iseed (FP64 contexts); LEPIC_MCC_PROD draws per particle from a
 * counter-based generator keyed by (seed, it). */
typedef struct {
  double jne;             /* |js| in mA/cm^2 */
  double Lg, Lh, xg1;     /* grid length, hole length, electrode face x (m) */
  double THm;             /* emitted temperature (eV) */
  double chmfrd_array[4]; /* chamfer x1, x2, y1, y2 (m) */
  int32_t n_holes, n_icp, ptype_fx, tag_b, flag_chmfrd, ns_Hm;
} lepic_emitter;
typedef struct {
  int64_t Nh;             /* particles emitted */
  int64_t created[8];     /* per species (index ptype-1) */
  int64_t cnt_b[2];       /* cnt_b(1:2): flat surface, chamfer */
  double P_loss4[8];      /* P_loss(4,ptype): kinetic energy emitted */
  double Lgrd, dNh;       /* effective grid length, expected count */
  int32_t seed, reserved; /* check mode: ran2 state after the call */
} lepic_emitter_info;
int lepic_flux_emitter(lepic_ctx *ctx, const lepic_emitter *e, int32_t it, int32_t mode, uint64_t seed,
                       int32_t *iseed, lepic_emitter_info *info);

/* ---- volume injection (part_injection.f90:1-237; the species loop of
 *      main.f90:1063-1071, every ns_inj steps between the push and the
 *      deposition) ----
 * For each species in turn: the count -- opt_inj = +-2: electrons
 * ni0(1)*(ions lost this step, tagged negative ions excepted), the tagged
 * negative ion its ni0 share of the same, every other ion its own losses
 * (N_inj, index ptype-1); otherwise I_inj*ns_inj*dt*ni0(ptype)/(qe*Nm(1)),
 * rounded by one draw; +-3 and +-4 (flag_nmn = 0) inject electrons only --
 * then the particles: positions uniform in [xr_pow,xl_pow] x [yr_pow,yl_pow]
 * (flag_c = 0) or the disk of radius dr about (xa, ymax/2) (flag_c = 1), or
 * Gaussian of width dr for opt_inj < 0, never in an all-wall cell;
 * velocities Maxwellian of vt0(ptype) (the electrons' from THm with
 * flag_Tp = 1), with a z beam sqrt(2 qe THm/m) for +-3, or a shifted
 * Maxwellian flux off x = 0 along th_B for +-4.  Appends after each species'
 * particles, counts ss2D(1) (LEPIC_GRID_SRC), and with deposit = 1 adds the
 * new charge to np (call it after lepic_part_mover(..., 1)).  Modes as for
 * the MCC. */
typedef struct {
  int32_t opt_inj, flag_c, flag_nmn, flag_Tp, flag_diffsrc, ns_inj;
  double I_inj;           /* A */
  double xl_pow, xr_pow, yl_pow, yr_pow, xa, dr;   /* m */
  double THm, th_B;       /* eV, rad */
  const double *ni0;      /* ni0(1:ntype), host memory */
  const double *vt0;      /* vt0(1:ntype) (m/s), host memory */
} lepic_injection;
typedef struct {
  int64_t injected[8];    /* per species (index ptype-1) */
  double P_loss4[8];      /* P_loss(4,ptype): kinetic energy injected */
  double P_loss1[8];      /* P_loss(1,ptype): -Nm q phi of the injected (flag_diffsrc = 1) */
  int32_t seed, reserved; /* check mode: ran2 state after the call */
} lepic_injection_info;
int lepic_part_injection(lepic_ctx *ctx, const lepic_injection *inj, const int64_t *N_inj, int32_t it,
                         int32_t mode, uint64_t seed, int32_t *iseed, int32_t deposit,
                         lepic_injection_info *info);

/* ---- synthetic input (load_part.f90:59-224 restated for a GPU) ----
 * Fills [xl,xr) x [0,ymax) of a species with ppc particles per cell
 * uniformly, velocities Maxwellian with thermal speed vt (vt0 of
 * main.f90:468), from a counter-based generator keyed by seed; cells whose
 * four corners are walls are skipped, like load_part.f90:115-116.  Appends. */
int lepic_load_uniform(lepic_ctx *ctx, int32_t ptype, int32_t ppc, double vt, double xl,
                       double xr, uint64_t seed);

/* ---- one whole step (main.f90:689-1090 order) ----
 * rhs -> Poisson -> E -> [sort if it%nsort==1] -> [dens_red+collisions if
 * it>1 and it%ns_coll==1] -> push+deposit of every species. */
typedef struct {
  double eps;             /* Poisson tolerance */
  int32_t ncycl;          /* V-cycle limit (20000 in main.f90:261) */
  int32_t nsort;          /* 10 */
  int32_t ns_coll;        /* 10 */
  int32_t mcc_mode;       /* LEPIC_MCC_* or -1 for no collisions */
  uint64_t seed;
  int32_t iseed;          /* check-mode ran2 state (updated) */
  int32_t start;          /* Poisson start */
  /* outputs */
  lepic_solve_info solve;
  double ms_rhs_poisson_e, ms_sort, ms_mcc, ms_push;   /* phase times (HIP events) */
  /* electron heating before the push when heat != NULL and it%ns_heat == 0
   * (main.f90:1033-1037), in mode heat_mode with the seeds above */
  const lepic_heating *heat;
  int32_t ns_heat, heat_mode;
  double P_heat;          /* output: the heating's energy change this step */
  /* surface emission after the push when emit != NULL, jne > 0 and
   * (ns_Hm == 1 or it%ns_Hm == 1) (main.f90:1092-1096), in mode heat_mode */
  const lepic_emitter *emit;
  int64_t emitted;        /* output: particles emitted this step */
  /* volume injection after the push when inj != NULL and it%ns_inj == 0
   * (main.f90:1063-1071), opt_inj = +-2 re-injecting this push's losses */
  const lepic_injection *inj;
  int64_t injected;       /* output: particles injected this step */
} lepic_step_args;
/* ---- learned warm start (no counterpart in the reference) ----
 * A U-Net (4 levels, 16-32-64-128 channels, BF16 MFMA kernels with FP32
 * accumulation) maps the change of rhs since the last solve and the interior
 * mask to the change of phi.  lepic_iterative_solver with LEPIC_START_NET
 * starts from the network's guess only when it lowers the initial residual
 * (else from the previous phi) and always iterates to the same tolerance.
 * Samples are the (drhs, dphi) pairs of consecutive converged solves,
 * recorded while recording is on.  A guess that does not lower the initial
 * residual by at least min_gain (default 2) makes the next 8, 16, ...
 * max_backoff (default 256) START_NET solves start from the previous phi
 * without evaluating the network; a guess that does resets the back-off.
 * max_backoff = 0 evaluates the network on every START_NET solve. */
int lepic_net_create(lepic_ctx *ctx, int32_t capacity, uint64_t seed);
int lepic_net_recording(lepic_ctx *ctx, int32_t on);
int lepic_net_samples(lepic_ctx *ctx, int32_t *n);
/* Offline training across runs: append src's recorded samples (oldest
 * first) to dst's sample store (same grid, both networks created; the
 * store keeps its newest `capacity` samples), and copy src's trained
 * network (weights and output scale) into dst, whose grid may differ. */
int lepic_net_copy_samples(lepic_ctx *dst, lepic_ctx *src, int32_t *copied);
int lepic_net_copy(lepic_ctx *dst, lepic_ctx *src);
int lepic_net_policy(lepic_ctx *ctx, float min_gain, int32_t max_backoff);
int lepic_net_train(lepic_ctx *ctx, int32_t iters, int32_t batch, float lr, float *loss_first,
                    float *loss_last);
/* parameters in the flat layout [layer][cout][tap][cin] + 16 output weights */
int lepic_net_set_params(lepic_ctx *ctx, const float *flat, int64_t n);
int lepic_net_get_params(lepic_ctx *ctx, float *flat, int64_t n);
/* forward pass, and the training loss with its gradient, on a host batch
 * x(B,H,W,2): masked MSE against target plus lam times the masked squared
 * Poisson residual of the prediction, A(mask*y) - x[..,0] with the 5-point
 * stencil cw*(W+E) + cn*(S+N) - 2(cw+cn)*centre (lepic_net_train uses the
 * finest level's coefficients in the network's units and lam = 1) */
int lepic_net_forward(lepic_ctx *ctx, const float *x, int32_t B, int32_t H, int32_t W, float *y);
int lepic_net_loss_grad(lepic_ctx *ctx, const float *x, const float *target, const float *mask, int32_t B,
                        int32_t H, int32_t W, float lam, float cw, float cn, double *loss, float *grad,
                        int64_t n);

/* ---- averaged moments (part_moments utils.f90:1-136, calc_avg utils.f90:138-265,
 *      phi/E averages main.f90:933-944) ----
 * Each call adds, on the grid coarsened by size_na (na = n/size_na), every
 * species' density, temperature, current densities and sink map to
 * data_pavg(7,na1+1,na2+1,ntype) = [n, Tp, jx, jy, jz, source, sink], phi
 * and E to phi_avg(3,0:n1+2,0:n2+2), counts the call (cnt_avg(1)) and
 * clears the source and sink maps ss2D (LEPIC_GRID_SRC, LEPIC_GRID_SINK:
 * counts of created/injected/emitted and of lost/killed particles), as the
 * reference does every navg steps.  Moments use the particles' half-step-
 * back positions and the density np of the last deposition.  size_na must
 * divide n1 and n2 and stay the same until a reset. */
int lepic_moments_accumulate(lepic_ctx *ctx, int32_t size_na);
int lepic_moments_get(lepic_ctx *ctx, double *data_pavg, double *phi_avg, int32_t *count);
int lepic_moments_reset(lepic_ctx *ctx);

/* ---- charge and energy totals (no counterpart routine in the reference,
 * whose P_loss tallies account for the energy flows; these are the stocks
 * the flows change, for drift checks) ----
 * lepic_species_totals: out[0] = particle count, out[1] = kinetic energy
 * 0.5*Nm*mass*sum|v|^2 (J), out[2] = charge Nm*charge*count (C).
 * lepic_field_energy: 0.5*eps0*sum|E|^2*h1*h2*zmax over the interior nodes
 * (bcnd == -1) of the current E grid (J).  Sums in a fixed order. */
int lepic_species_totals(lepic_ctx *ctx, int32_t ptype, double *out);
int lepic_field_energy(lepic_ctx *ctx, double *W);

/* ---- particle backup files (main.f90:1102-1132 writer, restart.f90:56-113 reader) ----
 * The reference's particlesNNN.bak layout: Fortran unformatted sequential
 * records (4-byte length markers, subrecords with signed markers above
 * 2^31-9 bytes) holding (time, gtype, Pabs_cor), np_tot(ntype,nproc), then
 * vxp(6,np_tot) per (iproc, ptype).  The host functions need no device;
 * vxp[ptype-1 + ntype*iproc] points at 6*np_tot doubles. */
int lepic_backup_write(const char *path, double time, int32_t gtype, double pabs_cor, int32_t ntype,
                       int32_t nproc, const int32_t *np_tot, const double *const *vxp);
int lepic_backup_header(const char *path, int32_t ntype, double *time, int32_t *gtype, double *pabs_cor,
                        int32_t *nproc, int32_t *np_tot, int32_t cap);
/* the particles of ptype from every partition of the file, in partition
 * order (restart with another thread count, restart.f90:66-113) */
int lepic_backup_read_species(const char *path, int32_t ntype, int32_t ptype, double *vxp, int64_t cap,
                              int64_t *np);
/* the context's species as one partition; loading replaces every species'
 * particles (the caller then redeposits, as restart_OMP does) */
int lepic_save_particles(lepic_ctx *ctx, const char *path, double time, int32_t gtype, double pabs_cor);
int lepic_load_particles(lepic_ctx *ctx, const char *path, double *time, int32_t *gtype, double *pabs_cor);

int lepic_step(lepic_ctx *ctx, int32_t it, lepic_step_args *a);

#ifdef __cplusplus
}
#endif
#endif /* LEPIC_H */
