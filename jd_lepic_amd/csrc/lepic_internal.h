// lepic_internal.h -- context state and device helpers shared by the HIP
// translation units of liblepic_hip.so (gfx950 only).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <vector>

#include "../../include/lepic.h"

namespace lepic {

struct Net;

constexpr int kMaxSpecies = 8;
constexpr int kMaxLevels = 12;
constexpr int kMccF64Head = 64;       // scalar sums at the head of the MCC FP64 scratch
constexpr int kRegionBlocks = 256;    // blocks of a region sum (collisions.f90 np_tot_rg)
constexpr int kMaxRegions = 5;
constexpr int kMaxReac = 64;
constexpr int kMaxTypes = 32;
constexpr int kBlock = 256;          // threads per block of the particle kernels
constexpr int kItems = 4;            // particles per thread in the mover
constexpr int kChunk = kBlock * kItems;  // slots per particle segment
constexpr int kSweepsPerLaunch = 5;  // red-black sweeps per launch on multi-block MG levels
constexpr int kRot = kSweepsPerLaunch + 1;  // rotating state buffers of a multi-block level
constexpr int kRing = 16;            // residual partial-sum slots (sweep index mod kRing)
constexpr int kPinnedDoubles = 1024; // pinned host scratch (doubles)

#define LEPIC_HIP(call)                                                              \
  do {                                                                               \
    hipError_t e_ = (call);                                                          \
    if (e_ != hipSuccess) return ctx->fail(LEPIC_ERR_HIP, std::string(#call) + ": " + \
                                           hipGetErrorString(e_));                   \
  } while (0)

// One MG level; level 1 aliases phi/rhs.
struct Level {
  int n1 = 0, n2 = 0;
  double h1 = 0, h2 = 0;
  double aw = 0, an = 0, ac = 0;     // mg.h coefficients (ae=aw, as=an)
  double* u = nullptr;               // (n1+3)(n2+3)
  double* b = nullptr;               // (n1+1)(n2+1)
  float* uf = nullptr;               // FP32 copies for mixed smoothing
  float* bf = nullptr;
  int8_t* bc = nullptr;              // (n1+3)(n2+3)
  // multi-block levels: rotating states (index 0 is u or uf) and the
  // residual partial sums of the sweep kernels (kRing slots x blocks)
  bool mb = false;
  bool small_tiles = false;          // multi-block sweeps on 32-wide tiles
  void* rot[kRot] = {};
  int nbx = 0, nby = 0;
  double* partials = nullptr;
  // FP64: sweeps of the first launch of the level's down and up visits, one
  // more than each visit took in the previous V-cycle; mixed: the visits'
  // fixed sweep counts
  int first_dn = kSweepsPerLaunch, first_up = kSweepsPerLaunch;
};

// Device-side state of a multi-block MG level across its visits.  A visit
// (down or up leg of one V-cycle) carries an epoch; visits of a level have
// consecutive epochs.  The sweep kernels record an early stop
// (done_epoch/ksor/res); the kernel consuming the visit's result resolves it
// and stores the buffer holding the result in cur[epoch & 1], which the
// level's next visit reads as cur[(epoch - 1) & 1].
struct LevelCtrl {
  int32_t done_epoch;
  int32_t ksor;
  int32_t cur[2];
  double res;
};
// Particles of a species live in segments of kChunk slots: segment s holds
// seg[cur][s] particles in its first slots, and the species' order is the
// segments' contents in segment order.  The mover compacts each segment in
// place of the other buffer, so survivors keep the reference's order
// (part_expmover.f90:496-503) without any cross-block prefix.  "Packed"
// means every segment but the last is full (a dense array of count slots);
// kernels that index particles globally run on packed arrays.
struct Species {
  double charge = 0, mass = 0, Nm = 0;
  int64_t cap = 0;
  int64_t ub = 0;                    // host-side upper bound of the count
  int cur = 0;                       // current particle buffer
  int32_t* seg[2] = {nullptr, nullptr};  // per-buffer segment counts
  int64_t nseg_cap = 0;
  bool packed = true;
  void* mem[2] = {nullptr, nullptr};
  double* d[2][6] = {};              // FP64: x y z vx vy vz
  int32_t* cell[2] = {};             // mixed: (iy<<16)|ix, 1-based cell
  float* f[2][6] = {};               // mixed: fx fy z vx vy vz
  int8_t* flag_dead = nullptr;
  // some flag_dead entry may be set: collisions or lepic_set_flag_dead ran
  // since the last push (a push consumes and clears every set flag)
  bool dead_maybe = true;
  double* np = nullptr;              // deposited density (nodes)
  int64_t* acc = nullptr;            // fixed-point deposition accumulator (nodes), scaled into np
  double* np_red = nullptr;
  double* sink = nullptr;            // ss2D(2): sink counts (nodes)
  double* src = nullptr;             // ss2D(1): source counts (nodes)
};

// Tally block per species on the device.
struct Tally {
  double P_loss[4];
  double cnt_dead;
  double p_mac[2][65];               // (count, energy) x igrid 0..64
};
// Copies of a species' tally: a mover block adds to copy blockIdx % kTallyCopies
// (atomics on one address serialise; a step loses ~1e4 electrons at the
// walls), and a read sums the copies in copy order.
constexpr int kTallyCopies = 64;

struct Reactions {
  bool set = false;
  int ntype_all = 0, ncol = 0, npt = 0, ns_coll = 10, n_cell = 0, n_rg = 0;
  int iyl_rg = 0, iyr_rg = 0, ixl_rg[5] = {}, ixr_rg[5] = {};
  int flag_ionrep = 0;
  std::vector<int32_t> col_info, p_ncol, sig_list;
  std::vector<double> sigv_mx, Eth, charge, mass, Nm, vt0, np_mx, nu_uplim;
  double* d_sig_Er = nullptr;        // [npt]
  double* d_sig = nullptr;           // [ncol*npt]
  float* d_sig_Er_f = nullptr;
  float* d_sig_f = nullptr;
  double* d_cnt_col = nullptr;       // [3*ncol*n_rg]
};

}  // namespace lepic

// Averaged moments (calc_avg, utils.f90:138-265): data_pavg(7,na1+1,na2+1,
// ntype), phi_avg(3,0:n1+2,0:n2+2), the number of accumulations, and the
// per-species part_moments scratch p_mts(4,na1+1,na2+1).
struct Moments {
  int size_na = 0;
  int count = 0;
  double* data = nullptr;
  double* phi = nullptr;
  double* pm = nullptr;
};

struct lepic_ctx {
  lepic_grid_desc g{};
  int device = 0;
  hipStream_t stream = nullptr;
  std::string err;
  int nsp = 0;
  lepic::Species sp[lepic::kMaxSpecies];

  // grids
  int64_t nodes = 0, rhs_n = 0;
  int8_t* bc = nullptr;              // level-1 bcnd (int8)
  int8_t* d_cellwall = nullptr;      // per node (ix,iy): cell (ix,iy) has four wall corners
  std::vector<int32_t> bc_host;      // int32 copy
  double* phi = nullptr;
  double* phi_given = nullptr;       // guess set through lepic_set_grid(PHI) for START_GIVEN
  double* rhs = nullptr;
  double* E = nullptr;               // 2 per node
  float* Ef = nullptr;               // mixed: 2 per node
  double* Bi = nullptr;              // 4 per B node
  float* Bf = nullptr;               // mixed: 4 per B node
  int nB1 = 1, nB2 = 1, flag_B = 0, flag_gridB = 0;
  double hB1 = 0, hB2 = 0, Bc[3] = {0, 0, 0};
  int efield_serial = 0;             // 1: wall extrapolation needs the reference's order
  bool boundary_set = false;

  // MG
  int ng = 1;
  lepic::Level lv[lepic::kMaxLevels + 1];
  int Lc = 1;                        // first level handled by the single-workgroup kernel
  int8_t* d_tail_bc = nullptr;       // LDS image of the tail levels' bc arrays
  size_t tail_bc_cap = 0;
  int tail_bc_key = -1;              // layout the image was built for (-1: rebuild)
  int32_t mg_epoch = 0;              // epoch of the last multi-block level visit
  lepic::LevelCtrl* d_ctrl = nullptr;   // [kMaxLevels+1]
  double* d_partials = nullptr;      // FP64 residual partial sums of the finest level
  void* d_erot = nullptr;            // mixed: device array of the level-1 FP32 correction buffers
  int64_t partials_cap = 0;
  double* d_scalar = nullptr;        // small scratch (res0 etc.)
  double* h_pinned = nullptr;        // pinned host scratch (kPinnedDoubles)
  double* d_resid64 = nullptr;       // FP64 residual grid for mixed refinement
  float* d_corr = nullptr;

  // device scalars
  int64_t* d_counts = nullptr;       // [kMaxSpecies]
  lepic::Tally* d_tally = nullptr;   // [kMaxSpecies][kTallyCopies]
  uint32_t* d_segoff = nullptr;      // segment offsets for packing (max nseg_cap)
  int64_t segoff_cap = 0;
  unsigned long long* d_ticket = nullptr;   // monotonic block ticket of the movers
  unsigned long long ticket_base = 0;       // tickets taken by earlier launches
  bool timing = false;                      // lepic_kernel_timing
  hipEvent_t ev_mover[lepic::kMaxSpecies][2] = {};
  hipEvent_t ev_sync = nullptr;      // polled by the solver's per-V-cycle read-back
  bool ev_recorded[lepic::kMaxSpecies] = {};
  int32_t* d_error = nullptr;        // device-side error flags

  // sort scratch
  uint32_t* s_key[2] = {nullptr, nullptr};
  uint32_t* s_val[2] = {nullptr, nullptr};
  uint32_t* s_hist = nullptr;
  int64_t s_cap = 0, s_hist_cap = 0;
  int32_t* d_plist = nullptr;        // cell-end list (pl_max+1)
  uint32_t* d_scan_tmp = nullptr;
  int64_t plist_cap = 0;           // entries of d_plist
  // segment sort: per-slot (run, rank) of the segment-local sort, the runs
  // of every segment (cell key, particle count) at a stride of kChunk, the
  // dense run counts, and the per-segment run counts / bases
  uint32_t* s_meta = nullptr;
  uint32_t* s_rkey = nullptr;
  uint32_t* s_rcnt = nullptr;
  uint32_t* s_rcntd = nullptr;
  uint32_t* s_nrun = nullptr;
  int64_t s_seg_cap = 0;           // slots covered by s_meta / s_rkey / s_rcnt / s_rcntd (segments x kChunk)
  double* d_diag = nullptr;          // per-block partials of the charge/energy totals
  // this device's copy of the static sort scratch array (device address
  // from hipGetSymbolAddress, resolved at creation)
  void* sort_static = nullptr;

  // collisions
  lepic::Reactions rx;
  uint32_t* d_claim = nullptr;       // candidate claim bitmap
  int64_t claim_cap = 0;
  void* d_mcc_stage = nullptr;       // staged new particles
  int64_t stage_cap = 0;
  int64_t* d_mcc_i64 = nullptr;      // [64] counters
  double* d_mcc_f64 = nullptr;       // [kMccF64Head] sums: [0..8] P_loss(3,ptype), [40..44] 1/np_tot_rg;
                                     // then kRegionBlocks block partials per region
  void* d_mcc_tab = nullptr;         // device copy of the reaction table

  // learned warm start (unet.hip)
  lepic::Net* net = nullptr;
  Moments mom;
  double* d_heat = nullptr;          // electron heating: vt table, power, seed (heating.hip)
  // particle sources (sources.hip): staged particles (x y z vx vy vz),
  // their 1-based species, per-species ranks, and the device tally
  double* d_stage6 = nullptr;
  int8_t* d_stype = nullptr;
  uint32_t* d_srank = nullptr;
  int64_t stage6_cap = 0;
  void* d_srctally = nullptr;
  // lepic_step with opt_inj = +-2: wall losses of every push since the last
  // injection, per species (N_inj of main.f90:1044-1056, reset after the
  // injection at :1063-1071)
  double inj_lost[lepic::kMaxSpecies] = {};

  int fail(int code, const std::string& msg) {
    err = msg;
    return code;
  }
};

namespace lepic {

// ---- host helpers implemented in lepic_ctx.hip ----
int ensure_packed(lepic_ctx* ctx, int p);        // 0-based species
int set_packed_counts(lepic_ctx* ctx, int p);    // segments from the device count
inline int64_t nseg_of(const Species& s) { return (s.ub + kChunk - 1) / kChunk; }
// in place; ws: optional workspace of scan_workspace_words(n) words for the
// block partials (none allocated then)
int scan_exclusive_u32(lepic_ctx* ctx, uint32_t* data, int64_t n, uint32_t* ws = nullptr);
int64_t scan_workspace_words(int64_t n);
int zero_async(lepic_ctx* ctx, void* p, size_t bytes);
int deposit_range_launch(lepic_ctx* ctx, int p, int64_t seg0, int64_t nseg);   // push.hip

// ---- phase entry points implemented per translation unit ----
int mover_launch(lepic_ctx* ctx, int ptype, int deposit);
int deposit_launch(lepic_ctx* ctx, int ptype);
int rhs_launch(lepic_ctx* ctx);
int efield_launch(lepic_ctx* ctx);
int dens_red_launch(lepic_ctx* ctx, double* np_mx_host);
int poisson_setup(lepic_ctx* ctx);
void poisson_teardown(lepic_ctx* ctx);
int poisson_solve(lepic_ctx* ctx, double eps, int ncycl, int start, lepic_solve_info* info,
                  int32_t* log, int32_t log_cap);
int sort_launch(lepic_ctx* ctx, int ptype, int32_t* Plist_host);
// the same sort with its scratch in the library's static device array
// (no allocation); sort_scratch_resolve finds that array's address on the
// current device (once per device, at context creation)
int sort_noalloc_launch(lepic_ctx* ctx, int ptype, int32_t* Plist_host);
int sort_scratch_resolve(lepic_ctx* ctx);
int mcc_run(lepic_ctx* ctx, int it, int mode, uint64_t seed, int32_t* iseed,
            lepic_mcc_info* info);
int load_uniform_launch(lepic_ctx* ctx, int ptype, int ppc, double vt, double xl, double xr,
                        uint64_t seed);
void net_destroy(lepic_ctx* ctx);
int net_after_solve(lepic_ctx* ctx);
int net_guess(lepic_ctx* ctx, double* out);   // 1: no guess available
bool net_probe_due(lepic_ctx* ctx);
void net_probe_result(lepic_ctx* ctx, double gain);

// ---- device helpers ----
__device__ __forceinline__ int64_t node_index(int ix, int iy, int n1) {
  return (int64_t)iy * (n1 + 3) + ix;
}

__device__ __forceinline__ void atomic_add_f64(double* p, double v) { unsafeAtomicAdd(p, v); }

// ---- wave reductions in a fixed order through DPP lane moves ----
// (VALU data-parallel primitives: row_shr:1,2,4,8 build a scan inside each
// row of 16 lanes, row_bcast:15 and row_bcast:31 carry it across rows; lane
// 63 ends with the total.)  Every lane must be active.
template <int CTRL, int ROWS>
__device__ __forceinline__ double dpp_move(double v) {
  const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), CTRL, ROWS, 0xf, false);
  const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), CTRL, ROWS, 0xf, false);
  return __hiloint2double(hi, lo);
}

__device__ __forceinline__ double wave_sum(double v) {
  v += dpp_move<0x111, 0xf>(v);
  v += dpp_move<0x112, 0xf>(v);
  v += dpp_move<0x114, 0xf>(v);
  v += dpp_move<0x118, 0xf>(v);
  v += dpp_move<0x142, 0xa>(v);
  v += dpp_move<0x143, 0xc>(v);
  const int lo = __builtin_amdgcn_readlane(__double2loint(v), 63);
  const int hi = __builtin_amdgcn_readlane(__double2hiint(v), 63);
  return __hiloint2double(hi, lo);
}

}  // namespace lepic

namespace lepic {

// Philox4x32-10 counter-based generator (Salmon, Moraes, Dror, Shaw, "Parallel
// random numbers: as easy as 1, 2, 3", SC'11): multipliers 0xD2511F53 and
// 0xCD9E8D57, Weyl key increments 0x9E3779B9 and 0xBB67AE85, ten rounds.
struct U4 { uint32_t x, y, z, w; };

__host__ __device__ inline uint32_t mulhi32(uint32_t a, uint32_t b) {
  return (uint32_t)(((uint64_t)a * (uint64_t)b) >> 32);
}

__host__ __device__ inline U4 philox4x32_10(U4 c, uint32_t k0, uint32_t k1) {
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    const uint32_t hi0 = mulhi32(0xD2511F53u, c.x), lo0 = 0xD2511F53u * c.x;
    const uint32_t hi1 = mulhi32(0xCD9E8D57u, c.z), lo1 = 0xCD9E8D57u * c.z;
    c = U4{hi1 ^ c.y ^ k0, lo1, hi0 ^ c.w ^ k1, lo0};
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
  return c;
}

// 53-bit uniform in [0,1) from two words
__host__ __device__ inline double u01(uint32_t hi, uint32_t lo) {
  return (double)((((uint64_t)hi << 32) | lo) >> 11) * (1.0 / 9007199254740992.0);
}

// The reference's generator (ran2.f90:1-15: Park-Miller minimal standard,
// multiplier 16807 modulo 2^31-1, Schrage's factorisation) and its
// Maxwellian pair (load_gauss.f90:14-21), used by the check-mode replays.
__host__ __device__ inline double ran2(int32_t& s) {
  const int32_t ia = 16807, im = 2147483647, iq = 127773, ir = 2836;
  int32_t k = s / iq;
  s = ia * (s - k * iq) - ir * k;
  if (s < 0) s += im;
  return (1.0 / (double)im) * (double)s;
}

__device__ inline void load_gauss(double& vx, double& vy, double vt, double r1, double r2, double pi) {
  double vp = vt * sqrt(-log(1.0 - r1));
  double th = 2.0 * pi * r2;
  vx = vp * cos(th);
  vy = vp * sin(th);
}

}  // namespace lepic
