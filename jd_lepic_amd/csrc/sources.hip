// sources.hip -- particles created outside the MCC: the surface emitter of
// the extraction electrode (part_flux_emitter.f90:1-217) and volume injection
// (part_injection.f90:1-237).
//
// Both write the new particles, in the order the reference creates them,
// into a staging area (x y z vx vy vz and the 1-based species), then
// append_staged places each species' share in fresh segments after the
// species' last segment: no repacking of the resident particles, and the
// species' order (residents, then the new particles in creation order) is the
// reference's np_tot+1, np_tot+2, ... order.  Source counts go to ss2D(1)
// (the source map) and the emitted energy to P_loss(4, ptype).
//
// Check mode replays the reference's ran2 stream on one lane with the FP64
// state; production mode gives every new particle its own Philox4x32-10
// stream keyed by (seed, it, index).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>

#include "lepic_internal.h"

using namespace lepic;

namespace {

struct SrcTally {
  double P_loss4[kMaxSpecies + 1];          // by ptype
  double P_loss1[kMaxSpecies + 1];          // by ptype (injection with flag_diffsrc)
  unsigned long long cnt_b[2];
  unsigned long long tot[kMaxSpecies];      // staged particles per species (index ptype-1)
  int32_t status;                           // 1: no admissible candidate, 2: staging full
  int32_t seed;                             // check mode: ran2 state after the call
};

// ---- append ----------------------------------------------------------------

__global__ void k_count_types(const int8_t* __restrict__ type, int64_t M, unsigned long long* tot) {
  __shared__ unsigned int h[kMaxSpecies];
  if (threadIdx.x < kMaxSpecies) h[threadIdx.x] = 0;
  __syncthreads();
  for (int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; j < M; j += (int64_t)gridDim.x * blockDim.x) {
    const int t = type[j];
    if (t >= 1 && t <= kMaxSpecies) atomicAdd(&h[t - 1], 1u);
  }
  __syncthreads();
  if (threadIdx.x < kMaxSpecies && h[threadIdx.x]) atomicAdd(&tot[threadIdx.x], (unsigned long long)h[threadIdx.x]);
}

__global__ void k_type_flags(const int8_t* __restrict__ type, int64_t M, int pt, uint32_t* __restrict__ out) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j < M) out[j] = type[j] == pt ? 1u : 0u;
}

struct AppendDst {
  double* d[6];
  int32_t* cell;
  float* f[6];
  int8_t* flag_dead;
  double h1, h2;
};

template <bool MIXED>
__global__ void k_append(const double* __restrict__ stage, const int8_t* __restrict__ type,
                         const uint32_t* __restrict__ rank, int64_t M, int pt, int64_t base, AppendDst o) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= M || type[j] != pt) return;
  const int64_t i = base + rank[j];
  const double* v = stage + 6 * j;
  if (MIXED) {
    // the conversion of lepic_upload_particles: 1-based cell INT(x/h)+1 and
    // the FP32 offset inside it
    const double sx = v[0] / o.h1, sy = v[1] / o.h2;
    int ix = (int)sx, iy = (int)sy;
    float ox = (float)(sx - ix), oy = (float)(sy - iy);
    if (ox >= 1.0f) { ox = 0.0f; ++ix; }
    if (oy >= 1.0f) { oy = 0.0f; ++iy; }
    ix = ix < -1 ? -1 : (ix > 30000 ? 30000 : ix);
    iy = iy < -1 ? -1 : (iy > 30000 ? 30000 : iy);
    o.cell[i] = ((iy + 1) << 16) | (ix + 1);
    o.f[0][i] = ox; o.f[1][i] = oy;
    if (o.f[2]) o.f[2][i] = (float)v[2];
    o.f[3][i] = (float)v[3]; o.f[4][i] = (float)v[4]; o.f[5][i] = (float)v[5];
  } else {
    for (int k = 0; k < 6; ++k) o.d[k][i] = v[k];
  }
  o.flag_dead[i] = 0;
}

// segments s0 .. s0+ns-1 hold slots up to end (the new particles end there)
__global__ void k_seg_fill(int32_t* seg, int64_t s0, int64_t ns, int64_t end) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= ns) return;
  seg[s0 + k] = (int32_t)min((int64_t)kChunk, end - (s0 + k) * kChunk);
}

__global__ void k_count_add(int64_t* counts, int q, int64_t add) {
  if (threadIdx.x == 0 && blockIdx.x == 0) counts[q] += add;
}

// ---- the surface emitter ---------------------------------------------------

struct EmitArgs {
  const int8_t* bc;
  int n1, n2, n_icp, ptype_fx, tag_b, chamfered;
  double h1, h2, ymax, zmax, pi, xg1, Lg, y1, span;
  double cham_c[2], cham_s[2];     // cos, sin of th_chrd and -th_chrd (host libm)
  double rot_c[3], rot_s[3];       // cos, sin of pi/2, th_chrd, pi/2+th_chrd
  double vt[kMaxSpecies + 1];      // sqrt(2 qe THm / |mass(ptype)|)
  double half_Nm_m[kMaxSpecies + 1];
  int64_t Nh, max_tries;
  double* stage;
  int8_t* type;
  double* src[kMaxSpecies];
  SrcTally* t;
  uint32_t k0, k1, it;
};

// part_flux_emitter.f90:140-144 chamfer test, on the current y_inj
__device__ __forceinline__ bool on_chamfer(const EmitArgs& a, double y) {
  return a.chamfered && ((y <= a.ymax / 2.0 && y > a.y1) || (y > a.ymax / 2.0 && fabs(a.ymax - y) > a.y1));
}

// y_inj from a uniform draw, then the chamfer displacement (lines 126-163)
__device__ __forceinline__ void place(const EmitArgs& a, double r, double& x, double& y, int& ptype) {
  x = a.xg1;
  y = a.n_icp == 1 ? (a.ymax - a.Lg) / 2.0 + a.Lg * r : a.ymax * (1.0 - r);
  ptype = a.ptype_fx;
  if (on_chamfer(a, y)) {
    if (a.tag_b > 0) ptype = a.tag_b;
    int k = 0;
    double d0 = y - a.y1, d1 = a.y1;
    if (y > a.ymax / 2.0) {
      k = 1;
      d0 = fabs(y - a.ymax) - a.y1;
      d1 = a.ymax - a.y1;
    }
    x = a.xg1 + d0 * a.cham_c[k];
    y = d1 + d0 * a.cham_s[k];
  }
}

__device__ __forceinline__ bool in_nodes(const EmitArgs& a, int ix, int iy) {
  return ix >= 1 && ix <= a.n1 + 1 && iy >= 1 && iy <= a.n2 + 1;
}

// lines 169-172: the cell touches the electrode on its left or right side
__device__ __forceinline__ bool on_electrode(const EmitArgs& a, int ix, int iy) {
  const int8_t* b = a.bc;
  const int64_t n0 = node_index(ix, iy, a.n1), n1 = node_index(ix, iy + 1, a.n1);
  return (b[n0] >= 1 && b[n1] >= 1) || (b[n0 + 1] >= 1 && b[n1 + 1] >= 1);
}

// velocity and position of an accepted particle (lines 199-235)
__device__ __forceinline__ void finish(const EmitArgs& a, int64_t i, int ptype, double dir, double x, double y,
                                       double z, double r_n, double r1, double r2, double r_t, bool atomics) {
  const double vt = a.vt[ptype];
  const double vx = dir * vt * sqrt(-log(1.0 - r_n));
  double vy, vz;
  load_gauss(vy, vz, vt, r1, r2, a.pi);
  int k = 0;
  if (on_chamfer(a, y)) k = y > a.ymax / 2.0 ? 2 : 1;
  const double c = a.rot_c[k], s = a.rot_s[k];
  const double v3 = -(vx * s) + vy * c, v4 = vx * c + vy * s, v5 = vz;
  const double dt_tmp = r_t * a.span;
  double* st = a.stage + 6 * i;
  st[0] = x + v3 * dt_tmp;
  st[1] = y + v4 * dt_tmp;
  st[2] = z + v5 * dt_tmp;
  st[3] = v3; st[4] = v4; st[5] = v5;
  a.type[i] = (int8_t)ptype;
  const double e = a.half_Nm_m[ptype] * (v3 * v3 + v4 * v4 + v5 * v5);
  if (atomics) atomic_add_f64(&a.t->P_loss4[ptype], e);
  else a.t->P_loss4[ptype] = a.t->P_loss4[ptype] + e;
}

// one lane, the reference's order and its cumulative direction counters
__global__ void k_emit_check(EmitArgs a) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  int32_t s = a.t->seed;
  int64_t cnt_lhs = 0, cnt_rhs = 0;
  for (int64_t i = 0; i < a.Nh; ++i) {
    double x, y, z, dir = 0.0;
    int ptype = 0, ix = 0, iy = 0;
    int64_t tries = 0;
    for (;;) {
      if (++tries > a.max_tries) {
        a.t->status = 1;
        a.t->seed = s;
        return;
      }
      place(a, ran2(s), x, y, ptype);
      ix = (int)(x / a.h1) + 1;
      iy = (int)(y / a.h2) + 1;
      z = a.zmax * (1.0 - ran2(s));
      if (!in_nodes(a, ix, iy) || !on_electrode(a, ix, iy)) continue;
      if (a.bc[node_index(ix - 1, iy, a.n1)] == -1) cnt_lhs = cnt_lhs + 1;
      if (a.bc[node_index(ix + 1, iy, a.n1)] == -1) cnt_rhs = cnt_rhs + 1;
      if (cnt_lhs != cnt_rhs) {
        dir = cnt_lhs > cnt_rhs ? 1.0 : -1.0;
        break;
      }
    }
    a.t->cnt_b[ptype == a.tag_b ? 1 : 0] += 1;
    a.src[ptype - 1][node_index(ix, iy, a.n1)] += 1.0;
    const double r_n = ran2(s), r1 = ran2(s), r2 = ran2(s), r_t = ran2(s);
    finish(a, i, ptype, dir, x, y, z, r_n, r1, r2, r_t, false);
  }
  a.t->seed = s;
}

// uniform doubles from a Philox stream, two per block of four words
struct Stream {
  U4 c;
  uint32_t k0, k1;
  U4 buf;
  int used = 2;
  __device__ double next() {
    if (used == 2) {
      buf = philox4x32_10(c, k0, k1);
      c.w += 1;
      used = 0;
    }
    return used++ == 0 ? u01(buf.x, buf.y) : u01(buf.z, buf.w);
  }
};

// one lane per particle; the direction comes from the particle's own cell
// (plasma on exactly one side), where the reference compares counters summed
// over the call's earlier candidates -- the same answer whenever each
// emitting cell faces the plasma on one side
__global__ void k_emit_prod(EmitArgs a) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.Nh) return;
  Stream r{U4{(uint32_t)i, (uint32_t)(i >> 32), a.it, 0u}, a.k0, a.k1, U4{0, 0, 0, 0}};
  double x = 0, y = 0, z = 0, dir = 0.0;
  int ptype = 0, ix = 0, iy = 0;
  bool ok = false;
  for (int64_t tries = 0; tries < a.max_tries && !ok; ++tries) {
    place(a, r.next(), x, y, ptype);
    ix = (int)(x / a.h1) + 1;
    iy = (int)(y / a.h2) + 1;
    z = a.zmax * (1.0 - r.next());
    if (!in_nodes(a, ix, iy) || !on_electrode(a, ix, iy)) continue;
    const bool lhs = a.bc[node_index(ix - 1, iy, a.n1)] == -1, rhs = a.bc[node_index(ix + 1, iy, a.n1)] == -1;
    if (lhs != rhs) {
      dir = lhs ? 1.0 : -1.0;
      ok = true;
    }
  }
  if (!ok) {
    a.type[i] = 0;
    atomicOr(&a.t->status, 1);
    return;
  }
  atomicAdd(&a.t->cnt_b[ptype == a.tag_b ? 1 : 0], 1ull);
  atomic_add_f64(&a.src[ptype - 1][node_index(ix, iy, a.n1)], 1.0);
  const double r_n = r.next(), r1 = r.next(), r2 = r.next(), r_t = r.next();
  finish(a, i, ptype, dir, x, y, z, r_n, r1, r2, r_t, true);
}

// ---- volume injection ------------------------------------------------------

struct InjArgs {
  const int8_t* bc;
  const double* phi;
  int n1, n2, nsp, opt, aopt, flag_c, flag_nmn, flag_Tp, flag_diffsrc, tag_neg;
  double h1, h2, xmax, ymax, zmax, pi, span;        // span = real(ns_inj)*dt
  double xl_pow, xr_pow, yl_pow, yr_pow, xa, dr;
  double cos_thB, sin_thB;                          // host libm
  double vt[kMaxSpecies + 1];                       // thermal speed used for the velocities
  double vb[kMaxSpecies + 1], fmax[kMaxSpecies + 1];  // opt_inj 3, 4 (lines 36-45, 268-286)
  double half_Nm_m[kMaxSpecies + 1], Nm_q[kMaxSpecies + 1];
  // check mode: the count inputs; production: the counts and stage offsets
  double ni0[kMaxSpecies + 1], dN_cur[kMaxSpecies + 1];
  int64_t N_inj[kMaxSpecies + 1];
  int64_t n[kMaxSpecies + 1], off[kMaxSpecies + 1], pair_off[kMaxSpecies + 2];
  int64_t stage_cap, max_tries;
  double* stage;
  int8_t* type;
  double* src[kMaxSpecies];
  SrcTally* t;
  double* ploss1;                                   // [kMaxSpecies + 1]
  uint32_t k0, k1, it;
};

__device__ __forceinline__ bool all_wall_cell(const InjArgs& a, int ix, int iy) {
  const int64_t n0 = node_index(ix, iy, a.n1), n1 = node_index(ix, iy + 1, a.n1);
  return a.bc[n0] >= 1 && a.bc[n0 + 1] >= 1 && a.bc[n1 + 1] >= 1 && a.bc[n1] >= 1;
}

__device__ __forceinline__ bool cell_ok(const InjArgs& a, int ix, int iy) {
  return ix >= 0 && ix <= a.n1 + 1 && iy >= 0 && iy <= a.n2 + 1 && !all_wall_cell(a, ix, iy);
}

// the stored particle, its source count and tallies (lines 188-228)
__device__ __forceinline__ void inj_store(const InjArgs& a, int64_t j, int ptype, int ix, int iy, double x, double y,
                                          double z, double vx, double vy, double vz, bool atomics) {
  double* st = a.stage + 6 * j;
  st[0] = x; st[1] = y; st[2] = z; st[3] = vx; st[4] = vy;
  if (a.aopt == 3) vz = vz + a.vb[ptype];
  st[5] = vz;
  a.type[j] = (int8_t)ptype;
  const double e = a.half_Nm_m[ptype] * (vx * vx + vy * vy + vz * vz);
  double e1 = 0.0;
  if (a.flag_diffsrc == 1) {
    const double px = (ix * a.h1 - x) / a.h1, py = (iy * a.h2 - y) / a.h2;
    const double k1 = px * py, k2 = (1.0 - px) * py, k3 = (1.0 - px) * (1.0 - py), k4 = px * (1.0 - py);
    const double phip = k1 * a.phi[node_index(ix, iy, a.n1)] + k2 * a.phi[node_index(ix + 1, iy, a.n1)] +
                        k3 * a.phi[node_index(ix + 1, iy + 1, a.n1)] + k4 * a.phi[node_index(ix, iy + 1, a.n1)];
    e1 = a.Nm_q[ptype] * phip;
  }
  if (atomics) {
    atomic_add_f64(&a.src[ptype - 1][node_index(ix, iy, a.n1)], 1.0);
    atomic_add_f64(&a.t->P_loss4[ptype], e);
    if (a.flag_diffsrc == 1) atomic_add_f64(&a.ploss1[ptype], -e1);
  } else {
    a.src[ptype - 1][node_index(ix, iy, a.n1)] += 1.0;
    a.t->P_loss4[ptype] = a.t->P_loss4[ptype] + e;
    if (a.flag_diffsrc == 1) a.ploss1[ptype] = a.ploss1[ptype] - e1;
  }
}

// One lane: the main loop's species loop (main.f90:1063-1071) over
// part_injection with the reference's ran2 order.
__global__ void k_inject_check(InjArgs a) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  int32_t s = a.t->seed;
  int64_t N_inj[kMaxSpecies + 1];
  for (int q = 0; q <= kMaxSpecies; ++q) N_inj[q] = a.N_inj[q];
  int64_t j = 0;
  for (int ptype = 1; ptype <= a.nsp; ++ptype) {
    double vz_sav = 0.0, x_sav = 0.0;
    int flag = 0;
    double dN = 0.0;
    if (a.aopt == 2) {
      if (ptype == 1 || (ptype == a.tag_neg && a.tag_neg > 0)) {
        flag = 1;
        int64_t tmp = 0;
        for (int q = 2; q <= a.nsp; ++q) tmp += N_inj[q];
        if (a.tag_neg > 0) tmp = tmp - N_inj[a.tag_neg];
        dN = a.ni0[ptype] * (double)tmp;
      }
    } else {
      flag = 1;
      dN = a.dN_cur[ptype];
    }
    if (flag == 1) {
      int64_t n = (int64_t)dN;
      if (ran2(s) <= dN - (double)n) n = n + 1;
      N_inj[ptype] = n;
    }
    if (a.aopt == 3 || (a.aopt == 4 && a.flag_nmn == 0))
      for (int q = 2; q <= a.nsp; ++q) N_inj[q] = 0;
    for (int64_t i = 1; i <= N_inj[ptype]; ++i) {
      double x = 0.0, y = 0.0, z = 0.0;
      int ix = 0, iy = 0;
      int64_t tries = 0;
      bool fail = false;
      for (;;) {  // label 70
        if (++tries > a.max_tries) { fail = true; break; }
        if (a.opt > 0) {
          for (;;) {  // label 60
            x = ran2(s) * (a.xl_pow - a.xr_pow) + a.xr_pow;
            if (a.flag_c != 1) break;
            y = ran2(s) * (2.0 * a.dr) + a.ymax / 2.0 - a.dr;
            if (!(((x - a.xa) * (x - a.xa) + (y - a.ymax / 2.0) * (y - a.ymax / 2.0)) > a.dr * a.dr)) break;
            if (++tries > a.max_tries) { fail = true; break; }
          }
        }
        if (fail) break;
        if (a.opt < 0) {
          if (a.flag_c == 0) {
            for (;;) {  // label 50
              if (x_sav == 0.0) {
                const double r1 = ran2(s), r2 = ran2(s);
                load_gauss(x, x_sav, a.dr, r1, r2, a.pi);
                x = x + a.xa;
                x_sav = x_sav + a.xa;
              } else {
                x = x_sav;
                x_sav = 0.0;
              }
              if (a.flag_nmn == 1 && a.xa == 0.0) x = fabs(x);
              if (!(x < (a.xa - 4.0 * a.dr) || x > (a.xa + 4.0 * a.dr) || x < 0.0 || x > a.xmax)) break;
              if (++tries > a.max_tries) { fail = true; break; }
            }
          } else {
            for (;;) {  // label 40
              const double r1 = ran2(s), r2 = ran2(s);
              load_gauss(x, y, a.dr, r1, r2, a.pi);
              x = x + a.xa;
              y = y + a.ymax / 2.0;
              if (!(((x - a.xa) * (x - a.xa) + (y - a.ymax / 2.0) * (y - a.ymax / 2.0)) > 4.0 * a.dr * a.dr ||
                    x < 0.0 || x > a.xmax || y < 0.0 || y > a.ymax))
                break;
              if (++tries > a.max_tries) { fail = true; break; }
            }
          }
        }
        if (fail) break;
        if (a.flag_c == 0) y = ran2(s) * (a.yl_pow - a.yr_pow) + a.yr_pow;
        z = ran2(s) * a.zmax;
        ix = (int)(x / a.h1) + 1;
        iy = (int)(y / a.h2) + 1;
        if (cell_ok(a, ix, iy)) break;
      }
      if (fail || j >= a.stage_cap) {
        a.t->status = fail ? 1 : 2;
        a.t->seed = s;
        return;
      }
      const double vt = a.vt[ptype];
      double vx, vy, vz;
      if (a.aopt == 4) {
        const double vb_tmp = a.vb[ptype] * a.cos_thB;
        const double vm = (vb_tmp - 4.0 * vt) > 0.0 ? vb_tmp - 4.0 * vt : 0.0, vp = vb_tmp + 4.0 * vt;
        int64_t t4 = 0;
        for (;;) {  // label 75
          if (++t4 > a.max_tries) { fail = true; break; }
          for (int64_t t5 = 0;; ++t5) {  // shifted_maxwellian_flux, lines 239-265
            if (t5 >= a.max_tries) { fail = true; break; }
            const double r1 = ran2(s), r2 = ran2(s);
            vx = vm + r1 * (vp - vm);
            const double f = vx * exp(-(vx - vb_tmp) * (vx - vb_tmp) / (vt * vt));
            if (!(r2 > f / a.fmax[ptype])) break;
          }
          if (fail) break;
          const double r1 = ran2(s), r2 = ran2(s);
          load_gauss(vy, vz, vt, r1, r2, a.pi);
          vy = vy + a.vb[ptype] * a.sin_thB;
          const double dt_tmp = ran2(s) * a.span;
          x = vx * dt_tmp;
          if (!(x < 0.0 || x > a.xmax)) break;
        }
        if (fail) {
          a.t->status = 1;
          a.t->seed = s;
          return;
        }
        ix = (int)floor(x / a.h1) + 1;
      } else {
        double r1 = ran2(s), r2 = ran2(s);
        load_gauss(vx, vy, vt, r1, r2, a.pi);
        if (vz_sav == 0.0) {
          r1 = ran2(s); r2 = ran2(s);
          load_gauss(vx, vy, vt, r1, r2, a.pi);
          vz = vx;
          vz_sav = vy;
        } else {
          vz = vz_sav;
          vz_sav = 0.0;
        }
      }
      inj_store(a, j, ptype, ix, iy, x, y, z, vx, vy, vz, false);
      ++j;
    }
  }
  a.t->seed = s;
}

// One lane per pair of particles of a species: the pair shares the velocity
// draws the way consecutive particles of part_injection share vz_sav
// (lines 181-193): the first takes vx = vz, the second the first's vy as vz.
__global__ void k_inject_prod(InjArgs a) {
  const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= a.pair_off[a.nsp + 1]) return;
  int ptype = 1;
  while (ptype < a.nsp && g >= a.pair_off[ptype + 1]) ++ptype;
  const int64_t k = g - a.pair_off[ptype];
  Stream r{U4{(uint32_t)g, (uint32_t)ptype, a.it, 0u}, a.k0, a.k1, U4{0, 0, 0, 0}};
  double carry = 0.0;
  for (int m = 0; m < 2; ++m) {
    const int64_t i = 2 * k + m;
    if (i >= a.n[ptype]) break;
    double x = 0.0, y = 0.0, z = 0.0;
    int ix = 0, iy = 0;
    bool ok = false;
    for (int64_t tries = 0; tries < a.max_tries && !ok; ++tries) {
      if (a.opt > 0) {
        x = r.next() * (a.xl_pow - a.xr_pow) + a.xr_pow;
        if (a.flag_c == 1) {
          y = r.next() * (2.0 * a.dr) + a.ymax / 2.0 - a.dr;
          if (((x - a.xa) * (x - a.xa) + (y - a.ymax / 2.0) * (y - a.ymax / 2.0)) > a.dr * a.dr) continue;
        }
      } else {
        double gx, gy;
        load_gauss(gx, gy, a.dr, r.next(), r.next(), a.pi);
        if (a.flag_c == 0) {
          x = gx + a.xa;
          if (a.flag_nmn == 1 && a.xa == 0.0) x = fabs(x);
          if (x < (a.xa - 4.0 * a.dr) || x > (a.xa + 4.0 * a.dr) || x < 0.0 || x > a.xmax) continue;
        } else {
          x = gx + a.xa;
          y = gy + a.ymax / 2.0;
          if (((x - a.xa) * (x - a.xa) + (y - a.ymax / 2.0) * (y - a.ymax / 2.0)) > 4.0 * a.dr * a.dr ||
              x < 0.0 || x > a.xmax || y < 0.0 || y > a.ymax)
            continue;
        }
      }
      if (a.flag_c == 0) y = r.next() * (a.yl_pow - a.yr_pow) + a.yr_pow;
      z = r.next() * a.zmax;
      ix = (int)(x / a.h1) + 1;
      iy = (int)(y / a.h2) + 1;
      ok = cell_ok(a, ix, iy);
    }
    const int64_t j = a.off[ptype] + i;
    if (!ok) {
      a.type[j] = 0;
      atomicOr(&a.t->status, 1);
      return;
    }
    const double vt = a.vt[ptype];
    double vx = 0.0, vy = 0.0, vz = 0.0;
    if (a.aopt == 4) {
      const double vb_tmp = a.vb[ptype] * a.cos_thB;
      const double vm = (vb_tmp - 4.0 * vt) > 0.0 ? vb_tmp - 4.0 * vt : 0.0, vp = vb_tmp + 4.0 * vt;
      ok = false;
      for (int64_t t4 = 0; t4 < a.max_tries && !ok; ++t4) {
        vx = vm + r.next() * (vp - vm);
        const double f = vx * exp(-(vx - vb_tmp) * (vx - vb_tmp) / (vt * vt));
        if (r.next() > f / a.fmax[ptype]) continue;
        load_gauss(vy, vz, vt, r.next(), r.next(), a.pi);
        vy = vy + a.vb[ptype] * a.sin_thB;
        x = vx * (r.next() * a.span);
        ok = !(x < 0.0 || x > a.xmax);
      }
      if (!ok) {
        a.type[j] = 0;
        atomicOr(&a.t->status, 1);
        return;
      }
      ix = (int)floor(x / a.h1) + 1;
    } else if (m == 0) {
      load_gauss(vx, vy, vt, r.next(), r.next(), a.pi);
      vz = vx;
      carry = vy;
    } else {
      load_gauss(vx, vy, vt, r.next(), r.next(), a.pi);
      vz = carry;
    }
    inj_store(a, j, ptype, ix, iy, x, y, z, vx, vy, vz, true);
  }
}

}  // namespace

namespace lepic {

// Staging capacity for M particles; the staging arrays are kept across calls.
int reserve_stage(lepic_ctx* ctx, int64_t M) {
  if (M <= ctx->stage6_cap) return 0;
  const int64_t cap = std::max<int64_t>(M, 4096);
  if (ctx->d_stage6) LEPIC_HIP(hipFree(ctx->d_stage6));
  if (ctx->d_stype) LEPIC_HIP(hipFree(ctx->d_stype));
  if (ctx->d_srank) LEPIC_HIP(hipFree(ctx->d_srank));
  ctx->d_stage6 = nullptr; ctx->d_stype = nullptr; ctx->d_srank = nullptr; ctx->stage6_cap = 0;
  LEPIC_HIP(hipMalloc(&ctx->d_stage6, (size_t)cap * 48));
  LEPIC_HIP(hipMalloc(&ctx->d_stype, (size_t)cap));
  LEPIC_HIP(hipMalloc(&ctx->d_srank, (size_t)cap * 4));
  ctx->stage6_cap = cap;
  return 0;
}

// Appends the M staged particles to their species (type 0: none).  seg0[q]
// receives the first segment holding species q's new particles and added[q]
// their number.  A species goes to fresh segments after its last one; when
// those would not fit its arrays it is packed first and continued densely.
int append_staged(lepic_ctx* ctx, int64_t M, int64_t* seg0, int64_t* added, bool* dense_out) {
  for (int q = 0; q < kMaxSpecies; ++q) { seg0[q] = 0; added[q] = 0; if (dense_out) dense_out[q] = false; }
  if (M <= 0) return 0;
  SrcTally* t = (SrcTally*)ctx->d_srctally;
  unsigned long long* tot = t->tot;
  if (int rc = zero_async(ctx, tot, sizeof(t->tot))) return rc;
  const unsigned nb = (unsigned)std::min<int64_t>((M + 255) / 256, 1024);
  hipLaunchKernelGGL(k_count_types, dim3(nb), dim3(256), 0, ctx->stream, (const int8_t*)ctx->d_stype, M, tot);
  LEPIC_HIP(hipGetLastError());
  LEPIC_HIP(hipMemcpyAsync(ctx->h_pinned, tot, sizeof(t->tot), hipMemcpyDeviceToHost, ctx->stream));
  LEPIC_HIP(hipStreamSynchronize(ctx->stream));
  unsigned long long n[kMaxSpecies];
  std::memcpy(n, ctx->h_pinned, sizeof(n));
  const bool mixed = ctx->g.precision == LEPIC_MIXED;
  const dim3 g((unsigned)((M + 255) / 256));
  for (int q = 0; q < ctx->nsp; ++q) {
    if (n[q] == 0) continue;
    Species& s = ctx->sp[q];
    const int64_t cnt = (int64_t)n[q];
    const int64_t slots = (s.nseg_cap - 1) * kChunk;   // whole segments the arrays hold
    int64_t base = nseg_of(s) * kChunk;
    bool dense = false;
    if (base + cnt > slots) {
      if (int rc = ensure_packed(ctx, q)) return rc;
      int64_t N = 0;
      if (int rc = lepic_get_count(ctx, q + 1, &N)) return rc;
      if (N + cnt > std::min<int64_t>(slots, s.cap))
        return ctx->fail(LEPIC_ERR_CAPACITY, "k > nmax: particle sources exceed the species capacity");
      base = N;
      dense = true;
    }
    hipLaunchKernelGGL(k_type_flags, g, dim3(256), 0, ctx->stream, (const int8_t*)ctx->d_stype, M, q + 1,
                       ctx->d_srank);
    LEPIC_HIP(hipGetLastError());
    if (int rc = scan_exclusive_u32(ctx, ctx->d_srank, M)) return rc;
    AppendDst o{};
    const int b = s.cur;
    for (int k = 0; k < 6; ++k) { o.d[k] = s.d[b][k]; o.f[k] = s.f[b][k]; }
    o.cell = s.cell[b];
    o.flag_dead = s.flag_dead;
    o.h1 = ctx->g.h1; o.h2 = ctx->g.h2;
    if (mixed)
      hipLaunchKernelGGL(k_append<true>, g, dim3(256), 0, ctx->stream, (const double*)ctx->d_stage6,
                         (const int8_t*)ctx->d_stype, (const uint32_t*)ctx->d_srank, M, q + 1, base, o);
    else
      hipLaunchKernelGGL(k_append<false>, g, dim3(256), 0, ctx->stream, (const double*)ctx->d_stage6,
                         (const int8_t*)ctx->d_stype, (const uint32_t*)ctx->d_srank, M, q + 1, base, o);
    const int64_t s0 = base / kChunk;
    const int64_t ns = (base + cnt - 1) / kChunk - s0 + 1;
    hipLaunchKernelGGL(k_seg_fill, dim3((unsigned)((ns + 255) / 256)), dim3(256), 0, ctx->stream, s.seg[b], s0, ns,
                       base + cnt);
    hipLaunchKernelGGL(k_count_add, dim3(1), dim3(64), 0, ctx->stream, ctx->d_counts, q, cnt);
    LEPIC_HIP(hipGetLastError());
    // dense appends keep a packed species packed; fresh segments after a
    // partial one do not (an empty species starts at slot 0 either way)
    s.packed = dense || (s.packed && base == 0);
    s.ub = base + cnt;
    seg0[q] = s0;
    added[q] = cnt;
    if (dense_out) dense_out[q] = dense;
  }
  return 0;
}

int reserve_srctally(lepic_ctx* ctx) {
  if (!ctx->d_srctally) LEPIC_HIP(hipMalloc(&ctx->d_srctally, sizeof(SrcTally)));
  return 0;
}

}  // namespace lepic

extern "C" {

int lepic_flux_emitter(lepic_ctx* ctx, const lepic_emitter* e, int32_t it, int32_t mode, uint64_t seed,
                       int32_t* iseed, lepic_emitter_info* info) {
  if (!ctx || !e) return LEPIC_ERR_ARG;
  if (mode != LEPIC_MCC_CHECK && mode != LEPIC_MCC_PROD) return LEPIC_ERR_ARG;
  if (mode == LEPIC_MCC_CHECK && !iseed) return LEPIC_ERR_ARG;
  if (!ctx->boundary_set) return ctx->fail(LEPIC_ERR_STATE, "flux emitter before lepic_set_boundary");
  if (e->ptype_fx < 1 || e->ptype_fx > ctx->nsp || e->tag_b < 0 || e->tag_b > ctx->nsp || e->n_icp < 1 ||
      e->ns_Hm < 1)
    return LEPIC_ERR_ARG;
  if (mode == LEPIC_MCC_CHECK && ctx->g.precision != LEPIC_FP64)
    return ctx->fail(LEPIC_ERR_UNSUPPORTED, "check-mode emission replays the reference on the FP64 state");
  const double qe = ctx->g.qe, pi = 4.0 * std::atan(1.0);
  const double ymax = ctx->g.n2 * ctx->g.h2, zmax = ctx->g.zmax, dt = ctx->g.dt;
  // part_flux_emitter.f90:21-66: effective grid length and particle count
  const bool chamfered = e->ptype_fx == ctx->g.tag_neg && e->flag_chmfrd == 1;
  const double js = e->jne * 10.0;
  double Lgrd = (double)e->n_icp * (e->Lg - e->Lh * (double)e->n_holes);
  double th_chrd = 0.0, y1 = 0.0;
  if (chamfered) {
    const double x1 = e->chmfrd_array[0], x2 = e->chmfrd_array[1], y2 = e->chmfrd_array[3];
    y1 = e->chmfrd_array[2];
    const double d_chrd = std::sqrt((x2 - x1) * (x2 - x1) + (y2 - y1) * (y2 - y1));
    th_chrd = std::acos((x2 - x1) / d_chrd);
    Lgrd = 2.0 * (y1 + d_chrd);
    if (Lgrd > ymax) return ctx->fail(LEPIC_ERR_ARG, "Lgrd>ymax, please correct...");
  }
  const double I_sec = js * zmax * Lgrd;
  const double span = (double)(float)e->ns_Hm * dt;
  const double dNh = I_sec * span / (ctx->sp[e->ptype_fx - 1].Nm * qe) / 1.0 / 1.0;
  int64_t Nh = (int64_t)dNh;
  int32_t s = mode == LEPIC_MCC_CHECK ? *iseed : 0;
  double r0;
  if (mode == LEPIC_MCC_CHECK) {
    r0 = ran2(s);   // the rounding draw, line 52, on the host
  } else {
    const U4 w = philox4x32_10(U4{0xFFFFFFFEu, 0u, (uint32_t)it, 0u}, (uint32_t)seed ^ 0x656d6974u,
                               (uint32_t)(seed >> 32));
    r0 = u01(w.x, w.y);
  }
  if (r0 <= dNh - (double)Nh) Nh = Nh + 1;
  if (int rc = reserve_srctally(ctx)) return rc;
  if (int rc = reserve_stage(ctx, std::max<int64_t>(Nh, 1))) return rc;
  SrcTally* t = (SrcTally*)ctx->d_srctally;
  // tally, status and seed start from zero / the host's seed
  {
    LEPIC_HIP(hipStreamSynchronize(ctx->stream));
    SrcTally h{};
    h.seed = s;
    std::memcpy(ctx->h_pinned, &h, sizeof(h));
    LEPIC_HIP(hipMemcpyAsync(t, ctx->h_pinned, sizeof(h), hipMemcpyHostToDevice, ctx->stream));
  }
  EmitArgs a{};
  a.bc = ctx->bc; a.n1 = ctx->g.n1; a.n2 = ctx->g.n2; a.n_icp = e->n_icp;
  a.ptype_fx = e->ptype_fx; a.tag_b = e->tag_b; a.chamfered = chamfered ? 1 : 0;
  a.h1 = ctx->g.h1; a.h2 = ctx->g.h2; a.ymax = ymax; a.zmax = zmax; a.pi = pi;
  a.xg1 = e->xg1; a.Lg = e->Lg; a.y1 = y1; a.span = span;
  a.cham_c[0] = std::cos(th_chrd); a.cham_s[0] = std::sin(th_chrd);
  a.cham_c[1] = std::cos(-th_chrd); a.cham_s[1] = std::sin(-th_chrd);
  const double rot[3] = {pi / 2.0, th_chrd, pi / 2.0 + th_chrd};
  for (int k = 0; k < 3; ++k) { a.rot_c[k] = std::cos(rot[k]); a.rot_s[k] = std::sin(rot[k]); }
  for (int q = 0; q < ctx->nsp; ++q) {
    const Species& sp = ctx->sp[q];
    a.vt[q + 1] = std::sqrt(2.0 * qe * e->THm / std::fabs(sp.mass));
    a.half_Nm_m[q + 1] = 0.5 * sp.Nm * sp.mass;
    a.src[q] = sp.src;
  }
  a.Nh = Nh;
  a.stage = ctx->d_stage6; a.type = ctx->d_stype; a.t = t;
  a.k0 = (uint32_t)seed ^ 0x656d6974u; a.k1 = (uint32_t)(seed >> 32); a.it = (uint32_t)it;
  if (Nh > 0) {
    if (mode == LEPIC_MCC_CHECK) {
      a.max_tries = 1 << 20;
      hipLaunchKernelGGL(k_emit_check, dim3(1), dim3(64), 0, ctx->stream, a);
    } else {
      a.max_tries = 1 << 16;
      hipLaunchKernelGGL(k_emit_prod, dim3((unsigned)((Nh + 255) / 256)), dim3(256), 0, ctx->stream, a);
    }
    LEPIC_HIP(hipGetLastError());
  }
  LEPIC_HIP(hipMemcpyAsync(ctx->h_pinned, t, sizeof(SrcTally), hipMemcpyDeviceToHost, ctx->stream));
  LEPIC_HIP(hipStreamSynchronize(ctx->stream));
  SrcTally h;
  std::memcpy(&h, ctx->h_pinned, sizeof(h));
  if (h.status) return ctx->fail(LEPIC_ERR_STATE, "flux emitter: no electrode cell facing the plasma at xg1");
  int64_t seg0[kMaxSpecies], added[kMaxSpecies];
  if (int rc = append_staged(ctx, Nh, seg0, added, nullptr)) return rc;
  if (mode == LEPIC_MCC_CHECK) *iseed = h.seed;
  if (info) {
    std::memset(info, 0, sizeof(*info));
    info->Nh = Nh;
    info->cnt_b[0] = (int64_t)h.cnt_b[0];
    info->cnt_b[1] = (int64_t)h.cnt_b[1];
    for (int q = 0; q < ctx->nsp && q < 8; ++q) {
      info->P_loss4[q] = h.P_loss4[q + 1];
      info->created[q] = added[q];
    }
    info->Lgrd = Lgrd;
    info->dNh = dNh;
    info->seed = h.seed;
  }
  return 0;
}

}  // extern "C"

extern "C" {

int lepic_part_injection(lepic_ctx* ctx, const lepic_injection* in, const int64_t* N_inj_lost, int32_t it,
                         int32_t mode, uint64_t seed, int32_t* iseed, int32_t deposit, lepic_injection_info* info) {
  if (!ctx || !in || !in->ni0 || !in->vt0) return LEPIC_ERR_ARG;
  if (mode != LEPIC_MCC_CHECK && mode != LEPIC_MCC_PROD) return LEPIC_ERR_ARG;
  if (mode == LEPIC_MCC_CHECK && !iseed) return LEPIC_ERR_ARG;
  const int aopt = in->opt_inj < 0 ? -in->opt_inj : in->opt_inj;
  if (aopt < 1 || aopt > 4 || in->ns_inj < 1 || (aopt == 2 && !N_inj_lost)) return LEPIC_ERR_ARG;
  if (!ctx->boundary_set) return ctx->fail(LEPIC_ERR_STATE, "injection before lepic_set_boundary");
  if (mode == LEPIC_MCC_CHECK && ctx->g.precision != LEPIC_FP64)
    return ctx->fail(LEPIC_ERR_UNSUPPORTED, "check-mode injection replays the reference on the FP64 state");
  const int nsp = ctx->nsp;
  const double qe = ctx->g.qe, dt = ctx->g.dt;
  InjArgs a{};
  a.bc = ctx->bc; a.phi = ctx->phi;
  a.n1 = ctx->g.n1; a.n2 = ctx->g.n2; a.nsp = nsp; a.opt = in->opt_inj; a.aopt = aopt;
  a.flag_c = in->flag_c; a.flag_nmn = in->flag_nmn; a.flag_Tp = in->flag_Tp; a.flag_diffsrc = in->flag_diffsrc;
  a.tag_neg = ctx->g.tag_neg;
  a.h1 = ctx->g.h1; a.h2 = ctx->g.h2; a.xmax = ctx->g.n1 * ctx->g.h1; a.ymax = ctx->g.n2 * ctx->g.h2;
  a.zmax = ctx->g.zmax; a.pi = 4.0 * std::atan(1.0); a.span = (double)(float)in->ns_inj * dt;
  a.xl_pow = in->xl_pow; a.xr_pow = in->xr_pow; a.yl_pow = in->yl_pow; a.yr_pow = in->yr_pow;
  a.xa = in->xa; a.dr = in->dr;
  a.cos_thB = std::cos(in->th_B); a.sin_thB = std::sin(in->th_B);
  int64_t N[kMaxSpecies + 1] = {};
  for (int p = 1; p <= nsp; ++p) {
    const Species& sp = ctx->sp[p - 1];
    const double m = std::fabs(sp.mass);
    a.vt[p] = in->vt0[p - 1];
    if (aopt <= 2 && in->flag_Tp == 1 && p == 1) a.vt[p] = std::sqrt(2.0 * qe * in->THm / m);  // lines 153-158
    if (aopt == 3 || aopt == 4) {
      a.vb[p] = std::sqrt(2.0 * qe * in->THm / m);
      if (aopt == 4) {
        if (p == 1 && in->flag_nmn == 1) a.vb[p] = 0.0;
        const double vt0 = in->vt0[p - 1], vb = a.vb[p];
        const double vmax = (vb + std::sqrt(vb * vb + 2.0 * vt0 * vt0)) / 2.0;   // lines 268-286
        a.fmax[p] = vmax * std::exp(-(vmax - vb) * (vmax - vb) / (vt0 * vt0));
      }
    }
    a.half_Nm_m[p] = 0.5 * sp.Nm * sp.mass;
    a.Nm_q[p] = sp.Nm * sp.charge;
    a.ni0[p] = in->ni0[p - 1];
    a.dN_cur[p] = in->I_inj * (double)in->ns_inj * dt * in->ni0[p - 1] / (qe * ctx->sp[0].Nm) / 1.0 / 1.0;
    N[p] = aopt == 2 ? N_inj_lost[p - 1] : 0;
    a.N_inj[p] = N[p];
    a.src[p - 1] = sp.src;
  }
  // the counts of lines 47-77: exact in production (the rounding draws come
  // from the counter-based stream), an upper bound in check mode (its draws
  // interleave with the particles' and happen on the device)
  const uint32_t k0 = (uint32_t)seed ^ 0x696e6a65u, k1 = (uint32_t)(seed >> 32);
  int64_t total = 0;
  for (int p = 1; p <= nsp; ++p) {
    int flag = 0;
    double dN = 0.0;
    if (aopt == 2) {
      if (p == 1 || (p == a.tag_neg && a.tag_neg > 0)) {
        flag = 1;
        int64_t tmp = 0;
        for (int q = 2; q <= nsp; ++q) tmp += N[q];
        if (a.tag_neg > 0) tmp = tmp - N[a.tag_neg];
        dN = a.ni0[p] * (double)tmp;
      }
    } else {
      flag = 1;
      dN = a.dN_cur[p];
    }
    if (flag == 1) {
      int64_t n = (int64_t)dN;
      if (mode == LEPIC_MCC_CHECK) {
        n = n + 1;
      } else {
        const U4 w = philox4x32_10(U4{0xFFFFFFFDu, (uint32_t)p, (uint32_t)it, 0u}, k0, k1);
        if (u01(w.x, w.y) <= dN - (double)n) n = n + 1;
      }
      N[p] = n;
    }
    if (mode == LEPIC_MCC_CHECK) {
      total += N[p];
    } else if (aopt == 3 || (aopt == 4 && in->flag_nmn == 0)) {
      for (int q = 2; q <= nsp; ++q) N[q] = 0;
    }
  }
  if (mode != LEPIC_MCC_CHECK) {
    a.pair_off[1] = 0;
    for (int p = 1; p <= nsp; ++p) {
      a.n[p] = N[p];
      a.off[p] = total;
      total += N[p];
      a.pair_off[p + 1] = a.pair_off[p] + (N[p] + 1) / 2;
    }
  }
  if (int rc = reserve_srctally(ctx)) return rc;
  if (int rc = reserve_stage(ctx, std::max<int64_t>(total, 1))) return rc;
  SrcTally* t = (SrcTally*)ctx->d_srctally;
  {
    LEPIC_HIP(hipStreamSynchronize(ctx->stream));
    SrcTally h{};
    h.seed = mode == LEPIC_MCC_CHECK ? *iseed : 0;
    std::memcpy(ctx->h_pinned, &h, sizeof(h));
    LEPIC_HIP(hipMemcpyAsync(t, ctx->h_pinned, sizeof(h), hipMemcpyHostToDevice, ctx->stream));
  }
  if (int rc = zero_async(ctx, ctx->d_stype, (size_t)std::max<int64_t>(total, 1))) return rc;
  a.stage = ctx->d_stage6; a.type = ctx->d_stype; a.t = t; a.ploss1 = t->P_loss1;
  a.stage_cap = total;
  a.k0 = k0; a.k1 = k1; a.it = (uint32_t)it;
  if (total > 0 || mode == LEPIC_MCC_CHECK) {
    if (mode == LEPIC_MCC_CHECK) {
      a.max_tries = 1 << 20;
      hipLaunchKernelGGL(k_inject_check, dim3(1), dim3(64), 0, ctx->stream, a);
    } else {
      a.max_tries = 1 << 16;
      const int64_t pairs = a.pair_off[nsp + 1];
      if (pairs > 0)
        hipLaunchKernelGGL(k_inject_prod, dim3((unsigned)((pairs + 255) / 256)), dim3(256), 0, ctx->stream, a);
    }
    LEPIC_HIP(hipGetLastError());
  }
  LEPIC_HIP(hipMemcpyAsync(ctx->h_pinned, t, sizeof(SrcTally), hipMemcpyDeviceToHost, ctx->stream));
  LEPIC_HIP(hipStreamSynchronize(ctx->stream));
  SrcTally h;
  std::memcpy(&h, ctx->h_pinned, sizeof(h));
  if (h.status == 1) return ctx->fail(LEPIC_ERR_STATE, "injection: no admissible position in the injection region");
  if (h.status) return ctx->fail(LEPIC_ERR_CAPACITY, "injection: staging overflow");
  int64_t seg0[kMaxSpecies], added[kMaxSpecies];
  bool dense[kMaxSpecies];
  if (int rc = append_staged(ctx, total, seg0, added, dense)) return rc;
  if (deposit) {
    // injection precedes the deposition (main.f90:1063-1085): the new
    // particles join this step's density
    for (int q = 0; q < nsp; ++q) {
      if (!added[q]) continue;
      if (dense[q]) {
        if (int rc = deposit_launch(ctx, q)) return rc;
      } else if (int rc = deposit_range_launch(ctx, q, seg0[q], (added[q] + kChunk - 1) / kChunk)) {
        return rc;
      }
    }
  }
  if (mode == LEPIC_MCC_CHECK) *iseed = h.seed;
  if (info) {
    std::memset(info, 0, sizeof(*info));
    for (int q = 0; q < nsp && q < 8; ++q) {
      info->injected[q] = added[q];
      info->P_loss4[q] = h.P_loss4[q + 1];
      info->P_loss1[q] = h.P_loss1[q + 1];
    }
    info->seed = h.seed;
  }
  return 0;
}

}  // extern "C"
