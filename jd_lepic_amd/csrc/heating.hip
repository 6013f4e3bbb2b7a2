// heating.hip -- electron heating, eheating (e_heating.f90:1-112).
//
// Check mode replays the reference on one lane: the particles in their
// order, ran2 drawn exactly where the reference draws it (the heating test,
// the Maxwellian pair, and a second pair whose other value is kept for the
// next heated particle's vz).  Production mode gives every particle slot its
// own Philox4x32-10 stream keyed by (seed, it): one block per 1024-slot
// segment as the mover leaves them, the absorbed power summed per block in a
// fixed order and added with one atomic per block.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>

#include "lepic_internal.h"

using namespace lepic;

namespace {

struct HeatArgs {
  int flag_c, flag_gpp, eheat_type, n_icp;
  double xa, dr, yl_pow, yr_pow, ymax, nudt, dt, Pabs_cor, half_Nm_m, pi, h1, h2;
  const double* vt;
};

// region test and rate of e_heating.f90:48-67; returns false outside
__device__ __forceinline__ bool heat_rate(const HeatArgs& a, double xp, double yp, double& nu) {
  double rp2;
  bool in;
  if (a.flag_c == 0) {
    rp2 = (xp - a.xa) * (xp - a.xa);
    in = rp2 <= a.dr * a.dr && yp >= a.yl_pow && yp <= a.yr_pow;
  } else {
    rp2 = (xp - a.xa) * (xp - a.xa) + (yp - a.ymax / 2.0) * (yp - a.ymax / 2.0);
    in = rp2 <= a.dr * a.dr;
  }
  nu = a.nudt / a.dt;
  if (in && a.flag_gpp == 1) nu = (a.nudt / a.dt) * 9.0 * a.Pabs_cor * exp(-9.0 * rp2 / (a.dr * a.dr));
  return in;
}

__global__ void k_heat_check(HeatArgs a, double* x, double* y, double* vx, double* vy, double* vz,
                             const int64_t* nptr, int32_t* iseed, double* dP) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  const int64_t N = *nptr;
  int32_t s = *iseed;
  double vz_sav = 0.0, P = 0.0;
  double nu_tmp = a.nudt / a.dt;
  for (int64_t i = 0; i < N; ++i) {
    const double xp = x[i], yp = y[i];
    double nu;
    if (!heat_rate(a, xp, yp, nu)) continue;
    if (a.flag_gpp == 1) nu_tmp = nu;     // the reference keeps the last value otherwise
    if (ran2(s) <= nu_tmp * a.dt) {
      const double v2old = vx[i] * vx[i] + vy[i] * vy[i] + vz[i] * vz[i];
      const int i_icp = min(max((int)(a.n_icp * yp / a.ymax) + 1, 1), a.n_icp);   // y = ymax would read vt(n_icp+1)
      const double vt = a.vt[i_icp - 1];
      double dvx, dvy, dvz;
      const double r1 = ran2(s), r2 = ran2(s);
      load_gauss(dvx, dvy, vt, r1, r2, a.pi);
      if (vz_sav == 0.0) {
        const double r3 = ran2(s), r4 = ran2(s);
        load_gauss(dvz, vz_sav, vt, r3, r4, a.pi);
      } else {
        dvz = vz_sav;
        vz_sav = 0.0;
      }
      if (a.eheat_type == 1) {
        vx[i] = dvx; vy[i] = dvy; vz[i] = dvz;
      } else {
        vx[i] = vx[i] + dvx; vy[i] = vy[i] + dvy; vz[i] = vz[i] + dvz;
      }
      P = P + a.half_Nm_m * (vx[i] * vx[i] + vy[i] * vy[i] + vz[i] * vz[i] - v2old);
    }
  }
  *iseed = s;
  *dP = P;
}

template <bool MIXED>
__global__ __launch_bounds__(kBlock) void k_heat_prod(HeatArgs a, const int32_t* seg, const double* x,
                                                      const double* y, double* vx, double* vy, double* vz,
                                                      const int32_t* cell, const float* fx, const float* fy,
                                                      float* fvx, float* fvy, float* fvz, uint32_t k0,
                                                      uint32_t k1, uint32_t it, double* dP) {
  __shared__ double ws[kBlock / 64];
  const int cnt = seg[blockIdx.x];
  double P = 0.0;
  for (int k = 0; k < kItems; ++k) {
    const int slot = k * kBlock + threadIdx.x;
    if (slot >= cnt) continue;
    const int64_t i = (int64_t)blockIdx.x * kChunk + slot;
    double xp, yp, ux, uy, uz;
    if (MIXED) {
      const int c = cell[i];
      xp = ((double)((c & 0xFFFF) - 1) + (double)fx[i]) * a.h1;
      yp = ((double)((c >> 16) - 1) + (double)fy[i]) * a.h2;
      ux = fvx[i]; uy = fvy[i]; uz = fvz[i];
    } else {
      xp = x[i]; yp = y[i];
      ux = vx[i]; uy = vy[i]; uz = vz[i];
    }
    double nu;
    if (!heat_rate(a, xp, yp, nu)) continue;
    const U4 c0{(uint32_t)i, (uint32_t)(i >> 32), it, 0u};
    const U4 r0 = philox4x32_10(c0, k0, k1);
    if (u01(r0.x, r0.y) > nu * a.dt) continue;
    const U4 r1 = philox4x32_10(U4{c0.x, c0.y, c0.z, 1u}, k0, k1);
    const int i_icp = min(max((int)(a.n_icp * yp / a.ymax) + 1, 1), a.n_icp);
    const double vt = a.vt[i_icp - 1];
    double dvx, dvy, dvz, unused;
    load_gauss(dvx, dvy, vt, u01(r0.z, r0.w), u01(r1.x, r1.y), a.pi);
    const U4 r2 = philox4x32_10(U4{c0.x, c0.y, c0.z, 2u}, k0, k1);
    load_gauss(dvz, unused, vt, u01(r1.z, r1.w), u01(r2.x, r2.y), a.pi);
    const double v2old = ux * ux + uy * uy + uz * uz;
    if (a.eheat_type == 1) {
      ux = dvx; uy = dvy; uz = dvz;
    } else {
      ux += dvx; uy += dvy; uz += dvz;
    }
    if (MIXED) {
      fvx[i] = (float)ux; fvy[i] = (float)uy; fvz[i] = (float)uz;
      ux = fvx[i]; uy = fvy[i]; uz = fvz[i];   // the tally sees the stored values
    } else {
      vx[i] = ux; vy[i] = uy; vz[i] = uz;
    }
    P += a.half_Nm_m * (ux * ux + uy * uy + uz * uz - v2old);
  }
  P = wave_sum(P);
  if ((threadIdx.x & 63) == 0) ws[threadIdx.x >> 6] = P;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int w = 0; w < kBlock / 64; ++w) t += ws[w];
    if (t != 0.0) atomic_add_f64(dP, t);
  }
}

}  // namespace

extern "C" {

int lepic_eheating(lepic_ctx* ctx, const lepic_heating* h, int32_t mode, uint64_t seed, int32_t it, int32_t* iseed,
                   double* dP_loss2) {
  if (!ctx || !h || !h->vt || h->n_icp < 1 || h->n_icp > 64) return LEPIC_ERR_ARG;
  if (mode != LEPIC_MCC_CHECK && mode != LEPIC_MCC_PROD) return LEPIC_ERR_ARG;
  if (mode == LEPIC_MCC_CHECK && !iseed) return LEPIC_ERR_ARG;
  const bool mixed = ctx->g.precision == LEPIC_MIXED;
  if (mode == LEPIC_MCC_CHECK && mixed)
    return ctx->fail(LEPIC_ERR_UNSUPPORTED, "check-mode heating replays the reference on the FP64 state");
  Species& s = ctx->sp[0];   // e_heating.f90:33: electrons only
  HeatArgs a{};
  a.flag_c = h->flag_c; a.flag_gpp = h->flag_gpp; a.eheat_type = h->eheat_type; a.n_icp = h->n_icp;
  a.xa = h->xa; a.dr = h->dr; a.yl_pow = h->yl_pow; a.yr_pow = h->yr_pow;
  a.ymax = ctx->g.n2 * ctx->g.h2;
  a.nudt = h->nudt; a.dt = ctx->g.dt; a.Pabs_cor = h->Pabs_cor;
  a.half_Nm_m = 0.5 * s.Nm * s.mass;
  a.pi = 4.0 * std::atan(1.0);
  a.h1 = ctx->g.h1; a.h2 = ctx->g.h2;
  // device scratch: vt(1:64), then the seed and the power, staged through
  // the pinned buffer
  if (!ctx->d_heat) LEPIC_HIP(hipMalloc(&ctx->d_heat, 66 * 8));
  double* d = ctx->d_heat;
  LEPIC_HIP(hipStreamSynchronize(ctx->stream));   // h_pinned may still be the target of a copy
  std::memcpy(ctx->h_pinned, h->vt, (size_t)h->n_icp * 8);
  int32_t s0[2] = {mode == LEPIC_MCC_CHECK ? *iseed : 0, 0};
  std::memcpy(ctx->h_pinned + 64, s0, 8);
  ctx->h_pinned[65] = 0.0;
  LEPIC_HIP(hipMemcpyAsync(d, ctx->h_pinned, 66 * 8, hipMemcpyHostToDevice, ctx->stream));
  a.vt = d;
  int32_t* ds = (int32_t*)(d + 64);
  double* dP = d + 65;
  if (mode == LEPIC_MCC_CHECK) {
    if (int rc = ensure_packed(ctx, 0)) return rc;
    const int b = s.cur;
    hipLaunchKernelGGL(k_heat_check, dim3(1), dim3(64), 0, ctx->stream, a, s.d[b][0], s.d[b][1], s.d[b][3],
                       s.d[b][4], s.d[b][5], (const int64_t*)(ctx->d_counts + 0), ds, dP);
  } else {
    const int b = s.cur;
    const int64_t nb = nseg_of(s);
    const uint32_t k0 = (uint32_t)seed ^ 0x68656174u, k1 = (uint32_t)(seed >> 32);
    if (nb > 0) {
      if (mixed)
        hipLaunchKernelGGL(k_heat_prod<true>, dim3((unsigned)nb), dim3(kBlock), 0, ctx->stream, a,
                           (const int32_t*)s.seg[b], (const double*)nullptr, (const double*)nullptr, (double*)nullptr,
                           (double*)nullptr, (double*)nullptr, (const int32_t*)s.cell[b], (const float*)s.f[b][0],
                           (const float*)s.f[b][1], s.f[b][3], s.f[b][4], s.f[b][5], k0, k1, (uint32_t)it, dP);
      else
        hipLaunchKernelGGL(k_heat_prod<false>, dim3((unsigned)nb), dim3(kBlock), 0, ctx->stream, a,
                           (const int32_t*)s.seg[b], (const double*)s.d[b][0], (const double*)s.d[b][1], s.d[b][3],
                           s.d[b][4], s.d[b][5], (const int32_t*)nullptr, (const float*)nullptr,
                           (const float*)nullptr, (float*)nullptr, (float*)nullptr, (float*)nullptr, k0, k1,
                           (uint32_t)it, dP);
    }
  }
  LEPIC_HIP(hipGetLastError());
  LEPIC_HIP(hipMemcpyAsync(ctx->h_pinned, d + 64, 16, hipMemcpyDeviceToHost, ctx->stream));
  LEPIC_HIP(hipStreamSynchronize(ctx->stream));
  double out[2];
  std::memcpy(out, ctx->h_pinned, 16);
  if (dP_loss2) *dP_loss2 = out[1];
  if (mode == LEPIC_MCC_CHECK) std::memcpy(iseed, &out[0], 4);
  return 0;
}

}  // extern "C"
