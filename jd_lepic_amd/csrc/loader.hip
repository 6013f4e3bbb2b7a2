// loader.hip -- synthetic particle loading (the role of load_part.f90:59-224)
// and the whole-step driver entry point (main.f90:689-1090 order).
//
// The loader fills every non-wall cell of [xl,xr) with ppc particles in
// y-major cell order (so a fresh load is already cell-sorted), positions
// uniform inside the cell and shared by all species loaded with the same
// seed (the reference puts every species at one site, load_part.f90:141-143),
// velocities drawn like load_gauss.f90:14-21 with
// thermal speed vt; the random numbers come from Philox keyed by the seed and
// counted by (cell, particle, species), so a load is reproducible and
// independent of the launch shape.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>

#include "lepic_internal.h"

using namespace lepic;

namespace {

__global__ void k_cell_flags(const int8_t* __restrict__ bc, int n1, int n2, double h1, double xl, double xr,
                             uint32_t* __restrict__ flag) {
  int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= (int64_t)n1 * n2) return;
  const int ix = (int)(c % n1) + 1, iy = (int)(c / n1) + 1;
  const double xc = (ix - 0.5) * h1;
  const int64_t a = node_index(ix, iy, n1), b = node_index(ix, iy + 1, n1);
  const bool walls = bc[a] >= 1 && bc[a + 1] >= 1 && bc[b + 1] >= 1 && bc[b] >= 1;  // load_part.f90:115-116
  flag[c] = (!walls && xc >= xl && xc < xr) ? 1u : 0u;
}

struct LoadArgs {
  const uint32_t* rank;          // exclusive scan of the flags
  const uint32_t* flag;
  int n1, n2, ppc, ptype, mixed, track_z;
  double h1, h2, zmax, vt, pi;
  int64_t base;                  // first index to write
  uint32_t k0, k1;
  double* d[6];
  int32_t* cell;
  float* f[6];
};

__global__ void k_load(LoadArgs A) {
  const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t c = g / A.ppc;
  const int k = (int)(g % A.ppc);
  if (c >= (int64_t)A.n1 * A.n2 || !A.flag[c]) return;
  const int ix = (int)(c % A.n1) + 1, iy = (int)(c / A.n1) + 1;
  const int64_t i = A.base + (int64_t)A.rank[c] * A.ppc + k;
  // positions do not depend on the species: every species loaded with one
  // seed shares its sites, like load_part.f90:141-143
  U4 r0 = philox4x32_10(U4{(uint32_t)c, (uint32_t)k, 0u, 0u}, A.k0, A.k1);
  U4 r1 = philox4x32_10(U4{(uint32_t)c, (uint32_t)k, (uint32_t)A.ptype, 1u}, A.k0, A.k1);
  U4 r2 = philox4x32_10(U4{(uint32_t)c, (uint32_t)k, (uint32_t)A.ptype, 2u}, A.k0, A.k1);
  const double ux = u01(r0.x, r0.y), uy = u01(r0.z, r0.w);
  double vx, vy, vz, unused;
  {
    double vp = A.vt * sqrt(-log(1.0 - u01(r1.x, r1.y)));
    double th = 2.0 * A.pi * u01(r1.z, r1.w);
    vx = vp * cos(th);
    vy = vp * sin(th);
    vp = A.vt * sqrt(-log(1.0 - u01(r2.x, r2.y)));
    th = 2.0 * A.pi * u01(r2.z, r2.w);
    vz = vp * cos(th);
    unused = vp * sin(th);
    (void)unused;
  }
  const double z = u01(r2.y ^ r2.z, r2.x ^ r2.w) * A.zmax;
  if (A.mixed) {
    float fx = (float)ux, fy = (float)uy;
    if (fx >= 1.0f) fx = 0.99999994f;
    if (fy >= 1.0f) fy = 0.99999994f;
    A.cell[i] = (iy << 16) | ix;
    A.f[0][i] = fx; A.f[1][i] = fy;
    if (A.track_z) A.f[2][i] = (float)z;
    A.f[3][i] = (float)vx; A.f[4][i] = (float)vy; A.f[5][i] = (float)vz;
  } else {
    A.d[0][i] = ((ix - 1) + ux) * A.h1;
    A.d[1][i] = ((iy - 1) + uy) * A.h2;
    A.d[2][i] = z;
    A.d[3][i] = vx; A.d[4][i] = vy; A.d[5][i] = vz;
  }
}

__global__ void k_bump_count(int64_t* counts, int p, const uint32_t* rank, const uint32_t* flag, int64_t nc,
                             int ppc) {
  if (threadIdx.x || blockIdx.x) return;
  counts[p] += (int64_t)(rank[nc - 1] + flag[nc - 1]) * ppc;
}

}  // namespace

namespace lepic {

int load_uniform_launch(lepic_ctx* ctx, int p, int ppc, double vt, double xl, double xr, uint64_t seed) {
  Species& s = ctx->sp[p];
  const int n1 = ctx->g.n1, n2 = ctx->g.n2;
  const int64_t nc = (int64_t)n1 * n2;
  int64_t N0 = 0;
  if (int rc = ensure_packed(ctx, p)) return rc;
  if (int rc = lepic_get_count(ctx, p + 1, &N0)) return rc;
  uint32_t* buf = nullptr;
  LEPIC_HIP(hipMalloc(&buf, 2 * nc * 4));
  uint32_t *flag = buf, *rank = buf + nc;
  hipLaunchKernelGGL(k_cell_flags, dim3((unsigned)((nc + 255) / 256)), dim3(256), 0, ctx->stream, ctx->bc, n1, n2,
                     ctx->g.h1, xl, xr, flag);
  LEPIC_HIP(hipMemcpyAsync(rank, flag, nc * 4, hipMemcpyDeviceToDevice, ctx->stream));
  if (int rc = scan_exclusive_u32(ctx, rank, nc)) return rc;
  uint32_t last[2];
  LEPIC_HIP(hipMemcpyAsync(&last[0], rank + nc - 1, 4, hipMemcpyDeviceToHost, ctx->stream));
  LEPIC_HIP(hipMemcpyAsync(&last[1], flag + nc - 1, 4, hipMemcpyDeviceToHost, ctx->stream));
  LEPIC_HIP(hipStreamSynchronize(ctx->stream));
  const int64_t add = (int64_t)(last[0] + last[1]) * ppc;
  if (N0 + add > s.cap) {
    hipFree(buf);
    return ctx->fail(LEPIC_ERR_CAPACITY, "load exceeds the species capacity");
  }
  LoadArgs A{};
  A.rank = rank; A.flag = flag; A.n1 = n1; A.n2 = n2; A.ppc = ppc; A.ptype = p + 1;
  A.mixed = ctx->g.precision == LEPIC_MIXED; A.track_z = ctx->g.track_z;
  A.h1 = ctx->g.h1; A.h2 = ctx->g.h2; A.zmax = ctx->g.zmax; A.vt = vt; A.pi = 4.0 * std::atan(1.0);
  A.base = N0; A.k0 = (uint32_t)seed; A.k1 = (uint32_t)(seed >> 32);
  const int b = s.cur;
  for (int k = 0; k < 6; ++k) { A.d[k] = s.d[b][k]; A.f[k] = s.f[b][k]; }
  A.cell = s.cell[b];
  const int64_t nthreads = nc * ppc;
  hipLaunchKernelGGL(k_load, dim3((unsigned)((nthreads + 255) / 256)), dim3(256), 0, ctx->stream, A);
  hipLaunchKernelGGL(k_bump_count, dim3(1), dim3(64), 0, ctx->stream, ctx->d_counts, p, (const uint32_t*)rank,
                     (const uint32_t*)flag, nc, ppc);
  LEPIC_HIP(hipGetLastError());
  LEPIC_HIP(hipStreamSynchronize(ctx->stream));
  LEPIC_HIP(hipFree(buf));
  s.ub = N0 + add;
  return set_packed_counts(ctx, p);
}

// SUM(p_mac(ptype, np_loss, 0:ngrid)) per species: particles lost at walls
int lost_counts(lepic_ctx* ctx, double* out) {
  for (int q = 0; q < ctx->nsp; ++q) {
    double p_mac[2 * 65];
    if (int rc = lepic_get_mover_tally(ctx, q + 1, nullptr, p_mac, nullptr)) return rc;
    double s = 0.0;
    for (int g = 0; g <= ctx->g.ngrid; ++g) s += p_mac[g];
    out[q] = s;
  }
  return 0;
}

}  // namespace lepic

extern "C" int lepic_step(lepic_ctx* ctx, int32_t it, lepic_step_args* a) {
  if (!ctx || !a) return LEPIC_ERR_ARG;
  hipEvent_t ev[5];
  for (auto& e : ev) LEPIC_HIP(hipEventCreate(&e));
  LEPIC_HIP(hipEventRecord(ev[0], ctx->stream));
  int rc = lepic_calc_rhs(ctx);  // main.f90:821
  if (!rc) rc = lepic_iterative_solver(ctx, a->eps, a->ncycl, a->start, &a->solve, nullptr, 0);  // :829
  if (!rc) rc = lepic_calc_efield(ctx);  // :875
  LEPIC_HIP(hipEventRecord(ev[1], ctx->stream));
  if (!rc && a->nsort > 0 && it % a->nsort == 1)  // :886-901
    for (int p = 1; p <= ctx->nsp && !rc; ++p) rc = lepic_part_sorting(ctx, p, nullptr);
  LEPIC_HIP(hipEventRecord(ev[2], ctx->stream));
  if (!rc && a->mcc_mode >= 0 && it > 1 && a->ns_coll > 0 && it % a->ns_coll == 1) {  // :970-979
    rc = lepic_dens_red(ctx, nullptr);
    lepic_mcc_info info;
    if (!rc) rc = lepic_collisions(ctx, it, a->mcc_mode, a->seed, &a->iseed, &info);
  }
  a->P_heat = 0.0;
  if (!rc && a->heat && a->ns_heat > 0 && it % a->ns_heat == 0)  // :1033-1037
    rc = lepic_eheating(ctx, a->heat, a->heat_mode, a->seed, it, &a->iseed, &a->P_heat);
  LEPIC_HIP(hipEventRecord(ev[3], ctx->stream));
  const bool inject = a->inj && a->inj->ns_inj > 0 && it % a->inj->ns_inj == 0;
  // opt_inj = +-2 re-injects the wall losses of every push since the last
  // injection: sav_np before and after each push (main.f90:1044-1056)
  const bool count_losses = a->inj && (a->inj->opt_inj == 2 || a->inj->opt_inj == -2);
  double lost0[kMaxSpecies] = {};
  if (!rc && count_losses) rc = lost_counts(ctx, lost0);
  for (int p = 1; p <= ctx->nsp && !rc; ++p) rc = lepic_part_mover(ctx, p, 1);  // :1041-1086
  if (!rc && count_losses) {
    double lost1[kMaxSpecies] = {};
    rc = lost_counts(ctx, lost1);
    for (int q = 0; q < ctx->nsp; ++q) ctx->inj_lost[q] += lost1[q] - lost0[q];
  }
  a->injected = 0;
  if (!rc && inject) {  // :1063-1071
    int64_t N_inj[kMaxSpecies] = {};
    for (int q = 0; q < ctx->nsp; ++q) N_inj[q] = (int64_t)ctx->inj_lost[q];
    lepic_injection_info info;
    rc = lepic_part_injection(ctx, a->inj, N_inj, it, a->heat_mode, a->seed, &a->iseed, 1, &info);
    if (!rc)
      for (int q = 0; q < ctx->nsp; ++q) a->injected += info.injected[q];
    for (int q = 0; q < ctx->nsp; ++q) ctx->inj_lost[q] = 0.0;   // "Reset counter", :1070
  }
  a->emitted = 0;
  if (!rc && a->emit && a->emit->jne > 0.0 && (a->emit->ns_Hm == 1 || it % a->emit->ns_Hm == 1)) {  // :1092-1096
    lepic_emitter_info info;
    rc = lepic_flux_emitter(ctx, a->emit, it, a->heat_mode, a->seed, &a->iseed, &info);
    if (!rc) a->emitted = info.Nh;
  }
  LEPIC_HIP(hipEventRecord(ev[4], ctx->stream));
  LEPIC_HIP(hipEventSynchronize(ev[4]));
  float ms[4];
  for (int k = 0; k < 4; ++k) LEPIC_HIP(hipEventElapsedTime(&ms[k], ev[k], ev[k + 1]));
  a->ms_rhs_poisson_e = ms[0]; a->ms_sort = ms[1]; a->ms_mcc = ms[2]; a->ms_push = ms[3];
  for (auto& e : ev) hipEventDestroy(e);
  return rc;
}
