// moments.hip -- averaged particle moments on the GPU: part_moments
// (utils.f90:1-136), calc_avg (utils.f90:138-265) and the phi/E averages
// of the main loop (main.f90:903-944).
//
// part_moments: one block per 1024-slot particle segment (the particle
// layout of the mover), thread t owning four consecutive slots; each
// particle adds k1*kq*w*(|v|^2, vx, vy, vz) to the four nodes around its
// half-step-back position on the grid coarsened by size_na.  Consecutive
// particles of one thread that share a coarse cell are summed in registers
// first (the species is cell-sorted every few steps); the sums go to the
// FP64 moment grid with atomics, so only the summation order differs from
// the reference's serial loop.  calc_avg and the phi/E averages are
// elementwise over nodes.  Diagnostics: called every navg steps, so these
// kernels are HBM/atomic bound without further tuning.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

#include "lepic_internal.h"

using namespace lepic;

namespace {

struct MomArgs {
  const int32_t* seg;            // particles per segment
  // FP64 form
  const double *x, *y, *vx, *vy, *vz;
  // mixed form
  const int32_t* cell;
  const float *fx, *fy, *fvx, *fvy, *fvz;
  const int8_t* bc;              // fine bcnd: kq = 2 where bc != -1
  int n1, size_na, na1, na2;
  double h1, h2, ha1, ha2, xmax, ymax, half_dt, k1;
  double* pm;                    // (na2+1)(na1+1) x 4: |v|^2, vx, vy, vz
};

__device__ __forceinline__ double kq_at(const MomArgs& a, int ix, int iy) {
  return a.bc[(int64_t)iy * (a.n1 + 3) + ix] != -1 ? 2.0 : 1.0;
}

template <bool MIXED>
__global__ __launch_bounds__(kBlock) void k_part_moments(MomArgs a) {
  const int cnt = a.seg[blockIdx.x];
  const int64_t i0 = (int64_t)blockIdx.x * kChunk + (int64_t)kItems * threadIdx.x;
  const int nv = min(max(cnt - kItems * (int)threadIdx.x, 0), kItems);
  // running sum of consecutive particles in one coarse cell
  int cur = -1, cix = 0, ciy = 0;
  double acc[4][4] = {};         // [corner][|v|^2, vx, vy, vz]
  auto flush = [&]() {
    if (cur < 0) return;
    const int cx[4] = {cix, cix + 1, cix + 1, cix}, cy[4] = {ciy, ciy, ciy + 1, ciy + 1};
    for (int c = 0; c < 4; ++c) {
      if (cx[c] > a.na1 + 1 || cy[c] > a.na2 + 1) continue;
      double* m = a.pm + 4 * ((int64_t)(cy[c] - 1) * (a.na1 + 1) + (cx[c] - 1));
      for (int q = 0; q < 4; ++q)
        if (acc[c][q] != 0.0) atomic_add_f64(m + q, acc[c][q]);
    }
  };
  for (int k = 0; k < nv; ++k) {
    const int64_t i = i0 + k;
    double xp, yp, vx, vy, vz;
    if (MIXED) {
      const int c = a.cell[i];
      vx = a.fvx[i]; vy = a.fvy[i]; vz = a.fvz[i];
      xp = ((double)((c & 0xFFFF) - 1) + (double)a.fx[i]) * a.h1;
      yp = ((double)((c >> 16) - 1) + (double)a.fy[i]) * a.h2;
    } else {
      vx = a.vx[i]; vy = a.vy[i]; vz = a.vz[i];
      xp = a.x[i];
      yp = a.y[i];
    }
    xp = xp - vx * a.half_dt;
    yp = yp - vy * a.half_dt;
    if (xp > a.xmax || xp < 0.0 || yp > a.ymax || yp < 0.0) continue;
    const double Eki = vx * vx + vy * vy + vz * vz;
    const int ix = (int)(xp / a.ha1) + 1;
    const double px = (ix * a.ha1 - xp) / a.ha1;
    const int iy = (int)(yp / a.ha2) + 1;
    const double py = (iy * a.ha2 - yp) / a.ha2;
    const int key = iy * (a.na1 + 3) + ix;
    if (key != cur) {
      flush();
      cur = key; cix = ix; ciy = iy;
      for (int c = 0; c < 4; ++c)
        for (int q = 0; q < 4; ++q) acc[c][q] = 0.0;
    }
    const double ki[4] = {a.k1 * px * py, a.k1 * (1.0 - px) * py, a.k1 * (1.0 - px) * (1.0 - py),
                          a.k1 * px * (1.0 - py)};
    const int ixf = a.size_na * ix - (a.size_na - 1), iyf = a.size_na * iy - (a.size_na - 1);
    // kq at the fine node next to the coarse node's, as utils.f90:103-135 indexes it
    const double kq[4] = {kq_at(a, ixf, iyf), kq_at(a, ixf + 1, iyf), kq_at(a, ixf + 1, iyf + 1),
                          kq_at(a, ixf, iyf + 1)};
    for (int c = 0; c < 4; ++c) {
      const double kt = kq[c] * ki[c];
      acc[c][0] += kt * Eki;
      acc[c][1] += kt * vx;
      acc[c][2] += kt * vy;
      acc[c][3] += kt * vz;
    }
  }
  flush();
}

struct AvgArgs {
  const double* np;              // fine density of the species
  const double* sink;            // ss2D(2) of the species
  const double* src;             // ss2D(1) of the species
  const double* pm;
  const int8_t* bc;
  int n1, size_na, na1, na2;
  double mass_over_qe, k;        // mass/qe, Nm/Vc
  double* out;                   // data_pavg of the species: (na2+1)(na1+1) x 7
};

__global__ void k_calc_avg(AvgArgs a) {
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= (int64_t)(a.na1 + 1) * (a.na2 + 1)) return;
  const int ix = (int)(q % (a.na1 + 1)) + 1, iy = (int)(q / (a.na1 + 1)) + 1;
  const int64_t f = (int64_t)(a.size_na * iy - (a.size_na - 1)) * (a.n1 + 3) + (a.size_na * ix - (a.size_na - 1));
  const double* m = a.pm + 4 * q;
  double* o = a.out + 7 * q;
  const double n = a.np[f], jx = m[1], jy = m[2], jz = m[3];
  o[0] = o[0] + n;
  o[2] = o[2] + jx;
  o[3] = o[3] + jy;
  o[4] = o[4] + jz;
  double Tp = 0.0;
  if (n > 0) Tp = a.mass_over_qe * (1.0 / 3.0) * (m[0] / n - (jx * jx + jy * jy + jz * jz) / (n * n));
  if (Tp > 0) o[1] = o[1] + Tp;
  o[5] = o[5] + a.src[f] * a.k;
  o[6] = o[6] + a.sink[f] * a.k;
}

__global__ void k_phi_avg(const double* __restrict__ phi, const double* __restrict__ E, int64_t nodes,
                          double* __restrict__ out) {
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= nodes) return;
  out[3 * q] += phi[q];
  out[3 * q + 1] += E[2 * q];
  out[3 * q + 2] += E[2 * q + 1];
}

unsigned nblk(int64_t n) { return (unsigned)((n + 255) / 256); }


// ---- charge and energy totals ----------------------------------------------
constexpr int kTotBlocks = 1024;   // per-block partials, summed on the host in block order

// Block sum in a fixed order: waves by xor-shuffles, then the four wave sums
// in wave order.
__device__ __forceinline__ double block_sum256(double v, double* s_w) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  if ((threadIdx.x & 63) == 0) s_w[threadIdx.x >> 6] = v;
  __syncthreads();
  return s_w[0] + s_w[1] + s_w[2] + s_w[3];
}

// sum |v|^2 over the particles of one species (either layout); blocks stride
// over the segments, threads over a segment's slots
__global__ __launch_bounds__(256) void k_species_v2(const int32_t* __restrict__ seg, int64_t nseg,
                                                    const double* __restrict__ vx, const double* __restrict__ vy,
                                                    const double* __restrict__ vz, const float* __restrict__ fvx,
                                                    const float* __restrict__ fvy, const float* __restrict__ fvz,
                                                    double* __restrict__ partial) {
  __shared__ double s_w[4];
  double s = 0.0;
  for (int64_t g = blockIdx.x; g < nseg; g += gridDim.x) {
    const int cnt = seg[g];
    for (int j = threadIdx.x; j < cnt; j += 256) {
      const int64_t i = g * kChunk + j;
      double a, b, c;
      if (vx) { a = vx[i]; b = vy[i]; c = vz[i]; }
      else { a = fvx[i]; b = fvy[i]; c = fvz[i]; }
      s += a * a + b * b + c * c;
    }
  }
  const double t = block_sum256(s, s_w);
  if (threadIdx.x == 0) partial[blockIdx.x] = t;
}

// sum |E|^2 over the interior nodes (bcnd == -1)
__global__ __launch_bounds__(256) void k_field_e2(const double* __restrict__ E, const int8_t* __restrict__ bc,
                                                  int64_t nodes, double* __restrict__ partial) {
  __shared__ double s_w[4];
  double s = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < nodes; i += (int64_t)gridDim.x * 256)
    if (bc[i] == -1) s += E[2 * i] * E[2 * i] + E[2 * i + 1] * E[2 * i + 1];
  const double t = block_sum256(s, s_w);
  if (threadIdx.x == 0) partial[blockIdx.x] = t;
}

// launch one of the sums above with nb blocks and add the partials in order
template <typename Launch>
int sum_partials(lepic_ctx* ctx, int nb, Launch launch, double* out) {
  static_assert(kTotBlocks <= kPinnedDoubles, "partials fit the pinned scratch");
  if (!ctx->d_diag) LEPIC_HIP(hipMalloc(&ctx->d_diag, kTotBlocks * 8));
  launch(ctx->d_diag);
  LEPIC_HIP(hipGetLastError());
  LEPIC_HIP(hipMemcpyAsync(ctx->h_pinned, ctx->d_diag, (size_t)nb * 8, hipMemcpyDeviceToHost, ctx->stream));
  LEPIC_HIP(hipStreamSynchronize(ctx->stream));
  double t = 0.0;
  for (int b = 0; b < nb; ++b) t += ctx->h_pinned[b];
  *out = t;
  return 0;
}
}  // namespace

extern "C" {

int lepic_moments_reset(lepic_ctx* ctx) {
  if (!ctx) return LEPIC_ERR_ARG;
  Moments& M = ctx->mom;
  if (M.data) LEPIC_HIP(hipFree(M.data));
  if (M.phi) LEPIC_HIP(hipFree(M.phi));
  if (M.pm) LEPIC_HIP(hipFree(M.pm));
  M = Moments{};
  return 0;
}

int lepic_moments_accumulate(lepic_ctx* ctx, int32_t size_na) {
  if (!ctx || size_na < 1) return LEPIC_ERR_ARG;
  if (!ctx->boundary_set) return ctx->fail(LEPIC_ERR_STATE, "lepic_set_boundary not called");
  if (ctx->g.n1 % size_na || ctx->g.n2 % size_na)
    return ctx->fail(LEPIC_ERR_ARG, "size_na must divide n1 and n2 (read_input.f90:109-111)");
  Moments& M = ctx->mom;
  if (M.size_na && M.size_na != size_na)
    return ctx->fail(LEPIC_ERR_STATE, "size_na changed since the last reset");
  const int na1 = ctx->g.n1 / size_na, na2 = ctx->g.n2 / size_na;
  const int64_t nc = (int64_t)(na1 + 1) * (na2 + 1);
  if (!M.size_na) {
    M.size_na = size_na;
    LEPIC_HIP(hipMalloc(&M.data, (size_t)nc * 7 * ctx->nsp * 8));
    LEPIC_HIP(hipMalloc(&M.phi, (size_t)ctx->nodes * 3 * 8));
    LEPIC_HIP(hipMalloc(&M.pm, (size_t)nc * 4 * 8));
    if (int rc = zero_async(ctx, M.data, (size_t)nc * 7 * ctx->nsp * 8)) return rc;
    if (int rc = zero_async(ctx, M.phi, (size_t)ctx->nodes * 3 * 8)) return rc;
  }
  const bool mixed = ctx->g.precision == LEPIC_MIXED;
  for (int p = 0; p < ctx->nsp; ++p) {
    Species& s = ctx->sp[p];
    if (int rc = zero_async(ctx, M.pm, (size_t)nc * 4 * 8)) return rc;
    MomArgs a{};
    const int b = s.cur;
    a.seg = s.seg[b];
    if (mixed) {
      a.cell = s.cell[b];
      a.fx = s.f[b][0]; a.fy = s.f[b][1]; a.fvx = s.f[b][3]; a.fvy = s.f[b][4]; a.fvz = s.f[b][5];
    } else {
      a.x = s.d[b][0]; a.y = s.d[b][1]; a.vx = s.d[b][3]; a.vy = s.d[b][4]; a.vz = s.d[b][5];
    }
    a.bc = ctx->bc;
    a.n1 = ctx->g.n1; a.size_na = size_na; a.na1 = na1; a.na2 = na2;
    a.h1 = ctx->g.h1; a.h2 = ctx->g.h2;
    a.ha1 = size_na * ctx->g.h1; a.ha2 = size_na * ctx->g.h2;
    a.xmax = ctx->g.n1 * ctx->g.h1; a.ymax = ctx->g.n2 * ctx->g.h2;
    a.half_dt = ctx->g.dt / 2.0;
    a.k1 = s.Nm / (a.ha1 * a.ha2 * ctx->g.zmax);
    a.pm = M.pm;
    const int64_t nb = nseg_of(s);
    if (nb > 0) {
      if (mixed) hipLaunchKernelGGL(k_part_moments<true>, dim3((unsigned)nb), dim3(kBlock), 0, ctx->stream, a);
      else hipLaunchKernelGGL(k_part_moments<false>, dim3((unsigned)nb), dim3(kBlock), 0, ctx->stream, a);
      LEPIC_HIP(hipGetLastError());
    }
    AvgArgs g{};
    g.np = s.np; g.sink = s.sink; g.src = s.src; g.pm = M.pm; g.bc = ctx->bc;
    g.n1 = ctx->g.n1; g.size_na = size_na; g.na1 = na1; g.na2 = na2;
    g.mass_over_qe = s.mass / ctx->g.qe;
    g.k = s.Nm / (ctx->g.h1 * ctx->g.h2 * ctx->g.zmax);
    g.out = M.data + (size_t)p * nc * 7;
    hipLaunchKernelGGL(k_calc_avg, dim3(nblk(nc)), dim3(256), 0, ctx->stream, g);
    LEPIC_HIP(hipGetLastError());
    // ss2D = 0 after the average (main.f90:931, plt_src = 0)
    if (int rc = zero_async(ctx, s.sink, ctx->nodes * 8)) return rc;
    if (int rc = zero_async(ctx, s.src, ctx->nodes * 8)) return rc;
  }
  hipLaunchKernelGGL(k_phi_avg, dim3(nblk(ctx->nodes)), dim3(256), 0, ctx->stream, (const double*)ctx->phi,
                     (const double*)ctx->E, ctx->nodes, M.phi);
  LEPIC_HIP(hipGetLastError());
  M.count += 1;
  return 0;
}

int lepic_moments_get(lepic_ctx* ctx, double* data_pavg, double* phi_avg, int32_t* count) {
  if (!ctx) return LEPIC_ERR_ARG;
  Moments& M = ctx->mom;
  if (count) *count = M.count;
  if (!M.size_na) return (data_pavg || phi_avg) ? ctx->fail(LEPIC_ERR_STATE, "no moments accumulated") : 0;
  const int64_t nc = (int64_t)(ctx->g.n1 / M.size_na + 1) * (ctx->g.n2 / M.size_na + 1);
  if (data_pavg) {
    LEPIC_HIP(hipMemcpyAsync(data_pavg, M.data, (size_t)nc * 7 * ctx->nsp * 8, hipMemcpyDeviceToHost, ctx->stream));
  }
  if (phi_avg) {
    LEPIC_HIP(hipMemcpyAsync(phi_avg, M.phi, (size_t)ctx->nodes * 3 * 8, hipMemcpyDeviceToHost, ctx->stream));
  }
  LEPIC_HIP(hipStreamSynchronize(ctx->stream));
  return 0;
}

int lepic_species_totals(lepic_ctx* ctx, int32_t ptype, double* out) {
  if (!ctx || !out || ptype < 1 || ptype > ctx->nsp) return LEPIC_ERR_ARG;
  int64_t n = 0;
  if (int rc = lepic_get_count(ctx, ptype, &n)) return rc;
  Species& s = ctx->sp[ptype - 1];
  const int64_t nseg = nseg_of(s);
  double v2 = 0.0;
  if (nseg > 0) {
    const int nb = (int)std::min<int64_t>(nseg, kTotBlocks);
    const int b = s.cur;
    const bool mixed = ctx->g.precision == LEPIC_MIXED;
    if (int rc = sum_partials(ctx, nb, [&](double* part) {
          hipLaunchKernelGGL(k_species_v2, dim3(nb), dim3(256), 0, ctx->stream, (const int32_t*)s.seg[b], nseg,
                             mixed ? nullptr : (const double*)s.d[b][3], mixed ? nullptr : (const double*)s.d[b][4],
                             mixed ? nullptr : (const double*)s.d[b][5], mixed ? (const float*)s.f[b][3] : nullptr,
                             mixed ? (const float*)s.f[b][4] : nullptr, mixed ? (const float*)s.f[b][5] : nullptr,
                             part);
        }, &v2))
      return rc;
  }
  out[0] = (double)n;
  out[1] = 0.5 * s.Nm * s.mass * v2;
  out[2] = s.Nm * s.charge * (double)n;
  return 0;
}

int lepic_field_energy(lepic_ctx* ctx, double* W) {
  if (!ctx || !W) return LEPIC_ERR_ARG;
  if (!ctx->boundary_set) return ctx->fail(LEPIC_ERR_STATE, "lepic_set_boundary not called");
  const int nb = (int)std::min<int64_t>((ctx->nodes + 255) / 256, kTotBlocks);
  double e2 = 0.0;
  if (int rc = sum_partials(ctx, nb, [&](double* part) {
        hipLaunchKernelGGL(k_field_e2, dim3(nb), dim3(256), 0, ctx->stream, (const double*)ctx->E,
                           (const int8_t*)ctx->bc, ctx->nodes, part);
      }, &e2))
    return rc;
  *W = 0.5 * ctx->g.eps0 * e2 * ctx->g.h1 * ctx->g.h2 * ctx->g.zmax;
  return 0;
}

}  // extern "C"

