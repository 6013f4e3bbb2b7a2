// lepic_ctx.hip -- context lifecycle, transfers, generic scans and the C ABI
// entry points of liblepic_hip.so.
#include <hip/hip_runtime.h>

#include <vector>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <new>

#include "lepic_internal.h"

using namespace lepic;

namespace {

__global__ void k_fill_u8(uint8_t* p, size_t n) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  size_t stride = (size_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) p[i] = 0;
}

__global__ void k_fill_u64(uint64_t* p, size_t n) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  size_t stride = (size_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) p[i] = 0;
}

// AoS (x,y,z,vx,vy,vz) -> FP64 SoA
__global__ void k_aos_to_soa64(const double* __restrict__ a, int64_t n, double* x, double* y,
                               double* z, double* vx, double* vy, double* vz) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double* p = a + 6 * i;
  x[i] = p[0]; y[i] = p[1]; z[i] = p[2]; vx[i] = p[3]; vy[i] = p[4]; vz[i] = p[5];
}

__global__ void k_soa64_to_aos(double* __restrict__ a, int64_t n, const double* x, const double* y,
                               const double* z, const double* vx, const double* vy,
                               const double* vz) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double* p = a + 6 * i;
  p[0] = x[i]; p[1] = y[i]; p[2] = z[i]; p[3] = vx[i]; p[4] = vy[i]; p[5] = vz[i];
}

// AoS -> mixed state: 1-based cell (INT(x/h)+1, part_expmover.f90:103) and
// the FP32 offset inside the cell.
__global__ void k_aos_to_mixed(const double* __restrict__ a, int64_t n, double h1, double h2,
                               int32_t* cell, float* fx, float* fy, float* z, float* vx, float* vy,
                               float* vz) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double* p = a + 6 * i;
  double sx = p[0] / h1, sy = p[1] / h2;
  int ix = (int)sx, iy = (int)sy;
  float ox = (float)(sx - ix), oy = (float)(sy - iy);
  if (ox >= 1.0f) { ox = 0.0f; ++ix; }
  if (oy >= 1.0f) { oy = 0.0f; ++iy; }
  ix = ix < -1 ? -1 : (ix > 30000 ? 30000 : ix);  // cells outside the box are clamped by users
  iy = iy < -1 ? -1 : (iy > 30000 ? 30000 : iy);
  cell[i] = ((iy + 1) << 16) | (ix + 1);
  fx[i] = ox; fy[i] = oy;
  if (z) z[i] = (float)p[2];
  vx[i] = (float)p[3]; vy[i] = (float)p[4]; vz[i] = (float)p[5];
}

__global__ void k_mixed_to_aos(double* __restrict__ a, int64_t n, double h1, double h2,
                               const int32_t* cell, const float* fx, const float* fy,
                               const float* z, const float* vx, const float* vy, const float* vz) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double* p = a + 6 * i;
  int c = cell[i];
  int ix = c & 0xFFFF, iy = c >> 16;
  p[0] = ((double)(ix - 1) + (double)fx[i]) * h1;
  p[1] = ((double)(iy - 1) + (double)fy[i]) * h2;
  p[2] = z ? (double)z[i] : 0.0;
  p[3] = vx[i]; p[4] = vy[i]; p[5] = vz[i];
}

// ---- segmented particle layout ----
// packed counts: seg[s] = clamp(N - s*kChunk, 0, kChunk)
__global__ void k_seg_from_count(int32_t* __restrict__ seg, int64_t nseg, const int64_t* __restrict__ count) {
  const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= nseg) return;
  const int64_t r = *count - s * kChunk;
  seg[s] = (int32_t)(r < 0 ? 0 : (r > kChunk ? kChunk : r));
}

__global__ void k_seg_copy(const int32_t* __restrict__ seg, int64_t nseg, uint32_t* __restrict__ out) {
  const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s < nseg) out[s] = (uint32_t)seg[s];
}

// move the particles of every segment to their dense positions
__global__ void k_pack64(const int32_t* __restrict__ seg, const uint32_t* __restrict__ off, int64_t nslots,
                         const double* a0, const double* a1, const double* a2, const double* a3,
                         const double* a4, const double* a5, double* b0, double* b1, double* b2, double* b3,
                         double* b4, double* b5) {
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= nslots) return;
  const int64_t sg = q / kChunk;
  const int k = (int)(q - sg * kChunk);
  if (k >= seg[sg]) return;
  const int64_t d = (int64_t)off[sg] + k;
  b0[d] = a0[q]; b1[d] = a1[q]; b2[d] = a2[q]; b3[d] = a3[q]; b4[d] = a4[q]; b5[d] = a5[q];
}

__global__ void k_pack_mixed(const int32_t* __restrict__ seg, const uint32_t* __restrict__ off, int64_t nslots,
                             const int32_t* c0, const float* a0, const float* a1, const float* a2,
                             const float* a3, const float* a4, const float* a5, int32_t* d0, float* b0,
                             float* b1, float* b2, float* b3, float* b4, float* b5) {
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= nslots) return;
  const int64_t sg = q / kChunk;
  const int k = (int)(q - sg * kChunk);
  if (k >= seg[sg]) return;
  const int64_t d = (int64_t)off[sg] + k;
  d0[d] = c0[q]; b0[d] = a0[q]; b1[d] = a1[q];
  if (b2) b2[d] = a2[q];
  b3[d] = a3[q]; b4[d] = a4[q]; b5[d] = a5[q];
}

// ---- generic exclusive scan (uint32, in place) ----
constexpr int kScanItems = 16;
constexpr int kScanTile = 256 * kScanItems;

__device__ inline uint32_t block_exclusive_u32(uint32_t v, uint32_t* total) {
  __shared__ uint32_t wsum[4];
  int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  uint32_t x = v;
  for (int o = 1; o < 64; o <<= 1) {
    uint32_t y = __shfl_up(x, o, 64);
    if (lane >= o) x += y;
  }
  if (lane == 63) wsum[w] = x;
  __syncthreads();
  uint32_t base = 0, tot = 0;
  for (int k = 0; k < 4; ++k) {
    if (k < w) base += wsum[k];
    tot += wsum[k];
  }
  __syncthreads();
  *total = tot;
  return base + x - v;
}

__global__ void k_scan_reduce(const uint32_t* __restrict__ d, int64_t n, uint32_t* part) {
  int64_t base = (int64_t)blockIdx.x * kScanTile;
  uint32_t s = 0;
  for (int k = 0; k < kScanItems; ++k) {
    int64_t i = base + (int64_t)threadIdx.x * kScanItems + k;
    if (i < n) s += d[i];
  }
  uint32_t tot;
  block_exclusive_u32(s, &tot);
  if (threadIdx.x == 0) part[blockIdx.x] = tot;
}

__global__ void k_scan_apply(uint32_t* __restrict__ d, int64_t n, const uint32_t* part) {
  int64_t base = (int64_t)blockIdx.x * kScanTile;
  uint32_t v[kScanItems];
  uint32_t s = 0;
  for (int k = 0; k < kScanItems; ++k) {
    int64_t i = base + (int64_t)threadIdx.x * kScanItems + k;
    v[k] = (i < n) ? d[i] : 0u;
    s += v[k];
  }
  uint32_t tot;
  uint32_t run = block_exclusive_u32(s, &tot) + (part ? part[blockIdx.x] : 0u);
  for (int k = 0; k < kScanItems; ++k) {
    int64_t i = base + (int64_t)threadIdx.x * kScanItems + k;
    if (i < n) d[i] = run;
    run += v[k];
  }
}

}  // namespace

namespace lepic {

int zero_async(lepic_ctx* ctx, void* p, size_t bytes) {
  if (!bytes) return 0;
  if ((bytes & 7) == 0 && ((uintptr_t)p & 7) == 0) {
    size_t n = bytes / 8;
    int grid = (int)std::min<size_t>((n + 255) / 256, 4096);
    hipLaunchKernelGGL(k_fill_u64, dim3(grid), dim3(256), 0, ctx->stream, (uint64_t*)p, n);
  } else {
    int grid = (int)std::min<size_t>((bytes + 255) / 256, 4096);
    hipLaunchKernelGGL(k_fill_u8, dim3(grid), dim3(256), 0, ctx->stream, (uint8_t*)p, bytes);
  }
  LEPIC_HIP(hipGetLastError());
  return 0;
}

int set_packed_counts(lepic_ctx* ctx, int p) {
  Species& s = ctx->sp[p];
  hipLaunchKernelGGL(k_seg_from_count, dim3((unsigned)((s.nseg_cap + 255) / 256)), dim3(256), 0, ctx->stream,
                     s.seg[s.cur], s.nseg_cap, (const int64_t*)(ctx->d_counts + p));
  LEPIC_HIP(hipGetLastError());
  s.packed = true;
  return 0;
}

int ensure_packed(lepic_ctx* ctx, int p) {
  Species& s = ctx->sp[p];
  if (s.packed) return 0;
  const int64_t nseg = std::min<int64_t>(nseg_of(s), s.nseg_cap);
  if (nseg > 0) {
    const int b = s.cur, o = 1 - s.cur;
    hipLaunchKernelGGL(k_seg_copy, dim3((unsigned)((nseg + 255) / 256)), dim3(256), 0, ctx->stream,
                       (const int32_t*)s.seg[b], nseg, ctx->d_segoff);
    LEPIC_HIP(hipGetLastError());
    if (int rc = scan_exclusive_u32(ctx, ctx->d_segoff, nseg)) return rc;
    const int64_t nslots = nseg * kChunk;
    const dim3 g((unsigned)((nslots + 255) / 256));
    if (ctx->g.precision == LEPIC_MIXED)
      hipLaunchKernelGGL(k_pack_mixed, g, dim3(256), 0, ctx->stream, (const int32_t*)s.seg[b],
                         (const uint32_t*)ctx->d_segoff, nslots, s.cell[b], s.f[b][0], s.f[b][1], s.f[b][2],
                         s.f[b][3], s.f[b][4], s.f[b][5], s.cell[o], s.f[o][0], s.f[o][1], s.f[o][2],
                         s.f[o][3], s.f[o][4], s.f[o][5]);
    else
      hipLaunchKernelGGL(k_pack64, g, dim3(256), 0, ctx->stream, (const int32_t*)s.seg[b],
                         (const uint32_t*)ctx->d_segoff, nslots, s.d[b][0], s.d[b][1], s.d[b][2], s.d[b][3],
                         s.d[b][4], s.d[b][5], s.d[o][0], s.d[o][1], s.d[o][2], s.d[o][3], s.d[o][4],
                         s.d[o][5]);
    LEPIC_HIP(hipGetLastError());
    s.cur = o;
  }
  return set_packed_counts(ctx, p);
}

int64_t scan_workspace_words(int64_t n) {
  int64_t w = 0;
  for (int64_t nb = (n + kScanTile - 1) / kScanTile; nb > 1; nb = (nb + kScanTile - 1) / kScanTile) w += nb;
  return w;
}

int scan_exclusive_u32(lepic_ctx* ctx, uint32_t* data, int64_t n, uint32_t* ws) {
  if (n <= 0) return 0;
  int64_t nb = (n + kScanTile - 1) / kScanTile;
  if (nb == 1) {
    hipLaunchKernelGGL(k_scan_apply, dim3(1), dim3(256), 0, ctx->stream, data, n,
                       (const uint32_t*)nullptr);
    LEPIC_HIP(hipGetLastError());
    return 0;
  }
  // block partials: from the caller's workspace (scan_workspace_words(n)
  // words, the levels one after another) or stream-ordered allocations
  uint32_t* part = ws;
  if (!ws) LEPIC_HIP(hipMallocAsync((void**)&part, nb * sizeof(uint32_t), ctx->stream));
  hipLaunchKernelGGL(k_scan_reduce, dim3((unsigned)nb), dim3(256), 0, ctx->stream, data, n, part);
  LEPIC_HIP(hipGetLastError());
  if (int rc = scan_exclusive_u32(ctx, part, nb, ws ? ws + nb : nullptr)) return rc;
  hipLaunchKernelGGL(k_scan_apply, dim3((unsigned)nb), dim3(256), 0, ctx->stream, data, n,
                     (const uint32_t*)part);
  LEPIC_HIP(hipGetLastError());
  if (!ws) LEPIC_HIP(hipFreeAsync(part, ctx->stream));
  return 0;
}

}  // namespace lepic

// ===========================================================================
// C ABI
// ===========================================================================
namespace {

int alloc_species(lepic_ctx* ctx, int p) {
  Species& s = ctx->sp[p];
  // arrays hold whole segments, so a kernel may load a segment's slots as
  // 16-byte vectors past the species' count
  const int64_t cap = (std::max<int64_t>(s.cap, 1) + kChunk - 1) / kChunk * kChunk;
  const bool mixed = ctx->g.precision == LEPIC_MIXED;
  for (int b = 0; b < 2; ++b) {
    size_t bytes = mixed ? (size_t)cap * (4 + 6 * 4) : (size_t)cap * 48;
    LEPIC_HIP(hipMalloc(&s.mem[b], bytes + 256));
    char* base = (char*)s.mem[b];
    if (mixed) {
      s.cell[b] = (int32_t*)base;
      for (int k = 0; k < 6; ++k) s.f[b][k] = (float*)(base + (size_t)cap * 4 + (size_t)k * cap * 4);
      if (!ctx->g.track_z) s.f[b][2] = nullptr;
    } else {
      for (int k = 0; k < 6; ++k) s.d[b][k] = (double*)(base + (size_t)k * cap * 8);
    }
  }
  LEPIC_HIP(hipMalloc(&s.flag_dead, cap));
  LEPIC_HIP(hipMemsetAsync(s.flag_dead, 0, cap, ctx->stream));
  s.nseg_cap = (cap + kChunk - 1) / kChunk + 1;
  for (int b = 0; b < 2; ++b) {
    LEPIC_HIP(hipMalloc(&s.seg[b], s.nseg_cap * 4));
    LEPIC_HIP(hipMemsetAsync(s.seg[b], 0, s.nseg_cap * 4, ctx->stream));
  }
  if (s.nseg_cap > ctx->segoff_cap) {
    if (ctx->d_segoff) LEPIC_HIP(hipFree(ctx->d_segoff));
    LEPIC_HIP(hipMalloc(&ctx->d_segoff, s.nseg_cap * 4));
    ctx->segoff_cap = s.nseg_cap;
  }
  s.packed = true;
  LEPIC_HIP(hipMalloc(&s.np, ctx->nodes * 8));
  LEPIC_HIP(hipMemsetAsync(s.np, 0, ctx->nodes * 8, ctx->stream));
  LEPIC_HIP(hipMalloc(&s.acc, ctx->nodes * 8));
  LEPIC_HIP(hipMemsetAsync(s.acc, 0, ctx->nodes * 8, ctx->stream));
  LEPIC_HIP(hipMalloc(&s.np_red, ctx->nodes * 8));
  LEPIC_HIP(hipMemsetAsync(s.np_red, 0, ctx->nodes * 8, ctx->stream));
  LEPIC_HIP(hipMalloc(&s.sink, ctx->nodes * 8));
  LEPIC_HIP(hipMemsetAsync(s.sink, 0, ctx->nodes * 8, ctx->stream));
  LEPIC_HIP(hipMalloc(&s.src, ctx->nodes * 8));
  LEPIC_HIP(hipMemsetAsync(s.src, 0, ctx->nodes * 8, ctx->stream));
  return 0;
}

bool valid_ptype(lepic_ctx* ctx, int32_t ptype) { return ptype >= 1 && ptype <= ctx->nsp; }

constexpr int64_t kStage = 1 << 20;  // particles per transfer chunk

}  // namespace

extern "C" {

int lepic_create(const lepic_grid_desc* g, const lepic_species_desc* sp, lepic_ctx** out) {
  if (!g || !sp || !out) return LEPIC_ERR_ARG;
  *out = nullptr;
  lepic_ctx* ctx = new (std::nothrow) lepic_ctx();
  if (!ctx) return LEPIC_ERR_ARG;
  ctx->g = *g;
  auto bad = [&](const char* m) {
    ctx->err = m;
    *out = ctx;  // returned so the caller can read the message, then destroy
    return LEPIC_ERR_ARG;
  };
  if (g->n1 < 2 || g->n2 < 2 || g->n1 > 32000 || g->n2 > 32000) return bad("grid size out of range");
  if (g->ng < 1 || g->ng > kMaxLevels) return bad("ng must be 1..12");
  if ((g->n1 % (1 << (g->ng - 1))) || (g->n2 % (1 << (g->ng - 1))))
    return bad("n1, n2 must be divisible by 2^(ng-1) (main.f90:264-277)");
  if (g->nspecies < 1 || g->nspecies > kMaxSpecies) return bad("nspecies must be 1..8");
  if (g->precision != LEPIC_FP64 && g->precision != LEPIC_MIXED) return bad("bad precision");
  if (g->ngrid < 1 || g->ngrid > 64) return bad("ngrid must be 1..64");
  if (!(g->h1 > 0) || !(g->h2 > 0) || !(g->dt > 0) || !(g->zmax > 0) || !(g->eps0 > 0))
    return bad("h, dt, zmax, eps0 must be positive");
  ctx->nsp = g->nspecies;
  ctx->ng = g->ng;
  ctx->nodes = (int64_t)(g->n1 + 3) * (g->n2 + 3);
  ctx->rhs_n = (int64_t)(g->n1 + 1) * (g->n2 + 1);
  if (hipGetDevice(&ctx->device) != hipSuccess) return bad("no HIP device");
  if (hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess)
    return bad("hipStreamCreate failed");
  for (int p = 0; p < ctx->nsp; ++p) {
    ctx->sp[p].charge = sp[p].charge;
    ctx->sp[p].mass = sp[p].mass;
    ctx->sp[p].Nm = sp[p].Nm;
    ctx->sp[p].cap = sp[p].capacity;
    if (sp[p].capacity < 0 || sp[p].capacity > (int64_t)1 << 37) return bad("bad capacity");
    if (sp[p].mass == 0.0) return bad("species mass must be nonzero");
    if (int rc = alloc_species(ctx, p)) { *out = ctx; return rc; }
  }
  int rc = 0;
  auto H = [&](hipError_t e, const char* what) {
    if (e != hipSuccess && !rc) rc = ctx->fail(LEPIC_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
  };
  H(hipMalloc(&ctx->bc, ctx->nodes), "bc");
  H(hipMemset(ctx->bc, 1, ctx->nodes), "bc");
  H(hipMalloc(&ctx->phi, ctx->nodes * 8), "phi");
  H(hipMemset(ctx->phi, 0, ctx->nodes * 8), "phi");
  H(hipMalloc(&ctx->phi_given, ctx->nodes * 8), "phi_given");
  H(hipMemset(ctx->phi_given, 0, ctx->nodes * 8), "phi_given");
  H(hipMalloc(&ctx->rhs, ctx->rhs_n * 8), "rhs");
  H(hipMemset(ctx->rhs, 0, ctx->rhs_n * 8), "rhs");
  H(hipMalloc(&ctx->E, ctx->nodes * 16), "E");
  H(hipMemset(ctx->E, 0, ctx->nodes * 16), "E");
  if (g->precision == LEPIC_MIXED) {
    H(hipMalloc(&ctx->Ef, ctx->nodes * 8), "Ef");
    H(hipMemset(ctx->Ef, 0, ctx->nodes * 8), "Ef");
  }
  H(hipMalloc(&ctx->Bi, 4 * 4 * 8), "Bi");
  H(hipMemset(ctx->Bi, 0, 4 * 4 * 8), "Bi");
  H(hipMalloc(&ctx->d_counts, kMaxSpecies * 8), "counts");
  H(hipMemset(ctx->d_counts, 0, kMaxSpecies * 8), "counts");
  H(hipMalloc(&ctx->d_tally, kMaxSpecies * kTallyCopies * sizeof(Tally)), "tally");
  H(hipMemset(ctx->d_tally, 0, kMaxSpecies * kTallyCopies * sizeof(Tally)), "tally");
  H(hipMalloc(&ctx->d_ticket, 8), "ticket");
  H(hipMemset(ctx->d_ticket, 0, 8), "ticket");
  H(hipMalloc(&ctx->d_error, 64), "error");
  H(hipMemset(ctx->d_error, 0, 64), "error");
  H(hipMalloc(&ctx->d_mcc_i64, 64 * 8), "mcc");
  H(hipMemset(ctx->d_mcc_i64, 0, 64 * 8), "mcc");
  H(hipMalloc(&ctx->d_mcc_f64, (kMccF64Head + kMaxRegions * kRegionBlocks) * 8), "mcc");
  H(hipMemset(ctx->d_mcc_f64, 0, (kMccF64Head + kMaxRegions * kRegionBlocks) * 8), "mcc");
  H(hipHostMalloc((void**)&ctx->h_pinned, kPinnedDoubles * 8, hipHostMallocDefault), "pinned");
  ctx->bc_host.assign(ctx->nodes, 1);
  if (!rc) rc = poisson_setup(ctx);
  if (!rc) rc = sort_scratch_resolve(ctx);
  *out = ctx;
  if (!rc) rc = (hipDeviceSynchronize() == hipSuccess) ? 0 : ctx->fail(LEPIC_ERR_HIP, "sync");
  return rc;
}

int lepic_destroy(lepic_ctx* ctx) {
  if (!ctx) return 0;
  if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
  poisson_teardown(ctx);
  net_destroy(ctx);
  lepic_moments_reset(ctx);
  for (int p = 0; p < kMaxSpecies; ++p)
    for (int k = 0; k < 2; ++k)
      if (ctx->ev_mover[p][k]) (void)hipEventDestroy(ctx->ev_mover[p][k]);
  if (ctx->ev_sync) (void)hipEventDestroy(ctx->ev_sync);
  for (int p = 0; p < ctx->nsp; ++p) {
    Species& s = ctx->sp[p];
    for (int b = 0; b < 2; ++b) (void)hipFree(s.mem[b]);
    (void)hipFree(s.flag_dead); (void)hipFree(s.np); (void)hipFree(s.acc); (void)hipFree(s.np_red); (void)hipFree(s.sink); (void)hipFree(s.src);
    (void)hipFree(s.seg[0]); (void)hipFree(s.seg[1]);
  }
  void* ptrs[] = {ctx->bc, ctx->phi, ctx->phi_given, ctx->rhs, ctx->E, ctx->Ef, ctx->Bi, ctx->Bf,
                  ctx->d_ctrl, ctx->d_scalar, ctx->d_resid64, ctx->d_corr,
                  ctx->d_counts, ctx->d_tally, ctx->d_segoff, ctx->d_ticket, ctx->d_error,
                  ctx->s_key[0], ctx->s_key[1], ctx->s_val[0], ctx->s_val[1], ctx->s_hist,
                  ctx->d_plist, ctx->d_scan_tmp, ctx->s_meta, ctx->s_rkey, ctx->s_rcnt, ctx->s_rcntd,
                  ctx->s_nrun, ctx->rx.d_sig_Er, ctx->rx.d_sig,
                  ctx->rx.d_sig_Er_f, ctx->rx.d_sig_f, ctx->rx.d_cnt_col, ctx->d_claim,
                  ctx->d_mcc_stage, ctx->d_mcc_i64, ctx->d_mcc_f64, ctx->d_mcc_tab, ctx->d_heat,
                  ctx->d_stage6, ctx->d_stype, ctx->d_srank, ctx->d_srctally, ctx->d_cellwall,
                  ctx->d_diag};
  for (void* p : ptrs) if (p) (void)hipFree(p);
  for (int l = 2; l <= kMaxLevels; ++l) {
    if (ctx->lv[l].u) (void)hipFree(ctx->lv[l].u);
    if (ctx->lv[l].b) (void)hipFree(ctx->lv[l].b);
    if (ctx->lv[l].bc) (void)hipFree(ctx->lv[l].bc);
  }
  for (int l = 1; l <= kMaxLevels; ++l) {
    if (ctx->lv[l].uf) (void)hipFree(ctx->lv[l].uf);
    if (ctx->lv[l].bf) (void)hipFree(ctx->lv[l].bf);
  }
  if (ctx->h_pinned) (void)hipHostFree(ctx->h_pinned);
  if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
  delete ctx;
  return 0;
}

const char* lepic_last_error(const lepic_ctx* ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int lepic_sync(lepic_ctx* ctx) {
  if (!ctx) return LEPIC_ERR_ARG;
  LEPIC_HIP(hipStreamSynchronize(ctx->stream));
  int32_t e = 0;
  LEPIC_HIP(hipMemcpyAsync(&e, ctx->d_error, 4, hipMemcpyDeviceToHost, ctx->stream));
  LEPIC_HIP(hipStreamSynchronize(ctx->stream));
  if (e) return ctx->fail(LEPIC_ERR_HIP, "device-side wait bound exceeded (look-back)");
  return 0;
}

int lepic_stream(lepic_ctx* ctx, void** stream) {
  if (!ctx || !stream) return LEPIC_ERR_ARG;
  *stream = (void*)ctx->stream;
  return 0;
}

int lepic_kernel_timing(lepic_ctx* ctx, int32_t enable) {
  if (!ctx) return LEPIC_ERR_ARG;
  if (enable && !ctx->ev_mover[0][0])
    for (int p = 0; p < kMaxSpecies; ++p)
      for (int k = 0; k < 2; ++k) LEPIC_HIP(hipEventCreate(&ctx->ev_mover[p][k]));
  ctx->timing = enable != 0;
  return 0;
}

int lepic_get_kernel_ms(lepic_ctx* ctx, int32_t ptype, float* ms) {
  if (!ctx || !ms || ptype < 1 || ptype > ctx->nsp) return LEPIC_ERR_ARG;
  if (!ctx->ev_recorded[ptype - 1]) return ctx->fail(LEPIC_ERR_STATE, "no timed mover launch for this species");
  LEPIC_HIP(hipEventSynchronize(ctx->ev_mover[ptype - 1][1]));
  LEPIC_HIP(hipEventElapsedTime(ms, ctx->ev_mover[ptype - 1][0], ctx->ev_mover[ptype - 1][1]));
  return 0;
}

int lepic_set_boundary(lepic_ctx* ctx, const int32_t* bcnd, const double* phi) {
  if (!ctx || !bcnd || !phi) return LEPIC_ERR_ARG;
  const int n1 = ctx->g.n1, n2 = ctx->g.n2;
  auto at = [&](int ix, int iy) { return bcnd[(int64_t)iy * (n1 + 3) + ix]; };
  for (int64_t i = 0; i < ctx->nodes; ++i) {
    int v = bcnd[i];
    if (v == 0 || v == -2)
      return ctx->fail(LEPIC_ERR_UNSUPPORTED, "periodic (0) and Neumann (-2) nodes are not supported by this library");
    if (v != -1 && (v < 1 || v > ctx->g.ngrid))
      return ctx->fail(LEPIC_ERR_ARG, "bcnd values must be -1 or a wall label 1..ngrid");
  }
  for (int iy = 0; iy <= n2 + 2; ++iy)
    for (int ix = 0; ix <= n1 + 2; ++ix) {
      bool edge = ix <= 1 || ix >= n1 + 1 || iy <= 1 || iy >= n2 + 1;
      if (edge && at(ix, iy) < 1)
        return ctx->fail(LEPIC_ERR_UNSUPPORTED, "the outer boundary (rows/columns 0,1,n+1,n+2) must be walls");
    }
  // Efield.f90:87-118 extrapolates a wall node from the nodes one and two
  // away; when the second one is itself a wall its value depends on the
  // reference's loop order, which a parallel kernel reproduces only by a
  // serial pass.  Detect that case.
  int serial = 0;
  for (int iy = 1; iy <= n2 + 1 && !serial; ++iy)
    for (int ix = 1; ix <= n1 + 1; ++ix) {
      if (at(ix, iy) < 1) continue;
      if ((at(ix + 1, iy) <= 0 && at(ix + 2, iy) >= 1) || (at(ix - 1, iy) <= 0 && at(ix - 2, iy) >= 1) ||
          (at(ix, iy + 1) <= 0 && at(ix, iy + 2) >= 1) || (at(ix, iy - 1) <= 0 && at(ix, iy - 2) >= 1)) {
        serial = 1;
        break;
      }
    }
  ctx->efield_serial = serial;
  ctx->bc_host.assign(bcnd, bcnd + ctx->nodes);
  std::vector<int8_t> b8(ctx->nodes);
  for (int64_t i = 0; i < ctx->nodes; ++i) b8[i] = (int8_t)bcnd[i];
  LEPIC_HIP(hipMemcpyAsync(ctx->bc, b8.data(), ctx->nodes, hipMemcpyHostToDevice, ctx->stream));
  LEPIC_HIP(hipStreamSynchronize(ctx->stream));
  // the particle-loss test of part_expmover.f90:254-258 as one byte per cell
  std::vector<int8_t> cw(ctx->nodes, 0);
  for (int iy = 0; iy <= n2 + 1; ++iy)
    for (int ix = 0; ix <= n1 + 1; ++ix)
      cw[(int64_t)iy * (n1 + 3) + ix] =
          (at(ix, iy) >= 1 && at(ix + 1, iy) >= 1 && at(ix + 1, iy + 1) >= 1 && at(ix, iy + 1) >= 1) ? 1 : 0;
  if (!ctx->d_cellwall) LEPIC_HIP(hipMalloc(&ctx->d_cellwall, ctx->nodes));
  LEPIC_HIP(hipMemcpyAsync(ctx->d_cellwall, cw.data(), ctx->nodes, hipMemcpyHostToDevice, ctx->stream));
  LEPIC_HIP(hipStreamSynchronize(ctx->stream));
  LEPIC_HIP(hipMemcpyAsync(ctx->phi, phi, ctx->nodes * 8, hipMemcpyHostToDevice, ctx->stream));
  LEPIC_HIP(hipStreamSynchronize(ctx->stream));
  // coarse-level masks by injection (mg.f90:343-348); corners copied from
  // the finer corners (the reference never writes them and never reads them)
  lepic::Level* lv = ctx->lv;
  std::vector<int32_t> fine(bcnd, bcnd + ctx->nodes);
  for (int l = 2; l <= ctx->ng; ++l) {
    int f1 = lv[l - 1].n1, f2 = lv[l - 1].n2, m1 = lv[l].n1, m2 = lv[l].n2;
    std::vector<int32_t> c((size_t)(m1 + 3) * (m2 + 3), 1);
    auto F = [&](int i, int j) { return fine[(size_t)j * (f1 + 3) + i]; };
    auto Cw = [&](int i, int j) -> int32_t& { return c[(size_t)j * (m1 + 3) + i]; };
    Cw(0, 0) = F(0, 0); Cw(m1 + 2, 0) = F(f1 + 2, 0);
    Cw(0, m2 + 2) = F(0, f2 + 2); Cw(m1 + 2, m2 + 2) = F(f1 + 2, f2 + 2);
    for (int j = 1; j <= m2 + 1; ++j)
      for (int i = 1; i <= m1 + 1; ++i) {
        int ir = 2 * i - 1, jr = 2 * j - 1;
        Cw(i, j) = F(ir, jr);
        if (ir == 1) Cw(0, j) = F(0, jr);
        if (ir == f1 + 1) Cw(m1 + 2, j) = F(f1 + 2, jr);
        if (jr == 1) Cw(i, 0) = F(ir, 0);
        if (jr == f2 + 1) Cw(i, m2 + 2) = F(ir, f2 + 2);
      }
    std::vector<int8_t> c8(c.size());
    for (size_t i = 0; i < c.size(); ++i) c8[i] = (int8_t)c[i];
    LEPIC_HIP(hipMemcpyAsync(lv[l].bc, c8.data(), c8.size(), hipMemcpyHostToDevice, ctx->stream));
  LEPIC_HIP(hipStreamSynchronize(ctx->stream));
    fine.swap(c);
  }
  ctx->boundary_set = true;
  ctx->tail_bc_key = -1;   // coarse bc changed: rebuild the tail's LDS image
  return 0;
}

int lepic_set_bfield(lepic_ctx* ctx, int32_t nB1, int32_t nB2, double hB1, double hB2,
                     const double* Bi) {
  if (!ctx || !Bi || nB1 < 1 || nB2 < 1) return LEPIC_ERR_ARG;
  if (nB1 > 1 && (!(hB1 > 0) || !(hB2 > 0))) return ctx->fail(LEPIC_ERR_ARG, "hB must be positive");
  const int64_t bn = (int64_t)(nB1 + 3) * (nB2 + 3);
  LEPIC_HIP(hipStreamSynchronize(ctx->stream));
  if (ctx->Bi) LEPIC_HIP(hipFree(ctx->Bi));
  if (ctx->Bf) LEPIC_HIP(hipFree(ctx->Bf));
  ctx->Bf = nullptr;
  LEPIC_HIP(hipMalloc(&ctx->Bi, bn * 32));
  LEPIC_HIP(hipMemcpyAsync(ctx->Bi, Bi, bn * 32, hipMemcpyHostToDevice, ctx->stream));
  LEPIC_HIP(hipStreamSynchronize(ctx->stream));
  if (ctx->g.precision == LEPIC_MIXED) {
    std::vector<float> f(bn * 4);
    for (int64_t i = 0; i < bn * 4; ++i) f[i] = (float)Bi[i];
    LEPIC_HIP(hipMalloc(&ctx->Bf, bn * 16));
    LEPIC_HIP(hipMemcpyAsync(ctx->Bf, f.data(), bn * 16, hipMemcpyHostToDevice, ctx->stream));
  LEPIC_HIP(hipStreamSynchronize(ctx->stream));
  }
  ctx->nB1 = nB1; ctx->nB2 = nB2; ctx->hB1 = hB1; ctx->hB2 = hB2;
  ctx->flag_B = 1;
  ctx->flag_gridB = (nB1 > 1 && (nB1 != ctx->g.n1 || nB2 != ctx->g.n2)) ? 1 : 0;
  const double* c = Bi + 4 * (1 * (nB1 + 3) + 1);
  ctx->Bc[0] = c[0]; ctx->Bc[1] = c[1]; ctx->Bc[2] = c[2];
  if (nB1 == 1 && c[0] == 0 && c[1] == 0 && c[2] == 0 && Bi[4 * (nB1 + 3 + 1) + 3] == 0) ctx->flag_B = 1;
  return 0;
}

int lepic_upload_particles(lepic_ctx* ctx, int32_t ptype, const double* vxp, int64_t np) {
  if (!ctx || !valid_ptype(ctx, ptype) || (np > 0 && !vxp) || np < 0) return LEPIC_ERR_ARG;
  Species& s = ctx->sp[ptype - 1];
  if (np > s.cap) return ctx->fail(LEPIC_ERR_CAPACITY, "np exceeds capacity");
  double* stage = nullptr;
  int64_t chunk = std::min<int64_t>(std::max<int64_t>(np, 1), kStage);
  LEPIC_HIP(hipMalloc(&stage, chunk * 48));
  const int b = s.cur;
  for (int64_t off = 0; off < np; off += chunk) {
    int64_t n = std::min(chunk, np - off);
    LEPIC_HIP(hipMemcpyAsync(stage, vxp + 6 * off, n * 48, hipMemcpyHostToDevice, ctx->stream));
    dim3 grid((unsigned)((n + 255) / 256));
    if (ctx->g.precision == LEPIC_MIXED) {
      hipLaunchKernelGGL(k_aos_to_mixed, grid, dim3(256), 0, ctx->stream, stage, n, ctx->g.h1, ctx->g.h2,
                         s.cell[b] + off, s.f[b][0] + off, s.f[b][1] + off,
                         s.f[b][2] ? s.f[b][2] + off : nullptr, s.f[b][3] + off, s.f[b][4] + off,
                         s.f[b][5] + off);
    } else {
      hipLaunchKernelGGL(k_aos_to_soa64, grid, dim3(256), 0, ctx->stream, stage, n, s.d[b][0] + off,
                         s.d[b][1] + off, s.d[b][2] + off, s.d[b][3] + off, s.d[b][4] + off,
                         s.d[b][5] + off);
    }
    LEPIC_HIP(hipGetLastError());
  }
  LEPIC_HIP(hipStreamSynchronize(ctx->stream));
  LEPIC_HIP(hipFree(stage));
  LEPIC_HIP(hipMemcpyAsync(ctx->d_counts + (ptype - 1), &np, 8, hipMemcpyHostToDevice, ctx->stream));
  LEPIC_HIP(hipStreamSynchronize(ctx->stream));
  s.ub = np;
  return set_packed_counts(ctx, ptype - 1);
}

int lepic_get_count(lepic_ctx* ctx, int32_t ptype, int64_t* np) {
  if (!ctx || !valid_ptype(ctx, ptype) || !np) return LEPIC_ERR_ARG;
  LEPIC_HIP(hipMemcpyAsync(ctx->h_pinned, ctx->d_counts + (ptype - 1), 8, hipMemcpyDeviceToHost, ctx->stream));
  LEPIC_HIP(hipStreamSynchronize(ctx->stream));
  std::memcpy(np, ctx->h_pinned, 8);
  // the segment extent only shrinks when the species is packed
  if (ctx->sp[ptype - 1].packed) ctx->sp[ptype - 1].ub = *np;
  return 0;
}

namespace {

// particles [first, first+n) of a packed species into host vxp(6,n)
int download_span(lepic_ctx* ctx, int32_t ptype, int64_t first, int64_t n, double* vxp) {
  Species& s = ctx->sp[ptype - 1];
  int64_t chunk = std::min<int64_t>(std::max<int64_t>(n, 1), kStage);
  double* stage = nullptr;
  LEPIC_HIP(hipMalloc(&stage, chunk * 48));
  const int b = s.cur;
  for (int64_t off = 0; off < n; off += chunk) {
    const int64_t m = std::min(chunk, n - off), g = first + off;
    dim3 grid((unsigned)((m + 255) / 256));
    if (ctx->g.precision == LEPIC_MIXED) {
      hipLaunchKernelGGL(k_mixed_to_aos, grid, dim3(256), 0, ctx->stream, stage, m, ctx->g.h1, ctx->g.h2,
                         s.cell[b] + g, s.f[b][0] + g, s.f[b][1] + g,
                         s.f[b][2] ? s.f[b][2] + g : nullptr, s.f[b][3] + g, s.f[b][4] + g,
                         s.f[b][5] + g);
    } else {
      hipLaunchKernelGGL(k_soa64_to_aos, grid, dim3(256), 0, ctx->stream, stage, m, s.d[b][0] + g,
                         s.d[b][1] + g, s.d[b][2] + g, s.d[b][3] + g, s.d[b][4] + g,
                         s.d[b][5] + g);
    }
    LEPIC_HIP(hipGetLastError());
    LEPIC_HIP(hipMemcpyAsync(vxp + 6 * off, stage, m * 48, hipMemcpyDeviceToHost, ctx->stream));
  }
  LEPIC_HIP(hipStreamSynchronize(ctx->stream));
  LEPIC_HIP(hipFree(stage));
  return 0;
}

}  // namespace

int lepic_download_particles(lepic_ctx* ctx, int32_t ptype, double* vxp, int64_t cap, int64_t* np) {
  if (!ctx || !valid_ptype(ctx, ptype) || !np) return LEPIC_ERR_ARG;
  if (int rc = ensure_packed(ctx, ptype - 1)) return rc;
  int64_t n = 0;
  if (int rc = lepic_get_count(ctx, ptype, &n)) return rc;
  *np = n;
  if (n > cap) return ctx->fail(LEPIC_ERR_CAPACITY, "host buffer too small");
  if (n > 0 && !vxp) return LEPIC_ERR_ARG;
  return download_span(ctx, ptype, 0, n, vxp);
}

int lepic_download_particles_range(lepic_ctx* ctx, int32_t ptype, int64_t first, int64_t count, double* vxp) {
  if (!ctx || !valid_ptype(ctx, ptype) || first < 0 || count < 0 || (count > 0 && !vxp)) return LEPIC_ERR_ARG;
  if (int rc = ensure_packed(ctx, ptype - 1)) return rc;
  int64_t n = 0;
  if (int rc = lepic_get_count(ctx, ptype, &n)) return rc;
  if (first + count > n) return ctx->fail(LEPIC_ERR_ARG, "range beyond the species' particles");
  return download_span(ctx, ptype, first, count, vxp);
}

int lepic_set_flag_dead(lepic_ctx* ctx, int32_t ptype, const int8_t* flags, int64_t np) {
  if (!ctx || !valid_ptype(ctx, ptype) || np < 0 || (np && !flags)) return LEPIC_ERR_ARG;
  if (np > ctx->sp[ptype - 1].cap) return ctx->fail(LEPIC_ERR_CAPACITY, "np exceeds capacity");
  if (int rc = ensure_packed(ctx, ptype - 1)) return rc;
  LEPIC_HIP(hipStreamSynchronize(ctx->stream));
  LEPIC_HIP(hipMemcpyAsync(ctx->sp[ptype - 1].flag_dead, flags, np, hipMemcpyHostToDevice, ctx->stream));
  LEPIC_HIP(hipStreamSynchronize(ctx->stream));
  ctx->sp[ptype - 1].dead_maybe = true;
  return 0;
}

int lepic_get_flag_dead(lepic_ctx* ctx, int32_t ptype, int8_t* flags, int64_t np) {
  if (!ctx || !valid_ptype(ctx, ptype) || np < 0 || (np && !flags)) return LEPIC_ERR_ARG;
  if (np > ctx->sp[ptype - 1].cap) return ctx->fail(LEPIC_ERR_CAPACITY, "np exceeds capacity");
  if (int rc = ensure_packed(ctx, ptype - 1)) return rc;
  LEPIC_HIP(hipStreamSynchronize(ctx->stream));
  LEPIC_HIP(hipMemcpyAsync(flags, ctx->sp[ptype - 1].flag_dead, np, hipMemcpyDeviceToHost, ctx->stream));
  LEPIC_HIP(hipStreamSynchronize(ctx->stream));
  return 0;
}

static int grid_ptr(lepic_ctx* ctx, int32_t which, int32_t ptype, double** p, int64_t* n) {
  switch (which) {
    case LEPIC_GRID_PHI: *p = ctx->phi; *n = ctx->nodes; return 0;
    case LEPIC_GRID_RHS: *p = ctx->rhs; *n = ctx->rhs_n; return 0;
    case LEPIC_GRID_E: *p = ctx->E; *n = 2 * ctx->nodes; return 0;
    case LEPIC_GRID_NP:
    case LEPIC_GRID_NPRED:
    case LEPIC_GRID_SINK:
    case LEPIC_GRID_SRC:
      if (!valid_ptype(ctx, ptype)) return LEPIC_ERR_ARG;
      *p = which == LEPIC_GRID_NP      ? ctx->sp[ptype - 1].np
           : which == LEPIC_GRID_NPRED ? ctx->sp[ptype - 1].np_red
           : which == LEPIC_GRID_SINK  ? ctx->sp[ptype - 1].sink
                                       : ctx->sp[ptype - 1].src;
      *n = ctx->nodes;
      return 0;
  }
  return LEPIC_ERR_ARG;
}

int lepic_get_grid(lepic_ctx* ctx, int32_t which, int32_t ptype, double* out) {
  if (!ctx || !out) return LEPIC_ERR_ARG;
  double* p; int64_t n;
  if (grid_ptr(ctx, which, ptype, &p, &n)) return ctx->fail(LEPIC_ERR_ARG, "bad grid id/ptype");
  LEPIC_HIP(hipStreamSynchronize(ctx->stream));
  LEPIC_HIP(hipMemcpyAsync(out, p, n * 8, hipMemcpyDeviceToHost, ctx->stream));
  LEPIC_HIP(hipStreamSynchronize(ctx->stream));
  return 0;
}

int lepic_set_grid(lepic_ctx* ctx, int32_t which, int32_t ptype, const double* in) {
  if (!ctx || !in) return LEPIC_ERR_ARG;
  double* p; int64_t n;
  if (grid_ptr(ctx, which, ptype, &p, &n)) return ctx->fail(LEPIC_ERR_ARG, "bad grid id/ptype");
  LEPIC_HIP(hipStreamSynchronize(ctx->stream));
  LEPIC_HIP(hipMemcpyAsync(p, in, n * 8, hipMemcpyHostToDevice, ctx->stream));
  LEPIC_HIP(hipStreamSynchronize(ctx->stream));
  if (which == LEPIC_GRID_PHI) LEPIC_HIP(hipMemcpyAsync(ctx->phi_given, in, n * 8, hipMemcpyHostToDevice, ctx->stream));
  LEPIC_HIP(hipStreamSynchronize(ctx->stream));
  if (which == LEPIC_GRID_E && ctx->Ef) {
    std::vector<float> f(n);
    for (int64_t i = 0; i < n; ++i) f[i] = (float)in[i];
    LEPIC_HIP(hipMemcpyAsync(ctx->Ef, f.data(), n * 4, hipMemcpyHostToDevice, ctx->stream));
  LEPIC_HIP(hipStreamSynchronize(ctx->stream));
  }
  return 0;
}

int lepic_part_mover(lepic_ctx* ctx, int32_t ptype, int32_t deposit) {
  if (!ctx || !valid_ptype(ctx, ptype)) return LEPIC_ERR_ARG;
  if (!ctx->boundary_set) return ctx->fail(LEPIC_ERR_STATE, "lepic_set_boundary not called");
  return mover_launch(ctx, ptype - 1, deposit);
}

int lepic_charge_deposition(lepic_ctx* ctx, int32_t ptype) {
  if (!ctx || !valid_ptype(ctx, ptype)) return LEPIC_ERR_ARG;
  if (!ctx->boundary_set) return ctx->fail(LEPIC_ERR_STATE, "lepic_set_boundary not called");
  return deposit_launch(ctx, ptype - 1);
}

int lepic_get_mover_tally(lepic_ctx* ctx, int32_t ptype, double* P_loss, double* p_mac, double* cnt_dead) {
  if (!ctx || !valid_ptype(ctx, ptype)) return LEPIC_ERR_ARG;
  std::vector<Tally> c(kTallyCopies);
  LEPIC_HIP(hipStreamSynchronize(ctx->stream));
  LEPIC_HIP(hipMemcpyAsync(c.data(), ctx->d_tally + (size_t)(ptype - 1) * kTallyCopies, kTallyCopies * sizeof(Tally),
                           hipMemcpyDeviceToHost, ctx->stream));
  LEPIC_HIP(hipStreamSynchronize(ctx->stream));
  Tally t{};                         // the copies summed in copy order
  for (const Tally& u : c) {
    for (int k = 0; k < 4; ++k) t.P_loss[k] += u.P_loss[k];
    t.cnt_dead += u.cnt_dead;
    for (int k = 0; k < 2; ++k)
      for (int gI = 0; gI < 65; ++gI) t.p_mac[k][gI] += u.p_mac[k][gI];
  }
  if (P_loss) std::memcpy(P_loss, t.P_loss, 32);
  if (p_mac)
    for (int k = 0; k < 2; ++k)
      for (int gI = 0; gI <= ctx->g.ngrid; ++gI) p_mac[k * (ctx->g.ngrid + 1) + gI] = t.p_mac[k][gI];
  if (cnt_dead) *cnt_dead = t.cnt_dead;
  return 0;
}

int lepic_reset_tallies(lepic_ctx* ctx) {
  if (!ctx) return LEPIC_ERR_ARG;
  if (int rc = zero_async(ctx, ctx->d_tally, kMaxSpecies * kTallyCopies * sizeof(Tally))) return rc;
  if (ctx->rx.d_cnt_col)
    if (int rc = zero_async(ctx, ctx->rx.d_cnt_col, 3 * (size_t)ctx->rx.ncol * std::max(ctx->rx.n_rg, 1) * 8)) return rc;
  if (int rc = zero_async(ctx, ctx->d_mcc_f64, 64 * 8)) return rc;
  for (int p = 0; p < ctx->nsp; ++p) {
    if (int rc = zero_async(ctx, ctx->sp[p].sink, ctx->nodes * 8)) return rc;
    if (int rc = zero_async(ctx, ctx->sp[p].src, ctx->nodes * 8)) return rc;
  }
  return 0;
}

int lepic_calc_rhs(lepic_ctx* ctx) {
  if (!ctx) return LEPIC_ERR_ARG;
  return rhs_launch(ctx);
}

int lepic_calc_efield(lepic_ctx* ctx) {
  if (!ctx) return LEPIC_ERR_ARG;
  if (!ctx->boundary_set) return ctx->fail(LEPIC_ERR_STATE, "lepic_set_boundary not called");
  return efield_launch(ctx);
}

int lepic_iterative_solver(lepic_ctx* ctx, double eps, int32_t ncycl, int32_t start,
                           lepic_solve_info* info, int32_t* sweep_log, int32_t log_cap) {
  if (!ctx || ncycl < 1 || !(eps > 0)) return LEPIC_ERR_ARG;
  if (!ctx->boundary_set) return ctx->fail(LEPIC_ERR_STATE, "lepic_set_boundary not called");
  lepic_solve_info tmp{};
  return poisson_solve(ctx, eps, ncycl, start, info ? info : &tmp, sweep_log, log_cap);
}

int lepic_part_sorting(lepic_ctx* ctx, int32_t ptype, int32_t* Plist) {
  if (!ctx || !valid_ptype(ctx, ptype)) return LEPIC_ERR_ARG;
  return sort_launch(ctx, ptype - 1, Plist);
}

int lepic_part_sorting_noalloc(lepic_ctx* ctx, int32_t ptype, int32_t* Plist) {
  if (!ctx || !valid_ptype(ctx, ptype)) return LEPIC_ERR_ARG;
  return sort_noalloc_launch(ctx, ptype - 1, Plist);
}

int lepic_dens_red(lepic_ctx* ctx, double* np_mx) {
  if (!ctx) return LEPIC_ERR_ARG;
  return dens_red_launch(ctx, np_mx);
}

int lepic_load_uniform(lepic_ctx* ctx, int32_t ptype, int32_t ppc, double vt, double xl, double xr,
                       uint64_t seed) {
  if (!ctx || !valid_ptype(ctx, ptype) || ppc < 1 || !(xr > xl)) return LEPIC_ERR_ARG;
  if (!ctx->boundary_set) return ctx->fail(LEPIC_ERR_STATE, "lepic_set_boundary not called");
  return load_uniform_launch(ctx, ptype - 1, ppc, vt, xl, xr, seed);
}

}  // extern "C"
