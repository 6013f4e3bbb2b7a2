// sort.hip -- cell sort of one species (sorting.f90:92-251): a stable LSD
// radix sort of the y-major cell index with 8-bit digits, then a gather of
// the particle payload into the other buffer, and the cell-end list Plist.
//
// With one thread the reference's counting sort is stable (particles keep
// their order inside a cell); LSD radix sort with stable per-digit scatter
// gives exactly that permutation.  Per pass: a digit histogram per 4096-key
// tile (wave-aggregated LDS counts), a device-wide exclusive scan of the
// digit-major histogram, and a scatter that ranks keys inside the tile in
// index order.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "lepic_internal.h"

using namespace lepic;

namespace {

constexpr int kSortItems = 16;
constexpr int kTile = 256 * kSortItems;

__device__ __forceinline__ uint64_t lanemask_lt() { return (1ull << (threadIdx.x & 63)) - 1ull; }

// Per-wave match of equal digits: returns my rank among earlier lanes with the
// same digit; the first lane of each digit group adds the group size to cnt.
__device__ __forceinline__ uint32_t wave_match(bool valid, uint32_t d, uint32_t* cnt) {
  const int lane = threadIdx.x & 63;
  uint64_t remaining = __ballot(valid);
  uint32_t rank = 0;
  while (remaining) {
    const int leader = __builtin_ctzll(remaining);
    const uint32_t dl = __shfl(d, leader, 64);
    const uint64_t mask = __ballot(valid && d == dl) & remaining;
    if (valid && d == dl) rank = (uint32_t)__popcll(mask & lanemask_lt());
    if (lane == leader) cnt[dl] += (uint32_t)__popcll(mask);
    remaining &= ~mask;
  }
  return rank;
}

// keys (0-based cell index) and per-cell counts
__global__ void k_keys64(const double* __restrict__ x, const double* __restrict__ y, const int64_t* nptr,
                         double h1, double h2, int n1, int n2, uint32_t* __restrict__ key,
                         uint32_t* __restrict__ cnt) {
  const int64_t N = *nptr;
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  bool valid = i < N;
  uint32_t k = 0;
  if (valid) {
    int ix = (int)(x[i] / h1), iy = (int)(y[i] / h2);  // INT(x/h) (sorting.f90:147-151)
    ix = ix < 0 ? 0 : (ix > n1 - 1 ? n1 - 1 : ix);
    iy = iy < 0 ? 0 : (iy > n2 - 1 ? n2 - 1 : iy);
    k = (uint32_t)ix + (uint32_t)n1 * (uint32_t)iy;
    key[i] = k;
  }
  // aggregate equal keys of the wave, one atomic per distinct key
  const int lane = threadIdx.x & 63;
  uint64_t remaining = __ballot(valid);
  while (remaining) {
    const int leader = __builtin_ctzll(remaining);
    const uint32_t kl = __shfl(k, leader, 64);
    const uint64_t mask = __ballot(valid && k == kl) & remaining;
    if (lane == leader) atomicAdd(&cnt[kl], (uint32_t)__popcll(mask));
    remaining &= ~mask;
  }
}

__global__ void k_keys_mixed(const int32_t* __restrict__ cell, const int64_t* nptr, int n1, int n2,
                             uint32_t* __restrict__ key, uint32_t* __restrict__ cnt) {
  const int64_t N = *nptr;
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  bool valid = i < N;
  uint32_t k = 0;
  if (valid) {
    int c = cell[i];
    int ix = (c & 0xFFFF) - 1, iy = (c >> 16) - 1;
    ix = ix < 0 ? 0 : (ix > n1 - 1 ? n1 - 1 : ix);
    iy = iy < 0 ? 0 : (iy > n2 - 1 ? n2 - 1 : iy);
    k = (uint32_t)ix + (uint32_t)n1 * (uint32_t)iy;
    key[i] = k;
  }
  const int lane = threadIdx.x & 63;
  uint64_t remaining = __ballot(valid);
  while (remaining) {
    const int leader = __builtin_ctzll(remaining);
    const uint32_t kl = __shfl(k, leader, 64);
    const uint64_t mask = __ballot(valid && k == kl) & remaining;
    if (lane == leader) atomicAdd(&cnt[kl], (uint32_t)__popcll(mask));
    remaining &= ~mask;
  }
}

__global__ __launch_bounds__(256) void k_hist(const uint32_t* __restrict__ key, const int64_t* nptr, int shift,
                                              uint32_t* __restrict__ hist, int nb) {
  __shared__ uint32_t cnt[256];
  const int64_t N = *nptr;
  cnt[threadIdx.x] = 0;
  __syncthreads();
  const int64_t base = (int64_t)blockIdx.x * kTile;
  for (int j = 0; j < kSortItems; ++j) {
    const int64_t i = base + (int64_t)j * 256 + threadIdx.x;
    const bool valid = i < N;
    const uint32_t d = valid ? (key[i] >> shift) & 255u : 0u;
    // waves add to the shared table one at a time per round (no LDS atomics)
    for (int w = 0; w < 4; ++w) {
      if ((threadIdx.x >> 6) == w) wave_match(valid, d, cnt);
      __syncthreads();
    }
  }
  hist[(int64_t)threadIdx.x * nb + blockIdx.x] = cnt[threadIdx.x];
}

__global__ __launch_bounds__(256) void k_scatter(const uint32_t* __restrict__ kin, const uint32_t* __restrict__ vin,
                                                 uint32_t* __restrict__ kout, uint32_t* __restrict__ vout,
                                                 const uint32_t* __restrict__ hist, const int64_t* nptr,
                                                 int shift, int nb) {
  __shared__ uint32_t run[256];
  __shared__ uint32_t wc[4][256];
  const int64_t N = *nptr;
  const int t = threadIdx.x, w = t >> 6;
  run[t] = hist[(int64_t)t * nb + blockIdx.x];
  for (int v = 0; v < 4; ++v) wc[v][t] = 0;
  __syncthreads();
  const int64_t base = (int64_t)blockIdx.x * kTile;
  for (int j = 0; j < kSortItems; ++j) {
    const int64_t i = base + (int64_t)j * 256 + t;
    const bool valid = i < N;
    const uint32_t k = valid ? kin[i] : 0u;
    const uint32_t v = valid ? (vin ? vin[i] : (uint32_t)i) : 0u;
    const uint32_t d = (k >> shift) & 255u;
    const uint32_t r = wave_match(valid, d, wc[w]);
    __syncthreads();
    if (valid) {
      uint32_t pos = run[d] + r;
      for (int q = 0; q < w; ++q) pos += wc[q][d];
      kout[pos] = k;
      vout[pos] = v;
    }
    __syncthreads();
    run[t] += wc[0][t] + wc[1][t] + wc[2][t] + wc[3][t];
    wc[0][t] = wc[1][t] = wc[2][t] = wc[3][t] = 0;
    __syncthreads();
  }
}

__global__ void k_gather64(const uint32_t* __restrict__ perm, const int64_t* nptr, const double* __restrict__ a0,
                           const double* __restrict__ a1, const double* __restrict__ a2, const double* __restrict__ a3,
                           const double* __restrict__ a4, const double* __restrict__ a5, double* b0, double* b1,
                           double* b2, double* b3, double* b4, double* b5) {
  const int64_t N = *nptr;
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  const uint32_t s = perm[i];
  b0[i] = a0[s]; b1[i] = a1[s]; b2[i] = a2[s]; b3[i] = a3[s]; b4[i] = a4[s]; b5[i] = a5[s];
}

__global__ void k_gather_mixed(const uint32_t* __restrict__ perm, const int64_t* nptr,
                               const int32_t* __restrict__ c0, const float* __restrict__ a0,
                               const float* __restrict__ a1, const float* __restrict__ a2,
                               const float* __restrict__ a3, const float* __restrict__ a4,
                               const float* __restrict__ a5, int32_t* d0, float* b0, float* b1, float* b2,
                               float* b3, float* b4, float* b5) {
  const int64_t N = *nptr;
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  const uint32_t s = perm[i];
  d0[i] = c0[s]; b0[i] = a0[s]; b1[i] = a1[s];
  if (b2) b2[i] = a2[s];
  b3[i] = a3[s]; b4[i] = a4[s]; b5[i] = a5[s];
}

int ensure_sort_scratch(lepic_ctx* ctx, int64_t n, int64_t nb, int64_t ncnt) {
  if (n > ctx->s_cap) {
    for (int b = 0; b < 2; ++b) {
      if (ctx->s_key[b]) LEPIC_HIP(hipFree(ctx->s_key[b]));
      if (ctx->s_val[b]) LEPIC_HIP(hipFree(ctx->s_val[b]));
      LEPIC_HIP(hipMalloc(&ctx->s_key[b], n * 4));
      LEPIC_HIP(hipMalloc(&ctx->s_val[b], n * 4));
    }
    ctx->s_cap = n;
  }
  if (256 * nb > ctx->s_hist_cap) {
    if (ctx->s_hist) LEPIC_HIP(hipFree(ctx->s_hist));
    LEPIC_HIP(hipMalloc(&ctx->s_hist, 256 * nb * 4));
    ctx->s_hist_cap = 256 * nb;
  }
  if (!ctx->d_plist || ncnt > ctx->scan_tmp_cap) {
    if (ctx->d_plist) LEPIC_HIP(hipFree(ctx->d_plist));
    LEPIC_HIP(hipMalloc(&ctx->d_plist, ncnt * 4));
    ctx->scan_tmp_cap = ncnt;
  }
  return 0;
}

}  // namespace

namespace lepic {

int sort_launch(lepic_ctx* ctx, int p, int32_t* Plist_host) {
  Species& s = ctx->sp[p];
  if (int rc = ensure_packed(ctx, p)) return rc;
  int64_t N = 0;
  if (int rc = lepic_get_count(ctx, p + 1, &N)) return rc;
  const int n1 = ctx->g.n1, n2 = ctx->g.n2;
  const int64_t ncell = (int64_t)n1 * n2;
  const int64_t pl_max = (int64_t)n1 + ncell + 1;
  const int64_t nb = std::max<int64_t>((N + kTile - 1) / kTile, 1);
  if (int rc = ensure_sort_scratch(ctx, std::max<int64_t>(N, 1), nb, pl_max + 1)) return rc;
  uint32_t* cnt = (uint32_t*)ctx->d_plist;
  if (int rc = zero_async(ctx, cnt, (pl_max + 1) * 4)) return rc;
  const int b = s.cur, o = 1 - s.cur;
  const int64_t* nptr = ctx->d_counts + p;
  const dim3 g1((unsigned)std::max<int64_t>((N + 255) / 256, 1));
  if (ctx->g.precision == LEPIC_MIXED)
    hipLaunchKernelGGL(k_keys_mixed, g1, dim3(256), 0, ctx->stream, s.cell[b], nptr, n1, n2, ctx->s_key[0], cnt);
  else
    hipLaunchKernelGGL(k_keys64, g1, dim3(256), 0, ctx->stream, s.d[b][0], s.d[b][1], nptr, ctx->g.h1,
                       ctx->g.h2, n1, n2, ctx->s_key[0], cnt);
  LEPIC_HIP(hipGetLastError());
  int bits = 1;
  while ((1ll << bits) < ncell) ++bits;
  const int passes = (bits + 7) / 8;
  int cur = 0;
  for (int pass = 0; pass < passes; ++pass) {
    const int shift = 8 * pass;
    hipLaunchKernelGGL(k_hist, dim3((unsigned)nb), dim3(256), 0, ctx->stream, ctx->s_key[cur], nptr, shift,
                       ctx->s_hist, (int)nb);
    LEPIC_HIP(hipGetLastError());
    if (int rc = scan_exclusive_u32(ctx, ctx->s_hist, 256 * nb)) return rc;
    hipLaunchKernelGGL(k_scatter, dim3((unsigned)nb), dim3(256), 0, ctx->stream, ctx->s_key[cur],
                       pass == 0 ? (const uint32_t*)nullptr : ctx->s_val[cur], ctx->s_key[1 - cur],
                       ctx->s_val[1 - cur], ctx->s_hist, nptr, shift, (int)nb);
    LEPIC_HIP(hipGetLastError());
    cur = 1 - cur;
  }
  const uint32_t* perm = ctx->s_val[cur];
  if (ctx->g.precision == LEPIC_MIXED)
    hipLaunchKernelGGL(k_gather_mixed, g1, dim3(256), 0, ctx->stream, perm, nptr, s.cell[b], s.f[b][0],
                       s.f[b][1], s.f[b][2], s.f[b][3], s.f[b][4], s.f[b][5], s.cell[o], s.f[o][0],
                       s.f[o][1], s.f[o][2], s.f[o][3], s.f[o][4], s.f[o][5]);
  else
    hipLaunchKernelGGL(k_gather64, g1, dim3(256), 0, ctx->stream, perm, nptr, s.d[b][0], s.d[b][1],
                       s.d[b][2], s.d[b][3], s.d[b][4], s.d[b][5], s.d[o][0], s.d[o][1], s.d[o][2],
                       s.d[o][3], s.d[o][4], s.d[o][5]);
  LEPIC_HIP(hipGetLastError());
  s.cur = o;
  if (int rc = set_packed_counts(ctx, p)) return rc;
  // Plist(c) = number of particles in cells 1..c (1-based): exclusive scan of
  // the 0-based counts, shifted by one
  if (int rc = scan_exclusive_u32(ctx, cnt, pl_max + 1)) return rc;
  if (Plist_host) {
    LEPIC_HIP(hipMemcpyAsync(Plist_host, cnt, (pl_max + 1) * 4, hipMemcpyDeviceToHost, ctx->stream));
    LEPIC_HIP(hipStreamSynchronize(ctx->stream));
  }
  return 0;
}

}  // namespace lepic
