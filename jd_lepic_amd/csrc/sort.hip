// sort.hip -- cell sort of one species (sorting.f90:92-251): a stable LSD
// radix sort of the y-major cell index in as many passes of at most 11 bits
// (2048 buckets) as the key needs -- two below 2^22 cells, three for
// 2048 x 2048 cells and beyond -- then a gather of the particle payload into
// the other buffer and, when the caller asks for it, the cell-end list
// Plist.  The sort reads the species as the mover leaves it, segment by
// segment: slots past a segment's count carry a key above every cell and
// sort behind the particles, so the output is packed without a separate
// packing pass.
//
// With one thread the reference's counting sort is stable (particles keep
// their order inside a cell); LSD radix sort with stable per-digit scatter
// gives exactly that permutation.  Per pass: a digit histogram per 16384-key
// tile (wave-aggregated LDS counts), a device-wide exclusive scan of the
// digit-major histogram, and a scatter that ranks keys inside the tile in
// index order.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "lepic_internal.h"

using namespace lepic;

namespace {

constexpr int kSortItems = 64;
constexpr int kTile = 256 * kSortItems;
constexpr int kMaxRadix = 2048;   // buckets of the widest digit (11 bits)
constexpr int kMaxDigitBits = 11;

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

// keys: 0-based cell index, kEmpty for an empty slot; slot i is a particle
// when i % kChunk < seg[i / kChunk]
constexpr uint32_t kEmpty = 0xFFFFFFFFu;

__global__ void k_keys64(const double* __restrict__ x, const double* __restrict__ y, const int32_t* __restrict__ seg,
                         int64_t nslots, double h1, double h2, int n1, int n2, uint32_t* __restrict__ key) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const bool valid = i < nslots && (int)(i & (kChunk - 1)) < seg[i / kChunk];
  uint32_t k = kEmpty;
  if (valid) {
    int ix = (int)(x[i] / h1), iy = (int)(y[i] / h2);  // INT(x/h) (sorting.f90:147-151)
    ix = ix < 0 ? 0 : (ix > n1 - 1 ? n1 - 1 : ix);
    iy = iy < 0 ? 0 : (iy > n2 - 1 ? n2 - 1 : iy);
    k = (uint32_t)ix + (uint32_t)n1 * (uint32_t)iy;
  }
  if (i < nslots) key[i] = k;
}

__global__ void k_keys_mixed(const int32_t* __restrict__ cell, const int32_t* __restrict__ seg, int64_t nslots,
                             int n1, int n2, uint32_t* __restrict__ key) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const bool valid = i < nslots && (int)(i & (kChunk - 1)) < seg[i / kChunk];
  uint32_t k = kEmpty;
  if (valid) {
    int c = cell[i];
    int ix = (c & 0xFFFF) - 1, iy = (c >> 16) - 1;
    ix = ix < 0 ? 0 : (ix > n1 - 1 ? n1 - 1 : ix);
    iy = iy < 0 ? 0 : (iy > n2 - 1 ? n2 - 1 : iy);
    k = (uint32_t)ix + (uint32_t)n1 * (uint32_t)iy;
  }
  if (i < nslots) key[i] = k;
}

// Plist from the sorted keys K[0..nslots) (empty slots, kEmpty, at the end):
// plist[c] = number of particles in cells before c = the first position
// whose key is at least c, for c = 0..pl_max.  Position i (0..nslots) owns
// the cells (K[i-1], K[i]], so every entry is written once; past the last
// particle the rest of the list takes the particle count.
__global__ void k_plist_from_keys(const uint32_t* __restrict__ K, int64_t nslots, int64_t pl_max,
                                  uint32_t* __restrict__ plist) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i > nslots) return;
  const int64_t prev = i == 0 ? -1 : (K[i - 1] == kEmpty ? pl_max : (int64_t)K[i - 1]);
  if (prev >= pl_max) return;
  const int64_t cur = (i == nslots || K[i] == kEmpty) ? pl_max : std::min((int64_t)K[i], pl_max);
  for (int64_t c = prev + 1; c <= cur; ++c) plist[c] = (uint32_t)i;
}

// Digit counts of a wave's kSortItems x 64 keys from base, added to cnt.  The
// keys are loaded kLoadBatch per lane at a time, so a wave waits for one
// memory latency per batch rather than one per 64 keys.
constexpr int kLoadBatch = 16;
static_assert(kSortItems % kLoadBatch == 0, "whole load batches per wave");

template <int R>
__device__ __forceinline__ void count_digits(const uint32_t* __restrict__ key, int64_t base, int64_t nslots,
                                             int shift, uint32_t* cnt) {
  const int lane = threadIdx.x & 63;
  for (int j0 = 0; j0 < kSortItems; j0 += kLoadBatch) {
    uint32_t d[kLoadBatch];
#pragma unroll
    for (int u = 0; u < kLoadBatch; ++u) {
      const int64_t i = base + (int64_t)(j0 + u) * 64 + lane;
      d[u] = i < nslots ? (key[i] >> shift) & (uint32_t)(R - 1) : 0u;
    }
#pragma unroll
    for (int u = 0; u < kLoadBatch; ++u) wave_match(base + (int64_t)(j0 + u) * 64 + lane < nslots, d[u], cnt);
  }
}

// Each wave owns a contiguous quarter of the tile (kTile/4 keys, read 64 at a
// time), so stability across waves is a per-digit prefix over the four
// waves and no barrier is needed inside the loops.
template <int R>   // buckets per pass
__global__ __launch_bounds__(256) void k_hist(const uint32_t* __restrict__ key, int64_t nslots, int shift,
                                              uint32_t* __restrict__ hist, int nb) {
  __shared__ uint32_t wc[4][R];
  const int t = threadIdx.x, w = t >> 6, lane = t & 63;
  for (int q = t; q < R; q += 256) wc[0][q] = wc[1][q] = wc[2][q] = wc[3][q] = 0;
  __syncthreads();
  const int64_t base = (int64_t)blockIdx.x * kTile + (int64_t)w * (kTile / 4);
  count_digits<R>(key, base, nslots, shift, wc[w]);
  __syncthreads();
  for (int q = t; q < R; q += 256) hist[(int64_t)q * nb + blockIdx.x] = wc[0][q] + wc[1][q] + wc[2][q] + wc[3][q];
}

template <int R>
__global__ __launch_bounds__(256) void k_scatter(const uint32_t* __restrict__ kin, const uint32_t* __restrict__ vin,
                                                 uint32_t* __restrict__ kout, uint32_t* __restrict__ vout,
                                                 const uint32_t* __restrict__ hist, int64_t nslots, int shift,
                                                 int nb) {
  __shared__ uint32_t wc[4][R];
  const int t = threadIdx.x, w = t >> 6, lane = t & 63;
  for (int q = t; q < R; q += 256) wc[0][q] = wc[1][q] = wc[2][q] = wc[3][q] = 0;
  __syncthreads();
  const int64_t base = (int64_t)blockIdx.x * kTile + (int64_t)w * (kTile / 4);
  // the wave's digit counts
  count_digits<R>(kin, base, nslots, shift, wc[w]);
  __syncthreads();
  // first output position of each digit for each wave
  for (int q = t; q < R; q += 256) {
    const uint32_t g = hist[(int64_t)q * nb + blockIdx.x];
    const uint32_t c0 = wc[0][q], c1 = wc[1][q], c2 = wc[2][q];
    wc[0][q] = g;
    wc[1][q] = g + c0;
    wc[2][q] = g + c0 + c1;
    wc[3][q] = g + c0 + c1 + c2;
  }
  __syncthreads();
  for (int j0 = 0; j0 < kSortItems; j0 += kLoadBatch) {
    uint32_t kk[kLoadBatch], vv[kLoadBatch];
#pragma unroll
    for (int u = 0; u < kLoadBatch; ++u) {
      const int64_t i = base + (int64_t)(j0 + u) * 64 + lane;
      const bool valid = i < nslots;
      kk[u] = valid ? kin[i] : 0u;
      vv[u] = valid ? (vin ? vin[i] : (uint32_t)i) : 0u;
    }
#pragma unroll
    for (int u = 0; u < kLoadBatch; ++u) {
      const bool valid = base + (int64_t)(j0 + u) * 64 + lane < nslots;
      const uint32_t d = (kk[u] >> shift) & (uint32_t)(R - 1);
      const uint32_t b = valid ? wc[w][d] : 0u;          // read before this round's counts are added
      const uint32_t r = wave_match(valid, d, wc[w]);
      if (valid) {
        kout[b + r] = kk[u];
        vout[b + r] = vv[u];
      }
    }
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
  if (kMaxRadix * nb > ctx->s_hist_cap) {
    if (ctx->s_hist) LEPIC_HIP(hipFree(ctx->s_hist));
    LEPIC_HIP(hipMalloc(&ctx->s_hist, kMaxRadix * nb * 4));
    ctx->s_hist_cap = kMaxRadix * nb;
  }
  if (ncnt > 0 && (!ctx->d_plist || ncnt > ctx->plist_cap)) {
    if (ctx->d_plist) LEPIC_HIP(hipFree(ctx->d_plist));
    LEPIC_HIP(hipMalloc(&ctx->d_plist, ncnt * 4));
    ctx->plist_cap = ncnt;
  }
  return 0;
}

}  // namespace

namespace lepic {

template <int R>
int radix_pass(lepic_ctx* ctx, int cur, int pass, int shift, int64_t nslots, int64_t nb) {
  hipLaunchKernelGGL(k_hist<R>, dim3((unsigned)nb), dim3(256), 0, ctx->stream, ctx->s_key[cur], nslots, shift,
                     ctx->s_hist, (int)nb);
  LEPIC_HIP(hipGetLastError());
  if (int rc = scan_exclusive_u32(ctx, ctx->s_hist, (int64_t)R * nb)) return rc;
  hipLaunchKernelGGL(k_scatter<R>, dim3((unsigned)nb), dim3(256), 0, ctx->stream, ctx->s_key[cur],
                     pass == 0 ? (const uint32_t*)nullptr : ctx->s_val[cur], ctx->s_key[1 - cur],
                     ctx->s_val[1 - cur], ctx->s_hist, nslots, shift, (int)nb);
  LEPIC_HIP(hipGetLastError());
  return 0;
}

int sort_launch(lepic_ctx* ctx, int p, int32_t* Plist_host) {
  Species& s = ctx->sp[p];
  const int n1 = ctx->g.n1, n2 = ctx->g.n2;
  const int64_t ncell = (int64_t)n1 * n2;
  if (ncell >= (1ll << 31)) return ctx->fail(LEPIC_ERR_UNSUPPORTED, "2^31 cells or more");
  const int64_t pl_max = (int64_t)n1 + ncell + 1;
  // every slot of every segment; empty slots sort behind the particles
  const int64_t nslots = nseg_of(s) * kChunk;
  const int64_t nb = std::max<int64_t>((nslots + kTile - 1) / kTile, 1);
  if (int rc = ensure_sort_scratch(ctx, std::max<int64_t>(nslots, 1), nb, Plist_host ? pl_max + 1 : 0)) return rc;
  const int b = s.cur, o = 1 - s.cur;
  const int64_t* nptr = ctx->d_counts + p;
  const dim3 gs((unsigned)std::max<int64_t>((nslots + 255) / 256, 1));
  if (ctx->g.precision == LEPIC_MIXED)
    hipLaunchKernelGGL(k_keys_mixed, gs, dim3(256), 0, ctx->stream, s.cell[b], (const int32_t*)s.seg[b], nslots,
                       n1, n2, ctx->s_key[0]);
  else
    hipLaunchKernelGGL(k_keys64, gs, dim3(256), 0, ctx->stream, s.d[b][0], s.d[b][1], (const int32_t*)s.seg[b],
                       nslots, ctx->g.h1, ctx->g.h2, n1, n2, ctx->s_key[0]);
  LEPIC_HIP(hipGetLastError());
  // bits of the largest cell index plus the empty-slot key's place above it,
  // split into the fewest digits of at most 11 bits, widest first
  int bits = 1;
  while ((1ll << bits) <= ncell) ++bits;
  const int npass = (bits + kMaxDigitBits - 1) / kMaxDigitBits;
  const int lo = (bits + npass - 1) / npass;
  int cur = 0;
  for (int pass = 0; pass < npass; ++pass) {
    const int shift = pass * lo, width = std::min(lo, bits - shift);
    int rc;
    if (width <= 8) rc = radix_pass<256>(ctx, cur, pass, shift, nslots, nb);
    else if (width <= 9) rc = radix_pass<512>(ctx, cur, pass, shift, nslots, nb);
    else if (width <= 10) rc = radix_pass<1024>(ctx, cur, pass, shift, nslots, nb);
    else rc = radix_pass<2048>(ctx, cur, pass, shift, nslots, nb);
    if (rc) return rc;
    cur = 1 - cur;
  }
  int64_t N = 0;
  if (int rc = lepic_get_count(ctx, p + 1, &N)) return rc;
  const dim3 g1((unsigned)std::max<int64_t>((N + 255) / 256, 1));
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
  if (Plist_host) {
    // Plist(c) = number of particles in cells 1..c (1-based), i.e. in the
    // 0-based cells below c: the lower bound of c in the sorted keys
    uint32_t* plist = (uint32_t*)ctx->d_plist;
    hipLaunchKernelGGL(k_plist_from_keys, dim3((unsigned)((nslots + 1 + 255) / 256)), dim3(256), 0, ctx->stream,
                       ctx->s_key[cur], nslots, pl_max, plist);
    LEPIC_HIP(hipGetLastError());
    LEPIC_HIP(hipMemcpyAsync(Plist_host, plist, (pl_max + 1) * 4, hipMemcpyDeviceToHost, ctx->stream));
    LEPIC_HIP(hipStreamSynchronize(ctx->stream));
  }
  return 0;
}

}  // namespace lepic
