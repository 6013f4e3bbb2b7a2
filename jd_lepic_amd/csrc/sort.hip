// sort.hip -- cell sort of one species (sorting.f90:92-251), written for
// particles that are nearly sorted already.
//
// The reference's counting sort with one thread is a stable sort of the
// particles by the y-major cell index ic.  The species is stored as
// segments of kChunk slots (see lepic_internal.h); since the last sort the
// particles of a segment have moved at most a few cells, so a segment
// covers few distinct cells.  The sort therefore works on runs:
//  1. k_seg_sort: one block per segment sorts the segment's keys stably in
//     LDS (LSD passes of 8-bit digits of the key minus the segment's
//     smallest key, as many as the segment's key span needs) and records,
//     for every slot, its run (a maximal group of equal keys in the sorted
//     segment) and its rank inside the run, and for every run its key and
//     particle count.
//  2. The runs of all segments, in segment order, are sorted stably by key
//     (global LSD radix passes of up to 11 bits over the runs only), and an
//     exclusive scan of their counts in sorted order gives every run its
//     first output position.
//  3. k_seg_scatter: one block per segment writes every particle to its
//     run's position plus its rank, into the other particle buffer.
// A particle's output position is then the number of particles in lower
// cells, plus those of its cell in earlier segments, plus those of its cell
// earlier in its own segment: exactly the stable sort's.  The payload moves
// once (read and written), the keys are read once.  The cell-end list
// Plist comes from the sorted run keys and offsets when the caller asks for
// it.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <map>
#include <mutex>
#include <string>

#include "lepic_internal.h"

using namespace lepic;

namespace {

constexpr int kSortItems = 64;
constexpr int kTile = 256 * kSortItems;
constexpr int kMaxRadix = 2048;   // buckets of the widest global digit (11 bits)
constexpr int kMaxDigitBits = 11;
constexpr int kSegThreads = 256;  // one block per segment: 4 waves x 4 rounds of 64 slots
constexpr int kSegRounds = kChunk / kSegThreads;
constexpr int kRankBits = 10;     // rank inside a run and run index inside a segment: < kChunk
static_assert(kChunk == (1 << kRankBits), "meta packing assumes 1024-slot segments");

__device__ __forceinline__ uint64_t lanemask_lt() { return (1ull << (threadIdx.x & 63)) - 1ull; }

// Lanes of the wave holding my digit, matched by one ballot per digit bit
// (cost independent of how many distinct digits the wave holds).  Returns my
// rank among the valid earlier lanes with my digit; the first lane of each
// digit group adds the group size to cnt[d] (the wave's own LDS row).
template <int NB>
__device__ __forceinline__ uint32_t wave_match(bool valid, uint32_t d, uint32_t* cnt) {
  uint64_t m = __ballot(valid);
#pragma unroll
  for (int b = 0; b < NB; ++b) {
    const bool bit = (d >> b) & 1u;
    const uint64_t ones = __ballot(bit);
    m &= bit ? ones : ~ones;
  }
  const uint32_t rank = (uint32_t)__popcll(m & lanemask_lt());
  if (valid && rank == 0) cnt[d] += (uint32_t)__popcll(m);
  return rank;
}

template <int R> struct Log2;
template <> struct Log2<256> { static constexpr int v = 8; };
template <> struct Log2<512> { static constexpr int v = 9; };
template <> struct Log2<1024> { static constexpr int v = 10; };
template <> struct Log2<2048> { static constexpr int v = 11; };

// ---------------------------------------------------------------------------
// 1. segment-local stable sort and runs
// ---------------------------------------------------------------------------
struct SegSortArgs {
  const int32_t* seg;         // particles per segment
  const int32_t* cell;        // mixed: (iy<<16)|ix, 1-based
  const double* x;            // FP64 positions
  const double* y;
  double h1, h2;
  int n1, n2;
  uint32_t* meta;             // [slot] (run << kRankBits) | rank in run
  uint32_t* rkey;             // [segment * kChunk + run] cell key of the run
  uint32_t* rcnt;             // [segment * kChunk + run] particles of the run
  uint32_t* nrun;             // [segment] runs of the segment
};

// 0-based cell (ix, iy) of slot i: INT(x/h) with the clamp of
// sorting.f90:147-151 (FP64), or the stored cell (mixed); the y-major sort
// key is ix + n1*iy
template <bool kMixed>
__device__ __forceinline__ void cell_of(const SegSortArgs& A, int64_t i, uint32_t& cx, uint32_t& cy) {
  int ix, iy;
  if (kMixed) {
    const int c = A.cell[i];
    ix = (c & 0xFFFF) - 1;
    iy = (c >> 16) - 1;
  } else {
    ix = (int)(A.x[i] / A.h1);
    iy = (int)(A.y[i] / A.h2);
  }
  ix = ix < 0 ? 0 : (ix > A.n1 - 1 ? A.n1 - 1 : ix);
  iy = iy < 0 ? 0 : (iy > A.n2 - 1 ? A.n2 - 1 : iy);
  cx = (uint32_t)ix;
  cy = (uint32_t)iy;
}

// exclusive prefix of v over the block's 256 threads in thread order; the
// block total in *total.  s_tot: 4 words of LDS.  Contains barriers.
__device__ __forceinline__ uint32_t block_excl_scan(uint32_t v, uint32_t* s_tot, uint32_t* total) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  uint32_t inc = v;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const uint32_t u = __shfl_up(inc, o, 64);
    if (lane >= o) inc += u;
  }
  if (lane == 63) s_tot[w] = inc;
  __syncthreads();
  uint32_t pre = inc - v, tot = 0;
#pragma unroll
  for (int q = 0; q < kSegThreads / 64; ++q) {
    const uint32_t c = s_tot[q];
    if (q < w) pre += c;
    tot += c;
  }
  *total = tot;
  __syncthreads();
  return pre;
}

template <bool kMixed>
__global__ __launch_bounds__(kSegThreads, 8) void k_seg_sort(SegSortArgs A) {
  __shared__ uint32_t s_k[2][kChunk];
  __shared__ uint16_t s_i[2][kChunk];
  __shared__ uint32_t s_wc[4][256];
  __shared__ uint16_t s_head[kChunk + 1];
  __shared__ uint32_t s_mm[4][4];
  __shared__ uint32_t s_tot[4];
  const int t = threadIdx.x, w = t >> 6, lane = t & 63;
  const int64_t sg = blockIdx.x;
  const int cnt = A.seg[sg];
  const int64_t base = sg * (int64_t)kChunk;
  if (cnt <= 0) {                       // block-uniform
    if (t == 0) A.nrun[sg] = 0;
    return;
  }
  // cells in slot order; wave w owns positions [256 w, 256 w + 256).  The
  // segment is sorted by the compact key (iy - ymin) * dxs + (ix - xmin),
  // dxs = xmax - xmin + 1, which orders cells as ix + n1*iy does but spans
  // far fewer values when the segment's particles sit in a few short pieces
  // of neighbouring rows (one 8-bit pass instead of two)
  uint32_t X[kSegRounds], Y[kSegRounds];
  uint32_t xmin = 0xFFFFFFFFu, xmax = 0u, ymin = 0xFFFFFFFFu, ymax = 0u;
#pragma unroll
  for (int r = 0; r < kSegRounds; ++r) {
    const int q = w * 256 + r * 64 + lane;
    X[r] = Y[r] = 0u;
    if (q < cnt) {
      cell_of<kMixed>(A, base + q, X[r], Y[r]);
      xmin = min(xmin, X[r]); xmax = max(xmax, X[r]);
      ymin = min(ymin, Y[r]); ymax = max(ymax, Y[r]);
    }
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    xmin = min(xmin, (uint32_t)__shfl_xor((int)xmin, o, 64));
    xmax = max(xmax, (uint32_t)__shfl_xor((int)xmax, o, 64));
    ymin = min(ymin, (uint32_t)__shfl_xor((int)ymin, o, 64));
    ymax = max(ymax, (uint32_t)__shfl_xor((int)ymax, o, 64));
  }
  if (lane == 0) { s_mm[0][w] = xmin; s_mm[1][w] = xmax; s_mm[2][w] = ymin; s_mm[3][w] = ymax; }
  __syncthreads();
  xmin = min(min(s_mm[0][0], s_mm[0][1]), min(s_mm[0][2], s_mm[0][3]));
  xmax = max(max(s_mm[1][0], s_mm[1][1]), max(s_mm[1][2], s_mm[1][3]));
  ymin = min(min(s_mm[2][0], s_mm[2][1]), min(s_mm[2][2], s_mm[2][3]));
  ymax = max(max(s_mm[3][0], s_mm[3][1]), max(s_mm[3][2], s_mm[3][3]));
  const uint32_t dxs = xmax - xmin + 1u;
#pragma unroll
  for (int r = 0; r < kSegRounds; ++r) {
    const int q = w * 256 + r * 64 + lane;
    if (q < cnt) {
      s_k[0][q] = (Y[r] - ymin) * dxs + (X[r] - xmin);
      s_i[0][q] = (uint16_t)q;
    }
  }
  const uint32_t kmin = 0u, span = (ymax - ymin) * dxs + (xmax - xmin);
  int npass = 0;
  while (npass < 4 && (span >> (8 * npass)) != 0u) ++npass;
  if (npass <= 1) {  // block-uniform
    // The compact key is a single 8-bit digit: its counts give every slot's
    // place directly.  The slot's run is the number of distinct keys below
    // its own and its rank the number of equal keys before it (the wave's
    // earlier rounds, earlier waves, then earlier lanes): the stable sort's
    // runs without moving the keys.
    for (int e = t; e < 4 * 256; e += kSegThreads) (&s_wc[0][0])[e] = 0u;
    __syncthreads();
    uint32_t d[kSegRounds], pos[kSegRounds];
#pragma unroll
    for (int r = 0; r < kSegRounds; ++r) {
      const int q = w * 256 + r * 64 + lane;
      const bool v = q < cnt;
      d[r] = v ? (Y[r] - ymin) * dxs + (X[r] - xmin) : 0u;
      const uint32_t before = v ? s_wc[w][d[r]] : 0u;
      pos[r] = before + wave_match<8>(v, d[r], s_wc[w]);
    }
    __syncthreads();
    // key t: its count and its run index
    const uint32_t c0 = s_wc[0][t], c1 = s_wc[1][t], c2 = s_wc[2][t], c3 = s_wc[3][t];
    const uint32_t ck = c0 + c1 + c2 + c3;
    uint32_t nrun;
    const uint32_t run = block_excl_scan(ck > 0u ? 1u : 0u, s_tot, &nrun);
    // the earlier waves' counts of key t, then the run index, in LDS
    s_wc[1][t] = c0;
    s_wc[2][t] = c0 + c1;
    s_wc[3][t] = c0 + c1 + c2;
    s_wc[0][t] = run;
    if (ck > 0u) {
      A.rkey[base + run] = (t % dxs + xmin) + (uint32_t)A.n1 * (t / dxs + ymin);
      A.rcnt[base + run] = ck;
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < kSegRounds; ++r) {
      const int q = w * 256 + r * 64 + lane;
      if (q < cnt) {
        const uint32_t rank = (w > 0 ? s_wc[w][d[r]] : 0u) + pos[r];
        A.meta[base + q] = (s_wc[0][d[r]] << kRankBits) | rank;
      }
    }
    if (t == 0) A.nrun[sg] = nrun;
    return;
  }
  // stable LSD passes over the compact key (two or more digits; each pass
  // starts with a barrier, which also orders the key stores above)
  int cur = 0;
  for (int p = 0; p < npass; ++p) {
    const int shift = 8 * p;
    for (int e = t; e < 4 * 256; e += kSegThreads) (&s_wc[0][0])[e] = 0u;
    __syncthreads();
    uint32_t d[kSegRounds], pos[kSegRounds];
    bool v[kSegRounds];
#pragma unroll
    for (int r = 0; r < kSegRounds; ++r) {
      const int q = w * 256 + r * 64 + lane;
      v[r] = q < cnt;
      d[r] = v[r] ? ((s_k[cur][q] - kmin) >> shift) & 255u : 0u;
      // the wave's items with my digit in earlier rounds (read before this
      // round's counts are added) plus those before me in this round
      const uint32_t before = v[r] ? s_wc[w][d[r]] : 0u;
      pos[r] = before + wave_match<8>(v[r], d[r], s_wc[w]);
    }
    __syncthreads();
    // digit t: first position of each wave's items with that digit
    const uint32_t c0 = s_wc[0][t], c1 = s_wc[1][t], c2 = s_wc[2][t], c3 = s_wc[3][t];
    uint32_t tot;
    const uint32_t pre = block_excl_scan(c0 + c1 + c2 + c3, s_tot, &tot);
    s_wc[0][t] = pre;
    s_wc[1][t] = pre + c0;
    s_wc[2][t] = pre + c0 + c1;
    s_wc[3][t] = pre + c0 + c1 + c2;
    __syncthreads();
#pragma unroll
    for (int r = 0; r < kSegRounds; ++r) {
      const int q = w * 256 + r * 64 + lane;
      if (v[r]) {
        const uint32_t dst = s_wc[w][d[r]] + pos[r];
        s_k[1 - cur][dst] = s_k[cur][q];
        s_i[1 - cur][dst] = s_i[cur][q];
      }
    }
    __syncthreads();
    cur = 1 - cur;
  }
  // runs of equal keys over the sorted positions; thread t owns 4t .. 4t+3
  uint32_t kq[4];
  uint32_t heads = 0, nh = 0;
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const int q = 4 * t + j;
    kq[j] = q < cnt ? s_k[cur][q] : 0u;
    const bool hd = q < cnt && (q == 0 || s_k[cur][q - 1] != kq[j]);
    heads |= hd ? 1u << j : 0u;
    nh += hd ? 1u : 0u;
  }
  uint32_t nrun;
  const uint32_t rid0 = block_excl_scan(nh, s_tot, &nrun);
  uint32_t rid[4];
  {
    uint32_t r = rid0;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      if ((heads >> j) & 1u) {
        s_head[r] = (uint16_t)(4 * t + j);
        ++r;
      }
      rid[j] = r - 1;   // the run of position 4t+j (its head is at or before it)
    }
  }
  if (t == 0) s_head[nrun] = (uint16_t)cnt;
  __syncthreads();
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const int q = 4 * t + j;
    if (q < cnt) A.meta[base + s_i[cur][q]] = (rid[j] << kRankBits) | (uint32_t)(q - s_head[rid[j]]);
  }
  for (uint32_t r = t; r < nrun; r += kSegThreads) {
    const int h = s_head[r];
    const uint32_t ck = s_k[cur][h];
    A.rkey[base + r] = (ck % dxs + xmin) + (uint32_t)A.n1 * (ck / dxs + ymin);
    A.rcnt[base + r] = (uint32_t)(s_head[r + 1] - h);
  }
  if (t == 0) A.nrun[sg] = nrun;
}

// the runs of every segment into dense arrays, segment order (run q of the
// species = base[s] + r); base is the exclusive scan of the run counts
__global__ __launch_bounds__(64) void k_runs_pack(const uint32_t* __restrict__ rkey, const uint32_t* __restrict__ rcnt,
                                                  const uint32_t* __restrict__ rbase, uint32_t* __restrict__ key,
                                                  uint32_t* __restrict__ cnt) {
  const int64_t sg = blockIdx.x;
  const uint32_t b = rbase[sg], n = min(rbase[sg + 1] - b, (uint32_t)kChunk);
  for (uint32_t r = threadIdx.x; r < n; r += 64) {
    key[b + r] = rkey[sg * (int64_t)kChunk + r];
    cnt[b + r] = rcnt[sg * (int64_t)kChunk + r];
  }
}

// ---------------------------------------------------------------------------
// 2. global stable radix sort of the runs by key
// ---------------------------------------------------------------------------
// Digit counts of a wave's kSortItems x 64 keys from base, added to cnt.  The
// keys are loaded kLoadBatch per lane at a time, so a wave waits for one
// memory latency per batch rather than one per 64 keys.
constexpr int kLoadBatch = 16;
static_assert(kSortItems % kLoadBatch == 0, "whole load batches per wave");

template <int R>
__device__ __forceinline__ void count_digits(const uint32_t* __restrict__ key, int64_t base, int64_t n, int shift,
                                             uint32_t* cnt) {
  const int lane = threadIdx.x & 63;
  for (int j0 = 0; j0 < kSortItems; j0 += kLoadBatch) {
    uint32_t d[kLoadBatch];
#pragma unroll
    for (int u = 0; u < kLoadBatch; ++u) {
      const int64_t i = base + (int64_t)(j0 + u) * 64 + lane;
      d[u] = i < n ? (key[i] >> shift) & (uint32_t)(R - 1) : 0u;
    }
#pragma unroll
    for (int u = 0; u < kLoadBatch; ++u)
      wave_match<Log2<R>::v>(base + (int64_t)(j0 + u) * 64 + lane < n, d[u], cnt);
  }
}

// Each wave owns a contiguous quarter of the tile (kTile/4 keys, read 64 at a
// time), so stability across waves is a per-digit prefix over the four
// waves and no barrier is needed inside the loops.
template <int R>   // buckets per pass
__global__ __launch_bounds__(256) void k_hist(const uint32_t* __restrict__ key, int64_t n, int shift,
                                              uint32_t* __restrict__ hist, int nb) {
  __shared__ uint32_t wc[4][R];
  const int t = threadIdx.x, w = t >> 6;
  for (int q = t; q < R; q += 256) wc[0][q] = wc[1][q] = wc[2][q] = wc[3][q] = 0;
  __syncthreads();
  const int64_t base = (int64_t)blockIdx.x * kTile + (int64_t)w * (kTile / 4);
  count_digits<R>(key, base, n, shift, wc[w]);
  __syncthreads();
  for (int q = t; q < R; q += 256) hist[(int64_t)q * nb + blockIdx.x] = wc[0][q] + wc[1][q] + wc[2][q] + wc[3][q];
}

template <int R>
__global__ __launch_bounds__(256) void k_scatter(const uint32_t* __restrict__ kin, const uint32_t* __restrict__ vin,
                                                 uint32_t* __restrict__ kout, uint32_t* __restrict__ vout,
                                                 const uint32_t* __restrict__ hist, int64_t n, int shift, int nb) {
  __shared__ uint32_t wc[4][R];
  const int t = threadIdx.x, w = t >> 6, lane = t & 63;
  for (int q = t; q < R; q += 256) wc[0][q] = wc[1][q] = wc[2][q] = wc[3][q] = 0;
  __syncthreads();
  const int64_t base = (int64_t)blockIdx.x * kTile + (int64_t)w * (kTile / 4);
  count_digits<R>(kin, base, n, shift, wc[w]);
  __syncthreads();
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
      const bool valid = i < n;
      kk[u] = valid ? kin[i] : 0u;
      vv[u] = valid ? (vin ? vin[i] : (uint32_t)i) : 0u;
    }
#pragma unroll
    for (int u = 0; u < kLoadBatch; ++u) {
      const bool valid = base + (int64_t)(j0 + u) * 64 + lane < n;
      const uint32_t d = (kk[u] >> shift) & (uint32_t)(R - 1);
      const uint32_t b = valid ? wc[w][d] : 0u;          // read before this round's counts are added
      const uint32_t r = wave_match<Log2<R>::v>(valid, d, wc[w]);
      if (valid) {
        kout[b + r] = kk[u];
        vout[b + r] = vv[u];
      }
    }
  }
}

// counts of the runs in sorted order (for the offset scan) and, after the
// scan, each run's first output position by its dense index
__global__ void k_runs_sorted_counts(const uint32_t* __restrict__ perm, const uint32_t* __restrict__ cnt,
                                     uint32_t* __restrict__ out, int64_t M) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j < M) out[j] = cnt[perm[j]];
}

__global__ void k_runs_offsets(const uint32_t* __restrict__ perm, const uint32_t* __restrict__ off,
                               uint32_t* __restrict__ roff, int64_t M) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j < M && perm[j] < M) roff[perm[j]] = off[j];
}

// Plist from the sorted run keys K[0..M) and their first positions off:
// plist[c] = number of particles in 0-based cells below c, for c = 0..pl_max.
// Position j (0..M) owns the cells (K[j-1], K[j]], past the last run the
// rest of the list takes the particle count.
__global__ void k_plist_from_runs(const uint32_t* __restrict__ K, const uint32_t* __restrict__ off, int64_t M,
                                  int64_t pl_max, const int64_t* __restrict__ nptr, uint32_t* __restrict__ plist) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j > M) return;
  const int64_t prev = j == 0 ? -1 : (int64_t)K[j - 1];
  if (prev >= pl_max) return;
  const int64_t cur = j == M ? pl_max : std::min((int64_t)K[j], pl_max);
  const uint32_t v = j == M ? (uint32_t)*nptr : off[j];
  for (int64_t c = prev + 1; c <= cur; ++c) plist[c] = v;
}

// ---------------------------------------------------------------------------
// 3. payload scatter
// ---------------------------------------------------------------------------
struct SegScatterArgs {
  const int32_t* seg;
  const uint32_t* meta;
  const uint32_t* rbase;      // first dense run of each segment
  const uint32_t* roff;       // first output position of each dense run
  const double* a[6];         // FP64 payload in
  double* b[6];               // FP64 payload out
  const int32_t* ci;          // mixed payload in: cell + 6 FP32 (z may be null)
  const float* f[6];
  int32_t* co;
  float* g[6];
  int64_t cap;                // slots of the output buffer
  int32_t* err;               // set when a destination falls outside it
};

template <bool kMixed>
__global__ __launch_bounds__(kSegThreads) void k_seg_scatter(SegScatterArgs A) {
  // kSegRounds slots per thread (slot t + r*kSegThreads, so every access of a
  // round is coalesced); the payload and run loads of all rounds are issued
  // before any store
  constexpr int R = kSegRounds;
  const int64_t sg = blockIdx.x;
  const int cnt = A.seg[sg];
  const int64_t base = sg * (int64_t)kChunk;
  const uint32_t rb = A.rbase[sg];
  const int t = threadIdx.x;
  bool v[R];
  uint32_t m[R];
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const int j = t + r * kSegThreads;
    v[r] = j < cnt;
    m[r] = v[r] ? A.meta[base + j] : 0u;
  }
  if (kMixed) {
    int32_t c[R];
    float f[6][R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const int64_t i = base + t + r * kSegThreads;
      c[r] = v[r] ? A.ci[i] : 0;
#pragma unroll
      for (int k = 0; k < 6; ++k) f[k][r] = (v[r] && (k != 2 || A.g[2])) ? A.f[k][i] : 0.f;
    }
    int64_t o[R];
#pragma unroll
    for (int r = 0; r < R; ++r) o[r] = v[r] ? (int64_t)A.roff[rb + (m[r] >> kRankBits)] + (m[r] & (kChunk - 1)) : 0;
#pragma unroll
    for (int r = 0; r < R; ++r) {
      if (!v[r]) continue;
      if (o[r] >= A.cap) {        // cannot happen for consistent runs; never write past the buffer
        *A.err = 1;
        continue;
      }
      A.co[o[r]] = c[r];
#pragma unroll
      for (int k = 0; k < 6; ++k)
        if (k != 2 || A.g[2]) A.g[k][o[r]] = f[k][r];
    }
  } else {
    double d[6][R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const int64_t i = base + t + r * kSegThreads;
#pragma unroll
      for (int k = 0; k < 6; ++k) d[k][r] = v[r] ? A.a[k][i] : 0.0;
    }
    int64_t o[R];
#pragma unroll
    for (int r = 0; r < R; ++r) o[r] = v[r] ? (int64_t)A.roff[rb + (m[r] >> kRankBits)] + (m[r] & (kChunk - 1)) : 0;
#pragma unroll
    for (int r = 0; r < R; ++r) {
      if (!v[r]) continue;
      if (o[r] >= A.cap) {
        *A.err = 1;
        continue;
      }
#pragma unroll
      for (int k = 0; k < 6; ++k) A.b[k][o[r]] = d[k][r];
    }
  }
}

template <typename T>
int grow(lepic_ctx* ctx, T** p, int64_t n) {
  if (*p) LEPIC_HIP(hipFree(*p));
  *p = nullptr;
  LEPIC_HIP(hipMalloc(p, std::max<int64_t>(n, 1) * sizeof(T)));
  return 0;
}

// Device buffers of one sort: keys and values of the run sort (two of each),
// the per-slot (run, rank), the per-segment runs (key, count) at a stride
// of kChunk, the dense run counts, the run bases, the digit histograms, the
// cell-end list and the scans' block partials (null: allocated by the scan).
struct SortWork {
  uint32_t* key[2];
  uint32_t* val[2];
  uint32_t* meta;
  uint32_t* rkey;
  uint32_t* rcnt;
  uint32_t* rcntd;
  uint32_t* nrun;
  uint32_t* hist;
  uint32_t* plist;
  uint32_t* scan_ws;
};

int ensure_sort_scratch(lepic_ctx* ctx, int64_t nslots, int64_t nb, int64_t ncnt) {
  if (nslots > ctx->s_cap) {
    for (int b = 0; b < 2; ++b) {
      if (int rc = grow(ctx, &ctx->s_key[b], nslots)) return rc;
      if (int rc = grow(ctx, &ctx->s_val[b], nslots)) return rc;
    }
    ctx->s_cap = nslots;
  }
  if (nslots > ctx->s_seg_cap) {
    if (int rc = grow(ctx, &ctx->s_meta, nslots)) return rc;
    if (int rc = grow(ctx, &ctx->s_rkey, nslots)) return rc;
    if (int rc = grow(ctx, &ctx->s_rcnt, nslots)) return rc;
    if (int rc = grow(ctx, &ctx->s_rcntd, nslots)) return rc;
    if (int rc = grow(ctx, &ctx->s_nrun, nslots / kChunk + 2)) return rc;
    ctx->s_seg_cap = nslots;
  }
  if (kMaxRadix * nb > ctx->s_hist_cap) {
    if (int rc = grow(ctx, &ctx->s_hist, kMaxRadix * nb)) return rc;
    ctx->s_hist_cap = kMaxRadix * nb;
  }
  if (ncnt > 0 && (!ctx->d_plist || ncnt > ctx->plist_cap)) {
    if (int rc = grow(ctx, (uint32_t**)&ctx->d_plist, ncnt)) return rc;
    ctx->plist_cap = ncnt;
  }
  return 0;
}

// The static scratch of lepic_part_sorting_noalloc: a species of at most
// kStaticSlots slots (kStaticSlots / kChunk segments) and a cell-end list of
// at most kStaticPlist entries (grids of up to 1024 x 1024 cells).  Every
// call on a device shares that device's copy; a call leaves its contents
// behind, and the sort writes every word it reads before reading it (the
// one counter it needs zeroed, the run-count terminator, is cleared by a
// fill kernel).
constexpr int64_t kStaticSlots = int64_t(1) << 19;
constexpr int64_t kStaticSeg = kStaticSlots / kChunk;
constexpr int64_t kStaticHist = int64_t(kMaxRadix) * (kStaticSlots / kTile);
constexpr int64_t kStaticPlist = (int64_t(1) << 20) + 4096;
constexpr int64_t kStaticScanWs = 2 * (kStaticSlots / 4096 + 64);   // >= scan_workspace_words of any scan here
constexpr int64_t kStaticWords = 8 * kStaticSlots + (kStaticSeg + 2) + kStaticHist + kStaticPlist + kStaticScanWs;
static_assert(kStaticSlots % kTile == 0, "whole histogram tiles");

static __device__ __align__(256) uint32_t scratch[kStaticWords];

}  // namespace

namespace lepic {

template <int R>
int radix_pass(lepic_ctx* ctx, const SortWork& W, int cur, int pass, int shift, int64_t n, int64_t nb) {
  hipLaunchKernelGGL(k_hist<R>, dim3((unsigned)nb), dim3(256), 0, ctx->stream, W.key[cur], n, shift, W.hist,
                     (int)nb);
  LEPIC_HIP(hipGetLastError());
  if (int rc = scan_exclusive_u32(ctx, W.hist, (int64_t)R * nb, W.scan_ws)) return rc;
  hipLaunchKernelGGL(k_scatter<R>, dim3((unsigned)nb), dim3(256), 0, ctx->stream, W.key[cur],
                     pass == 0 ? (const uint32_t*)nullptr : W.val[cur], W.key[1 - cur], W.val[1 - cur], W.hist, n,
                     shift, (int)nb);
  LEPIC_HIP(hipGetLastError());
  return 0;
}

// The sort of species p with its device buffers in W (sized for nseg
// segments and, with Plist_host, pl_max + 1 list entries).
int sort_core(lepic_ctx* ctx, int p, int32_t* Plist_host, const SortWork& W, int64_t nseg, int64_t pl_max) {
  Species& s = ctx->sp[p];
  const int n1 = ctx->g.n1, n2 = ctx->g.n2;
  const int64_t ncell = (int64_t)n1 * n2;
  const bool mixed = ctx->g.precision == LEPIC_MIXED;
  const int b = s.cur, o = 1 - s.cur;
  int64_t M = 0;
  if (nseg > 0) {
    // 1. segment-local sorts and their runs
    SegSortArgs A{};
    A.seg = s.seg[b];
    A.cell = s.cell[b];
    A.x = s.d[b][0];
    A.y = s.d[b][1];
    A.h1 = ctx->g.h1; A.h2 = ctx->g.h2; A.n1 = n1; A.n2 = n2;
    A.meta = W.meta; A.rkey = W.rkey; A.rcnt = W.rcnt; A.nrun = W.nrun;
    if (mixed) hipLaunchKernelGGL(k_seg_sort<true>, dim3((unsigned)nseg), dim3(kSegThreads), 0, ctx->stream, A);
    else hipLaunchKernelGGL(k_seg_sort<false>, dim3((unsigned)nseg), dim3(kSegThreads), 0, ctx->stream, A);
    LEPIC_HIP(hipGetLastError());
    // run bases (exclusive scan over the segments; entry nseg = M)
    if (int rc = zero_async(ctx, W.nrun + nseg, 4)) return rc;
    if (int rc = scan_exclusive_u32(ctx, W.nrun, nseg + 1, W.scan_ws)) return rc;
    LEPIC_HIP(hipMemcpyAsync(ctx->h_pinned, W.nrun + nseg, 4, hipMemcpyDeviceToHost, ctx->stream));
    LEPIC_HIP(hipStreamSynchronize(ctx->stream));
    uint32_t m32;
    std::memcpy(&m32, ctx->h_pinned, 4);
    M = m32;
    hipLaunchKernelGGL(k_runs_pack, dim3((unsigned)nseg), dim3(64), 0, ctx->stream, (const uint32_t*)W.rkey,
                       (const uint32_t*)W.rcnt, (const uint32_t*)W.nrun, W.key[0], W.rcntd);
    LEPIC_HIP(hipGetLastError());
  }
  // 2. the runs sorted by key (stable), in the fewest digits of at most 11 bits
  int cur = 0;
  if (M > 0) {
    const int64_t nb = (M + kTile - 1) / kTile;
    int bits = 1;
    while ((1ll << bits) < ncell) ++bits;
    const int npass = (bits + kMaxDigitBits - 1) / kMaxDigitBits;
    const int lo = (bits + npass - 1) / npass;
    for (int pass = 0; pass < npass; ++pass) {
      const int shift = pass * lo, width = std::min(lo, bits - shift);
      int rc;
      if (width <= 8) rc = radix_pass<256>(ctx, W, cur, pass, shift, M, nb);
      else if (width <= 9) rc = radix_pass<512>(ctx, W, cur, pass, shift, M, nb);
      else if (width <= 10) rc = radix_pass<1024>(ctx, W, cur, pass, shift, M, nb);
      else rc = radix_pass<2048>(ctx, W, cur, pass, shift, M, nb);
      if (rc) return rc;
      cur = 1 - cur;
    }
    // first output position of every run: scan of the counts in sorted order
    const uint32_t* perm = W.val[cur];
    uint32_t* off = W.val[1 - cur];
    const dim3 gm((unsigned)((M + 255) / 256));
    hipLaunchKernelGGL(k_runs_sorted_counts, gm, dim3(256), 0, ctx->stream, perm, (const uint32_t*)W.rcntd, off, M);
    LEPIC_HIP(hipGetLastError());
    if (int rc = scan_exclusive_u32(ctx, off, M, W.scan_ws)) return rc;
    hipLaunchKernelGGL(k_runs_offsets, gm, dim3(256), 0, ctx->stream, perm, (const uint32_t*)off, W.rkey, M);
    LEPIC_HIP(hipGetLastError());
    // 3. the payload, once
    SegScatterArgs S{};
    S.seg = s.seg[b]; S.meta = W.meta; S.rbase = W.nrun; S.roff = W.rkey;
    for (int k = 0; k < 6; ++k) {
      S.a[k] = s.d[b][k]; S.b[k] = s.d[o][k];
      S.f[k] = s.f[b][k]; S.g[k] = s.f[o][k];
    }
    S.ci = s.cell[b]; S.co = s.cell[o];
    S.cap = (s.nseg_cap - 1) * kChunk;   // alloc_species: nseg_cap = slots / kChunk + 1
    S.err = ctx->d_error + 12;
    if (mixed) hipLaunchKernelGGL(k_seg_scatter<true>, dim3((unsigned)nseg), dim3(kSegThreads), 0, ctx->stream, S);
    else hipLaunchKernelGGL(k_seg_scatter<false>, dim3((unsigned)nseg), dim3(kSegThreads), 0, ctx->stream, S);
    LEPIC_HIP(hipGetLastError());
  }
  s.cur = o;
  if (int rc = set_packed_counts(ctx, p)) return rc;
  if (M > 0) {
    int32_t err = 0;
    LEPIC_HIP(hipMemcpyAsync(ctx->h_pinned, ctx->d_error + 12, 4, hipMemcpyDeviceToHost, ctx->stream));
    LEPIC_HIP(hipStreamSynchronize(ctx->stream));
    std::memcpy(&err, ctx->h_pinned, 4);
    if (err) return ctx->fail(LEPIC_ERR_STATE, "cell sort: inconsistent runs");
  }
  if (Plist_host) {
    hipLaunchKernelGGL(k_plist_from_runs, dim3((unsigned)((M + 1 + 255) / 256)), dim3(256), 0, ctx->stream,
                       (const uint32_t*)W.key[cur], (const uint32_t*)W.val[1 - cur], M, pl_max,
                       (const int64_t*)(ctx->d_counts + p), W.plist);
    LEPIC_HIP(hipGetLastError());
    LEPIC_HIP(hipMemcpyAsync(Plist_host, W.plist, (pl_max + 1) * 4, hipMemcpyDeviceToHost, ctx->stream));
    LEPIC_HIP(hipStreamSynchronize(ctx->stream));
  }
  return 0;
}

int sort_launch(lepic_ctx* ctx, int p, int32_t* Plist_host) {
  Species& s = ctx->sp[p];
  const int64_t ncell = (int64_t)ctx->g.n1 * ctx->g.n2;
  if (ncell >= (1ll << 31)) return ctx->fail(LEPIC_ERR_UNSUPPORTED, "2^31 cells or more");
  const int64_t pl_max = (int64_t)ctx->g.n1 + ncell + 1;
  const int64_t nseg = std::min<int64_t>(nseg_of(s), s.nseg_cap);
  const int64_t nslots = std::max<int64_t>(nseg, 1) * kChunk;
  if (int rc = ensure_sort_scratch(ctx, nslots, std::max<int64_t>((nslots + kTile - 1) / kTile, 1),
                                   Plist_host ? pl_max + 1 : 0))
    return rc;
  SortWork W{};
  for (int k = 0; k < 2; ++k) { W.key[k] = ctx->s_key[k]; W.val[k] = ctx->s_val[k]; }
  W.meta = ctx->s_meta; W.rkey = ctx->s_rkey; W.rcnt = ctx->s_rcnt; W.rcntd = ctx->s_rcntd;
  W.nrun = ctx->s_nrun; W.hist = ctx->s_hist; W.plist = (uint32_t*)ctx->d_plist; W.scan_ws = nullptr;
  return sort_core(ctx, p, Plist_host, W, nseg, pl_max);
}

int sort_scratch_resolve(lepic_ctx* ctx) {
  // one lookup per device, kept for the process (each device holds its own
  // copy of the array at its own address)
  static std::mutex mu;
  static std::map<int, void*> addr;
  int dev = -1;
  LEPIC_HIP(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lock(mu);
  auto it = addr.find(dev);
  if (it == addr.end()) {
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr)
      return ctx->fail(LEPIC_ERR_HIP, std::string("hipGetSymbolAddress(sort scratch): ") + hipGetErrorString(e));
    it = addr.emplace(dev, ptr).first;
  }
  ctx->sort_static = it->second;
  return 0;
}

int sort_noalloc_launch(lepic_ctx* ctx, int p, int32_t* Plist_host) {
  Species& s = ctx->sp[p];
  if (!ctx->sort_static) return ctx->fail(LEPIC_ERR_STATE, "static sort scratch not resolved on this device");
  const int64_t ncell = (int64_t)ctx->g.n1 * ctx->g.n2;
  const int64_t pl_max = (int64_t)ctx->g.n1 + ncell + 1;
  const int64_t nseg = std::min<int64_t>(nseg_of(s), s.nseg_cap);
  if (nseg > kStaticSeg)
    return ctx->fail(LEPIC_ERR_CAPACITY, "lepic_part_sorting_noalloc: the species spans " + std::to_string(nseg) +
                                             " segments of 1024 slots, the static scratch holds " +
                                             std::to_string(kStaticSeg) + " (524288 particles)");
  if (Plist_host && pl_max + 1 > kStaticPlist)
    return ctx->fail(LEPIC_ERR_CAPACITY, "lepic_part_sorting_noalloc: the cell-end list needs " +
                                             std::to_string(pl_max + 1) + " entries, the static scratch holds " +
                                             std::to_string(kStaticPlist));
  uint32_t* w = (uint32_t*)ctx->sort_static;
  SortWork W{};
  for (int k = 0; k < 2; ++k) {
    W.key[k] = w; w += kStaticSlots;
    W.val[k] = w; w += kStaticSlots;
  }
  W.meta = w; w += kStaticSlots;
  W.rkey = w; w += kStaticSlots;
  W.rcnt = w; w += kStaticSlots;
  W.rcntd = w; w += kStaticSlots;
  W.nrun = w; w += kStaticSeg + 2;
  W.hist = w; w += kStaticHist;
  W.plist = w; w += kStaticPlist;
  W.scan_ws = w;
  return sort_core(ctx, p, Plist_host, W, nseg, pl_max);
}

}  // namespace lepic
