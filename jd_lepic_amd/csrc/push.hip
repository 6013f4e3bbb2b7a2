// push.hip -- particle mover (part_expmover.f90:1-517) fused with charge
// deposition (part_expmover.f90:520-610), FP64 and mixed precision.
//
// One block handles one segment of kChunk = 1024 slots.  Mixed precision:
// 256 threads, thread t owning the four consecutive slots 4t .. 4t+3 (16-byte
// vector loads and stores); FP64: 512 threads, thread t owning slots t and
// t + 512.  Survivors are compacted in particle order (part_expmover.f90:
// 495-503) by a block ballot scan into the same segment of the other particle
// buffer; the segments' contents in segment order are the species'
// particles, so no prefix across blocks is needed and blocks never wait for
// each other.  The mixed mover reads E, B and the cell wall flags from a
// field tile of 10 x 32 nodes staged in LDS around the segment's first cell
// (global memory for particles outside it).  Deposition accumulates in an
// LDS window of 16 rows x 64 columns of nodes around the segment's first
// cell and flushes the window to the species' global accumulator; nodes
// outside the window go to the accumulator directly.  Equal cells are
// merged first among a thread's four consecutive particles (register sums in
// select form) and then over runs of adjacent lanes (segmented DPP sums).
// Window and accumulator hold
// fixed-point integers (CIC weight x 2^kFixShift64 in FP64 contexts,
// x 2^kFixShiftMx in mixed ones): integer addition is associative, so the
// density is the same bit for bit whatever order the atomics land in, the
// same as a flush in a fixed order.  k_np_from_acc then scales the
// accumulator to np = Nm/(h1 h2 zmax) * kq * sum (charge_deposition,
// part_expmover.f90:555-584).  The FP64 path computes exactly the
// reference's expressions (no contraction: the library is built with
// -ffp-contract=off).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>

#include "lepic_internal.h"

using namespace lepic;

namespace {

// LDS deposition window of a segment: kWinRows x kWinCols nodes from
// kWinX0 columns left and kWinY0 rows below the segment's first cell (the
// particles of a segment sit in a few consecutive cells of one row when
// sorted, and drift a few cells between sorts); nodes outside go straight to
// the global density with FP64 atomics
constexpr int kWinRows = 16;
constexpr int kWinCols = 64;
constexpr int kWinX0 = 8;
constexpr int kWinY0 = 7;
// rows are padded by one word: with a stride of 64 the nodes of one column
// in different rows share an LDS bank, and the atomics of lanes in
// neighbouring rows serialise
constexpr int kWinStride = kWinCols + 1;
constexpr int kWin = kWinRows * kWinStride;
// Field tile of the mixed mover: E, B and the cell wall flags of kTileRows x
// kTileCols nodes staged in LDS per segment, from kTileX0 columns left and
// kTileY0 rows below the segment's first cell.  A sorted segment of 1024
// particles at 100 per cell spans about 10 cells of one row, and particles
// drift a few cells between sorts; particles whose cell leaves the tile
// gather from global memory instead.
constexpr int kTileRows = 10;
constexpr int kTileCols = 32;
constexpr int kTileX0 = 6;
constexpr int kTileY0 = 4;
constexpr int kFieldTile = kTileRows * kTileCols;
// Fixed-point units of the deposition: a particle's four CIC weights sum to
// 1.  FP64 contexts round each run sum to 2^-42 of a particle (an absolute
// error below 1.2e-13 of one particle's charge per run; a node holds up to
// 2^21 particles; the runs of an FP64 segment are fixed by the particle
// order alone).  Mixed contexts round every particle's weights to 2^-20
// before any sum (finer than FP32 accumulation once a node holds more than
// 8 particles), so no run, merge or atomic order changes a bit; an int32
// LDS window node holds up to 2^11 particles, more than a 1024-slot segment
// can bring.
constexpr int kFixShift64 = 42;
constexpr int kFixShiftMx = 20;
constexpr double kFix64 = 4398046511104.0;   // 2^42
constexpr float kFixMx = 1048576.0f;         // 2^20
// threads per mover block; a block owns one segment of kChunk slots
constexpr int kMoverThreadsMixed = kBlock;
constexpr int kMoverThreads64 = 2 * kBlock;   // two particles per thread: FP64 state of 4 would cost occupancy

struct MoverCommon {
  const int8_t* bc;
  const int8_t* cellwall;       // 1 where all four corners of cell (ix,iy) are walls
  int n1, n2;
  double h1, h2, dt, k1;        // k1 = dt*q/(2m)
  double Nm, mass, NmQ;         // NmQ = Nm*charge for cnt_dead
  int tagneg;
  int flagB, gridB, Bconst, magnetize;
  int nB1, nB2;
  double hB1, hB2, Bc0, Bc1, Bc2;
  int deposit;
  int check_dead;               // 0: no flag_dead entry is set, the push skips reading them
  unsigned long long* acc;      // fixed-point density accumulator (int64 per node)
  double k1dep;                 // Nm/(h1*h2*zmax)
  Tally* tally;
  double* sink;
  int ngrid;
  const int32_t* seg_in;        // particles per input segment
  int32_t* seg_out;             // survivors per output segment
  unsigned long long* count;    // species count, reduced by the lost particles
  int32_t* err;
};

struct Soa64 { const double *x, *y, *z, *vx, *vy, *vz; };
struct Soa64W { double *x, *y, *z, *vx, *vy, *vz; };
struct SoaMx { const int32_t* cell; const float *fx, *fy, *z, *vx, *vy, *vz; };
struct SoaMxW { int32_t* cell; float *fx, *fy, *z, *vx, *vy, *vz; };

__device__ __forceinline__ uint64_t lanemask_lt() {
  return (1ull << (threadIdx.x & 63)) - 1ull;
}

// ---- deposition helpers ----------------------------------------------------
__device__ __forceinline__ long long fix64(double w) { return __double2ll_rn(w * kFix64); }
__device__ __forceinline__ int fixmx(float w) { return __float2int_rn(w * kFixMx); }
// A mixed particle's four CIC weights (corners 00,10,11,01 of
// part_expmover.f90:571-584) from its in-cell offsets, each rounded to
// fixed point before any summation.
__device__ __forceinline__ void cic_fix(float fx, float fy, int& q0, int& q1, int& q2, int& q3) {
  const float px = 1.0f - fx, py = 1.0f - fy;
  q0 = fixmx(px * py);
  q1 = fixmx(fx * py);
  q2 = fixmx(fx * fy);
  q3 = fixmx(px * fy);
}
__device__ __forceinline__ void acc_add(unsigned long long* p, long long v) {
  atomicAdd(p, (unsigned long long)v);
}

// Deposit the four corner sums of FP64 cell (ix,iy) (corners 00,10,11,01 in
// the order of part_expmover.f90:571-584), as fixed-point integers.  kq = 2
// on non-interior nodes (main.f90:334-344) is applied by k_np_from_acc (a
// factor 2 is exact).
__device__ __forceinline__ void deposit_cell64(const MoverCommon& a, unsigned long long* win, int wx0, int wy0,
                                               int ix, int iy, const double w[4]) {
  const long long q[4] = {fix64(w[0]), fix64(w[1]), fix64(w[2]), fix64(w[3])};
  const int cx = ix - wx0, cy = iy - wy0;
  if (cx >= 0 && cx + 1 < kWinCols && cy >= 0 && cy + 1 < kWinRows) {
    atomicAdd(&win[cy * kWinStride + cx], (unsigned long long)q[0]);
    atomicAdd(&win[cy * kWinStride + cx + 1], (unsigned long long)q[1]);
    atomicAdd(&win[(cy + 1) * kWinStride + cx + 1], (unsigned long long)q[2]);
    atomicAdd(&win[(cy + 1) * kWinStride + cx], (unsigned long long)q[3]);
  } else {
    const int64_t n00 = node_index(ix, iy, a.n1), n01 = node_index(ix, iy + 1, a.n1);
    acc_add(&a.acc[n00], q[0]);
    acc_add(&a.acc[n00 + 1], q[1]);
    acc_add(&a.acc[n01 + 1], q[2]);
    acc_add(&a.acc[n01], q[3]);
  }
}

// The window's nonzero nodes into the global accumulator (int64 for FP64
// windows, int32 sign-extended for mixed ones).
template <typename T>
__device__ __forceinline__ void flush_window(const MoverCommon& a, const T* win, int wx0, int wy0) {
  for (int e = threadIdx.x; e < kWin; e += blockDim.x) {
    const long long v = (long long)win[e];
    if (v != 0) {
      const int ix = wx0 + e % kWinStride, iy = wy0 + e / kWinStride;
      acc_add(&a.acc[node_index(ix, iy, a.n1)], v);
    }
  }
}

__device__ __forceinline__ void clamp_cell(int& ix, int& iy, int n1, int n2) {
  // check_outofbounds, part_expmover.f90:612-627
  ix = ix < 0 ? 0 : (ix > n1 + 1 ? n1 + 1 : ix);
  iy = iy < 0 ? 0 : (iy > n2 + 1 ? n2 + 1 : iy);
}

// Lost-particle tallies, part_expmover.f90:314-332, 424-427.
__device__ inline void tally_lost(const MoverCommon& a, double xp, double yp, int ix, int iy,
                                  double Eki) {
  Tally* t = a.tally + blockIdx.x % kTallyCopies;
  atomic_add_f64(&t->P_loss[0], Eki);
  int ixw = (int)llround(xp / a.h1) + 1, iyw = (int)llround(yp / a.h2) + 1;
  clamp_cell(ixw, iyw, a.n1, a.n2);
  atomic_add_f64(&a.sink[node_index(ixw, iyw, a.n1)], 1.0);
  int igrid = a.bc[node_index(ix, iy, a.n1)];
  atomic_add_f64(&t->p_mac[0][igrid], 1.0);
  atomic_add_f64(&t->p_mac[1][igrid], Eki);
}

__device__ inline void tally_dead(const MoverCommon& a, double Eki) {
  Tally* t = a.tally + blockIdx.x % kTallyCopies;
  atomic_add_f64(&t->P_loss[2], -Eki);
  if (a.tagneg) atomic_add_f64(&t->cnt_dead, a.NmQ);
}

// Order-preserving offsets of the survivors inside the block's segment;
// records the segment's survivor count and takes the lost ones off the
// species count.  Returns each item's offset in off[].
template <int BT, int IT = kChunk / BT>
__device__ inline void compact_offsets(const MoverCommon& a, int64_t bid, int cnt_in, const bool surv[IT],
                                       uint32_t off[IT]) {
  __shared__ uint32_t s_cnt[IT][BT / 64];
  const int w = threadIdx.x >> 6;
  uint64_t mask[IT];
#pragma unroll
  for (int k = 0; k < IT; ++k) {
    mask[k] = __ballot(surv[k]);
    if ((threadIdx.x & 63) == 0) s_cnt[k][w] = (uint32_t)__popcll(mask[k]);
  }
  __syncthreads();
  uint32_t run = 0;
#pragma unroll
  for (int k = 0; k < IT; ++k) {
    uint32_t before = run;
    for (int v = 0; v < BT / 64; ++v) {
      if (v < w) before += s_cnt[k][v];
      run += s_cnt[k][v];
    }
    off[k] = before + (uint32_t)__popcll(mask[k] & lanemask_lt());
  }
  if (threadIdx.x == 0) {
    a.seg_out[bid] = (int32_t)run;
    const int lost = cnt_in - (int)run;
    if (lost > 0) atomicAdd(a.count, (unsigned long long)(-(long long)lost));
  }
}

// Offsets for thread-major items (thread t holds slots kItems*t + k): the
// same order-preserving compaction as compact_offsets.  Returns the number
// of survivors of the segment.
template <int BT>
__device__ inline uint32_t compact_offsets_tmajor(const MoverCommon& a, int64_t bid, int cnt_in,
                                                  const bool surv[kItems], uint32_t off[kItems]) {
  __shared__ uint32_t s_wave[BT / 64];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  uint32_t mine = 0;
#pragma unroll
  for (int k = 0; k < kItems; ++k) mine += surv[k] ? 1u : 0u;
  // inclusive scan of the per-thread counts over the wave
  uint32_t incl = mine;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const uint32_t o = __shfl_up(incl, d, 64);
    if (lane >= d) incl += o;
  }
  if (lane == 63) s_wave[w] = incl;
  __syncthreads();
  uint32_t before = incl - mine, total = 0;
#pragma unroll
  for (int v = 0; v < BT / 64; ++v) {
    const uint32_t c = s_wave[v];
    if (v < w) before += c;
    total += c;
  }
#pragma unroll
  for (int k = 0; k < kItems; ++k) {
    off[k] = before;
    before += surv[k] ? 1u : 0u;
  }
  if (threadIdx.x == 0) {
    a.seg_out[bid] = (int32_t)total;
    const int lost = cnt_in - (int)total;
    if (lost > 0) atomicAdd(a.count, (unsigned long long)(-(long long)lost));
  }
  return total;
}

// DPP lane move of a 32-bit integer (row shifts 0x111..0x118, row
// broadcasts 0x142/0x143 with their row masks)
template <int CTRL, int ROWS>
__device__ __forceinline__ int dpp_i32(int v) {
  return __builtin_amdgcn_update_dpp(0, v, CTRL, ROWS, 0xf, false);
}

// Segmented inclusive sum of w[0..3] over runs of equal key in adjacent
// lanes: afterwards the last lane of each run holds the run's sums.  s =
// first lane of this lane's run.  Every lane must be active.  The mixed
// deposition sums fixed-point integers, so the result does not depend on
// which particles share a run or in what order they arrive.
__device__ __forceinline__ void seg_sum4(int w[4], int s) {
  const int lane = threadIdx.x & 63;
  const bool a1 = lane - 1 >= s, a2 = lane - 2 >= s, a4 = lane - 4 >= s, a8 = lane - 8 >= s;
  const bool b15 = (lane & 16) && (lane & ~15) - 1 >= s;   // rows 1 and 3
  const bool b31 = lane >= 32 && s <= 31;                  // rows 2 and 3
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    int v = w[c], o;
    o = dpp_i32<0x111, 0xf>(v); if (a1) v += o;
    o = dpp_i32<0x112, 0xf>(v); if (a2) v += o;
    o = dpp_i32<0x114, 0xf>(v); if (a4) v += o;
    o = dpp_i32<0x118, 0xf>(v); if (a8) v += o;
    o = dpp_i32<0x142, 0xa>(v); if (b15) v += o;
    o = dpp_i32<0x143, 0xc>(v); if (b31) v += o;
    w[c] = v;
  }
}

// The same for FP64 weights, each moved as two 32-bit DPP halves.
template <int CTRL, int ROWS>
__device__ __forceinline__ double dpp_f64(double v) {
  const long long bits = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_update_dpp(0, (int)bits, CTRL, ROWS, 0xf, false);
  const int hi = __builtin_amdgcn_update_dpp(0, (int)(bits >> 32), CTRL, ROWS, 0xf, false);
  return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}

__device__ __forceinline__ void seg_sum4(double w[4], int s) {
  const int lane = threadIdx.x & 63;
  const bool a1 = lane - 1 >= s, a2 = lane - 2 >= s, a4 = lane - 4 >= s, a8 = lane - 8 >= s;
  const bool b15 = (lane & 16) && (lane & ~15) - 1 >= s;
  const bool b31 = lane >= 32 && s <= 31;
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    double v = w[c], o;
    o = dpp_f64<0x111, 0xf>(v); if (a1) v += o;
    o = dpp_f64<0x112, 0xf>(v); if (a2) v += o;
    o = dpp_f64<0x114, 0xf>(v); if (a4) v += o;
    o = dpp_f64<0x118, 0xf>(v); if (a8) v += o;
    o = dpp_f64<0x142, 0xa>(v); if (b15) v += o;
    o = dpp_f64<0x143, 0xc>(v); if (b31) v += o;
    w[c] = v;
  }
}

// FP64 deposition of one item slot: runs of equal node in adjacent lanes are
// summed by the segmented DPP scan and the run's last lane deposits them
// (waves whose particles all sit in different cells skip the scan).  Every
// lane of the wave must call it; ok marks a lane with a particle, key its
// node index.
__device__ __forceinline__ void deposit_runs64(const MoverCommon& a, unsigned long long* win, int wx0, int wy0, bool ok,
                                               int key, int ix, int iy, double w[4]) {
  const int lane = threadIdx.x & 63;
  const int kk = ok ? key : -2 - lane;
  const int kp = __shfl_up(kk, 1, 64);
  const uint64_t heads = __ballot(lane == 0 || kp != kk);
  const uint64_t act = __ballot(ok);
  const int s = 63 - __builtin_clzll(heads & ((2ull << lane) - 1ull));
  if (__popcll(heads & act) != __popcll(act)) seg_sum4(w, s);
  const bool tail = lane == 63 || ((heads >> (lane + 1)) & 1ull);
  if (tail && ok) deposit_cell64(a, win, wx0, wy0, ix, iy, w);
}

// Mixed weights of one cell straight to the global accumulator (nodes
// outside the LDS window).
__device__ __forceinline__ void deposit_global(const MoverCommon& a, int ix, int iy, int q0, int q1, int q2,
                                            int q3) {
  const int64_t n00 = node_index(ix, iy, a.n1), n01 = node_index(ix, iy + 1, a.n1);
  acc_add(&a.acc[n00], q0);
  acc_add(&a.acc[n00 + 1], q1);
  acc_add(&a.acc[n01 + 1], q2);
  acc_add(&a.acc[n01], q3);
}

// Window keys of a thread's items: key[k] = window-local cell (row stride
// kWinStride) or -1; cells whose nodes leave the window get gkey[k] = packed
// cell (iy<<16)|ix, else -1.
__device__ __forceinline__ void window_keys(const MoverCommon& a, int wx0, int wy0, const int C[kItems],
                                            const bool ok[kItems], int key[kItems], int gkey[kItems]) {
#pragma unroll
  for (int k = 0; k < kItems; ++k) {
    int ix = C[k] & 0xFFFF, iy = C[k] >> 16;
    clamp_cell(ix, iy, a.n1, a.n2);
    const int cx = ix - wx0, cy = iy - wy0;
    const bool in = (unsigned)cx < (unsigned)(kWinCols - 1) && (unsigned)cy < (unsigned)(kWinRows - 1);
    key[k] = ok[k] && in ? cy * kWinStride + cx : -1;
    gkey[k] = ok[k] && !in ? (iy << 16) | ix : -1;
  }
}

// CIC deposition of a thread's four consecutive items into the LDS window
// while the segment's particles are in cell order.  Runs of equal cell
// within the thread are summed in registers in select form (no branches per
// item); a run that a cell change ends inside the thread is added by its
// lane directly (rare: a few lanes per wave), and the run that ends the
// thread is summed over runs of equal cell in adjacent lanes (segmented DPP
// sums) and added once per run.  key[k]: window-local cell or -1.
__device__ __forceinline__ void deposit_window_merged(int* win, const int key[kItems], const float FX[kItems],
                                                      const float FY[kItems]) {
  const int lane = threadIdx.x & 63;
  int cur = -1;
  int acc[4] = {0, 0, 0, 0};
  int ek[kItems];          // the run that ends just before item k (k >= 1), or -1
  int ew[kItems][4];
#pragma unroll
  for (int k = 0; k < kItems; ++k) {
    int q[4];
    cic_fix(FX[k], FY[k], q[0], q[1], q[2], q[3]);
    const bool in = key[k] >= 0;
    const bool start = in && key[k] != cur;
    ek[k] = start ? cur : -1;
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      ew[k][c] = acc[c];
      const int qc = in ? q[c] : 0;
      acc[c] = start ? qc : acc[c] + qc;
    }
    cur = start ? key[k] : cur;
  }
#pragma unroll
  for (int k = 1; k < kItems; ++k) {
    if (__ballot(ek[k] >= 0) == 0) continue;      // wave-uniform
    if (ek[k] >= 0) {
      const int kk = ek[k];
      atomicAdd(&win[kk], ew[k][0]);
      atomicAdd(&win[kk + 1], ew[k][1]);
      atomicAdd(&win[kk + kWinStride + 1], ew[k][2]);
      atomicAdd(&win[kk + kWinStride], ew[k][3]);
    }
  }
  const uint64_t act = __ballot(cur >= 0);
  if (act == 0) return;                           // wave-uniform
  // lanes without a run take a key no active lane has, distinct from neighbours
  const int kk = cur >= 0 ? cur : -2 - lane;
  const int kp = __shfl_up(kk, 1, 64);
  const uint64_t heads = __ballot(lane == 0 || kp != kk);
  if (__popcll(heads & act) != __popcll(act)) {
    const int s = 63 - __builtin_clzll(heads & ((2ull << lane) - 1ull));
    seg_sum4(acc, s);
  }
  const bool tail = lane == 63 || ((heads >> (lane + 1)) & 1ull);
  if (tail && cur >= 0) {
    atomicAdd(&win[cur], acc[0]);
    atomicAdd(&win[cur + 1], acc[1]);
    atomicAdd(&win[cur + kWinStride + 1], acc[2]);
    atomicAdd(&win[cur + kWinStride], acc[3]);
  }
}

// Items whose nodes leave the LDS window (a segment that continues on the
// next row after the end of one, or particles far from their segment):
// merged within a thread, then over runs of equal cell in adjacent lanes,
// and added to the global density per run.  gk: packed cell (iy<<16)|ix or
// -1.  Skipped by waves that have none.
__device__ __forceinline__ void deposit_global_runs(const MoverCommon& a, const int gk[kItems],
                                                    const float FX[kItems], const float FY[kItems]) {
  bool any = false;
#pragma unroll
  for (int k = 0; k < kItems; ++k) any = any || gk[k] >= 0;
  if (__ballot(any) == 0) return;                 // wave-uniform
  const int lane = threadIdx.x & 63;
  int key[kItems];
  int w[kItems][4];
  int prev = -1;
#pragma unroll
  for (int k = 0; k < kItems; ++k) {
    key[k] = -1;
    w[k][0] = w[k][1] = w[k][2] = w[k][3] = 0;
    if (gk[k] < 0) continue;
    int q0, q1, q2, q3;
    cic_fix(FX[k], FY[k], q0, q1, q2, q3);
    if (gk[k] == prev) {
#pragma unroll
      for (int j = k - 1; j >= 0; --j)
        if (key[j] == gk[k]) { w[j][0] += q0; w[j][1] += q1; w[j][2] += q2; w[j][3] += q3; break; }
    } else {
      key[k] = gk[k];
      w[k][0] = q0; w[k][1] = q1; w[k][2] = q2; w[k][3] = q3;
      prev = gk[k];
    }
  }
#pragma unroll
  for (int k = 0; k < kItems; ++k) {
    if (__ballot(key[k] >= 0) == 0) continue;     // wave-uniform
    const int kk = key[k] >= 0 ? key[k] : -2 - lane;
    const int kp = __shfl_up(kk, 1, 64);
    const uint64_t heads = __ballot(lane == 0 || kp != kk);
    const int s = 63 - __builtin_clzll(heads & ((2ull << lane) - 1ull));
    int v[4] = {w[k][0], w[k][1], w[k][2], w[k][3]};
    seg_sum4(v, s);
    const bool tail = lane == 63 || ((heads >> (lane + 1)) & 1ull);
    if (tail && key[k] >= 0) deposit_global(a, key[k] & 0xFFFF, key[k] >> 16, v[0], v[1], v[2], v[3]);
  }
}

// ---------------------------------------------------------------------------
// FP64 mover
// ---------------------------------------------------------------------------
template <int BT>
__global__ __launch_bounds__(BT) void k_mover64(MoverCommon a, Soa64 in, Soa64W out,
                                                    int8_t* flag_dead, const double* __restrict__ E,
                                                    const double* __restrict__ B) {
  constexpr int IT = kChunk / BT;   // particles per thread
  __shared__ unsigned long long win[kWin];
  __shared__ int s_wx0, s_wy0;
  const int64_t bid = blockIdx.x;
  const int cnt_in = a.seg_in[bid];
  const int64_t base = bid * (int64_t)(BT * IT);
  const int t = threadIdx.x;
  const int n1 = a.n1, n2 = a.n2;
  const double h1 = a.h1, h2 = a.h2, dt = a.dt, k1 = a.k1;

  if (a.deposit) {
    for (int e = t; e < kWin; e += BT) win[e] = 0ull;
    if (t == 0) {
      int ix0 = 1, iy0 = 1;
      if (cnt_in > 0) { ix0 = (int)(in.x[base] / h1) + 1; iy0 = (int)(in.y[base] / h2) + 1; }
      s_wx0 = ix0 - kWinX0 < 0 ? 0 : ix0 - kWinX0;
      s_wy0 = iy0 - kWinY0 < 0 ? 0 : iy0 - kWinY0;
    }
  }

  double X[IT], Y[IT], Z[IT], VX[IT], VY[IT], VZ[IT];
  bool surv[IT];
#pragma unroll
  for (int k = 0; k < IT; ++k) {
    const int64_t i = base + (int64_t)k * BT + t;
    surv[k] = false;
    if (k * BT + t >= cnt_in) continue;
    double xp = in.x[i], yp = in.y[i], zp = in.z[i];
    double vx = in.vx[i], vy = in.vy[i], vz = in.vz[i];
    if (a.check_dead && flag_dead[i] == 1) {  // part_expmover.f90:84-98
      flag_dead[i] = 0;
      tally_dead(a, 0.5 * a.Nm * a.mass * (vx * vx + vy * vy + vz * vz));
      continue;
    }
    int ix = (int)(xp / h1) + 1;  // INT(), lines 103-104
    int iy = (int)(yp / h2) + 1;
    clamp_cell(ix, iy, n1, n2);   // identity for positions inside the box
    double px = (ix * h1 - xp) / h1;
    double py = (iy * h2 - yp) / h2;
    double k0 = px * py, kk1 = (1.0 - px) * py, kk2 = (1.0 - px) * (1.0 - py), kk3 = px * (1.0 - py);
    const int64_t n00 = node_index(ix, iy, n1), n01 = node_index(ix, iy + 1, n1);
    const double2 e00 = *(const double2*)(E + 2 * n00), e10 = *(const double2*)(E + 2 * (n00 + 1));
    const double2 e11 = *(const double2*)(E + 2 * (n01 + 1)), e01 = *(const double2*)(E + 2 * n01);
    double Exp = k0 * e00.x + kk1 * e10.x + kk2 * e11.x + kk3 * e01.x;
    double Eyp = k0 * e00.y + kk1 * e10.y + kk2 * e11.y + kk3 * e01.y;
    bool boris = false;
    double Bx = 0, By = 0, Bz = 0;
    if (a.flagB) {  // lines 131-182
      if (a.Bconst) {
        Bx = a.Bc0; By = a.Bc1; Bz = a.Bc2;
        boris = true;
      } else {
        int bx = ix, by = iy;
        if (a.gridB) {
          bx = (int)(xp / a.hB1) + 1; by = (int)(yp / a.hB2) + 1;
          bx = bx < 0 ? 0 : (bx > a.nB1 + 1 ? a.nB1 + 1 : bx);
          by = by < 0 ? 0 : (by > a.nB2 + 1 ? a.nB2 + 1 : by);
        }
        const int64_t bs = a.nB1 + 3;
        const double* B00 = B + 4 * ((int64_t)by * bs + bx);
        if (B00[3] > 1.0e-5) {
          if (a.gridB) {
            px = (bx * a.hB1 - xp) / a.hB1;
            py = (by * a.hB2 - yp) / a.hB2;
            k0 = px * py; kk1 = (1.0 - px) * py; kk2 = (1.0 - px) * (1.0 - py); kk3 = px * (1.0 - py);
          }
          const double* B10 = B00 + 4;
          const double* B01 = B + 4 * ((int64_t)(by + 1) * bs + bx);
          const double* B11 = B01 + 4;
          Bx = k0 * B00[0] + kk1 * B10[0] + kk2 * B11[0] + kk3 * B01[0];
          By = k0 * B00[1] + kk1 * B10[1] + kk2 * B11[1] + kk3 * B01[1];
          Bz = k0 * B00[2] + kk1 * B10[2] + kk2 * B11[2] + kk3 * B01[2];
          boris = true;
        }
      }
    }
    if (!a.magnetize) boris = false;  // line 185
    if (boris) {                       // lines 187-206
      double Bt = sqrt(Bx * Bx + By * By + Bz * Bz);
      double k2 = k1 * (2.0 / (1.0 + (k1 * Bt) * (k1 * Bt)));
      double v1x = vx + k1 * Exp, v1y = vy + k1 * Eyp, v1z = vz;
      double v3x = v1x + k1 * (v1y * Bz - v1z * By);
      double v3y = v1y + k1 * (v1z * Bx - v1x * Bz);
      double v3z = v1z + k1 * (v1x * By - v1y * Bx);
      double v2x = v1x + k2 * (v3y * Bz - v3z * By);
      double v2y = v1y + k2 * (v3z * Bx - v3x * Bz);
      double v2z = v1z + k2 * (v3x * By - v3y * Bx);
      vx = v2x + k1 * Exp;
      vy = v2y + k1 * Eyp;
      vz = v2z;
    } else {  // lines 210-211
      vx = vx + 2.0 * k1 * Exp;
      vy = vy + 2.0 * k1 * Eyp;
    }
    xp = xp + dt * vx;  // lines 216-219
    yp = yp + dt * vy;
    zp = zp + dt * vz;
    ix = (int)floor(xp / h1) + 1;  // FLOOR(), lines 244-245
    iy = (int)floor(yp / h2) + 1;
    clamp_cell(ix, iy, n1, n2);
    if (a.cellwall[node_index(ix, iy, n1)]) {  // lines 254-258, 314-446
      tally_lost(a, xp, yp, ix, iy, 0.5 * a.Nm * a.mass * (vx * vx + vy * vy + vz * vz));
      continue;
    }
    X[k] = xp; Y[k] = yp; Z[k] = zp; VX[k] = vx; VY[k] = vy; VZ[k] = vz;
    surv[k] = true;
  }

  uint32_t off[IT];
  compact_offsets<BT>(a, bid, cnt_in, surv, off);  // contains __syncthreads
  const int64_t gbase = base;
#pragma unroll
  for (int k = 0; k < IT; ++k) {
    if (!surv[k]) continue;
    const int64_t o = gbase + off[k];
    out.x[o] = X[k]; out.y[o] = Y[k]; out.z[o] = Z[k];
    out.vx[o] = VX[k]; out.vy[o] = VY[k]; out.vz[o] = VZ[k];
  }
  if (!a.deposit) return;
  const int wx0 = s_wx0, wy0 = s_wy0;
#pragma unroll
  for (int k = 0; k < IT; ++k) {
    // charge_deposition, part_expmover.f90:555-584, at the new position
    int64_t key = -2 - (int64_t)t;
    double w[4] = {0, 0, 0, 0};
    int ix = 0, iy = 0;
    if (surv[k]) {
      ix = (int)(X[k] / h1) + 1;
      iy = (int)(Y[k] / h2) + 1;
      double px = (ix * h1 - X[k]) / h1;
      double py = (iy * h2 - Y[k]) / h2;
      w[0] = px * py;
      w[1] = (1.0 - px) * py;
      w[2] = (1.0 - px) * (1.0 - py);
      w[3] = px * (1.0 - py);
      key = node_index(ix, iy, n1);
    }
    deposit_runs64(a, win, wx0, wy0, surv[k], (int)key, ix, iy, w);
  }
  __syncthreads();
  flush_window(a, win, wx0, wy0);
}

// ---------------------------------------------------------------------------
// Mixed-precision mover: int32 cell + FP32 in-cell offset, FP32 velocity,
// FP32 E/B gathers; deposition of fixed-point weights (int32 run sums) into
// an int32 LDS window, flushed to the int64 accumulator (scaled to the FP64
// density by k_np_from_acc with the FP64 factor Nm/(h1 h2 zmax)).
// ---------------------------------------------------------------------------
struct MixedParams {
  float k1f, dtoh1, dtoh2, dtf;
  float hB1r, hB2r;   // h/hB ratios for a B map on another grid
};

// at most 80 VGPRs (6 waves per SIMD instead of 5): the electrons' global
// gathers and atomics late in a sort cycle wait less (measured 1.09 -> 1.06 ms
// per 1024^2 electron push, ions 0.93 -> 0.92 ms; 8 waves would spill)
template <int BT>
__global__ __launch_bounds__(BT, 6) void k_mover_mixed(MoverCommon a, MixedParams m, SoaMx in,
                                                        SoaMxW out, int8_t* flag_dead,
                                                        const float* __restrict__ E,
                                                        const float* __restrict__ B) {
  __shared__ int win[kWin];
  __shared__ float2 sE[kFieldTile];
  __shared__ float4 sB[kFieldTile];
  __shared__ int8_t sW[kFieldTile];
  const int64_t bid = blockIdx.x;
  const int cnt_in = a.seg_in[bid];
  const int64_t base = bid * (int64_t)(BT * kItems);
  const int t = threadIdx.x;
  const int n1 = a.n1, n2 = a.n2;
  const float k1 = m.k1f;
  // window anchors from the segment's first cell, computed by every thread
  const int c0 = cnt_in > 0 ? in.cell[base] : ((1 << 16) | 1);
  const int wx0 = max((c0 & 0xFFFF) - kWinX0, 0), wy0 = max((c0 >> 16) - kWinY0, 0);
  const int fx0 = max(min((c0 & 0xFFFF) - kTileX0, n1 + 3 - kTileCols), 0);
  const int fy0 = max(min((c0 >> 16) - kTileY0, n2 + 3 - kTileRows), 0);
  const bool tileB = a.flagB && !a.Bconst && !a.gridB;
  if (a.deposit)
    for (int e = t; e < kWin; e += BT) win[e] = 0;
  // thread t owns the kItems consecutive slots i0 .. i0+3 of the segment;
  // the arrays are padded to whole segments, so the 16-byte loads stay inside
  // them also past the segment's count
  const int64_t i0 = base + (int64_t)kItems * t;
  const int nv = min(max(cnt_in - kItems * t, 0), kItems);
  int C[kItems] = {};
  float FX[kItems] = {}, FY[kItems] = {}, Z[kItems] = {}, VX[kItems] = {}, VY[kItems] = {}, VZ[kItems] = {};
  bool surv[kItems];
  uint32_t dead4 = 0;
  if (nv > 0) {
    const int4 c4 = *(const int4*)(in.cell + i0);
    const float4 fx4 = *(const float4*)(in.fx + i0), fy4 = *(const float4*)(in.fy + i0);
    const float4 vx4 = *(const float4*)(in.vx + i0), vy4 = *(const float4*)(in.vy + i0);
    const float4 vz4 = *(const float4*)(in.vz + i0);
    const float4 z4 = in.z ? *(const float4*)(in.z + i0) : make_float4(0.f, 0.f, 0.f, 0.f);
    dead4 = a.check_dead ? *(const uint32_t*)(flag_dead + i0) : 0u;
    C[0] = c4.x; C[1] = c4.y; C[2] = c4.z; C[3] = c4.w;
    FX[0] = fx4.x; FX[1] = fx4.y; FX[2] = fx4.z; FX[3] = fx4.w;
    FY[0] = fy4.x; FY[1] = fy4.y; FY[2] = fy4.z; FY[3] = fy4.w;
    VX[0] = vx4.x; VX[1] = vx4.y; VX[2] = vx4.z; VX[3] = vx4.w;
    VY[0] = vy4.x; VY[1] = vy4.y; VY[2] = vy4.z; VY[3] = vy4.w;
    VZ[0] = vz4.x; VZ[1] = vz4.y; VZ[2] = vz4.z; VZ[3] = vz4.w;
    Z[0] = z4.x; Z[1] = z4.y; Z[2] = z4.z; Z[3] = z4.w;
  }
  // the field tile (nodes beyond the grid of a small domain are never read)
  for (int e = t; e < kFieldTile; e += BT) {
    const int gx = fx0 + e % kTileCols, gy = fy0 + e / kTileCols;
    float2 ev = make_float2(0.f, 0.f);
    float4 bv = make_float4(0.f, 0.f, 0.f, 0.f);
    int8_t wv = 1;
    if (gx <= n1 + 2 && gy <= n2 + 2) {
      const int64_t nd = node_index(gx, gy, n1);
      ev = *(const float2*)(E + 2 * nd);
      if (tileB) bv = *(const float4*)(B + 4 * nd);
      wv = a.cellwall[nd];
    }
    sE[e] = ev;
    sB[e] = bv;
    sW[e] = wv;
  }
  __syncthreads();
  // particles flagged dead by the MCC leave before the push
  // (part_expmover.f90:84-98); rare, so one branch per thread
  uint32_t deadm = 0;
  if (dead4 != 0) {
#pragma unroll
    for (int k = 0; k < kItems; ++k)
      if (k < nv && ((dead4 >> (8 * k)) & 0xFF) == 1) {
        flag_dead[i0 + k] = 0;
        tally_dead(a, 0.5 * a.Nm * a.mass *
                          ((double)VX[k] * VX[k] + (double)VY[k] * VY[k] + (double)VZ[k] * VZ[k]));
        deadm |= 1u << k;
      }
  }
  // the push as straight-line code for every item slot; slots without a live
  // particle are pushed from a harmless state and dropped
  uint32_t lostm = 0;
  int IXU[kItems], IYU[kItems];
#pragma unroll
  for (int k = 0; k < kItems; ++k) {
    const bool live = k < nv && !((deadm >> k) & 1u);
    float fx = live ? FX[k] : 0.5f, fy = live ? FY[k] : 0.5f;
    float vx = live ? VX[k] : 0.f, vy = live ? VY[k] : 0.f, vz = live ? VZ[k] : 0.f, zp = live ? Z[k] : 0.f;
    int ix = live ? (C[k] & 0xFFFF) : 2, iy = live ? (C[k] >> 16) : 2;
    clamp_cell(ix, iy, n1, n2);
    const float px = 1.0f - fx, py = 1.0f - fy;
    const float w0 = px * py, w1 = fx * py, w2 = fx * fy, w3 = px * fy;
    // the cell's four nodes from the tile when it holds them
    const int tx = ix - fx0, ty = iy - fy0;
    const bool in_tile = (unsigned)tx < (unsigned)(kTileCols - 1) && (unsigned)ty < (unsigned)(kTileRows - 1);
    const int q = ty * kTileCols + tx;
    float2 e00, e10, e11, e01;
    if (in_tile) {
      e00 = sE[q]; e10 = sE[q + 1]; e11 = sE[q + kTileCols + 1]; e01 = sE[q + kTileCols];
    } else {
      const int64_t n00 = node_index(ix, iy, n1), n01 = node_index(ix, iy + 1, n1);
      e00 = *(const float2*)(E + 2 * n00); e10 = *(const float2*)(E + 2 * (n00 + 1));
      e11 = *(const float2*)(E + 2 * (n01 + 1)); e01 = *(const float2*)(E + 2 * n01);
    }
    float Ex = fmaf(w3, e01.x, fmaf(w2, e11.x, fmaf(w1, e10.x, w0 * e00.x)));
    float Ey = fmaf(w3, e01.y, fmaf(w2, e11.y, fmaf(w1, e10.y, w0 * e00.y)));
    bool boris = false;
    float Bx = 0.f, By = 0.f, Bz = 0.f;
    if (a.flagB) {
      if (a.Bconst) {
        Bx = (float)a.Bc0; By = (float)a.Bc1; Bz = (float)a.Bc2;
        boris = true;
      } else if (tileB && in_tile) {
        const float4 B00 = sB[q], B10 = sB[q + 1], B11 = sB[q + kTileCols + 1], B01 = sB[q + kTileCols];
        if (B00.w > 1.0e-5f) {
          Bx = fmaf(w3, B01.x, fmaf(w2, B11.x, fmaf(w1, B10.x, w0 * B00.x)));
          By = fmaf(w3, B01.y, fmaf(w2, B11.y, fmaf(w1, B10.y, w0 * B00.y)));
          Bz = fmaf(w3, B01.z, fmaf(w2, B11.z, fmaf(w1, B10.z, w0 * B00.z)));
          boris = true;
        }
      } else {
        int bx = ix, by = iy;
        float b0 = w0, b1 = w1, b2 = w2, b3 = w3;
        if (a.gridB) {
          float gx = ((float)(ix - 1) + fx) * m.hB1r, gy = ((float)(iy - 1) + fy) * m.hB2r;
          float fbx = floorf(gx), fby = floorf(gy);
          bx = (int)fbx + 1; by = (int)fby + 1;
          bx = bx < 0 ? 0 : (bx > a.nB1 + 1 ? a.nB1 + 1 : bx);
          by = by < 0 ? 0 : (by > a.nB2 + 1 ? a.nB2 + 1 : by);
          float ox = gx - fbx, oy = gy - fby, qx = 1.0f - ox, qy = 1.0f - oy;
          b0 = qx * qy; b1 = ox * qy; b2 = ox * oy; b3 = qx * oy;
        }
        const int64_t bs = a.nB1 + 3;
        const float4 B00 = *(const float4*)(B + 4 * ((int64_t)by * bs + bx));
        if (B00.w > 1.0e-5f) {
          const float4 B10 = *(const float4*)(B + 4 * ((int64_t)by * bs + bx + 1));
          const float4 B01 = *(const float4*)(B + 4 * ((int64_t)(by + 1) * bs + bx));
          const float4 B11 = *(const float4*)(B + 4 * ((int64_t)(by + 1) * bs + bx + 1));
          Bx = fmaf(b3, B01.x, fmaf(b2, B11.x, fmaf(b1, B10.x, b0 * B00.x)));
          By = fmaf(b3, B01.y, fmaf(b2, B11.y, fmaf(b1, B10.y, b0 * B00.y)));
          Bz = fmaf(b3, B01.z, fmaf(b2, B11.z, fmaf(b1, B10.z, b0 * B00.z)));
          boris = true;
        }
      }
    }
    if (!a.magnetize) boris = false;
    if (boris) {
      float Bt2 = Bx * Bx + By * By + Bz * Bz;
      float k2 = (2.0f * k1) * __builtin_amdgcn_rcpf(1.0f + k1 * k1 * Bt2);
      float v1x = fmaf(k1, Ex, vx), v1y = fmaf(k1, Ey, vy), v1z = vz;
      float v3x = fmaf(k1, v1y * Bz - v1z * By, v1x);
      float v3y = fmaf(k1, v1z * Bx - v1x * Bz, v1y);
      float v3z = fmaf(k1, v1x * By - v1y * Bx, v1z);
      vx = fmaf(k1, Ex, fmaf(k2, v3y * Bz - v3z * By, v1x));
      vy = fmaf(k1, Ey, fmaf(k2, v3z * Bx - v3x * Bz, v1y));
      vz = fmaf(k2, v3x * By - v3y * Bx, v1z);
    } else {
      vx = fmaf(2.0f * k1, Ex, vx);
      vy = fmaf(2.0f * k1, Ey, vy);
    }
    fx = fmaf(m.dtoh1, vx, fx);
    fy = fmaf(m.dtoh2, vy, fy);
    zp = fmaf(m.dtf, vz, zp);
    float sx = floorf(fx), sy = floorf(fy);
    fx -= sx; fy -= sy;
    if (fx >= 1.0f) { fx -= 1.0f; sx += 1.0f; }
    if (fy >= 1.0f) { fy -= 1.0f; sy += 1.0f; }
    // the shift is bounded by the CFL of the run; clamp before int conversion
    sx = fminf(fmaxf(sx, -4096.f), 4096.f);
    sy = fminf(fmaxf(sy, -4096.f), 4096.f);
    ix += (int)sx;
    iy += (int)sy;
    int cx = ix, cy = iy;
    clamp_cell(cx, cy, n1, n2);
    // part_expmover.f90:254-258, from the tile's wall flags when it holds the cell
    const int wx = cx - fx0, wy = cy - fy0;
    const bool w_tile = (unsigned)wx < (unsigned)kTileCols && (unsigned)wy < (unsigned)kTileRows;
    const int8_t wall = w_tile ? sW[wy * kTileCols + wx] : a.cellwall[node_index(cx, cy, n1)];
    const bool lost = live && wall != 0;
    lostm |= lost ? 1u << k : 0u;
    IXU[k] = ix; IYU[k] = iy;
    C[k] = (iy << 16) | ix;
    FX[k] = fx; FY[k] = fy; Z[k] = zp; VX[k] = vx; VY[k] = vy; VZ[k] = vz;
    surv[k] = live && !lost;
  }
  if (lostm) {
#pragma unroll
    for (int k = 0; k < kItems; ++k)
      if ((lostm >> k) & 1u) {
        int cx = IXU[k], cy = IYU[k];
        clamp_cell(cx, cy, n1, n2);
        const double xd = ((double)(IXU[k] - 1) + FX[k]) * a.h1, yd = ((double)(IYU[k] - 1) + FY[k]) * a.h2;
        tally_lost(a, xd, yd, cx, cy,
                   0.5 * a.Nm * a.mass * ((double)VX[k] * VX[k] + (double)VY[k] * VY[k] + (double)VZ[k] * VZ[k]));
      }
  }
  uint32_t off[kItems];
  const uint32_t kept = compact_offsets_tmajor<BT>(a, bid, cnt_in, surv, off);
  if (kept == (uint32_t)cnt_in && nv == kItems) {
    // nothing of the segment was lost: every survivor keeps its slot
    *(int4*)(out.cell + i0) = make_int4(C[0], C[1], C[2], C[3]);
    *(float4*)(out.fx + i0) = make_float4(FX[0], FX[1], FX[2], FX[3]);
    *(float4*)(out.fy + i0) = make_float4(FY[0], FY[1], FY[2], FY[3]);
    if (out.z) *(float4*)(out.z + i0) = make_float4(Z[0], Z[1], Z[2], Z[3]);
    *(float4*)(out.vx + i0) = make_float4(VX[0], VX[1], VX[2], VX[3]);
    *(float4*)(out.vy + i0) = make_float4(VY[0], VY[1], VY[2], VY[3]);
    *(float4*)(out.vz + i0) = make_float4(VZ[0], VZ[1], VZ[2], VZ[3]);
  } else {
#pragma unroll
    for (int k = 0; k < kItems; ++k) {
      if (!surv[k]) continue;
      const int64_t o = base + off[k];
      out.cell[o] = C[k]; out.fx[o] = FX[k]; out.fy[o] = FY[k];
      if (out.z) out.z[o] = Z[k];
      out.vx[o] = VX[k]; out.vy[o] = VY[k]; out.vz[o] = VZ[k];
    }
  }
  if (!a.deposit) return;
  // the deposition at the new positions: the window's cells merged in
  // registers and over adjacent lanes, cells leaving the window to the
  // global accumulator
  int dkey[kItems], gkey[kItems];
  window_keys(a, wx0, wy0, C, surv, dkey, gkey);
  deposit_global_runs(a, gkey, FX, FY);
  deposit_window_merged(win, dkey, FX, FY);
  __syncthreads();
  flush_window(a, win, wx0, wy0);
}

// ---------------------------------------------------------------------------
// Deposition only (charge_deposition without a push)
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kBlock) void k_deposit64(MoverCommon a, Soa64 in) {
  __shared__ unsigned long long win[kWin];
  __shared__ int s_wx0, s_wy0;
  const int cnt_in = a.seg_in[blockIdx.x];
  const int64_t base = (int64_t)blockIdx.x * kChunk;
  const int t = threadIdx.x;
  for (int e = t; e < kWin; e += kBlock) win[e] = 0ull;
  if (t == 0) {
    int ix0 = 1, iy0 = 1;
    if (cnt_in > 0) { ix0 = (int)(in.x[base] / a.h1) + 1; iy0 = (int)(in.y[base] / a.h2) + 1; }
    s_wx0 = ix0 - kWinX0 < 0 ? 0 : ix0 - kWinX0;
    s_wy0 = iy0 - kWinY0 < 0 ? 0 : iy0 - kWinY0;
  }
  __syncthreads();
  const int wx0 = s_wx0, wy0 = s_wy0;
#pragma unroll
  for (int k = 0; k < kItems; ++k) {
    const int64_t i = base + (int64_t)k * kBlock + t;
    int64_t key = -2 - (int64_t)t;
    double w[4] = {0, 0, 0, 0};
    int ix = 0, iy = 0;
    bool ok = k * kBlock + t < cnt_in;
    if (ok) {
      double xp = in.x[i], yp = in.y[i];
      ix = (int)(xp / a.h1) + 1;
      iy = (int)(yp / a.h2) + 1;
      clamp_cell(ix, iy, a.n1, a.n2);
      double px = (ix * a.h1 - xp) / a.h1;
      double py = (iy * a.h2 - yp) / a.h2;
      w[0] = px * py;
      w[1] = (1.0 - px) * py;
      w[2] = (1.0 - px) * (1.0 - py);
      w[3] = px * (1.0 - py);
      key = node_index(ix, iy, a.n1);
    }
    deposit_runs64(a, win, wx0, wy0, ok, (int)key, ix, iy, w);
  }
  __syncthreads();
  flush_window(a, win, wx0, wy0);
}

__global__ __launch_bounds__(kBlock) void k_deposit_mixed(MoverCommon a, SoaMx in) {
  __shared__ int win[kWin];
  __shared__ int s_wx0, s_wy0;
  const int cnt_in = a.seg_in[blockIdx.x];
  const int64_t base = (int64_t)blockIdx.x * kChunk;
  const int t = threadIdx.x;
  for (int e = t; e < kWin; e += kBlock) win[e] = 0;
  if (t == 0) {
    int c0 = cnt_in > 0 ? in.cell[base] : ((1 << 16) | 1);
    s_wx0 = (c0 & 0xFFFF) - kWinX0 < 0 ? 0 : (c0 & 0xFFFF) - kWinX0;
    s_wy0 = (c0 >> 16) - kWinY0 < 0 ? 0 : (c0 >> 16) - kWinY0;
  }
  const int64_t i0 = base + (int64_t)kItems * t;
  const int nv = min(max(cnt_in - kItems * t, 0), kItems);
  int C[kItems] = {0, 0, 0, 0};
  float FX[kItems] = {0.f, 0.f, 0.f, 0.f}, FY[kItems] = {0.f, 0.f, 0.f, 0.f};
  bool ok[kItems];
  if (nv > 0) {
    const int4 c4 = *(const int4*)(in.cell + i0);
    const float4 fx4 = *(const float4*)(in.fx + i0), fy4 = *(const float4*)(in.fy + i0);
    C[0] = c4.x; C[1] = c4.y; C[2] = c4.z; C[3] = c4.w;
    FX[0] = fx4.x; FX[1] = fx4.y; FX[2] = fx4.z; FX[3] = fx4.w;
    FY[0] = fy4.x; FY[1] = fy4.y; FY[2] = fy4.z; FY[3] = fy4.w;
  }
#pragma unroll
  for (int k = 0; k < kItems; ++k) ok[k] = k < nv;
  __syncthreads();
  int key[kItems], gkey[kItems];
  window_keys(a, s_wx0, s_wy0, C, ok, key, gkey);
  deposit_global_runs(a, gkey, FX, FY);
  deposit_window_merged(win, key, FX, FY);
  __syncthreads();
  flush_window(a, win, s_wx0, s_wy0);
}

// np = Nm/(h1 h2 zmax) * kq * accumulator (kq = 2 on non-interior nodes,
// main.f90:334-344); scale = Nm/(h1 h2 zmax) * 2^-shift.
__global__ __launch_bounds__(256) void k_np_from_acc(const unsigned long long* __restrict__ acc,
                                                     const int8_t* __restrict__ bc, double* __restrict__ np,
                                                     int64_t n, double scale) {
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) {
    const double kq = bc[i] != -1 ? 2.0 : 1.0;
    np[i] = kq * ((double)(long long)acc[i] * scale);
  }
}

MoverCommon common_args(lepic_ctx* ctx, int p) {
  Species& s = ctx->sp[p];
  MoverCommon a{};
  a.bc = ctx->bc;
  a.cellwall = ctx->d_cellwall;
  a.n1 = ctx->g.n1; a.n2 = ctx->g.n2;
  a.h1 = ctx->g.h1; a.h2 = ctx->g.h2; a.dt = ctx->g.dt;
  a.k1 = ctx->g.dt * s.charge / (2.0 * s.mass);  // part_expmover.f90:61
  a.Nm = s.Nm; a.mass = s.mass; a.NmQ = s.Nm * s.charge;
  a.tagneg = (ctx->g.tag_neg == p + 1);
  a.flagB = ctx->flag_B; a.gridB = ctx->flag_gridB; a.Bconst = (ctx->nB1 == 1);
  a.magnetize = !(ctx->g.flag_B_pos == 1 && s.charge > 0.0);
  a.nB1 = ctx->nB1; a.nB2 = ctx->nB2; a.hB1 = ctx->hB1; a.hB2 = ctx->hB2;
  a.Bc0 = ctx->Bc[0]; a.Bc1 = ctx->Bc[1]; a.Bc2 = ctx->Bc[2];
  a.acc = (unsigned long long*)s.acc;
  a.k1dep = s.Nm / (ctx->g.h1 * ctx->g.h2 * ctx->g.zmax);  // part_expmover.f90:564
  a.tally = ctx->d_tally + (size_t)p * kTallyCopies;
  a.sink = s.sink;
  a.ngrid = ctx->g.ngrid;
  a.seg_in = s.seg[s.cur];
  a.seg_out = s.seg[1 - s.cur];
  a.count = (unsigned long long*)(ctx->d_counts + p);
  a.err = ctx->d_error;
  return a;
}

// np(:,:,ptype) from the species' fixed-point accumulator
int np_from_acc(lepic_ctx* ctx, int p) {
  Species& s = ctx->sp[p];
  const int shift = ctx->g.precision == LEPIC_MIXED ? kFixShiftMx : kFixShift64;
  const double scale = std::ldexp(s.Nm / (ctx->g.h1 * ctx->g.h2 * ctx->g.zmax), -shift);
  const int grid = (int)std::min<int64_t>((ctx->nodes + 255) / 256, 2048);
  hipLaunchKernelGGL(k_np_from_acc, dim3(grid), dim3(256), 0, ctx->stream, (const unsigned long long*)s.acc,
                     (const int8_t*)ctx->bc, s.np, ctx->nodes, scale);
  LEPIC_HIP(hipGetLastError());
  return 0;
}

}  // namespace

namespace lepic {

int mover_launch(lepic_ctx* ctx, int p, int deposit) {
  Species& s = ctx->sp[p];
  MoverCommon a = common_args(ctx, p);
  a.deposit = deposit ? 1 : 0;
  a.check_dead = s.dead_maybe ? 1 : 0;
  const bool mixed = ctx->g.precision == LEPIC_MIXED;
  const int64_t nb = std::max<int64_t>(nseg_of(s), 1);  // one block per segment
  if (deposit)
    if (int rc = zero_async(ctx, s.acc, ctx->nodes * 8)) return rc;
  const int b = s.cur, o = 1 - s.cur;
  if (ctx->timing) LEPIC_HIP(hipEventRecord(ctx->ev_mover[p][0], ctx->stream));
  if (mixed) {
    MixedParams m;
    m.k1f = (float)a.k1;
    m.dtoh1 = (float)(ctx->g.dt / ctx->g.h1);
    m.dtoh2 = (float)(ctx->g.dt / ctx->g.h2);
    m.dtf = (float)ctx->g.dt;
    m.hB1r = ctx->flag_gridB ? (float)(ctx->g.h1 / ctx->hB1) : 1.0f;
    m.hB2r = ctx->flag_gridB ? (float)(ctx->g.h2 / ctx->hB2) : 1.0f;
    SoaMx in{s.cell[b], s.f[b][0], s.f[b][1], s.f[b][2], s.f[b][3], s.f[b][4], s.f[b][5]};
    SoaMxW out{s.cell[o], s.f[o][0], s.f[o][1], s.f[o][2], s.f[o][3], s.f[o][4], s.f[o][5]};
    hipLaunchKernelGGL(k_mover_mixed<kMoverThreadsMixed>, dim3((unsigned)nb), dim3(kMoverThreadsMixed), 0,
                       ctx->stream, a, m, in, out,
                       s.flag_dead, (const float*)ctx->Ef, (const float*)(ctx->Bf ? ctx->Bf : (float*)ctx->Ef));
  } else {
    Soa64 in{s.d[b][0], s.d[b][1], s.d[b][2], s.d[b][3], s.d[b][4], s.d[b][5]};
    Soa64W out{s.d[o][0], s.d[o][1], s.d[o][2], s.d[o][3], s.d[o][4], s.d[o][5]};
    hipLaunchKernelGGL(k_mover64<kMoverThreads64>, dim3((unsigned)nb), dim3(kMoverThreads64), 0, ctx->stream,
                       a, in, out,
                       s.flag_dead, (const double*)ctx->E, (const double*)ctx->Bi);
  }
  LEPIC_HIP(hipGetLastError());
  if (ctx->timing) {
    LEPIC_HIP(hipEventRecord(ctx->ev_mover[p][1], ctx->stream));
    ctx->ev_recorded[p] = true;
  }
  s.cur = o;
  s.packed = false;
  s.dead_maybe = false;   // the push cleared every flag it consumed
  if (deposit)
    if (int rc = np_from_acc(ctx, p)) return rc;
  return 0;
}

int deposit_launch(lepic_ctx* ctx, int p) {
  Species& s = ctx->sp[p];
  MoverCommon a = common_args(ctx, p);
  a.deposit = 1;
  if (int rc = zero_async(ctx, s.acc, ctx->nodes * 8)) return rc;
  const int64_t nb = std::max<int64_t>(nseg_of(s), 1);
  const int b = s.cur;
  if (ctx->g.precision == LEPIC_MIXED) {
    SoaMx in{s.cell[b], s.f[b][0], s.f[b][1], s.f[b][2], s.f[b][3], s.f[b][4], s.f[b][5]};
    hipLaunchKernelGGL(k_deposit_mixed, dim3((unsigned)nb), dim3(kBlock), 0, ctx->stream, a, in);
  } else {
    Soa64 in{s.d[b][0], s.d[b][1], s.d[b][2], s.d[b][3], s.d[b][4], s.d[b][5]};
    hipLaunchKernelGGL(k_deposit64, dim3((unsigned)nb), dim3(kBlock), 0, ctx->stream, a, in);
  }
  LEPIC_HIP(hipGetLastError());
  return np_from_acc(ctx, p);
}

// Adds the charge of segments [seg0, seg0+nseg) of the current buffer to
// the accumulator of the last deposition without clearing it, and rescales
// np(:,:,ptype) from it: particles appended after the mover
// (volume injection, part_injection.f90, runs between the push and the
// deposition, main.f90:1063-1071) join the step's density this way.
int deposit_range_launch(lepic_ctx* ctx, int p, int64_t seg0, int64_t nseg) {
  if (nseg <= 0) return 0;
  Species& s = ctx->sp[p];
  MoverCommon a = common_args(ctx, p);
  a.deposit = 1;
  a.seg_in = s.seg[s.cur] + seg0;
  const int b = s.cur;
  const int64_t o = seg0 * kChunk;
  if (ctx->g.precision == LEPIC_MIXED) {
    SoaMx in{s.cell[b] + o, s.f[b][0] + o, s.f[b][1] + o, s.f[b][2] ? s.f[b][2] + o : nullptr,
             s.f[b][3] + o, s.f[b][4] + o, s.f[b][5] + o};
    hipLaunchKernelGGL(k_deposit_mixed, dim3((unsigned)nseg), dim3(kBlock), 0, ctx->stream, a, in);
  } else {
    Soa64 in{s.d[b][0] + o, s.d[b][1] + o, s.d[b][2] + o, s.d[b][3] + o, s.d[b][4] + o, s.d[b][5] + o};
    hipLaunchKernelGGL(k_deposit64, dim3((unsigned)nseg), dim3(kBlock), 0, ctx->stream, a, in);
  }
  LEPIC_HIP(hipGetLastError());
  return np_from_acc(ctx, p);
}

}  // namespace lepic
