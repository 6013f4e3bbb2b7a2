// poisson.hip -- iterative_solver (iterative_solver.f90:1-69), the MG
// V-cycle (mg.f90:1-232) and red-black SOR with adaptive sweep control
// (sors.f90:1-435).
//
// A kernel launch costs about 3 us on MI355X whatever it does, and a grid-wide
// software barrier costs more, so the solver is organised to launch few
// kernels and to need no grid synchronisation:
//  * Fine ("multi-block") levels run k_sweeps: each block owns a 64-wide tile,
//    stages the tile plus a halo of 2S+1 nodes in LDS and performs S red-black
//    sweeps on it, recomputing the halo redundantly instead of exchanging it.
//    After every sweep it writes the tile to that sweep's state buffer (S+1
//    rotating buffers per level) and its share of the L1 residual.  The exit
//    test of sors.f90:371-377 for those sweeps is taken at the start of the
//    next launch (or by the kernel that consumes the level), by every block
//    from the same partial sums in the same order, so all blocks agree and no
//    atomics or host round trips are needed.  A visit with the reference's
//    limit of 10 sweeps takes two launches.
//  * Coarse levels (at most mg_tail_nodes nodes) run in k_tail, one workgroup
//    of 1024 threads that performs the whole lower part of the V-cycle (down
//    sweeps, restrictions, the coarsest solve, prolongations and up sweeps),
//    with the levels that fit in its LDS (up to 158 KB) held there.
//  * Grid transfers between multi-block levels: FP64 runs them as kernels
//    between the visits (k_restrict, LDS-tiled; k_prolong).  In mixed mode a
//    down visit's last sweep launch computes the restriction of its tile from
//    the region still in LDS (one node of extra halo), and an up visit's
//    first launch adds the prolonged coarse correction to its region as it
//    loads it.
// FP64 mode computes the reference's expressions in their order (no
// contraction), so a sweep is bit-identical to sors.f90; only the order of
// the L1 residual sum differs.  Mixed mode runs the same V-cycle in FP32 on
// the correction equation inside an FP64 iterative-refinement loop whose
// FP64 residual keeps the reference's convergence definition.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <type_traits>
#include <vector>

#include "lepic_internal.h"

using namespace lepic;

namespace {

constexpr int kVisitLog = 2 * kMaxLevels;     // ig = 1 .. 2ng-1
constexpr size_t kTailLds = 160 * 1024 - 2048;  // dynamic LDS of k_tail (static arrays take < 2 KB)
constexpr int kTailThreads = 1024;
constexpr int kS = kSweepsPerLaunch;          // most sweeps of one launch (kRot - 1 state buffers ahead)
// halo of a launch of at most KS sweeps: validity shrinks by one node per colour pass
template <int KS> struct Halo { static constexpr int v = 2 * KS + 1; };
constexpr int kTX = 64;                       // wide tiles: 64 x TileY nodes, 1024 threads
constexpr int kSmallTX = 32;                  // small tiles: 32 x SmallTileY nodes, 512 threads
constexpr int kSmallThreads = 512;
constexpr int kSmallTileBelow = 64;           // a level with fewer wide tiles than this takes small ones
constexpr int kSweepThreads = 1024;
constexpr int kDefaultTailNodes = 5000;
constexpr int kResidBlocks = 512;             // blocks of the FP64 finest-level residual (their partial sums fit the pinned read-back)
// byte offset of the residual partials behind the scalars, visit log and status
constexpr size_t kPartialsOff = (64 * 8 + (kVisitLog + 1) * sizeof(int) + 7) / 8 * 8;
static_assert(kPartialsOff - 32 * 8 + kResidBlocks * 8 <= kPinnedDoubles * 8, "read-back fits the pinned buffer");
constexpr int kNr = 10;                       // sweep limit of sors.f90:63-64
// Mixed form: the FP32 correction V-cycle is a preconditioner inside FP64
// refinement to the unchanged tolerance, so its schedule is chosen, not
// taken from the reference's exit test (sors.f90:371-377).  The finest level
// (when multi-block) runs a fixed 4 down / 3 up red-black sweeps at omega
// 1.15: the reference's omega 1.4 (iterative_solver.f90:27) damps the
// highest frequencies less, and its exit test stops these visits after 2
// sweeps on plasma solves and runs them to 10 near electrodes.  The other
// multi-block levels and the tail levels above the coarsest run the 10 down
// / 2 up sweeps the test gives them on plasma solves, without the per-sweep
// residuals and sums the test needs; the coarsest level keeps the test.
// Measured: 1024^2 argon box 4 -> 3 V-cycles per solve.
constexpr int kFineDn = 4, kFineUp = 3;
constexpr float kFineOmega = 1.15f;
constexpr int kFixDn = kNr, kFixUp = 2;
constexpr int kWaveNodes = 320;               // tail levels with at most this many interior nodes run in one wave
constexpr int kRowBlocks = 1024;              // row-loop blocks of the grid-transfer kernels (one row each up to 1024 rows)

template <typename T> struct TileY;
template <> struct TileY<float> { static constexpr int v = 32; };
template <> struct TileY<double> { static constexpr int v = 16; };
template <typename T> struct SmallTileY;
template <> struct SmallTileY<float> { static constexpr int v = 16; };
template <> struct SmallTileY<double> { static constexpr int v = 8; };

// ---- SOR update of one node (sors.f90:108-110 / 238-240) ----
// FP64 divides like the reference; the FP32 smoother of the mixed form
// multiplies by the reciprocal (the FP64 refinement sets the final accuracy)
template <typename T>
__device__ __forceinline__ T sor_update(T u, T b, T uW, T uN, T uS, T uE, T aw, T an, T ac, T omega) {
  if constexpr (sizeof(T) == 4)
    return (1.0f - omega) * u + omega * (b - aw * uW - an * uN - an * uS - aw * uE) * (1.0f / ac);
  else
    return ((T)1.0 - omega) * u + omega * (b - aw * uW - an * uN - an * uS - aw * uE) / ac;
}

// ---- residual of one node, sors.f90:348-349 order ----
template <typename T>
__device__ __forceinline__ double sor_resid(T u, T b, T uW, T uN, T uS, T uE, T aw, T an, T ac) {
  return (double)(b - (aw * uW + an * uN + an * uS + aw * uE + ac * u));
}

// ---- getres, mg.f90:512-547 order (aw, ae, an, as, ac) ----
template <typename T, typename PU, typename PB>
__device__ __forceinline__ T getres(PU u, PB b, int ir, int jr, int n1, T aw, T an, T ac) {
  const int64_t W = n1 + 3;
  const int64_t c = (int64_t)jr * W + ir;
  T uij = u[c], un = u[c + W], us = u[c - W], ue = u[c + 1], uw = u[c - 1];
  T urhs = b[(int64_t)(jr - 1) * (n1 + 1) + (ir - 1)];
  return urhs - (aw * uw + aw * ue + an * un + an * us + ac * uij);
}

// getres of the node at index x of an LDS region with rows of RW (same
// expression and order)
template <typename T>
__device__ __forceinline__ T lds_getres(const T* su, const T* sb, int x, int RW, T aw, T an, T ac) {
  return sb[x] - (aw * su[x - 1] + aw * su[x + 1] + an * su[x + RW] + an * su[x - RW] + ac * su[x]);
}

// mg.f90:386-414: the weighted sum of the centre residual g9 and the eight
// around it (g1 north-east, then clockwise), each counted when its flag holds
template <typename T>
__device__ __forceinline__ T full_weight(T g9, T g1, T g2, T g3, T g4, T g5, T g6, T g7, T g8, bool f1, bool f2,
                                         bool f3, bool f4, bool f5, bool f6, bool f7, bool f8) {
  const T r9 = g9;
  const T r1 = f1 ? g1 : (T)0, r2 = f2 ? g2 : (T)0, r3 = f3 ? g3 : (T)0, r4 = f4 ? g4 : (T)0;
  const T r5 = f5 ? g5 : (T)0, r6 = f6 ? g6 : (T)0, r7 = f7 ? g7 : (T)0, r8 = f8 ? g8 : (T)0;
  return ((T)4.0 * r9 + (T)2.0 * (r2 + r4 + r6 + r8) + r1 + r3 + r5 + r7) / (T)16.0;
}

// full-weighted residual of coarse node (i,j), mg.f90:386-414.  The nine
// fine residuals are evaluated unconditionally (at a position clamped into
// the interior when the reference skips one), so all loads issue at once;
// the reference's conditions select which of them count.
template <typename T, typename PU, typename PB, typename PC>
__device__ __forceinline__ T restrict_node(PU u, PB b, PC bc, int i, int j, int n1, int n2, T aw, T an, T ac) {
  const int ir = 2 * i - 1, jr = 2 * j - 1;
  const int64_t W = n1 + 3;
  auto B = [&](int a, int c) { return bc[(int64_t)c * W + a]; };
  auto G = [&](int a, int c) {
    a = a < 1 ? 1 : (a > n1 + 1 ? n1 + 1 : a);
    c = c < 1 ? 1 : (c > n2 + 1 ? n2 + 1 : c);
    return getres(u, b, a, c, n1, aw, an, ac);
  };
  const bool xe = ir != n1 + 1, xw = ir != 1, yn = jr != n2 + 1, ys = jr != 1;
  const bool f1 = xe && yn && B(ir + 1, jr + 1) <= 0;
  const bool f2 = xe && B(ir + 1, jr) <= 0;
  const bool f3 = B(ir + 1, jr - 1) <= 0 && xe && ys;
  const bool f4 = B(ir, jr - 1) <= 0 && ys;
  const bool f5 = B(ir - 1, jr - 1) <= 0 && xw && ys;
  const bool f6 = B(ir - 1, jr) <= 0 && xw;
  const bool f7 = B(ir - 1, jr + 1) <= 0 && xw && yn;
  const bool f8 = B(ir, jr + 1) <= 0 && yn;
  const T g9 = G(ir, jr), g1 = G(ir + 1, jr + 1), g2 = G(ir + 1, jr), g3 = G(ir + 1, jr - 1);
  const T g4 = G(ir, jr - 1), g5 = G(ir - 1, jr - 1), g6 = G(ir - 1, jr), g7 = G(ir - 1, jr + 1);
  const T g8 = G(ir, jr + 1);
  return full_weight<T>(g9, g1, g2, g3, g4, g5, g6, g7, g8, f1, f2, f3, f4, f5, f6, f7, f8);
}

// bilinear add of coarse node (i,j)'s correction to its four fine nodes,
// mg.f90:585-611
template <typename T, typename PU, typename PC, typename PE>
__device__ __forceinline__ void prolong_node(PU u, PC bc, int n1, PE e2h, int m1, int i, int j) {
  const int64_t Wc = m1 + 3, W = n1 + 3;
  T r1 = e2h[(int64_t)j * Wc + i], r2 = e2h[(int64_t)(j - 1) * Wc + i];
  T r3 = e2h[(int64_t)(j - 1) * Wc + i - 1], r4 = e2h[(int64_t)j * Wc + i - 1];
  const int ir = 2 * i - 1, jr = 2 * j - 1;
  const int64_t c = (int64_t)jr * W + ir;
  if (bc[c] <= 0) u[c] = u[c] + r1;
  if (bc[c - W] <= 0) u[c - W] = u[c - W] + (r1 + r2) / (T)2.0;
  if (bc[c - W - 1] <= 0) u[c - W - 1] = u[c - W - 1] + (r1 + r2 + r3 + r4) / (T)4.0;
  if (bc[c - 1] <= 0) u[c - 1] = u[c - 1] + (r1 + r4) / (T)2.0;
}

// The correction prolong_node adds to fine node (I, J) (1..n1+1, 1..n2+1):
// the coarse node (I/2 + 1, J/2 + 1) and its west / south neighbours, with
// prolong_node's expressions.
template <typename T>
__device__ __forceinline__ T prolong_at(const T* e2h, int m1, int I, int J) {
  const int64_t Wc = m1 + 3;
  const int ci = I / 2 + 1, cj = J / 2 + 1;
  const T r1 = e2h[(int64_t)cj * Wc + ci];
  if (I & 1) {
    if (J & 1) return r1;
    return (r1 + e2h[(int64_t)(cj - 1) * Wc + ci]) / (T)2.0;
  }
  const T r4 = e2h[(int64_t)cj * Wc + ci - 1];
  if (J & 1) return (r1 + r4) / (T)2.0;
  const T r2 = e2h[(int64_t)(cj - 1) * Wc + ci], r3 = e2h[(int64_t)(cj - 1) * Wc + ci - 1];
  return (r1 + r2 + r3 + r4) / (T)4.0;
}

// ---------------------------------------------------------------------------
// visit control of multi-block levels
// ---------------------------------------------------------------------------
struct VisitRef {
  LevelCtrl* ctrl;
  const double* partials;
  int nb;                  // blocks of the level's sweep kernel
  int epoch, dig, nr, ig;
  double eps;
  int nexec;               // sweeps launched in the visit
  int klast0;              // sweeps before the visit's last launch
  int fixed;               // the visit runs exactly nexec sweeps (mixed form)
  int* visit;
  double* vres;
};

// L1 residuals after sweeps kf-1 .. kt of the current visit (kt-kf+2 <= kS+1
// values): the blocks' partial sums in a fixed order (lane strides, then a
// shuffle tree).  All loads are issued before any sum, so the evaluation
// costs one memory round trip.  Wave 0, all lanes; r[0] is sweep kf-1 (zero
// when kf = 1).
__device__ void ring_res(const double* partials, int nb, int kf, int kt, double* r) {
  const int lane = threadIdx.x & 63;
  double s[kS + 1];
#pragma unroll
  for (int m = 0; m <= kS; ++m) {
    const int k = kf - 1 + m;
    s[m] = 0.0;
    if (k >= 1 && k <= kt) {
      const double* p = partials + (size_t)(k % kRing) * nb;
      for (int q = lane; q < nb; q += 64) s[m] += p[q];
    }
  }
#pragma unroll
  for (int m = 0; m <= kS; ++m) r[m] = wave_sum(s[m]);
}

// sors.f90:371-377 for sweeps kf..kt (at most kS of them): the first sweep
// after which the reference stops (0 if none); res_s is the previous
// sweep's residual, zero before the first sweep of a visit.
__device__ int first_stop(const double* partials, int nb, int kf, int kt, int dig, double eps, int nr,
                          double* r_out) {
  double r[kS + 1];
  ring_res(partials, nb, kf, kt, r);
  double prev = r[0];
  for (int k = kf; k <= kt; ++k) {
    const double rk = r[k - kf + 1];
    bool stop = (dig == 0) ? (rk <= eps) : (prev > 0.0 && rk / prev <= 0.6);
    if (k >= nr) stop = true;
    if (stop) {
      *r_out = rk;
      return k;
    }
    prev = rk;
  }
  return 0;
}

struct Resolved {
  int ks;                  // sweeps of the visit (0: not finished)
  int buf;                 // rotating buffer holding the state after them
  double res;
};

// Outcome of a visit after all its launches; every block of the consuming
// kernel evaluates it (wave 0) and block 0 records it.  `canon` records
// buffer 0 as the level's state (the caller copies the result there).
__device__ Resolved resolve_visit(const VisitRef& V, bool writer, bool canon) {
  LevelCtrl* c = V.ctrl;
  const int c0 = c->cur[(V.epoch - 1) & 1];
  Resolved out{0, -1, 0.0};
  if (V.fixed) {
    // exactly nexec sweeps ran and the state after them is the result (the
    // host logs the count)
    out.ks = V.nexec;
    out.buf = (c0 + V.nexec) % kRot;
    if (writer && threadIdx.x == 0) c->cur[V.epoch & 1] = canon ? 0 : out.buf;
    return out;
  }
  if (c->done_epoch == V.epoch) {
    out.ks = c->ksor;
    out.res = c->res;
  } else {
    out.ks = first_stop(V.partials, V.nb, V.klast0 + 1, V.nexec, V.dig, V.eps, V.nr, &out.res);
  }
  if (out.ks) out.buf = (c0 + out.ks) % kRot;
  if (writer && out.ks && threadIdx.x == 0) {
    c->cur[V.epoch & 1] = canon ? 0 : out.buf;
    V.visit[V.ig] = out.ks;
    V.vres[V.ig] = out.res;
  }
  return out;
}

// ---------------------------------------------------------------------------
// multi-block levels
// ---------------------------------------------------------------------------
template <typename T>
struct SweepArgs {
  T* rot[kRot];
  const T* b;
  const int8_t* bc;
  int n1, n2;
  T aw, an, ac, omega;
  LevelCtrl* ctrl;
  double* partials;
  int nbx, nby;
  int epoch, dig, nr;
  int fixed;               // run the launch's sweeps without the exit test (mixed form)
  int nores;               // no residuals (a visit whose sweep count nothing reads back): only the last state is written
  double eps;
  // mixed form, first launch of an up visit: the coarse level's correction
  // is added to the region as it is loaded (the prolongation of
  // mg.f90:585-611, no kernel of its own); pcrot / pvc as in k_prolong
  int pro, pro_mb, pm1;
  T* pcrot[kRot];
  VisitRef pvc;
  // mixed form, last launch of a down visit: the restriction onto the level
  // below (k_restrict's work, from the final state in LDS; the launch's halo
  // is one sweep wider than its sweeps need) and the fine visit's record
  int rst;
  T* r2h;
  const int8_t* bc2h;
  T* rcrot[kRot];
  LevelCtrl* rcctrl;       // null when the coarse level runs in k_tail
};

// Sweeps k0+1 .. k0+kexec (kexec <= KS) of a visit on one TX x TY tile with
// NT threads (see the file comment).  A launch that continues a visit first
// takes the reference's exit test for the sweeps of the visit's previous
// launch, kchk .. k0.
template <typename T, int kTX, int TY, int kSweepThreads, int KS, bool RST = false>
__global__ __launch_bounds__(kSweepThreads, RST ? 8 : 1) void k_sweeps(SweepArgs<T> A, int k0, int kexec, int kchk) {
  constexpr int kHalo = Halo<KS>::v;
  // a block's tile is kTX x TY nodes, or one node wider (taller) in the last
  // column (row) of tiles when the level's node count leaves one node over
  // (2^k + 1 nodes per side); the region has room for that and stays even
  constexpr int kTP = kTX + 1, kTYP = TY + 1;
  constexpr int RW = kTX + 2 + 2 * kHalo, RH = kTYP + 2 * kHalo, RN = RW * RH;
  constexpr int HW = (RW - 2) / 2;  // nodes of one colour per inner region row
  static_assert((RW & 1) == 0, "region width must be even");
  __shared__ T su[RN];
  __shared__ T sb[RN];
  __shared__ int8_t sbc[RN];
  __shared__ int s_go, s_c0, s_pb, s_rb;
  __shared__ double ws[kSweepThreads / 64];
  const int nb = A.nbx * A.nby;
  const int bid = blockIdx.y * A.nbx + blockIdx.x;
  const bool pro = A.pro && k0 == 0;
  if (threadIdx.x < 64) {
    LevelCtrl* c = A.ctrl;
    int go = 1;
    int pb = 0;
    if (pro && A.pro_mb) {  // the coarse level's up visit, as k_prolong resolves it
      Resolved r = resolve_visit(A.pvc, bid == 0, false);
      pb = r.buf < 0 ? 0 : r.buf;
    }
    const int c0 = c->cur[(A.epoch - 1) & 1];
    if (A.fixed) {
      go = 1;
    } else if (__hip_atomic_load(&c->done_epoch, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == A.epoch) {
      go = 0;
    } else if (k0 > 0) {
      double r = 0.0;
      const int ks = first_stop(A.partials, nb, kchk, k0, A.dig, A.eps, A.nr, &r);
      if (ks) {
        go = 0;
        if (bid == 0 && threadIdx.x == 0) {
          c->ksor = ks;
          c->res = r;
          __hip_atomic_store(&c->done_epoch, A.epoch, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
      }
    }
    if (threadIdx.x == 0) {
      s_go = go;
      s_c0 = c0;
      s_pb = pb;
      s_rb = (A.rst && A.rcctrl) ? A.rcctrl->cur[(A.epoch - 1) & 1] : 0;   // k_restrict's coarse buffer
    }
  }
  __syncthreads();
  if (!s_go) return;

  const int n1 = A.n1, n2 = A.n2;
  const int64_t W = n1 + 3;
  const int tw = (int)blockIdx.x == A.nbx - 1 ? n1 + 1 - (int)blockIdx.x * kTX : kTX;
  const int th = (int)blockIdx.y == A.nby - 1 ? n2 + 1 - (int)blockIdx.y * TY : TY;
  const int gi0 = 1 + (int)blockIdx.x * kTX - kHalo;  // global i of region column 0
  const int gj0 = 1 + (int)blockIdx.y * TY - kHalo;
  const T* uin = A.rot[(s_c0 + k0) % kRot];
  const T* e2h = pro ? A.pcrot[s_pb] : nullptr;
  {
    // every load of the region is issued before any LDS store: one memory
    // latency for the whole region instead of one per loop trip
    constexpr int kIt = (RN + kSweepThreads - 1) / kSweepThreads;
    T lu[kIt], lb[kIt], lp[kIt];
    int8_t lc[kIt];
#pragma unroll
    for (int it = 0; it < kIt; ++it) {
      const int q = threadIdx.x + it * kSweepThreads;
      const int r = q / RW, cc = q - r * RW;
      const int i = gi0 + cc, j = gj0 + r;
      const bool in = q < RN && i >= 1 && i <= n1 + 1 && j >= 1 && j <= n2 + 1;
      lu[it] = in ? uin[(int64_t)j * W + i] : (T)0;
      lc[it] = in ? A.bc[(int64_t)j * W + i] : (int8_t)1;
      lb[it] = in ? A.b[(int64_t)(j - 1) * (n1 + 1) + (i - 1)] : (T)0;
      lp[it] = e2h ? prolong_at<T>(e2h, A.pm1, in ? i : 1, in ? j : 1) : (T)0;
    }
#pragma unroll
    for (int it = 0; it < kIt; ++it) {
      const int q = threadIdx.x + it * kSweepThreads;
      if (e2h && lc[it] <= 0) lu[it] = lu[it] + lp[it];   // prolong_node: bcnd <= 0 nodes
      if (q < RN) { su[q] = lu[it]; sbc[q] = lc[it]; sb[q] = lb[it]; }
    }
  }
  __syncthreads();
  const T aw = A.aw, an = A.an, ac = A.ac, om = A.omega;
  constexpr int NB = 4;  // slots per thread whose loads are issued together
  constexpr int kPass = (RH - 2) * HW;
  for (int s = 1; s <= kexec; ++s) {
    for (int color = 1; color >= 0; --color) {  // red (i+j odd), then black
      // nodes of one colour never read each other, so a thread gathers all
      // its slots before it writes any
      for (int q0 = threadIdx.x; q0 < kPass; q0 += NB * kSweepThreads) {
        int x[NB];
        int8_t m[NB];
        T c[NB], w[NB], e[NB], n[NB], so[NB], bb[NB];
#pragma unroll
        for (int k = 0; k < NB; ++k) {
          const int q = q0 + k * kSweepThreads;
          const int qq = q < kPass ? q : 0;
          const int r = 1 + qq / HW;
          x[k] = r * RW + 1 + 2 * (qq % HW) + ((color ^ ((gi0 + gj0 + r + 1) & 1)) & 1);
          m[k] = sbc[x[k]];
          if (q >= kPass) m[k] = 1;
          c[k] = su[x[k]]; w[k] = su[x[k] - 1]; e[k] = su[x[k] + 1];
          n[k] = su[x[k] + RW]; so[k] = su[x[k] - RW]; bb[k] = sb[x[k]];
        }
#pragma unroll
        for (int k = 0; k < NB; ++k)
          if (m[k] <= 0) su[x[k]] = sor_update(c[k], bb[k], w[k], n[k], so[k], e[k], aw, an, ac, om);
      }
      __syncthreads();
    }
    T* uout = A.rot[(s_c0 + k0 + s) % kRot];
    if (A.nores) {
      if (s == kexec)
        for (int q = threadIdx.x; q < kTP * kTYP; q += kSweepThreads) {
          const int tr = q / kTP, tc = q - tr * kTP;
          if (tc < tw && tr < th)
            uout[(int64_t)(gj0 + kHalo + tr) * W + gi0 + kHalo + tc] = su[(kHalo + tr) * RW + kHalo + tc];
        }
      continue;
    }
    double rr = 0.0;
    for (int q0 = threadIdx.x; q0 < kTP * kTYP; q0 += NB * kSweepThreads) {
      int x[NB], gi[NB], gj[NB];
      int8_t m[NB];
      T c[NB], w[NB], e[NB], n[NB], so[NB], bb[NB];
#pragma unroll
      for (int k = 0; k < NB; ++k) {
        const int q = q0 + k * kSweepThreads;
        const int qq = q < kTP * kTYP ? q : 0;
        const int tr = qq / kTP, tc = qq - tr * kTP;
        x[k] = (kHalo + tr) * RW + kHalo + tc;
        gi[k] = (q < kTP * kTYP && tc < tw && tr < th) ? gi0 + kHalo + tc : n1 + 2;   // n1 + 2: not this tile's
        gj[k] = gj0 + kHalo + tr;
        m[k] = sbc[x[k]];
        c[k] = su[x[k]]; w[k] = su[x[k] - 1]; e[k] = su[x[k] + 1];
        n[k] = su[x[k] + RW]; so[k] = su[x[k] - RW]; bb[k] = sb[x[k]];
      }
#pragma unroll
      for (int k = 0; k < NB; ++k) {
        if (gi[k] > n1 + 1 || gj[k] > n2 + 1) continue;
        if (m[k] == -1) rr += fabs(sor_resid(c[k], bb[k], w[k], n[k], so[k], e[k], aw, an, ac));
        uout[(int64_t)gj[k] * W + gi[k]] = c[k];
      }
    }
    rr = wave_sum(rr);
    if ((threadIdx.x & 63) == 0) ws[threadIdx.x >> 6] = rr;
    __syncthreads();
    if (threadIdx.x < 64) {  // fixed-order sum of the 16 wave sums
      const double t = wave_sum(threadIdx.x < kSweepThreads / 64 ? ws[threadIdx.x] : 0.0);
      if (threadIdx.x == 0) A.partials[(size_t)((k0 + s) % kRing) * nb + bid] = t;
    }
  }
  if (RST && A.rst) {  // the last colour pass ended with a barrier
    // the fine visit's record (resolve_visit of a fixed-length visit)
    if (bid == 0 && threadIdx.x == 0) A.ctrl->cur[A.epoch & 1] = (s_c0 + k0 + kexec) % kRot;
    // the coarse nodes whose fine centre (2ci-1, 2cj-1) lies in this tile,
    // plus the coarse border next to the domain's edge for the zeroing
    const int m1 = n1 / 2;
    const int ti0 = 1 + (int)blockIdx.x * kTX, tj0 = 1 + (int)blockIdx.y * TY;
    const int ci0 = (ti0 + 1) / 2, ci1 = (ti0 + tw) / 2, cj0 = (tj0 + 1) / 2, cj1 = (tj0 + th) / 2;
    const int zi0 = ci0 - (blockIdx.x == 0), zi1 = ci1 + ((int)blockIdx.x == A.nbx - 1);
    const int zj0 = cj0 - (blockIdx.y == 0), zj1 = cj1 + ((int)blockIdx.y == A.nby - 1);
    const int zw = zi1 - zi0 + 1, zn = zw * (zj1 - zj0 + 1);
    T* e2h = A.rcrot[0];  // selected without a dynamic index into the argument
#pragma unroll
    for (int r = 1; r < kRot; ++r) e2h = s_rb == r ? A.rcrot[r] : e2h;
    const int64_t Wc = m1 + 3;
    for (int q = threadIdx.x; q < zn; q += kSweepThreads) {
      const int ci = zi0 + q % zw, cj = zj0 + q / zw;
      e2h[(int64_t)cj * Wc + ci] = (T)0;
      if (ci < ci0 || ci > ci1 || cj < cj0 || cj > cj1) continue;
      T out = (T)0;
      if (A.bc2h[(int64_t)cj * Wc + ci] <= 0) {
        const int ir = 2 * ci - 1, jr = 2 * cj - 1;
        // region index of fine node (ir + dx, jr + dy), unclamped (bc flags)
        // and clamped into the interior (getres, as restrict_node evaluates it)
        const int xr = (jr - gj0) * RW + (ir - gi0);
        const int dxe = ir + 1 <= n1 + 1 ? 1 : 0, dxw = ir - 1 >= 1 ? -1 : 0;
        const int dyn = jr + 1 <= n2 + 1 ? RW : 0, dys = jr - 1 >= 1 ? -RW : 0;
        const bool xe = ir != n1 + 1, xw = ir != 1, yn = jr != n2 + 1, ys = jr != 1;
        const bool f1 = xe && yn && sbc[xr + 1 + RW] <= 0;
        const bool f2 = xe && sbc[xr + 1] <= 0;
        const bool f3 = sbc[xr + 1 - RW] <= 0 && xe && ys;
        const bool f4 = sbc[xr - RW] <= 0 && ys;
        const bool f5 = sbc[xr - 1 - RW] <= 0 && xw && ys;
        const bool f6 = sbc[xr - 1] <= 0 && xw;
        const bool f7 = sbc[xr - 1 + RW] <= 0 && xw && yn;
        const bool f8 = sbc[xr + RW] <= 0 && yn;
        out = full_weight<T>(lds_getres<T>(su, sb, xr, RW, aw, an, ac),
                             lds_getres<T>(su, sb, xr + dxe + dyn, RW, aw, an, ac),
                             lds_getres<T>(su, sb, xr + dxe, RW, aw, an, ac),
                             lds_getres<T>(su, sb, xr + dxe + dys, RW, aw, an, ac),
                             lds_getres<T>(su, sb, xr + dys, RW, aw, an, ac),
                             lds_getres<T>(su, sb, xr + dxw + dys, RW, aw, an, ac),
                             lds_getres<T>(su, sb, xr + dxw, RW, aw, an, ac),
                             lds_getres<T>(su, sb, xr + dxw + dyn, RW, aw, an, ac),
                             lds_getres<T>(su, sb, xr + dyn, RW, aw, an, ac), f1, f2, f3, f4, f5, f6, f7, f8);
      }
      A.r2h[(int64_t)(cj - 1) * (m1 + 1) + (ci - 1)] = out;
    }
  }
}

// restriction, mg.f90:335-444, from a multi-block level: resolves the fine
// level's down visit, writes the coarse right-hand side and zeroes the
// coarse correction in the buffer the coarse level's visit starts from.
template <typename T>
struct RestrictArgs {
  VisitRef vf;
  T* frot[kRot];
  const T* fb;
  const int8_t* fbc;
  int n1, n2;
  T aw, an, ac;
  T* r2h;
  const int8_t* bc2h;
  T* crot[kRot];
  LevelCtrl* cctrl;        // null when the coarse level runs in k_tail
  int cepoch;
};

// A block restricts a tile of kRcx x kRcy coarse nodes: the fine residuals
// its nodes weight (2 kRcx + 1 by 2 kRcy + 1 positions) are evaluated once
// each into LDS, with their bc flags, and every coarse node then combines
// nine of them; the residual at a position is getres at the position clamped
// into the interior, as restrict_node evaluates it.  Each block also zeroes
// a slice of the coarse correction.
constexpr int kRcx = 32, kRcy = 8;
constexpr int kRrx = 2 * kRcx + 1, kRry = 2 * kRcy + 1;
constexpr int kRestrictThreads = kRcx * kRcy;

template <typename T>
__global__ __launch_bounds__(kRestrictThreads) void k_restrict(RestrictArgs<T> A) {
  constexpr int kRN = kRrx * kRry;
  constexpr int kIt = (kRN + kRestrictThreads - 1) / kRestrictThreads;
  __shared__ int s_fb, s_cb;
  __shared__ T sr[kRN];
  __shared__ int8_t sf[kRN];
  const int t = threadIdx.x;
  if (t < 64) {
    const bool writer = blockIdx.x == 0 && blockIdx.y == 0;
    Resolved r = resolve_visit(A.vf, writer, false);
    if (t == 0) {
      s_fb = r.buf < 0 ? 0 : r.buf;
      s_cb = A.cctrl ? A.cctrl->cur[(A.cepoch - 1) & 1] : 0;
    }
  }
  __syncthreads();
  const T* u = A.frot[s_fb];
  T* e2h = A.crot[s_cb];
  const int n1 = A.n1, n2 = A.n2, m1 = n1 / 2, m2 = n2 / 2;
  const int64_t W = n1 + 3;
  {
    const int64_t nc = (int64_t)(m1 + 3) * (m2 + 3);
    const int64_t nbk = (int64_t)gridDim.x * gridDim.y, bk = (int64_t)blockIdx.y * gridDim.x + blockIdx.x;
    const int64_t per = (nc + nbk - 1) / nbk, q1 = min(nc, (bk + 1) * per);
    for (int64_t q = bk * per + t; q < q1; q += kRestrictThreads) e2h[q] = (T)0;
  }
  const int ci0 = 1 + (int)blockIdx.x * kRcx, cj0 = 1 + (int)blockIdx.y * kRcy;
  const int fa0 = 2 * ci0 - 2, fc0 = 2 * cj0 - 2;   // fine position of region (0, 0)
  {
    T g[kIt];
    int8_t fl[kIt];
#pragma unroll
    for (int k = 0; k < kIt; ++k) {
      const int q = t + k * kRestrictThreads;
      const int qq = q < kRN ? q : 0;
      const int y = qq / kRrx, x = qq - y * kRrx;
      const int a = fa0 + x, c = fc0 + y;
      const int ac = a < 1 ? 1 : (a > n1 + 1 ? n1 + 1 : a);
      const int cc = c < 1 ? 1 : (c > n2 + 1 ? n2 + 1 : c);
      g[k] = getres(u, A.fb, ac, cc, n1, A.aw, A.an, A.ac);
      fl[k] = (a <= n1 + 2 && c <= n2 + 2) ? A.fbc[(int64_t)c * W + a] : (int8_t)1;
    }
#pragma unroll
    for (int k = 0; k < kIt; ++k) {
      const int q = t + k * kRestrictThreads;
      if (q < kRN) { sr[q] = g[k]; sf[q] = fl[k]; }
    }
  }
  __syncthreads();
  const int ti = t % kRcx, tj = t / kRcx;
  const int i = ci0 + ti, j = cj0 + tj;
  if (i > m1 + 1 || j > m2 + 1) return;
  T out = (T)0;
  if (A.bc2h[(int64_t)j * (m1 + 3) + i] <= 0) {
    const int ir = 2 * i - 1, jr = 2 * j - 1;
    const int x0 = 2 * ti + 1, y0 = 2 * tj + 1;      // region position of (ir, jr)
    auto R = [&](int dx, int dy) { return sr[(y0 + dy) * kRrx + x0 + dx]; };
    auto B = [&](int dx, int dy) { return sf[(y0 + dy) * kRrx + x0 + dx]; };
    const bool xe = ir != n1 + 1, xw = ir != 1, yn = jr != n2 + 1, ys = jr != 1;
    const bool f1 = xe && yn && B(1, 1) <= 0;
    const bool f2 = xe && B(1, 0) <= 0;
    const bool f3 = B(1, -1) <= 0 && xe && ys;
    const bool f4 = B(0, -1) <= 0 && ys;
    const bool f5 = B(-1, -1) <= 0 && xw && ys;
    const bool f6 = B(-1, 0) <= 0 && xw;
    const bool f7 = B(-1, 1) <= 0 && xw && yn;
    const bool f8 = B(0, 1) <= 0 && yn;
    out = full_weight<T>(R(0, 0), R(1, 1), R(1, 0), R(1, -1), R(0, -1), R(-1, -1), R(-1, 0), R(-1, 1), R(0, 1),
                         f1, f2, f3, f4, f5, f6, f7, f8);
  }
  A.r2h[(int64_t)(j - 1) * (m1 + 1) + (i - 1)] = out;
}

// prolongation into a multi-block level from the level below (resolving the
// coarse level's up visit when it is multi-block itself)
template <typename T>
struct ProlongArgs {
  T* frot[kRot];
  LevelCtrl* fctrl;
  int fepoch;              // epoch of the fine level's coming up visit
  const int8_t* bc;
  int n1, n2;
  T* crot[kRot];
  int coarse_mb;
  VisitRef vc;
};

template <typename T>
__global__ __launch_bounds__(128) void k_prolong(ProlongArgs<T> A) {
  __shared__ int s_fb, s_cb;
  if (threadIdx.x < 64) {
    int cb = 0;
    if (A.coarse_mb) {
      const bool writer = blockIdx.x == 0 && blockIdx.y == 0;
      Resolved r = resolve_visit(A.vc, writer, false);
      cb = r.buf < 0 ? 0 : r.buf;
    }
    if (threadIdx.x == 0) {
      s_cb = cb;
      s_fb = A.fctrl->cur[(A.fepoch - 1) & 1];
    }
  }
  __syncthreads();
  const int m1 = A.n1 / 2, m2 = A.n2 / 2;
  const int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
  if (i > m1 + 1) return;
  for (int j = blockIdx.y + 1; j <= m2 + 1; j += gridDim.y)
    prolong_node<T>(A.frot[s_fb], A.bc, A.n1, (const T*)A.crot[s_cb], m1, i, j);
}

// FP64 mode, end of a V-cycle on a multi-block finest level: resolve the up
// visit and copy its result into phi (buffer 0) when it lies elsewhere.
struct Finish64Args {
  VisitRef v;
  double* rot[kRot];
  int n1, n2;
  int* status;             // sweeps of the visit, 0 if not finished
};

__global__ __launch_bounds__(256) void k_finish64(Finish64Args A) {
  __shared__ int s_buf, s_ks;
  if (threadIdx.x < 64) {
    const bool writer = blockIdx.x == 0 && blockIdx.y == 0;
    Resolved r = resolve_visit(A.v, writer, true);
    if (threadIdx.x == 0) {
      s_buf = r.buf;
      s_ks = r.ks;
      if (writer) *A.status = r.ks;
    }
  }
  __syncthreads();
  if (s_ks == 0 || s_buf == 0) return;
  const int64_t W = A.n1 + 3;
  for (int j = blockIdx.y + 1; j <= A.n2 + 1; j += gridDim.y)
    for (int i = blockIdx.x * blockDim.x + threadIdx.x + 1; i <= A.n1 + 1; i += gridDim.x * blockDim.x)
      A.rot[0][(int64_t)j * W + i] = A.rot[s_buf][(int64_t)j * W + i];
}

// Mixed mode, end of a V-cycle: phi += e (bcnd <= 0 nodes) with the FP32
// correction of the up visit, which is then zeroed for the next V-cycle.
struct AddCorrArgs {
  double* phi;
  float* rot[kRot];
  const int8_t* bc;
  int n1, n2;
  int mb;
  VisitRef v;
};

__global__ __launch_bounds__(256) void k_add_corr(AddCorrArgs A) {
  __shared__ int s_buf;
  if (threadIdx.x < 64) {
    int buf = 0;
    if (A.mb) {
      const bool writer = blockIdx.x == 0 && blockIdx.y == 0;
      Resolved r = resolve_visit(A.v, writer, false);
      buf = r.buf < 0 ? 0 : r.buf;
    }
    if (threadIdx.x == 0) s_buf = buf;
  }
  __syncthreads();
  float* e = A.rot[s_buf];
  const int64_t W = A.n1 + 3;
  for (int j = blockIdx.y + 1; j <= A.n2 + 1; j += gridDim.y)
    for (int i = blockIdx.x * blockDim.x + threadIdx.x + 1; i <= A.n1 + 1; i += gridDim.x * blockDim.x) {
      const int64_t q = (int64_t)j * W + i;
      if (A.bc[q] <= 0) A.phi[q] = A.phi[q] + (double)e[q];
      e[q] = 0.0f;
    }
}

// FP64 L1 residual of phi on the finest level (the reference's definition,
// sors.f90:348-349 order); in mixed mode it also becomes the FP32 right-hand
// side of the next correction equation.  One partial sum per block; blocks
// walk rows j = blockIdx.x + 1 + m*gridDim.x.  erot also clears the level-1
// FP32 correction buffer the next visit starts from.
__global__ __launch_bounds__(256) void k_resid64(const double* __restrict__ u, const double* __restrict__ b,
                                                 const int8_t* __restrict__ bc, int n1, int n2, double aw,
                                                 double an, double ac, float* __restrict__ bf,
                                                 float* const* __restrict__ erot, const LevelCtrl* ctrl,
                                                 int epoch_next, double* __restrict__ partials) {
  __shared__ double ws[4];
  __shared__ int s_zb;
  if (threadIdx.x == 0) s_zb = erot ? (ctrl ? ctrl->cur[(epoch_next - 1) & 1] : 0) : -1;
  __syncthreads();
  float* ez = s_zb >= 0 ? erot[s_zb] : nullptr;
  const int64_t W = n1 + 3;
  double r = 0.0;
  for (int j = blockIdx.x + 1; j <= n2 + 1; j += gridDim.x) {
    for (int i = threadIdx.x + 1; i <= n1 + 1; i += blockDim.x) {
      const int64_t c = (int64_t)j * W + i;
      const int64_t q = (int64_t)(j - 1) * (n1 + 1) + (i - 1);
      double rij = 0.0;
      if (bc[c] == -1)
        rij = b[q] - (aw * u[c - 1] + an * u[c + W] + an * u[c - W] + aw * u[c + 1] + ac * u[c]);
      if (bf) bf[q] = (float)rij;
      if (ez) ez[c] = 0.0f;
      r += fabs(rij);
    }
  }
  r = wave_sum(r);
  if ((threadIdx.x & 63) == 0) ws[threadIdx.x >> 6] = r;
  __syncthreads();
  if (threadIdx.x == 0) partials[blockIdx.x] = ((ws[0] + ws[1]) + ws[2]) + ws[3];
}

__global__ void k_cold_start(double* __restrict__ u, const int8_t* __restrict__ bc, int64_t n) {
  int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q < n && bc[q] == -1) u[q] = 0.0;
}

__global__ void k_copy_interior(double* __restrict__ u, const double* __restrict__ g,
                                const int8_t* __restrict__ bc, int64_t n) {
  int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q < n && bc[q] == -1) u[q] = g[q];
}

// ---------------------------------------------------------------------------
// single-workgroup kernel for the coarse levels Ls..ng
// ---------------------------------------------------------------------------
template <typename T>
struct TLevel {
  int n1, n2;
  T aw, an, ac;
  T* gu;                   // global arrays
  T* gb;
  const int8_t* gbc;
  int lds;                 // 1: held in LDS at the offsets below
  uint32_t ou, ob, obc;
};

template <typename T>
struct TailArgs {
  int Ls, ng, k;
  int fix_dn, fix_up;      // levels above the coarsest: fixed down / up sweep counts (0: the exit test)
  T omega;
  double eps;
  const int8_t* bc_img;    // bc arrays of the LDS-resident levels, in their LDS layout at offset 0
  int bc_bytes;
  int* visit;
  double* vres;
  TLevel<T> lv[kMaxLevels + 1];
};

// Sum over the block in a fixed order (wave trees, then a tree over the wave
// sums that every wave evaluates the same way); the caller guarantees a
// barrier since the previous call's reads of ws.
__device__ inline double block_sum_1024(double v, double* ws) {
  v = wave_sum(v);
  if ((threadIdx.x & 63) == 0) ws[threadIdx.x >> 6] = v;
  __syncthreads();
  const int lane = threadIdx.x & 63, nw = blockDim.x >> 6;
  return wave_sum(lane < nw ? ws[lane] : 0.0);
}

// explicit address spaces, so LDS-resident levels use ds_ instructions and
// global ones global_ instructions instead of flat accesses
template <typename T> using lds_p = __attribute__((address_space(3))) T*;
template <typename T> using glb_p = __attribute__((address_space(1))) T*;

// red-black SOR at one tail level; lanes walk columns, waves walk rows
template <typename T, typename PU, typename PB, typename PC>
__device__ void t_smooth(PU u, PB b, PC bc, int n1, int n2, T aw, T an, T ac, T omega, double eps, int dig,
                         int nr, double* ws, int* ksor_out, double* res_out) {
  const int W = n1 + 3;
  const int hw = (n1 + 2) / 2;   // slots of one colour per row
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, nw = blockDim.x >> 6;
  double res_s = 0.0, res = 0.0;
  int ksor = 0;
  constexpr int NB = sizeof(T) == 8 ? 1 : 2;  // rows per wave whose loads are issued together
  for (int k = 1; k <= nr; ++k) {
    for (int color = 1; color >= 0; --color) {  // red (i+j odd) then black
      for (int j0 = 1 + wv; j0 <= n2 + 1; j0 += NB * nw) {
        for (int mm = lane; mm < hw; mm += 64) {
          int x[NB], xb[NB];
          bool act[NB];
          T c[NB], w[NB], e[NB], n[NB], so[NB], bb[NB];
#pragma unroll
          for (int q = 0; q < NB; ++q) {
            const int j = j0 + q * nw;
            const int i = 1 + 2 * mm + ((color ^ ((j + 1) & 1)) & 1);
            act[q] = j <= n2 + 1 && i <= n1 + 1;
            x[q] = act[q] ? j * W + i : W + 1;
            xb[q] = act[q] ? (j - 1) * (n1 + 1) + (i - 1) : 0;
            const int8_t mq = bc[x[q]];
            act[q] = act[q] & (mq <= 0);
            c[q] = u[x[q]]; w[q] = u[x[q] - 1]; e[q] = u[x[q] + 1];
            n[q] = u[x[q] + W]; so[q] = u[x[q] - W]; bb[q] = b[xb[q]];
          }
#pragma unroll
          for (int q = 0; q < NB; ++q)
            if (act[q]) u[x[q]] = sor_update(c[q], bb[q], w[q], n[q], so[q], e[q], aw, an, ac, omega);
        }
      }
      __syncthreads();
    }
    double r = 0.0;
    for (int j0 = 1 + wv; j0 <= n2 + 1; j0 += NB * nw)
      for (int i = 1 + lane; i <= n1 + 1; i += 64) {
        int x[NB], xb[NB];
        bool act[NB];
        T c[NB], w[NB], e[NB], n[NB], so[NB], bb[NB];
#pragma unroll
        for (int q = 0; q < NB; ++q) {
          const int j = j0 + q * nw;
          act[q] = j <= n2 + 1;
          x[q] = act[q] ? j * W + i : W + 1;
          xb[q] = act[q] ? (j - 1) * (n1 + 1) + (i - 1) : 0;
          const int8_t mq = bc[x[q]];
          act[q] = act[q] & (mq == -1);
          c[q] = u[x[q]]; w[q] = u[x[q] - 1]; e[q] = u[x[q] + 1];
          n[q] = u[x[q] + W]; so[q] = u[x[q] - W]; bb[q] = b[xb[q]];
        }
#pragma unroll
        for (int q = 0; q < NB; ++q)
          if (act[q]) r += fabs(sor_resid(c[q], bb[q], w[q], n[q], so[q], e[q], aw, an, ac));
      }
    res = block_sum_1024(r, ws);
    ksor += 1;
    bool stop;
    if (dig == 0) stop = res <= eps;
    else stop = res_s > 0.0 && (res / res_s) <= 0.6;
    if (stop) break;
    res_s = res;
  }
  *ksor_out = ksor;
  *res_out = res;
}

// Red-black sweeps of one tail level with the residual folded into the
// sweeps.  Interior node q = (j-1)(n1+1) + (i-1) is red (i+j odd) exactly
// when q is odd (n1+1 is odd), so a participant handles red nodes 2m+1 and
// black nodes 2m for m = tid + P*r, r < R, with the indices computed once.
// The residual after sweep k is the black nodes' residual, taken right after
// their update (their red neighbours are final), plus the red nodes'
// residual, taken in the red pass of sweep k+1 from the same loads as that
// pass's update; the update is stored only if the exit test (sors.f90:371-377)
// lets sweep k+1 run.  Expressions and exit tests are those of t_smooth.
// WAVE: wave 0 alone (P = 64, no block barriers); else the whole block.
template <bool WAVE>
__device__ __forceinline__ void tail_sync() {
  if (WAVE) {
    // a level may live in global memory: wait for the wave's stores and
    // order the later loads after them, as a block barrier would
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
  } else {
    __syncthreads();
  }
}

// sum over the participants, the same value in every participant; ws holds
// two banks of 16 wave sums used alternately (par), so a bank is rewritten
// only after the barriers of a whole sweep
template <bool WAVE>
__device__ __forceinline__ double tail_sum(double v, double* ws, int par) {
  v = wave_sum(v);
  if (WAVE) return v;
  const int w = threadIdx.x >> 6, nw = blockDim.x >> 6;
  double* bank = ws + 16 * par;
  if ((threadIdx.x & 63) == 0) bank[w] = v;
  __syncthreads();
  double t = 0.0;
  for (int q = 0; q < nw; ++q) t += bank[q];
  return t;
}

template <typename T, int R, bool WAVE, typename PU, typename PB, typename PC>
__device__ __forceinline__ void t_smooth_fused(PU u, PB b, PC bc, int n1, int n2, T aw, T an, T ac, T omega, double eps,
                               int dig, int nr, double* ws, int* ksor_out, double* res_out, int nfix) {
  const int W = n1 + 3, n1p = n1 + 1;
  const int nint = n1p * (n2 + 1);
  const int P = WAVE ? 64 : (int)blockDim.x;
  const int tid = WAVE ? (int)(threadIdx.x & 63) : (int)threadIdx.x;
  int xr[R], qr[R], xk[R], qk[R];
  bool ur[R], rr[R], uk[R], rk[R];
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const int m = tid + P * r;
    int q = 2 * m + 1;
    bool ok = q < nint;
    q = ok ? q : 0;
    int x = (q / n1p + 1) * W + q % n1p + 1;
    int8_t mq = bc[x];
    xr[r] = x; qr[r] = q; ur[r] = ok && mq <= 0; rr[r] = ok && mq == -1;
    q = 2 * m;
    ok = q < nint;
    q = ok ? q : 0;
    x = (q / n1p + 1) * W + q % n1p + 1;
    mq = bc[x];
    xk[r] = x; qk[r] = q; uk[r] = ok && mq <= 0; rk[r] = ok && mq == -1;
  }
  // in block mode the last item of a level is often held by a few waves
  // only: waves that hold none skip it (a wave-uniform branch after the
  // other items' loads, so those still issue together)
  bool wr[R], wk[R];
#pragma unroll
  for (int r = 0; r < R; ++r) {
    wr[r] = WAVE || r + 1 < R || __ballot(2 * (tid + P * r) + 1 < nint) != 0;
    wk[r] = WAVE || r + 1 < R || __ballot(2 * (tid + P * r) < nint) != 0;
  }
  T cand[R];
  auto red_pass = [&](bool want_res) {
    double rs = 0.0;
    T c[R], w[R], e[R], nn[R], so[R], bb[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
      if (!wr[r]) continue;
      c[r] = u[xr[r]]; w[r] = u[xr[r] - 1]; e[r] = u[xr[r] + 1];
      nn[r] = u[xr[r] + W]; so[r] = u[xr[r] - W]; bb[r] = b[qr[r]];
    }
#pragma unroll
    for (int r = 0; r < R; ++r) {
      if (!wr[r]) continue;
      if (want_res && rr[r]) rs += fabs(sor_resid(c[r], bb[r], w[r], nn[r], so[r], e[r], aw, an, ac));
      cand[r] = sor_update(c[r], bb[r], w[r], nn[r], so[r], e[r], aw, an, ac, omega);
    }
    return rs;
  };
  auto store_red = [&]() {
#pragma unroll
    for (int r = 0; r < R; ++r)
      if (ur[r]) u[xr[r]] = cand[r];
  };
  auto black_pass = [&]() {
    double rs = 0.0;
    T c[R], w[R], e[R], nn[R], so[R], bb[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
      if (!wk[r]) continue;
      c[r] = u[xk[r]]; w[r] = u[xk[r] - 1]; e[r] = u[xk[r] + 1];
      nn[r] = u[xk[r] + W]; so[r] = u[xk[r] - W]; bb[r] = b[qk[r]];
    }
#pragma unroll
    for (int r = 0; r < R; ++r) {
      if (!wk[r]) continue;
      T v = c[r];
      if (uk[r]) {
        v = sor_update(c[r], bb[r], w[r], nn[r], so[r], e[r], aw, an, ac, omega);
        u[xk[r]] = v;
      }
      if (rk[r]) rs += fabs(sor_resid(v, bb[r], w[r], nn[r], so[r], e[r], aw, an, ac));
    }
    return rs;
  };
  if (nfix > 0) {  // a fixed count without residuals or exit test (mixed form)
    for (int k = 1; k <= nfix; ++k) {
      red_pass(false);
      store_red();
      tail_sync<WAVE>();
      black_pass();
      tail_sync<WAVE>();
    }
    *ksor_out = nfix;
    *res_out = 0.0;
    return;
  }
  // sweep 1
  red_pass(false);
  store_red();
  tail_sync<WAVE>();
  double rb = black_pass();
  tail_sync<WAVE>();
  int k = 1;
  double res_s = 0.0, res = 0.0;
  for (;;) {
    const double rs = red_pass(true);
    res = tail_sum<WAVE>(rs + rb, ws, k & 1);
    bool stop;
    if (dig == 0) stop = res <= eps;
    else stop = res_s > 0.0 && (res / res_s) <= 0.6;
    if (stop || k >= nr) break;
    res_s = res;
    store_red();
    tail_sync<WAVE>();
    rb = black_pass();
    tail_sync<WAVE>();
    ++k;
  }
  // the last red pass only read; the next phase may overwrite nothing it read
  tail_sync<WAVE>();
  *ksor_out = k;
  *res_out = res;
}

// items per participant and colour of a level, 0 if above kFusedMaxR
constexpr int kFusedMaxR = 3;
static_assert((kWaveNodes + 1) / 2 <= 64 * kFusedMaxR, "wave-mode levels must fit the fused smoother");
__device__ __forceinline__ int fused_items(int n1, int n2, int P) {
  const int half = ((n1 + 1) * (n2 + 1) + 1) / 2;
  const int R = (half + P - 1) / P;
  return R <= kFusedMaxR ? R : 0;
}

// LDS-resident levels run the fused smoother with one or three items per
// participant (two-item levels take the three-item form with one masked);
// levels in global memory, or larger than three items, run t_smooth.  Few
// instantiations keep the kernel's code small: it is fetched anew by the CU
// that runs each launch.
template <typename T, bool WAVE, typename PU, typename PB, typename PC>
__device__ __forceinline__ bool t_smooth_fused_any(PU u, PB b, PC bc, int n1, int n2, T aw, T an, T ac, T omega,
                                                   double eps, int dig, int nr, double* ws, int* ks, double* res,
                                                   int nfix) {
  const int R = fused_items(n1, n2, WAVE ? 64 : (int)blockDim.x);
  if (R == 1) {
    t_smooth_fused<T, 1, WAVE>(u, b, bc, n1, n2, aw, an, ac, omega, eps, dig, nr, ws, ks, res, nfix);
    return true;
  }
  if (R == 2 || R == 3) {
    t_smooth_fused<T, 3, WAVE>(u, b, bc, n1, n2, aw, an, ac, omega, eps, dig, nr, ws, ks, res, nfix);
    return true;
  }
  return false;
}

// nfix > 0: exactly nfix sweeps, no residuals (LDS-resident levels of the
// mixed form); 0: the reference's exit test
template <typename T, bool WAVE = false>
__device__ __forceinline__ void smooth_level(const TLevel<T>& L, T* u, T* b, const int8_t* bc, T omega, double eps,
                                             int dig, int nr, double* ws, int* ks, double* res, int nfix = 0) {
  if (L.lds && t_smooth_fused_any<T, WAVE>((lds_p<T>)u, (lds_p<const T>)b, (lds_p<const int8_t>)bc, L.n1, L.n2,
                                           L.aw, L.an, L.ac, omega, eps, dig, nr, ws, ks, res, nfix))
    return;
  if constexpr (!WAVE) {
    if (L.lds)
      t_smooth<T>((lds_p<T>)u, (lds_p<const T>)b, (lds_p<const int8_t>)bc, L.n1, L.n2, L.aw, L.an, L.ac, omega,
                  eps, dig, nr, ws, ks, res);
    else
      t_smooth<T>((glb_p<T>)u, (glb_p<const T>)b, (glb_p<const int8_t>)bc, L.n1, L.n2, L.aw, L.an, L.ac, omega,
                  eps, dig, nr, ws, ks, res);
  }
}

// restriction of level l onto l+1 and prolongation of l+1 onto l by the
// participants (tid, P); LDS-resident pairs of levels use LDS-typed pointers
template <typename T, typename PU, typename PB, typename PC, typename PUC, typename PBC, typename PCC>
__device__ __forceinline__ void tail_restrict_p(const TLevel<T>& F, const TLevel<T>& C, PU uf, PB bf, PC bcf,
                                                PUC e2h, PBC r2h, PCC bcc, int tid, int P) {
  for (int q = tid; q < (C.n1 + 3) * (C.n2 + 3); q += P) e2h[q] = (T)0;
  for (int q = tid; q < (C.n1 + 1) * (C.n2 + 1); q += P) {
    const int i = q % (C.n1 + 1) + 1, j = q / (C.n1 + 1) + 1;
    T out = (T)0;
    if (bcc[j * (C.n1 + 3) + i] <= 0) out = restrict_node<T>(uf, bf, bcf, i, j, F.n1, F.n2, F.aw, F.an, F.ac);
    r2h[q] = out;
  }
}

template <typename T>
__device__ __forceinline__ void tail_restrict(const TLevel<T>& F, const TLevel<T>& C, T* uf, T* bf,
                                              const int8_t* bcf, T* e2h, T* r2h, const int8_t* bcc, int tid,
                                              int P) {
  if (F.lds && C.lds)
    tail_restrict_p<T>(F, C, (lds_p<const T>)uf, (lds_p<const T>)bf, (lds_p<const int8_t>)bcf, (lds_p<T>)e2h,
                       (lds_p<T>)r2h, (lds_p<const int8_t>)bcc, tid, P);
  else
    tail_restrict_p<T>(F, C, (const T*)uf, (const T*)bf, bcf, e2h, r2h, bcc, tid, P);
}

template <typename T, typename PU, typename PC, typename PE>
__device__ __forceinline__ void tail_prolong_p(const TLevel<T>& F, const TLevel<T>& C, PU uf, PC bcf, PE uc,
                                               int tid, int P) {
  for (int q = tid; q < (C.n1 + 1) * (C.n2 + 1); q += P)
    prolong_node<T>(uf, bcf, F.n1, uc, C.n1, q % (C.n1 + 1) + 1, q / (C.n1 + 1) + 1);
}

template <typename T>
__device__ __forceinline__ void tail_prolong(const TLevel<T>& F, const TLevel<T>& C, T* uf, const int8_t* bcf,
                                             const T* uc, int tid, int P) {
  if (F.lds && C.lds)
    tail_prolong_p<T>(F, C, (lds_p<T>)uf, (lds_p<const int8_t>)bcf, (lds_p<const T>)uc, tid, P);
  else
    tail_prolong_p<T>(F, C, uf, bcf, uc, tid, P);
}

// Block copies of `bytes` bytes between 16-byte aligned global and LDS
// buffers: every thread issues its 16-byte loads (up to four) before it
// stores, so a copy costs one memory latency instead of one per loop trip.
typedef int v4i __attribute__((ext_vector_type(4)));

template <typename VD, typename VS, typename BD, typename BS>
__device__ __forceinline__ void copy_block(VD d4, VS s4, BD d1, BS s1, int bytes) {
  const int nv = bytes >> 4;
  for (int q0 = threadIdx.x; q0 < nv; q0 += 4 * (int)blockDim.x) {
    v4i v[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const int q = q0 + k * (int)blockDim.x;
      if (q < nv) v[k] = s4[q];
    }
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const int q = q0 + k * (int)blockDim.x;
      if (q < nv) d4[q] = v[k];
    }
  }
  for (int q = (nv << 4) + (int)threadIdx.x; q < bytes; q += blockDim.x) d1[q] = s1[q];
}

__device__ __forceinline__ void copy_g2l(char* lds_dst, const void* src, int bytes) {
  copy_block((lds_p<v4i>)lds_dst, (glb_p<const v4i>)src, (lds_p<char>)lds_dst, (glb_p<const char>)src, bytes);
}

__device__ __forceinline__ void copy_l2g(void* dst, const char* lds_src, int bytes) {
  copy_block((glb_p<v4i>)dst, (lds_p<const v4i>)lds_src, (glb_p<char>)dst, (lds_p<const char>)lds_src, bytes);
}

// three global-to-LDS copies whose loads are all issued before any store
// (segments of up to 4 x 16 x blockDim bytes take one memory latency)
__device__ __forceinline__ void copy3_g2l(char* d0, const void* s0, int n0, char* d1, const void* s1, int n1,
                                          char* d2, const void* s2, int n2) {
  lds_p<v4i> d[3] = {(lds_p<v4i>)d0, (lds_p<v4i>)d1, (lds_p<v4i>)d2};
  glb_p<const v4i> sv[3] = {(glb_p<const v4i>)s0, (glb_p<const v4i>)s1, (glb_p<const v4i>)s2};
  const int nb[3] = {n0, n1, n2};
  const int P = blockDim.x, t = threadIdx.x;
  v4i v[3][4];
#pragma unroll
  for (int g = 0; g < 3; ++g)
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const int q = t + k * P;
      if (q < (nb[g] >> 4)) v[g][k] = sv[g][q];
    }
#pragma unroll
  for (int g = 0; g < 3; ++g)
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const int q = t + k * P;
      if (q < (nb[g] >> 4)) d[g][q] = v[g][k];
    }
  // what is left: vectors beyond 4 per thread, then the last bytes
  char* dc[3] = {d0, d1, d2};
  const void* sc[3] = {s0, s1, s2};
#pragma unroll
  for (int g = 0; g < 3; ++g) {
    for (int q = t + 4 * P; q < (nb[g] >> 4); q += P) d[g][q] = sv[g][q];
    for (int q = (nb[g] & ~15) + t; q < nb[g]; q += P)
      ((lds_p<char>)dc[g])[q] = ((glb_p<const char>)sc[g])[q];
  }
}

template <typename T>
__global__ __launch_bounds__(kTailThreads) void k_tail(TailArgs<T> A) {
  extern __shared__ __align__(16) char smem[];
  __shared__ double ws[32];
  __shared__ T* sU[kMaxLevels + 1];
  __shared__ T* sB[kMaxLevels + 1];
  __shared__ const int8_t* sBC[kMaxLevels + 1];
  const int Ls = A.Ls, ng = A.ng;
  if (threadIdx.x == 0) {
    for (int l = Ls; l <= ng; ++l) {
      const TLevel<T>& L = A.lv[l];
      sU[l] = L.lds ? (T*)(smem + L.ou) : L.gu;
      sB[l] = L.lds ? (T*)(smem + L.ob) : L.gb;
      sBC[l] = L.lds ? (const int8_t*)(smem + L.obc) : L.gbc;
    }
  }
  __syncthreads();
  {
    // one memory latency for all of it: the bc image of the resident levels
    // and, when resident, u and b of the first level
    const TLevel<T>& L = A.lv[Ls];
    const int ub = L.lds ? (L.n1 + 3) * (L.n2 + 3) * (int)sizeof(T) : 0;
    const int bb = L.lds ? (L.n1 + 1) * (L.n2 + 1) * (int)sizeof(T) : 0;
    copy3_g2l(smem, A.bc_img, A.bc_bytes, smem + L.ou, L.gu, ub, smem + L.ob, L.gb, bb);
  }
  __syncthreads();
  int ks;
  double res = 0.0;
  // levels from Lw down are run by wave 0 alone (no block barriers)
  int Lw = ng + 1;
  for (int l = ng; l >= Ls && A.lv[l].lds && (A.lv[l].n1 + 1) * (A.lv[l].n2 + 1) <= kWaveNodes; --l) Lw = l;
  const int tid = threadIdx.x, P = blockDim.x;
  for (int l = Ls; l < ng && l < Lw; ++l) {  // down leg: ig = l
    const TLevel<T>& F = A.lv[l];
    smooth_level<T>(F, sU[l], sB[l], sBC[l], A.omega, A.eps, l - ng, kNr, ws, &ks, &res, A.fix_dn);
    if (tid == 0) { A.visit[l] = ks; A.vres[l] = res; }
    tail_restrict<T>(F, A.lv[l + 1], sU[l], sB[l], sBC[l], sU[l + 1], sB[l + 1], sBC[l + 1], tid, P);
    __syncthreads();
  }
  if (Lw <= ng) {
    if (tid < 64) {
      for (int l = Lw; l < ng; ++l) {
        const TLevel<T>& F = A.lv[l];
        smooth_level<T, true>(F, sU[l], sB[l], sBC[l], A.omega, A.eps, l - ng, kNr, ws, &ks, &res, A.fix_dn);
        if (tid == 0) { A.visit[l] = ks; A.vres[l] = res; }
        tail_restrict<T>(F, A.lv[l + 1], sU[l], sB[l], sBC[l], sU[l + 1], sB[l + 1], sBC[l + 1], tid, 64);
        tail_sync<true>();
      }
      smooth_level<T, true>(A.lv[ng], sU[ng], sB[ng], sBC[ng], A.omega, A.eps, 0, A.k == 1 ? 1000000 : kNr, ws,
                            &ks, &res);
      if (tid == 0) { A.visit[ng] = ks; A.vres[ng] = res; }
      for (int l = ng - 1; l >= Lw; --l) {
        const TLevel<T>& F = A.lv[l];
        tail_prolong<T>(F, A.lv[l + 1], sU[l], sBC[l], (const T*)sU[l + 1], tid, 64);
        tail_sync<true>();
        smooth_level<T, true>(F, sU[l], sB[l], sBC[l], A.omega, A.eps, ng - l, kNr, ws, &ks, &res, A.fix_up);
        if (tid == 0) { A.visit[2 * ng - l] = ks; A.vres[2 * ng - l] = res; }
      }
    }
    __syncthreads();
  } else {  // coarsest: ig = ng; the first V-cycle iterates to eps (sors.f90:63-64)
    smooth_level<T>(A.lv[ng], sU[ng], sB[ng], sBC[ng], A.omega, A.eps, 0, A.k == 1 ? 1000000 : kNr, ws, &ks,
                    &res);
    if (tid == 0) { A.visit[ng] = ks; A.vres[ng] = res; }
  }
  for (int l = std::min(Lw, ng) - 1; l >= Ls; --l) {  // up leg: ig = 2ng - l
    const TLevel<T>& F = A.lv[l];
    tail_prolong<T>(F, A.lv[l + 1], sU[l], sBC[l], (const T*)sU[l + 1], tid, P);
    __syncthreads();
    smooth_level<T>(F, sU[l], sB[l], sBC[l], A.omega, A.eps, ng - l, kNr, ws, &ks, &res, A.fix_up);
    if (tid == 0) { A.visit[2 * ng - l] = ks; A.vres[2 * ng - l] = res; }
  }
  const TLevel<T>& L = A.lv[Ls];
  if (L.lds) copy_l2g(L.gu, smem + L.ou, (L.n1 + 3) * (L.n2 + 3) * (int)sizeof(T));
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
template <typename T>
T* rot_of(Level& L, int r) { return (T*)L.rot[r]; }

template <typename T>
T* u_of(Level& L, bool f32) { return f32 ? (T*)L.uf : (T*)L.u; }

template <typename T>
T* b_of(Level& L, bool f32) { return f32 ? (T*)L.bf : (T*)L.b; }

int* d_visit_of(lepic_ctx* ctx) { return (int*)(ctx->d_scalar + 64); }
double* d_vres_of(lepic_ctx* ctx) { return ctx->d_scalar + 32; }
int* d_status_of(lepic_ctx* ctx) { return (int*)(ctx->d_scalar + 64) + kVisitLog; }

VisitRef visit_ref(lepic_ctx* ctx, int l, int epoch, int dig, int nr, int ig, double eps, int nexec,
                   int klast0 = -1, int fixed = 0) {
  Level& L = ctx->lv[l];
  VisitRef v;
  v.ctrl = ctx->d_ctrl + l;
  v.nb = L.nbx * L.nby;
  v.partials = L.partials;
  v.epoch = epoch; v.dig = dig; v.nr = nr; v.ig = ig; v.eps = eps;
  v.nexec = nexec;
  v.klast0 = klast0 >= 0 ? klast0 : ((nexec - 1) / kS) * kS;
  v.fixed = fixed;
  v.visit = d_visit_of(ctx);
  v.vres = d_vres_of(ctx);
  return v;
}

// Start of the last launch of a visit whose launches cover sweeps
// k0+1 .. k_end, the first of them `first` sweeps and the others kS.
int last_launch_start(int k0, int k_end, int first) {
  int k = k0, last = k0;
  for (int size = std::min(first, kS); k < k_end; size = kS) {
    last = k;
    k += std::min(size, k_end - k);
  }
  return last;
}

// How a multi-block level's visit runs.  FP64: the reference's exit test
// (sors.f90:371-377) decides, the first launch `first` sweeps and
// continuations up to kNr.  Mixed: the visit runs its fixed count
// (kFineDn / kFineUp on the finest level, kFixDn / kFixUp below) and no
// continuation launch tests anything.
struct VisitPlan {
  int k_end, first, klast0, fixed;
};

VisitPlan visit_plan(const Level& L, bool up, bool f32) {
  const int f = up ? L.first_up : L.first_dn;
  if (!f32) return VisitPlan{kNr, f, last_launch_start(0, kNr, f), 0};
  const int first = std::min(f, kS);
  return VisitPlan{f, first, last_launch_start(0, f, first), 1};
}

// the launch size KS (halo 2 KS + 1) is kexec, or kexec + 1 for a launch
// that restricts (its full weighting reads the state two nodes past the tile)
template <typename T, int kTX, int TY, int NT>
void launch_sweeps(lepic_ctx* ctx, const Level& L, const SweepArgs<T>& A, int k, int kexec, int kchk) {
  const dim3 g(L.nbx, L.nby);
  if (A.rst) {
    // mixed form only: FP64 restricts in its own kernel
    if constexpr (sizeof(T) == 4) {
      const int ks = kexec + 1;
      if (ks <= 3) hipLaunchKernelGGL((k_sweeps<T, kTX, TY, NT, 3, true>), g, dim3(NT), 0, ctx->stream, A, k, kexec, kchk);
      else if (ks <= 4) hipLaunchKernelGGL((k_sweeps<T, kTX, TY, NT, 4, true>), g, dim3(NT), 0, ctx->stream, A, k, kexec, kchk);
      else if (ks <= kS) hipLaunchKernelGGL((k_sweeps<T, kTX, TY, NT, kS, true>), g, dim3(NT), 0, ctx->stream, A, k, kexec, kchk);
      else hipLaunchKernelGGL((k_sweeps<T, kTX, TY, NT, kS + 1, true>), g, dim3(NT), 0, ctx->stream, A, k, kexec, kchk);
    }
    return;
  }
  if (kexec <= 2) hipLaunchKernelGGL((k_sweeps<T, kTX, TY, NT, 2>), g, dim3(NT), 0, ctx->stream, A, k, kexec, kchk);
  else if (kexec <= 3) hipLaunchKernelGGL((k_sweeps<T, kTX, TY, NT, 3>), g, dim3(NT), 0, ctx->stream, A, k, kexec, kchk);
  else if (kexec <= 4) hipLaunchKernelGGL((k_sweeps<T, kTX, TY, NT, 4>), g, dim3(NT), 0, ctx->stream, A, k, kexec, kchk);
  else hipLaunchKernelGGL((k_sweeps<T, kTX, TY, NT, kS>), g, dim3(NT), 0, ctx->stream, A, k, kexec, kchk);
}

// launches of one visit of a multi-block level: sweeps k0+1 .. k_end, the
// first launch `first` sweeps (one more than the same visit took in the
// previous V-cycle: most visits then finish in one launch of a small
// halo), the others kS; each continuation first tests the previous launch's
// sweeps and returns at once when the visit has stopped there
template <typename T>
int launch_visit(lepic_ctx* ctx, int l, bool f32, int epoch, int dig, int nr, double eps, int k0, int k_end,
                 int first = kS, int fixed = 0, bool pro = false, bool rst = false) {
  Level& L = ctx->lv[l];
  SweepArgs<T> A;
  A.rst = 0;
  if (rst) {  // the down visit's last launch restricts onto level l + 1
    Level& C = ctx->lv[l + 1];
    A.r2h = b_of<T>(C, f32);
    A.bc2h = C.bc;
    for (int r = 0; r < kRot; ++r) A.rcrot[r] = C.mb ? rot_of<T>(C, r) : u_of<T>(C, f32);
    A.rcctrl = C.mb ? ctx->d_ctrl + (l + 1) : nullptr;
  } else {
    A.r2h = nullptr;
    A.bc2h = nullptr;
    for (int r = 0; r < kRot; ++r) A.rcrot[r] = nullptr;
    A.rcctrl = nullptr;
  }
  A.pro = pro ? 1 : 0;
  A.pro_mb = 0;
  A.pm1 = L.n1 / 2;
  A.pvc = VisitRef{};
  if (pro) {  // the up visit's first launch takes the prolongation from level l + 1
    Level& C = ctx->lv[l + 1];
    for (int r = 0; r < kRot; ++r) A.pcrot[r] = C.mb ? rot_of<T>(C, r) : u_of<T>(C, f32);
    A.pro_mb = C.mb ? 1 : 0;
    if (C.mb) {
      const VisitPlan pc = visit_plan(C, true, f32);
      A.pvc = visit_ref(ctx, l + 1, epoch, ctx->ng - (l + 1), kNr, 2 * ctx->ng - (l + 1), eps, pc.k_end, pc.klast0,
                        pc.fixed);
    }
  } else {
    for (int r = 0; r < kRot; ++r) A.pcrot[r] = nullptr;
  }
  for (int r = 0; r < kRot; ++r) A.rot[r] = rot_of<T>(L, r);
  A.b = b_of<T>(L, f32);
  A.bc = L.bc;
  A.n1 = L.n1; A.n2 = L.n2;
  A.aw = (T)L.aw; A.an = (T)L.an; A.ac = (T)L.ac;
  A.omega = (f32 && l == 1) ? (T)kFineOmega : (T)1.4;   // iterative_solver.f90:27
  A.ctrl = ctx->d_ctrl + l;
  A.partials = L.partials;
  A.nbx = L.nbx; A.nby = L.nby;
  A.epoch = epoch; A.dig = dig; A.nr = nr; A.eps = eps;
  A.fixed = fixed;
  // the mixed form's visits run fixed counts: their residuals feed nothing
  A.nores = (f32 && fixed) ? 1 : 0;
  int kchk = 0;
  for (int k = k0, size = std::min(std::max(first, 1), kS); k < k_end; size = kS) {
    const int kexec = std::min(size, k_end - k);
    A.rst = (rst && k + kexec == k_end) ? 1 : 0;
    if (L.small_tiles) launch_sweeps<T, kSmallTX, SmallTileY<T>::v, kSmallThreads>(ctx, L, A, k, kexec, kchk);
    else launch_sweeps<T, kTX, TileY<T>::v, kSweepThreads>(ctx, L, A, k, kexec, kchk);
    kchk = k + 1;
    k += kexec;
  }
  LEPIC_HIP(hipGetLastError());
  return 0;
}

template <typename T>
int launch_restrict(lepic_ctx* ctx, int l, bool f32, int epoch, double eps) {
  const int ng = ctx->ng;
  Level& F = ctx->lv[l];
  Level& C = ctx->lv[l + 1];
  RestrictArgs<T> A;
  const VisitPlan pf = visit_plan(F, false, std::is_same<T, float>::value);
  A.vf = visit_ref(ctx, l, epoch, l - ng, kNr, l, eps, pf.k_end, pf.klast0, pf.fixed);
  for (int r = 0; r < kRot; ++r) A.frot[r] = rot_of<T>(F, r);
  A.fb = b_of<T>(F, f32);
  A.fbc = F.bc;
  A.n1 = F.n1; A.n2 = F.n2;
  A.aw = (T)F.aw; A.an = (T)F.an; A.ac = (T)F.ac;
  A.r2h = b_of<T>(C, f32);
  A.bc2h = C.bc;
  for (int r = 0; r < kRot; ++r) A.crot[r] = C.mb ? rot_of<T>(C, r) : u_of<T>(C, f32);
  A.cctrl = C.mb ? ctx->d_ctrl + (l + 1) : nullptr;
  A.cepoch = epoch;
  const dim3 g((unsigned)((C.n1 + 1 + kRcx - 1) / kRcx), (unsigned)((C.n2 + 1 + kRcy - 1) / kRcy));
  hipLaunchKernelGGL(k_restrict<T>, g, dim3(kRestrictThreads), 0, ctx->stream, A);
  LEPIC_HIP(hipGetLastError());
  return 0;
}

template <typename T>
int launch_prolong(lepic_ctx* ctx, int l, bool f32, int epoch_up, double eps) {
  const int ng = ctx->ng;
  Level& F = ctx->lv[l];
  Level& C = ctx->lv[l + 1];
  ProlongArgs<T> A;
  for (int r = 0; r < kRot; ++r) A.frot[r] = rot_of<T>(F, r);
  A.fctrl = ctx->d_ctrl + l;
  A.fepoch = epoch_up;
  A.bc = F.bc;
  A.n1 = F.n1; A.n2 = F.n2;
  for (int r = 0; r < kRot; ++r) A.crot[r] = C.mb ? rot_of<T>(C, r) : u_of<T>(C, f32);
  A.coarse_mb = C.mb ? 1 : 0;
  if (C.mb)
  {
    const VisitPlan pc = visit_plan(C, true, std::is_same<T, float>::value);
    A.vc = visit_ref(ctx, l + 1, epoch_up, ng - (l + 1), kNr, 2 * ng - (l + 1), eps, pc.k_end, pc.klast0, pc.fixed);
  }
  else A.vc = VisitRef{};
  dim3 g((C.n1 + 1 + 127) / 128, std::min(C.n2 + 1, kRowBlocks));
  hipLaunchKernelGGL(k_prolong<T>, g, dim3(128), 0, ctx->stream, A);
  LEPIC_HIP(hipGetLastError());
  return 0;
}

template <typename T>
int launch_tail(lepic_ctx* ctx, bool f32, int k, double eps) {
  TailArgs<T> A{};
  A.Ls = ctx->Lc; A.ng = ctx->ng; A.k = k; A.omega = (T)1.4; A.eps = eps;
  // mixed form: the tail levels above the coarsest run the counts the exit
  // test gives them on plasma solves, without the per-sweep sums the test
  // needs
  A.fix_dn = f32 ? kFixDn : 0;
  A.fix_up = f32 ? kFixUp : 0;
  A.visit = d_visit_of(ctx);
  A.vres = d_vres_of(ctx);
  auto al = [](size_t x) { return (x + 15) & ~size_t(15); };
  // residency, coarsest first, until the LDS budget is spent; the LDS holds
  // the bc arrays of the resident levels first, then their u and b
  int Lr = ctx->ng + 1;   // finest resident level
  size_t need_all = 0;
  for (int l = ctx->ng; l >= ctx->Lc; --l) {
    const Level& L = ctx->lv[l];
    const size_t need = al((size_t)(L.n1 + 3) * (L.n2 + 3) * sizeof(T)) +
                        al((size_t)(L.n1 + 1) * (L.n2 + 1) * sizeof(T)) + al((size_t)(L.n1 + 3) * (L.n2 + 3));
    if (need_all + need > kTailLds) break;
    need_all += need;
    Lr = l;
  }
  size_t off = 0;
  for (int l = ctx->ng; l >= Lr; --l) {
    const Level& L = ctx->lv[l];
    A.lv[l].obc = (uint32_t)off;
    off = al(off + (size_t)(L.n1 + 3) * (L.n2 + 3));
  }
  const size_t bc_bytes = off;
  for (int l = ctx->ng; l >= ctx->Lc; --l) {
    Level& L = ctx->lv[l];
    TLevel<T>& C = A.lv[l];
    C.n1 = L.n1; C.n2 = L.n2; C.aw = (T)L.aw; C.an = (T)L.an; C.ac = (T)L.ac;
    C.gu = u_of<T>(L, f32); C.gb = b_of<T>(L, f32); C.gbc = L.bc;
    C.lds = l >= Lr ? 1 : 0;
    if (C.lds) {
      C.ou = (uint32_t)off; off = al(off + (size_t)(L.n1 + 3) * (L.n2 + 3) * sizeof(T));
      C.ob = (uint32_t)off; off = al(off + (size_t)(L.n1 + 1) * (L.n2 + 1) * sizeof(T));
    }
  }
  // the bc image, rebuilt when the boundary or the layout changes
  const int key = Lr * 2 + (f32 ? 1 : 0);
  if (ctx->tail_bc_key != key) {
    if (bc_bytes > ctx->tail_bc_cap) {
      if (ctx->d_tail_bc) LEPIC_HIP(hipFree(ctx->d_tail_bc));
      LEPIC_HIP(hipMalloc(&ctx->d_tail_bc, bc_bytes));
      ctx->tail_bc_cap = bc_bytes;
    }
    for (int l = ctx->ng; l >= Lr; --l) {
      const Level& L = ctx->lv[l];
      LEPIC_HIP(hipMemcpyAsync(ctx->d_tail_bc + A.lv[l].obc, L.bc, (size_t)(L.n1 + 3) * (L.n2 + 3),
                               hipMemcpyDeviceToDevice, ctx->stream));
    }
    ctx->tail_bc_key = key;
  }
  A.bc_img = ctx->d_tail_bc;
  A.bc_bytes = (int)bc_bytes;
  static bool attr_set = false;   // per instantiation; the attribute is per kernel
  if (!attr_set) {
    LEPIC_HIP(hipFuncSetAttribute((const void*)k_tail<T>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                  (int)kTailLds));
    attr_set = true;
  }
  hipLaunchKernelGGL(k_tail<T>, dim3(1), dim3(kTailThreads), off, ctx->stream, A);
  LEPIC_HIP(hipGetLastError());
  return 0;
}

// One V-cycle (mg.f90:85-225) at precision T: levels 1..Ls-1 multi-block,
// Ls..ng in k_tail; then the end-of-cycle kernel of the mode.
template <typename T>
int vcycle(lepic_ctx* ctx, bool f32, int k, double eps) {
  const int ng = ctx->ng, Ls = ctx->Lc;
  const int e_dn = ctx->mg_epoch + 1, e_up = ctx->mg_epoch + 2;
  ctx->mg_epoch += 2;
  for (int l = 1; l < Ls; ++l) {  // down leg
    // FP64: the restriction kernel; mixed: folded into the visit's last launch
    const VisitPlan pl = visit_plan(ctx->lv[l], false, f32);
    if (int rc = launch_visit<T>(ctx, l, f32, e_dn, l - ng, kNr, eps, 0, pl.k_end, pl.first, pl.fixed, false, f32))
      return rc;
    if constexpr (sizeof(T) == 8)
      if (int rc = launch_restrict<T>(ctx, l, f32, e_dn, eps)) return rc;
  }
  if (int rc = launch_tail<T>(ctx, f32, k, eps)) return rc;
  for (int l = Ls - 1; l >= 1; --l) {  // up leg
    // FP64: the prolongation kernel; mixed: folded into the up visit's first launch
    if constexpr (sizeof(T) == 8)
      if (int rc = launch_prolong<T>(ctx, l, f32, e_up, eps)) return rc;
    const VisitPlan pl = visit_plan(ctx->lv[l], true, f32);
    if (int rc = launch_visit<T>(ctx, l, f32, e_up, ng - l, kNr, eps, 0, pl.k_end, pl.first, pl.fixed, f32)) return rc;
  }
  Level& L1 = ctx->lv[1];
  const unsigned gx = (unsigned)((L1.n1 + 1 + 255) / 256);
  if (!f32) {
    if (L1.mb) {
      Finish64Args A;
      const VisitPlan p1 = visit_plan(L1, true, false);
      A.v = visit_ref(ctx, 1, e_up, ng - 1, kNr, 2 * ng - 1, eps, p1.k_end, p1.klast0, p1.fixed);
      for (int r = 0; r < kRot; ++r) A.rot[r] = rot_of<double>(L1, r);
      A.n1 = L1.n1; A.n2 = L1.n2;
      A.status = d_status_of(ctx);
      hipLaunchKernelGGL(k_finish64, dim3(gx, std::min(L1.n2 + 1, kRowBlocks)), dim3(256), 0, ctx->stream, A);
    }
  } else {
    AddCorrArgs A;
    A.phi = ctx->phi;
    for (int r = 0; r < kRot; ++r) A.rot[r] = L1.mb ? rot_of<float>(L1, r) : L1.uf;
    A.bc = ctx->bc;
    A.n1 = L1.n1; A.n2 = L1.n2;
    A.mb = L1.mb ? 1 : 0;
    const VisitPlan p1 = visit_plan(L1, true, true);
    A.v = L1.mb ? visit_ref(ctx, 1, e_up, ng - 1, kNr, 2 * ng - 1, eps, p1.k_end, p1.klast0, p1.fixed) : VisitRef{};
    hipLaunchKernelGGL(k_add_corr, dim3(gx, std::min(L1.n2 + 1, kRowBlocks)), dim3(256), 0, ctx->stream, A);
  }
  LEPIC_HIP(hipGetLastError());
  return 0;
}

double alpha_of(lepic_ctx* ctx) {
  // iterative_solver.f90:22-23
  return std::fabs((double)((ctx->g.n1 + 1) * (ctx->g.n2 + 1)) * ctx->lv[1].ac) * 1.0;
}

int resid_blocks(lepic_ctx* ctx) { return std::min(ctx->g.n2 + 1, kResidBlocks); }

// FP64 residual of phi; fills the FP32 right-hand side (and clears the
// correction the next visit starts from) when fill_f32
int launch_resid64(lepic_ctx* ctx, bool fill_f32, const double* u = nullptr) {
  Level& L = ctx->lv[1];
  float* const* erot = nullptr;
  if (fill_f32) erot = (float* const*)ctx->d_erot;
  hipLaunchKernelGGL(k_resid64, dim3(resid_blocks(ctx)), dim3(256), 0, ctx->stream, u ? u : (const double*)ctx->phi,
                     (const double*)ctx->rhs, (const int8_t*)ctx->bc, L.n1, L.n2, L.aw, L.an, L.ac,
                     fill_f32 ? L.bf : nullptr, erot, L.mb ? (const LevelCtrl*)(ctx->d_ctrl + 1) : nullptr,
                     ctx->mg_epoch + 1, ctx->d_partials);
  LEPIC_HIP(hipGetLastError());
  return 0;
}

double sum_resid_partials(const double* p, int nb) {
  double s = 0.0;
  for (int q = 0; q < nb; ++q) s += p[q];
  return s;
}

// read back the visit log, the per-visit residuals and (mixed) the FP64
// residual partial sums of the finest level
// The V-cycle's visit residuals and log (and the FP64 residual's partials)
// in one copy; the host then waits for the stream by polling an event, which
// returns as soon as the copy lands instead of after a blocking wake-up.
int read_back(lepic_ctx* ctx, int* visit, double* vres, double* resid_sum) {
  double* h = ctx->h_pinned;
  const int nb = resid_blocks(ctx);
  const size_t bytes = resid_sum ? kPartialsOff - 32 * 8 + (size_t)nb * 8 : 32 * 8 + (kVisitLog + 1) * sizeof(int);
  LEPIC_HIP(hipMemcpyAsync(h, ctx->d_scalar + 32, bytes, hipMemcpyDeviceToHost, ctx->stream));
  LEPIC_HIP(hipEventRecord(ctx->ev_sync, ctx->stream));
  hipError_t q;
  while ((q = hipEventQuery(ctx->ev_sync)) == hipErrorNotReady) {
  }
  LEPIC_HIP(q);
  std::memcpy(vres, h, kVisitLog * 8);
  std::memcpy(visit, (char*)h + 32 * 8, kVisitLog * sizeof(int));
  if (resid_sum) *resid_sum = sum_resid_partials((const double*)((const char*)h + kPartialsOff - 32 * 8), nb);
  return 0;
}

}  // namespace

namespace lepic {

// tiles of t nodes along a side of `nodes` nodes; a single node left over
// joins the last tile (k_sweeps' tiles can be one node wider)
static int tiles_of(int nodes, int t) {
  int n = (nodes + t - 1) / t;
  if (n > 1 && nodes - (n - 1) * t == 1) --n;
  return n;
}

int poisson_setup(lepic_ctx* ctx) {
  const int ng = ctx->ng;
  const double eps0 = ctx->g.eps0;
  const bool mixed = ctx->g.precision == LEPIC_MIXED;
  for (int l = 1; l <= ng; ++l) {
    Level& L = ctx->lv[l];
    L.n1 = ctx->g.n1 >> (l - 1);
    L.n2 = ctx->g.n2 >> (l - 1);
    const double f = (double)(1 << (l - 1));  // h2=2.d0*h, h4=4.d0*h ... (mg.f90:59-80)
    L.h1 = f * ctx->g.h1;
    L.h2 = f * ctx->g.h2;
    L.aw = eps0 / (L.h1 * L.h1);  // mg.h:11-15
    L.an = eps0 / (L.h2 * L.h2);
    L.ac = -(L.an + L.an + L.aw + L.aw);
    const size_t nn = (size_t)(L.n1 + 3) * (L.n2 + 3), nr = (size_t)(L.n1 + 1) * (L.n2 + 1);
    if (l == 1) {
      L.u = ctx->phi;
      L.b = ctx->rhs;
      L.bc = ctx->bc;
    } else {
      LEPIC_HIP(hipMalloc(&L.u, nn * 8));
      LEPIC_HIP(hipMemsetAsync(L.u, 0, nn * 8, ctx->stream));
      LEPIC_HIP(hipMalloc(&L.b, nr * 8));
      LEPIC_HIP(hipMemsetAsync(L.b, 0, nr * 8, ctx->stream));
      LEPIC_HIP(hipMalloc(&L.bc, nn));
      LEPIC_HIP(hipMemsetAsync(L.bc, 1, nn, ctx->stream));
    }
    if (mixed) {
      LEPIC_HIP(hipMalloc(&L.uf, nn * 4));
      LEPIC_HIP(hipMemsetAsync(L.uf, 0, nn * 4, ctx->stream));
      LEPIC_HIP(hipMalloc(&L.bf, nr * 4));
      LEPIC_HIP(hipMemsetAsync(L.bf, 0, nr * 4, ctx->stream));
    }
  }
  // levels with more nodes than the tail threshold run multi-block; the
  // coarsest level always runs in k_tail (unless ng = 1)
  const int64_t tail_nodes = ctx->g.mg_tail_nodes > 0 ? ctx->g.mg_tail_nodes : kDefaultTailNodes;
  ctx->Lc = ng;
  for (int l = 1; l <= ng; ++l) {
    const Level& L = ctx->lv[l];
    if ((int64_t)(L.n1 + 1) * (L.n2 + 1) <= tail_nodes) { ctx->Lc = l; break; }
  }
  if (ng == 1 && (int64_t)(ctx->lv[1].n1 + 1) * (ctx->lv[1].n2 + 1) > tail_nodes) ctx->Lc = 2;
  const size_t esz = mixed ? 4 : 8;
  const int ty = mixed ? TileY<float>::v : TileY<double>::v;
  for (int l = 1; l < ctx->Lc && l <= ng; ++l) {
    Level& L = ctx->lv[l];
    L.mb = true;
    // the mixed form's visits run fixed counts; FP64 first launches start at kS
    L.first_dn = mixed ? (l == 1 ? kFineDn : kFixDn) : kS;
    L.first_up = mixed ? (l == 1 ? kFineUp : kFixUp) : kS;
    const size_t nn = (size_t)(L.n1 + 3) * (L.n2 + 3);
    L.rot[0] = mixed ? (void*)L.uf : (void*)L.u;
    for (int r = 1; r < kRot; ++r) {
      LEPIC_HIP(hipMalloc(&L.rot[r], nn * esz));
      LEPIC_HIP(hipMemsetAsync(L.rot[r], 0, nn * esz, ctx->stream));
    }
    L.nbx = tiles_of(L.n1 + 1, kTX);
    L.nby = tiles_of(L.n2 + 1, ty);
    // few wide tiles leave most CUs idle: the level takes small tiles, whose
    // launch latency is lower (each block has less than half the nodes)
    L.small_tiles = L.nbx * L.nby < kSmallTileBelow;
    if (L.small_tiles) {
      const int sty = mixed ? SmallTileY<float>::v : SmallTileY<double>::v;
      L.nbx = tiles_of(L.n1 + 1, kSmallTX);
      L.nby = tiles_of(L.n2 + 1, sty);
    }
    const size_t np = (size_t)kRing * L.nbx * L.nby;
    LEPIC_HIP(hipMalloc(&L.partials, np * 8));
    LEPIC_HIP(hipMemsetAsync(L.partials, 0, np * 8, ctx->stream));
  }
  if (mixed && ctx->lv[1].mb) {
    LEPIC_HIP(hipMalloc(&ctx->d_erot, kRot * sizeof(float*)));
    LEPIC_HIP(hipMemcpyAsync(ctx->d_erot, ctx->lv[1].rot, kRot * sizeof(float*), hipMemcpyHostToDevice, ctx->stream));
  LEPIC_HIP(hipStreamSynchronize(ctx->stream));
  } else if (mixed) {
    std::vector<float*> e(kRot, ctx->lv[1].uf);
    LEPIC_HIP(hipMalloc(&ctx->d_erot, kRot * sizeof(float*)));
    LEPIC_HIP(hipMemcpyAsync(ctx->d_erot, e.data(), kRot * sizeof(float*), hipMemcpyHostToDevice, ctx->stream));
  LEPIC_HIP(hipStreamSynchronize(ctx->stream));
  }
  LEPIC_HIP(hipMalloc(&ctx->d_ctrl, (kMaxLevels + 1) * sizeof(LevelCtrl)));
  LEPIC_HIP(hipMemsetAsync(ctx->d_ctrl, 0, (kMaxLevels + 1) * sizeof(LevelCtrl), ctx->stream));
  ctx->mg_epoch = 0;
  // one allocation: 64 doubles of scalars (the visit residuals from 32),
  // the visit log and status, then the FP64 residual's block partials, so a
  // V-cycle's results come back in one copy
  LEPIC_HIP(hipMalloc(&ctx->d_scalar, kPartialsOff + kResidBlocks * 8));
  LEPIC_HIP(hipMemsetAsync(ctx->d_scalar, 0, kPartialsOff + kResidBlocks * 8, ctx->stream));
  ctx->d_partials = (double*)((char*)ctx->d_scalar + kPartialsOff);
  ctx->partials_cap = kResidBlocks;
  if (!ctx->ev_sync) LEPIC_HIP(hipEventCreateWithFlags(&ctx->ev_sync, hipEventDisableTiming));
  return 0;
}

void poisson_teardown(lepic_ctx* ctx) {
  for (int l = 1; l <= kMaxLevels; ++l) {
    Level& L = ctx->lv[l];
    for (int r = 1; r < kRot; ++r)
      if (L.rot[r]) (void)hipFree(L.rot[r]);
    if (L.partials) (void)hipFree(L.partials);
  }
  if (ctx->d_erot) (void)hipFree(ctx->d_erot);
  if (ctx->d_tail_bc) (void)hipFree(ctx->d_tail_bc);
  ctx->d_tail_bc = nullptr;
  ctx->tail_bc_cap = 0;
  ctx->tail_bc_key = -1;
}

int poisson_solve(lepic_ctx* ctx, double eps, int ncycl, int start, lepic_solve_info* info,
                  int32_t* log, int32_t log_cap) {
  const int ng = ctx->ng;
  const bool mixed = ctx->g.precision == LEPIC_MIXED;
  const int64_t n = ctx->nodes;
  const dim3 g1((unsigned)((n + 255) / 256));
  if (ng == 1 && mixed) return ctx->fail(LEPIC_ERR_UNSUPPORTED, "ng=1 runs in FP64 only");
  const double alpha = alpha_of(ctx);
  int visit[kVisitLog];
  double vres[kVisitLog];
  double res = 0.0, res_s = 0.0, ktot = 0.0;
  info->nlog = 0;
  info->start_used = start;
  info->res0_alt = -1.0;
  if (start == LEPIC_START_NET) {
    // the network's guess replaces the previous phi only if it lowers the
    // initial residual; otherwise the solve starts from the previous phi
    info->start_used = LEPIC_START_PREV;
    const int g = net_probe_due(ctx) ? net_guess(ctx, ctx->phi_given) : 1;
    if (g < 0) return g;
    if (g == 0) {
      double r_prev = 0.0, r_net = 0.0;
      if (int rc = launch_resid64(ctx, false)) return rc;
      if (int rc = read_back(ctx, visit, vres, &r_prev)) return rc;
      if (int rc = launch_resid64(ctx, false, ctx->phi_given)) return rc;
      if (int rc = read_back(ctx, visit, vres, &r_net)) return rc;
      net_probe_result(ctx, r_net > 0.0 ? r_prev / r_net : HUGE_VAL);
      if (r_net < r_prev) {
        hipLaunchKernelGGL(k_copy_interior, g1, dim3(256), 0, ctx->stream, ctx->phi,
                           (const double*)ctx->phi_given, (const int8_t*)ctx->bc, n);
        info->start_used = LEPIC_START_NET;
        info->res0_alt = r_prev / alpha;
      } else {
        info->res0_alt = r_net / alpha;
      }
    }
  } else if (start == LEPIC_START_COLD) {
    hipLaunchKernelGGL(k_cold_start, g1, dim3(256), 0, ctx->stream, ctx->phi, (const int8_t*)ctx->bc, n);
  } else if (start == LEPIC_START_GIVEN) {
    hipLaunchKernelGGL(k_copy_interior, g1, dim3(256), 0, ctx->stream, ctx->phi,
                       (const double*)ctx->phi_given, (const int8_t*)ctx->bc, n);
  }
  LEPIC_HIP(hipGetLastError());
  // residual of the starting guess; in mixed mode it also fills the FP32
  // right-hand side of the first correction equation
  double r0 = 0.0;
  if (int rc = launch_resid64(ctx, mixed)) return rc;
  if (int rc = read_back(ctx, visit, vres, &r0)) return rc;
  info->res0 = r0 / alpha;
  int k = 1;
  info->status = 0;
  for (;;) {
    if (ng == 1 && ctx->lv[1].mb) {
      // red-black SOR alone on a multi-block grid (mg.f90:197-198 with
      // ng=1): the first call iterates to eps (up to 10^6 sweeps), later
      // calls do 10; launches go in batches with a host check in between
      const int nr = (k == 1) ? 1000000 : kNr;
      const int e = ++ctx->mg_epoch;
      int k0 = 0, ks = 0;
      while (k0 < nr) {
        const int k_end = std::min(nr, k0 + 16 * kS);
        if (int rc = launch_visit<double>(ctx, 1, false, e, 0, nr, alpha * eps, k0, k_end)) return rc;
        k0 = k_end;
        Level& L1 = ctx->lv[1];
        Finish64Args A;
        A.v = visit_ref(ctx, 1, e, 0, nr, 1, alpha * eps, k0);
        for (int r = 0; r < kRot; ++r) A.rot[r] = rot_of<double>(L1, r);
        A.n1 = L1.n1; A.n2 = L1.n2;
        A.status = d_status_of(ctx);
        hipLaunchKernelGGL(k_finish64, dim3((unsigned)((L1.n1 + 1 + 255) / 256), std::min(L1.n2 + 1, kRowBlocks)),
                           dim3(256), 0, ctx->stream, A);
        LEPIC_HIP(hipGetLastError());
        LEPIC_HIP(hipMemcpyAsync(ctx->h_pinned, d_status_of(ctx), 4, hipMemcpyDeviceToHost, ctx->stream));
        LEPIC_HIP(hipStreamSynchronize(ctx->stream));
        std::memcpy(&ks, ctx->h_pinned, 4);
        if (ks) break;
      }
      if (!ks) return ctx->fail(LEPIC_ERR_STATE, "SOR visit did not finish");
      if (int rc = read_back(ctx, visit, vres, nullptr)) return rc;
    } else if (!mixed) {
      if (int rc = vcycle<double>(ctx, false, k, alpha * eps)) return rc;
      if (int rc = read_back(ctx, visit, vres, nullptr)) return rc;
    } else {
      // FP32 correction V-cycle on the FP64 residual, then phi += e; the new
      // FP64 residual is both the convergence test and the next right-hand side
      if (int rc = vcycle<float>(ctx, true, k, alpha * eps)) return rc;
      if (int rc = launch_resid64(ctx, true)) return rc;
      double rs = 0.0;
      if (int rc = read_back(ctx, visit, vres, &rs)) return rc;
      vres[2 * ng - 1] = rs;
    }
    const int nvis = 2 * ng - 1;
    if (ng > 1)
      for (int l = 1; l < ctx->Lc; ++l) {
        Level& L = ctx->lv[l];
        if (!L.mb) continue;
        if (!mixed) {
          // one sweep more than last time: a spare sweep costs less than a
          // continuation launch when the count grows by one
          L.first_dn = std::min(std::max(visit[l] + 1, 2), kS);
          L.first_up = std::min(std::max(visit[2 * ng - l] + 1, 2), kS);
        } else {   // the mixed form's fixed counts
          visit[l] = L.first_dn;
          visit[2 * ng - l] = L.first_up;
        }
      }
    for (int ig = 1; ig <= nvis; ++ig) {
      int l = ig < ng ? ig : (ig > ng ? 2 * ng - ig : ng);
      ktot += (double)visit[ig] / std::pow(2.0, 2 * (l - 1));
      if (log && info->nlog < log_cap) log[info->nlog] = visit[ig];
      info->nlog++;
    }
    res = vres[nvis] / alpha;
    if (k == 1) res_s = res;
    else if (std::fabs(res / res_s) > 10.0) {
      info->vcycles = k;
      info->res = res;
      return ctx->fail(LEPIC_ERR_DIVERGED, "PDE solver is diverging (iterative_solver.f90:42-49)");
    }
    if (res <= eps || k == ncycl) break;
    ++k;
  }
  info->vcycles = k;
  info->status = (res <= eps) ? 0 : 1;
  info->res = res;
  info->ktot = ktot;
  return net_after_solve(ctx);
}

}  // namespace lepic
