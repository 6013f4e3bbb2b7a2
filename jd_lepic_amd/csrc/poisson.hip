// poisson.hip -- iterative_solver (iterative_solver.f90:1-69), the MG
// V-cycle (mg.f90:1-232) and red-black SOR with adaptive sweep control
// (sors.f90:1-435).
//
// Levels with many nodes run node-parallel kernels (one red, one black and
// one residual kernel per sweep); the residual kernel's last block sums the
// block partials in a fixed order and takes the reference's exit decision
// on the device, so later sweeps of the same level visit exit at once and
// the host never waits inside a V-cycle.  The coarse levels whose arrays fit
// in 64 KB of LDS run in a single workgroup kernel that performs the whole
// lower part of the V-cycle (down-leg sweeps and restrictions, the coarsest
// solve, prolongations and up-leg sweeps) without leaving the CU.
// FP64 mode computes the reference's expressions in their order (no
// contraction), so a sweep is bit-identical to sors.f90; only the order of
// the L1 residual sum differs.  Mixed mode runs the same V-cycle in FP32 on
// the correction equation inside an FP64 iterative-refinement loop whose
// FP64 residual keeps the reference's convergence definition.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "lepic_internal.h"

using namespace lepic;

namespace {

constexpr int kVisitLog = 2 * kMaxLevels;    // ig = 1 .. 2ng-1
constexpr size_t kCoarseLds = 64 * 1024;
constexpr int kCoarseThreads = 1024;

template <typename T>
struct Lv {
  T* u;
  const T* b;
  const int8_t* bc;
  int n1, n2;
  T aw, an, ac;
};

template <typename T>
Lv<T> lv_of(lepic_ctx* ctx, int l, bool f32) {
  Level& L = ctx->lv[l];
  Lv<T> r;
  r.u = f32 ? (T*)L.uf : (T*)L.u;
  r.b = f32 ? (const T*)L.bf : (const T*)L.b;
  r.bc = L.bc;
  r.n1 = L.n1; r.n2 = L.n2;
  r.aw = (T)L.aw; r.an = (T)L.an; r.ac = (T)L.ac;
  return r;
}

// ---- SOR update of one node (sors.f90:108-110 / 238-240) ----
template <typename T>
__device__ __forceinline__ T sor_update(T u, T b, T uW, T uN, T uS, T uE, T aw, T an, T ac, T omega) {
  return ((T)1.0 - omega) * u + omega * (b - aw * uW - an * uN - an * uS - aw * uE) / ac;
}

// ---- residual of one node, sors.f90:348-349 order ----
template <typename T>
__device__ __forceinline__ double sor_resid(T u, T b, T uW, T uN, T uS, T uE, T aw, T an, T ac) {
  return (double)(b - (aw * uW + an * uN + an * uS + aw * uE + ac * u));
}

// ---- getres, mg.f90:512-547 order (aw, ae, an, as, ac) ----
template <typename T>
__device__ __forceinline__ T getres(const T* u, const T* b, int ir, int jr, int n1, T aw, T an, T ac) {
  const int64_t W = n1 + 3;
  const int64_t c = (int64_t)jr * W + ir;
  T uij = u[c], un = u[c + W], us = u[c - W], ue = u[c + 1], uw = u[c - 1];
  T urhs = b[(int64_t)(jr - 1) * (n1 + 1) + (ir - 1)];
  return urhs - (aw * uw + aw * ue + an * un + an * us + ac * uij);
}

// ---------------------------------------------------------------------------
// multi-block kernels
// ---------------------------------------------------------------------------
template <typename T>
__global__ void k_sor_color(Lv<T> L, T omega, int color, const LevelCtrl* __restrict__ ctrl) {
  if (__hip_atomic_load(&ctrl->done, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) return;
  const int j = blockIdx.y + 1;
  const int is = color ? ((j & 1) ? 2 : 1) : ((j & 1) ? 1 : 2);  // red: i+j odd
  const int i = is + 2 * (int)(blockIdx.x * blockDim.x + threadIdx.x);
  if (i > L.n1 + 1) return;
  const int64_t W = L.n1 + 3, c = (int64_t)j * W + i;
  if (L.bc[c] > 0) return;
  T* u = L.u;
  u[c] = sor_update(u[c], L.b[(int64_t)(j - 1) * (L.n1 + 1) + (i - 1)], u[c - 1], u[c + W], u[c - W],
                    u[c + 1], L.aw, L.an, L.ac, omega);
}

// L1 residual over interior nodes and the exit test of sors.f90:338-379.
template <typename T>
__global__ __launch_bounds__(256) void k_residual(Lv<T> L, LevelCtrl* __restrict__ ctrl,
                                                  double* __restrict__ partials, int dig, double eps,
                                                  int nr, int* __restrict__ visit, int ig,
                                                  double* __restrict__ vres) {
  if (__hip_atomic_load(&ctrl->done, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) return;
  __shared__ double ws[4];
  __shared__ bool last;
  const int j = blockIdx.y + 1;
  const int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
  double r = 0.0;
  if (i <= L.n1 + 1) {
    const int64_t W = L.n1 + 3, c = (int64_t)j * W + i;
    if (L.bc[c] == -1) {
      const T* u = L.u;
      r = fabs(sor_resid(u[c], L.b[(int64_t)(j - 1) * (L.n1 + 1) + (i - 1)], u[c - 1], u[c + W], u[c - W],
                         u[c + 1], L.aw, L.an, L.ac));
    }
  }
  for (int o = 32; o > 0; o >>= 1) r += __shfl_down(r, o, 64);
  if ((threadIdx.x & 63) == 0) ws[threadIdx.x >> 6] = r;
  __syncthreads();
  const unsigned nb = gridDim.x * gridDim.y;
  const unsigned bid = blockIdx.y * gridDim.x + blockIdx.x;
  if (threadIdx.x == 0) {
    double s = ((ws[0] + ws[1]) + ws[2]) + ws[3];
    __hip_atomic_store(&partials[bid], s, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __threadfence();
    unsigned prev = atomicAdd(&ctrl->arrive, 1u);
    last = (prev == nb - 1);
  }
  __syncthreads();
  if (!last) return;
  __threadfence();
  double s = 0.0;
  for (unsigned k = threadIdx.x; k < nb; k += blockDim.x)
    s += __hip_atomic_load(&partials[k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  for (int o = 32; o > 0; o >>= 1) s += __shfl_down(s, o, 64);
  __syncthreads();
  if ((threadIdx.x & 63) == 0) ws[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    double res = ((ws[0] + ws[1]) + ws[2]) + ws[3];
    ctrl->arrive = 0;
    ctrl->ksor += 1;
    ctrl->iter += 1;
    ctrl->res = res;
    bool stop;
    if (dig == 0) stop = res <= eps;
    else stop = ctrl->res_s > 0.0 && (res / ctrl->res_s) <= 0.6;
    if (ctrl->ksor >= nr) stop = true;
    ctrl->res_s = res;
    if (stop) {
      visit[ig] = ctrl->ksor;
      *vres = res;
      __hip_atomic_store(&ctrl->done, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
  }
}

__global__ void k_ctrl_reset(LevelCtrl* c) {
  c->res = 0.0; c->res_s = 0.0; c->ksor = 0; c->done = 0; c->arrive = 0;
}

// restriction, mg.f90:335-444: full-weighted residual onto the coarse grid
// and a zero coarse correction (the coarse mask is precomputed).
template <typename T>
__global__ void k_restrict(Lv<T> F, T* __restrict__ r2h, T* __restrict__ e2h,
                           const int8_t* __restrict__ bc2h) {
  const int m1 = F.n1 / 2, m2 = F.n2 / 2;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y;
  if (i > m1 + 2) return;
  e2h[(int64_t)j * (m1 + 3) + i] = (T)0;
  if (i < 1 || i > m1 + 1 || j < 1 || j > m2 + 1) return;
  const int ir = 2 * i - 1, jr = 2 * j - 1, n1 = F.n1, n2 = F.n2;
  const int64_t W = n1 + 3;
  const int8_t* bc = F.bc;
  T out = (T)0;
  if (bc2h[(int64_t)j * (m1 + 3) + i] <= 0) {
    auto B = [&](int a, int b) { return bc[(int64_t)b * W + a]; };
    T r1 = 0, r2 = 0, r3 = 0, r4 = 0, r5 = 0, r6 = 0, r7 = 0, r8 = 0;
    T r9 = getres(F.u, F.b, ir, jr, n1, F.aw, F.an, F.ac);
    if (ir != n1 + 1 && jr != n2 + 1 && B(ir + 1, jr + 1) <= 0) r1 = getres(F.u, F.b, ir + 1, jr + 1, n1, F.aw, F.an, F.ac);
    if (ir != n1 + 1 && B(ir + 1, jr) <= 0) r2 = getres(F.u, F.b, ir + 1, jr, n1, F.aw, F.an, F.ac);
    if (B(ir + 1, jr - 1) <= 0 && ir != n1 + 1 && jr != 1) r3 = getres(F.u, F.b, ir + 1, jr - 1, n1, F.aw, F.an, F.ac);
    if (B(ir, jr - 1) <= 0 && jr != 1) r4 = getres(F.u, F.b, ir, jr - 1, n1, F.aw, F.an, F.ac);
    if (B(ir - 1, jr - 1) <= 0 && ir != 1 && jr != 1) r5 = getres(F.u, F.b, ir - 1, jr - 1, n1, F.aw, F.an, F.ac);
    if (B(ir - 1, jr) <= 0 && ir != 1) r6 = getres(F.u, F.b, ir - 1, jr, n1, F.aw, F.an, F.ac);
    if (B(ir - 1, jr + 1) <= 0 && ir != 1 && jr != n2 + 1) r7 = getres(F.u, F.b, ir - 1, jr + 1, n1, F.aw, F.an, F.ac);
    if (B(ir, jr + 1) <= 0 && jr != n2 + 1) r8 = getres(F.u, F.b, ir, jr + 1, n1, F.aw, F.an, F.ac);
    out = ((T)4.0 * r9 + (T)2.0 * (r2 + r4 + r6 + r8) + r1 + r3 + r5 + r7) / (T)16.0;
  }
  r2h[(int64_t)(j - 1) * (m1 + 1) + (i - 1)] = out;
}

// prolongation, mg.f90:585-611: bilinear add of the coarse correction
template <typename T>
__global__ void k_prolong(T* __restrict__ u, const int8_t* __restrict__ bc, int n1, int n2,
                          const T* __restrict__ e2h) {
  const int m1 = n1 / 2, m2 = n2 / 2;
  const int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
  const int j = blockIdx.y + 1;
  if (i > m1 + 1 || j > m2 + 1) return;
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

// ---------------------------------------------------------------------------
// single-workgroup kernel for the coarse levels Lc..ng
// ---------------------------------------------------------------------------
template <typename T>
struct CLevel {
  int n1, n2;
  T aw, an, ac;
  uint32_t ou, ob, obc;       // byte offsets in LDS
  const int8_t* gbc;
};

template <typename T>
struct CoarseArgs {
  int Lc, ng, k;
  T omega;
  double eps;
  T* gu;                      // level Lc correction/potential (global)
  const T* gb;                // level Lc right-hand side (global)
  int* visit;
  double* vres;
  CLevel<T> lv[kMaxLevels + 1];
};

__device__ inline double block_sum_1024(double v, double* ws) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  __syncthreads();
  if ((threadIdx.x & 63) == 0) ws[threadIdx.x >> 6] = v;
  __syncthreads();
  double s = 0.0;
  for (int w = 0; w < (int)(blockDim.x >> 6); ++w) s += ws[w];  // same order in every thread
  return s;
}

template <typename T>
__device__ void c_smooth(const CLevel<T>& L, char* smem, T omega, double eps, int dig, int nr,
                         double* ws, int* ksor_out, double* res_out) {
  T* u = (T*)(smem + L.ou);
  const T* b = (const T*)(smem + L.ob);
  const int8_t* bc = (const int8_t*)(smem + L.obc);
  const int n1 = L.n1, n2 = L.n2;
  const int W = n1 + 3;
  const int cnt = (n1 + 1) * (n2 + 1);
  double res_s = 0.0, res = 0.0;
  int ksor = 0;
  for (int k = 1; k <= nr; ++k) {
    for (int color = 1; color >= 0; --color) {  // red (i+j odd) then black
      for (int q = threadIdx.x; q < cnt; q += blockDim.x) {
        int i = q % (n1 + 1) + 1, j = q / (n1 + 1) + 1;
        if (((i + j) & 1) != color) continue;
        int c = j * W + i;
        if (bc[c] > 0) continue;
        u[c] = sor_update(u[c], b[q], u[c - 1], u[c + W], u[c - W], u[c + 1], L.aw, L.an, L.ac, omega);
      }
      __syncthreads();
    }
    double r = 0.0;
    for (int q = threadIdx.x; q < cnt; q += blockDim.x) {
      int i = q % (n1 + 1) + 1, j = q / (n1 + 1) + 1;
      int c = j * W + i;
      if (bc[c] == -1) r += fabs(sor_resid(u[c], b[q], u[c - 1], u[c + W], u[c - W], u[c + 1], L.aw, L.an, L.ac));
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

template <typename T>
__device__ void c_restrict(const CLevel<T>& F, const CLevel<T>& C, char* smem) {
  const T* u = (const T*)(smem + F.ou);
  const T* b = (const T*)(smem + F.ob);
  const int8_t* bc = (const int8_t*)(smem + F.obc);
  T* r2h = (T*)(smem + C.ob);
  T* e2h = (T*)(smem + C.ou);
  const int8_t* bc2 = (const int8_t*)(smem + C.obc);
  const int n1 = F.n1, n2 = F.n2, m1 = C.n1, m2 = C.n2, W = n1 + 3;
  for (int q = threadIdx.x; q < (m1 + 3) * (m2 + 3); q += blockDim.x) e2h[q] = (T)0;
  for (int q = threadIdx.x; q < (m1 + 1) * (m2 + 1); q += blockDim.x) {
    int i = q % (m1 + 1) + 1, j = q / (m1 + 1) + 1;
    int ir = 2 * i - 1, jr = 2 * j - 1;
    T out = (T)0;
    if (bc2[j * (m1 + 3) + i] <= 0) {
      auto B = [&](int a, int c) { return bc[c * W + a]; };
      T r1 = 0, r2 = 0, r3 = 0, r4 = 0, r5 = 0, r6 = 0, r7 = 0, r8 = 0;
      T r9 = getres(u, b, ir, jr, n1, F.aw, F.an, F.ac);
      if (ir != n1 + 1 && jr != n2 + 1 && B(ir + 1, jr + 1) <= 0) r1 = getres(u, b, ir + 1, jr + 1, n1, F.aw, F.an, F.ac);
      if (ir != n1 + 1 && B(ir + 1, jr) <= 0) r2 = getres(u, b, ir + 1, jr, n1, F.aw, F.an, F.ac);
      if (B(ir + 1, jr - 1) <= 0 && ir != n1 + 1 && jr != 1) r3 = getres(u, b, ir + 1, jr - 1, n1, F.aw, F.an, F.ac);
      if (B(ir, jr - 1) <= 0 && jr != 1) r4 = getres(u, b, ir, jr - 1, n1, F.aw, F.an, F.ac);
      if (B(ir - 1, jr - 1) <= 0 && ir != 1 && jr != 1) r5 = getres(u, b, ir - 1, jr - 1, n1, F.aw, F.an, F.ac);
      if (B(ir - 1, jr) <= 0 && ir != 1) r6 = getres(u, b, ir - 1, jr, n1, F.aw, F.an, F.ac);
      if (B(ir - 1, jr + 1) <= 0 && ir != 1 && jr != n2 + 1) r7 = getres(u, b, ir - 1, jr + 1, n1, F.aw, F.an, F.ac);
      if (B(ir, jr + 1) <= 0 && jr != n2 + 1) r8 = getres(u, b, ir, jr + 1, n1, F.aw, F.an, F.ac);
      out = ((T)4.0 * r9 + (T)2.0 * (r2 + r4 + r6 + r8) + r1 + r3 + r5 + r7) / (T)16.0;
    }
    r2h[q] = out;
  }
  __syncthreads();
}

template <typename T>
__device__ void c_prolong(const CLevel<T>& F, const CLevel<T>& C, char* smem) {
  T* u = (T*)(smem + F.ou);
  const int8_t* bc = (const int8_t*)(smem + F.obc);
  const T* e2h = (const T*)(smem + C.ou);
  const int n1 = F.n1, m1 = C.n1, m2 = C.n2, W = n1 + 3, Wc = m1 + 3;
  for (int q = threadIdx.x; q < (m1 + 1) * (m2 + 1); q += blockDim.x) {
    int i = q % (m1 + 1) + 1, j = q / (m1 + 1) + 1;
    T r1 = e2h[j * Wc + i], r2 = e2h[(j - 1) * Wc + i], r3 = e2h[(j - 1) * Wc + i - 1], r4 = e2h[j * Wc + i - 1];
    int ir = 2 * i - 1, jr = 2 * j - 1, c = jr * W + ir;
    if (bc[c] <= 0) u[c] = u[c] + r1;
    if (bc[c - W] <= 0) u[c - W] = u[c - W] + (r1 + r2) / (T)2.0;
    if (bc[c - W - 1] <= 0) u[c - W - 1] = u[c - W - 1] + (r1 + r2 + r3 + r4) / (T)4.0;
    if (bc[c - 1] <= 0) u[c - 1] = u[c - 1] + (r1 + r4) / (T)2.0;
  }
  __syncthreads();
}

template <typename T>
__global__ __launch_bounds__(kCoarseThreads) void k_coarse(CoarseArgs<T> A) {
  extern __shared__ __align__(16) char smem[];
  __shared__ double ws[16];
  const int Lc = A.Lc, ng = A.ng;
  for (int l = Lc; l <= ng; ++l) {
    const CLevel<T>& L = A.lv[l];
    int8_t* bc = (int8_t*)(smem + L.obc);
    for (int q = threadIdx.x; q < (L.n1 + 3) * (L.n2 + 3); q += blockDim.x) bc[q] = L.gbc[q];
  }
  {
    const CLevel<T>& L = A.lv[Lc];
    T* u = (T*)(smem + L.ou);
    T* b = (T*)(smem + L.ob);
    for (int q = threadIdx.x; q < (L.n1 + 3) * (L.n2 + 3); q += blockDim.x) u[q] = A.gu[q];
    for (int q = threadIdx.x; q < (L.n1 + 1) * (L.n2 + 1); q += blockDim.x) b[q] = A.gb[q];
  }
  __syncthreads();
  int ks;
  double res;
  for (int l = Lc; l < ng; ++l) {  // down leg: ig = l
    c_smooth(A.lv[l], smem, A.omega, A.eps, l - ng, 10, ws, &ks, &res);
    if (threadIdx.x == 0) A.visit[l] = ks;
    c_restrict(A.lv[l], A.lv[l + 1], smem);
  }
  c_smooth(A.lv[ng], smem, A.omega, A.eps, 0, A.k == 1 ? 1000000 : 10, ws, &ks, &res);  // ig = ng
  if (threadIdx.x == 0) A.visit[ng] = ks;
  for (int l = ng - 1; l >= Lc; --l) {  // up leg: ig = 2ng - l
    c_prolong(A.lv[l], A.lv[l + 1], smem);
    c_smooth(A.lv[l], smem, A.omega, A.eps, ng - l, 10, ws, &ks, &res);
    if (threadIdx.x == 0) A.visit[2 * ng - l] = ks;
  }
  {
    const CLevel<T>& L = A.lv[Lc];
    const T* u = (const T*)(smem + L.ou);
    for (int q = threadIdx.x; q < (L.n1 + 3) * (L.n2 + 3); q += blockDim.x) A.gu[q] = u[q];
  }
  if (threadIdx.x == 0) *A.vres = res;
}

// ---------------------------------------------------------------------------
// mixed-precision helpers: FP64 residual of phi (reference definition) and
// its FP32 copy as the right-hand side of the correction equation
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_resid64(const double* __restrict__ u, const double* __restrict__ b,
                                                 const int8_t* __restrict__ bc, int n1, int n2, double aw,
                                                 double an, double ac, float* __restrict__ bf,
                                                 float* __restrict__ ef, double* __restrict__ partials,
                                                 unsigned* __restrict__ arrive, double* __restrict__ out) {
  __shared__ double ws[4];
  __shared__ bool last;
  const int j = blockIdx.y + 1;
  const int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
  double r = 0.0;
  if (i <= n1 + 1) {
    const int64_t W = n1 + 3, c = (int64_t)j * W + i;
    const int64_t q = (int64_t)(j - 1) * (n1 + 1) + (i - 1);
    double rij = 0.0;
    if (bc[c] == -1)
      rij = b[q] - (aw * u[c - 1] + an * u[c + W] + an * u[c - W] + aw * u[c + 1] + ac * u[c]);
    if (bf) bf[q] = (float)rij;
    r = fabs(rij);
  }
  if (ef) {  // zero the FP32 correction on this block's nodes (and ghosts)
    const int64_t W = n1 + 3;
    if (i <= n1 + 1) ef[(int64_t)j * W + i] = 0.0f;
    if (i == 1) { ef[(int64_t)j * W] = 0.0f; ef[(int64_t)j * W + n1 + 2] = 0.0f; }
    if (j == 1 && i <= n1 + 1) { ef[i] = 0.0f; ef[(int64_t)(n2 + 2) * W + i] = 0.0f; }
    if (i == 1 && j == 1) { ef[0] = 0.0f; ef[n1 + 2] = 0.0f; ef[(int64_t)(n2 + 2) * W] = 0.0f; ef[(int64_t)(n2 + 2) * W + n1 + 2] = 0.0f; }
  }
  for (int o = 32; o > 0; o >>= 1) r += __shfl_down(r, o, 64);
  if ((threadIdx.x & 63) == 0) ws[threadIdx.x >> 6] = r;
  __syncthreads();
  const unsigned nb = gridDim.x * gridDim.y;
  const unsigned bid = blockIdx.y * gridDim.x + blockIdx.x;
  if (threadIdx.x == 0) {
    __hip_atomic_store(&partials[bid], ((ws[0] + ws[1]) + ws[2]) + ws[3], __ATOMIC_RELAXED,
                       __HIP_MEMORY_SCOPE_AGENT);
    __threadfence();
    last = atomicAdd(arrive, 1u) == nb - 1;
  }
  __syncthreads();
  if (!last) return;
  __threadfence();
  double s = 0.0;
  for (unsigned k = threadIdx.x; k < nb; k += blockDim.x)
    s += __hip_atomic_load(&partials[k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  for (int o = 32; o > 0; o >>= 1) s += __shfl_down(s, o, 64);
  __syncthreads();
  if ((threadIdx.x & 63) == 0) ws[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    *out = ((ws[0] + ws[1]) + ws[2]) + ws[3];
    *arrive = 0;
  }
}

__global__ void k_add_correction(double* __restrict__ u, const float* __restrict__ e,
                                 const int8_t* __restrict__ bc, int64_t n) {
  int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q < n && bc[q] <= 0) u[q] = u[q] + (double)e[q];
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
// host side
// ---------------------------------------------------------------------------
struct Plan {
  int* d_visit;                // [kVisitLog]
  double* d_vres;
};

template <typename T>
int launch_sweeps(lepic_ctx* ctx, int l, bool f32, int dig, double eps, int nr, int ig, int count,
                  int* visit, double* vres) {
  Lv<T> L = lv_of<T>(ctx, l, f32);
  LevelCtrl* c = ctx->d_ctrl + l;
  const int hw = (L.n1 + 2) / 2 + 1;
  dim3 gc((hw + 127) / 128, L.n2 + 1), gr((L.n1 + 1 + 255) / 256, L.n2 + 1);
  const T omega = (T)1.4;
  for (int s = 0; s < count; ++s) {
    hipLaunchKernelGGL(k_sor_color<T>, gc, dim3(128), 0, ctx->stream, L, omega, 1, (const LevelCtrl*)c);
    hipLaunchKernelGGL(k_sor_color<T>, gc, dim3(128), 0, ctx->stream, L, omega, 0, (const LevelCtrl*)c);
    hipLaunchKernelGGL(k_residual<T>, gr, dim3(256), 0, ctx->stream, L, c, ctx->d_partials, dig, eps, nr,
                       visit, ig, vres);
  }
  LEPIC_HIP(hipGetLastError());
  return 0;
}

template <typename T>
int launch_coarse(lepic_ctx* ctx, bool f32, int k, double eps, int* visit, double* vres) {
  CoarseArgs<T> A{};
  A.Lc = ctx->Lc; A.ng = ctx->ng; A.k = k; A.omega = (T)1.4; A.eps = eps;
  A.visit = visit; A.vres = vres;
  uint32_t off = 0;
  auto al = [](uint32_t x) { return (x + 15u) & ~15u; };
  for (int l = ctx->Lc; l <= ctx->ng; ++l) {
    Level& L = ctx->lv[l];
    CLevel<T>& C = A.lv[l];
    C.n1 = L.n1; C.n2 = L.n2; C.aw = (T)L.aw; C.an = (T)L.an; C.ac = (T)L.ac;
    C.ou = off; off = al(off + (uint32_t)((L.n1 + 3) * (L.n2 + 3) * sizeof(T)));
    C.ob = off; off = al(off + (uint32_t)((L.n1 + 1) * (L.n2 + 1) * sizeof(T)));
    C.obc = off; off = al(off + (uint32_t)((L.n1 + 3) * (L.n2 + 3)));
    C.gbc = L.bc;
  }
  if (off > kCoarseLds) return ctx->fail(LEPIC_ERR_ARG, "coarse levels exceed the LDS budget");
  Level& L = ctx->lv[ctx->Lc];
  A.gu = f32 ? (T*)L.uf : (T*)L.u;
  A.gb = f32 ? (const T*)L.bf : (const T*)L.b;
  hipLaunchKernelGGL(k_coarse<T>, dim3(1), dim3(kCoarseThreads), off, ctx->stream, A);
  LEPIC_HIP(hipGetLastError());
  return 0;
}

// One V-cycle (mg.f90:85-225) at precision T.  Levels below Lc use the
// multi-block kernels, Lc..ng the single-workgroup kernel.
template <typename T>
int vcycle(lepic_ctx* ctx, bool f32, int k, double eps, int* visit, double* vres) {
  const int ng = ctx->ng, Lc = ctx->Lc;
  for (int l = 1; l < Lc; ++l) {  // down leg
    hipLaunchKernelGGL(k_ctrl_reset, dim3(1), dim3(1), 0, ctx->stream, ctx->d_ctrl + l);
    if (int rc = launch_sweeps<T>(ctx, l, f32, l - ng, eps, 10, l, 10, visit, vres)) return rc;
    Lv<T> F = lv_of<T>(ctx, l, f32);
    Level& C = ctx->lv[l + 1];
    T* r2 = f32 ? (T*)C.bf : (T*)C.b;
    T* e2 = f32 ? (T*)C.uf : (T*)C.u;
    dim3 g((C.n1 + 3 + 127) / 128, C.n2 + 3);
    hipLaunchKernelGGL(k_restrict<T>, g, dim3(128), 0, ctx->stream, F, r2, e2, (const int8_t*)C.bc);
  }
  if (int rc = launch_coarse<T>(ctx, f32, k, eps, visit, vres)) return rc;
  for (int l = Lc - 1; l >= 1; --l) {  // up leg
    Level& F = ctx->lv[l];
    Level& C = ctx->lv[l + 1];
    T* u = f32 ? (T*)F.uf : (T*)F.u;
    const T* e2 = f32 ? (const T*)C.uf : (const T*)C.u;
    dim3 g((C.n1 + 1 + 127) / 128, C.n2 + 1);
    hipLaunchKernelGGL(k_prolong<T>, g, dim3(128), 0, ctx->stream, u, (const int8_t*)F.bc, F.n1, F.n2, e2);
    hipLaunchKernelGGL(k_ctrl_reset, dim3(1), dim3(1), 0, ctx->stream, ctx->d_ctrl + l);
    if (int rc = launch_sweeps<T>(ctx, l, f32, ng - l, eps, 10, 2 * ng - l, 10, visit, vres)) return rc;
  }
  LEPIC_HIP(hipGetLastError());
  return 0;
}

int read_back(lepic_ctx* ctx, int* visit_h, double* vres_h, int* d_visit, double* d_vres) {
  LEPIC_HIP(hipMemcpyAsync(ctx->h_pinned, d_vres, 8, hipMemcpyDeviceToHost, ctx->stream));
  LEPIC_HIP(hipMemcpyAsync(ctx->h_pinned + 1, d_visit, kVisitLog * sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
  LEPIC_HIP(hipStreamSynchronize(ctx->stream));
  std::memcpy(vres_h, ctx->h_pinned, 8);
  std::memcpy(visit_h, ctx->h_pinned + 1, kVisitLog * sizeof(int));
  return 0;
}

double alpha_of(lepic_ctx* ctx) {
  // iterative_solver.f90:22-23
  return std::fabs((double)((ctx->g.n1 + 1) * (ctx->g.n2 + 1)) * ctx->lv[1].ac) * 1.0;
}

int resid64(lepic_ctx* ctx, bool fill_f32, double* res_h) {
  Level& L = ctx->lv[1];
  dim3 g((L.n1 + 1 + 255) / 256, L.n2 + 1);
  double* out = ctx->d_scalar;
  unsigned* arrive = (unsigned*)(ctx->d_scalar + 8);
  hipLaunchKernelGGL(k_resid64, g, dim3(256), 0, ctx->stream, (const double*)ctx->phi,
                     (const double*)ctx->rhs, (const int8_t*)ctx->bc, L.n1, L.n2, L.aw, L.an, L.ac,
                     fill_f32 ? L.bf : nullptr, fill_f32 ? L.uf : nullptr, ctx->d_partials, arrive, out);
  LEPIC_HIP(hipGetLastError());
  if (res_h) {
    LEPIC_HIP(hipMemcpyAsync(ctx->h_pinned, out, 8, hipMemcpyDeviceToHost, ctx->stream));
    LEPIC_HIP(hipStreamSynchronize(ctx->stream));
    std::memcpy(res_h, ctx->h_pinned, 8);
  }
  return 0;
}

}  // namespace

namespace lepic {

int poisson_setup(lepic_ctx* ctx) {
  const int ng = ctx->ng;
  const double eps0 = ctx->g.eps0;
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
      LEPIC_HIP(hipMemset(L.u, 0, nn * 8));
      LEPIC_HIP(hipMalloc(&L.b, nr * 8));
      LEPIC_HIP(hipMemset(L.b, 0, nr * 8));
      LEPIC_HIP(hipMalloc(&L.bc, nn));
      LEPIC_HIP(hipMemset(L.bc, 1, nn));
    }
    if (ctx->g.precision == LEPIC_MIXED) {
      LEPIC_HIP(hipMalloc(&L.uf, nn * 4));
      LEPIC_HIP(hipMemset(L.uf, 0, nn * 4));
      LEPIC_HIP(hipMalloc(&L.bf, nr * 4));
      LEPIC_HIP(hipMemset(L.bf, 0, nr * 4));
    }
  }
  // coarse kernel takes the finest level from which all coarser levels fit
  // in the LDS budget at FP64 (the FP32 copy then fits as well)
  auto bytes = [&](int l) {
    const Level& L = ctx->lv[l];
    auto al = [](size_t x) { return (x + 15) & ~size_t(15); };
    return al((size_t)(L.n1 + 3) * (L.n2 + 3) * 8) + al((size_t)(L.n1 + 1) * (L.n2 + 1) * 8) +
           al((size_t)(L.n1 + 3) * (L.n2 + 3));
  };
  size_t tot = 0;
  ctx->Lc = ng + 1;
  for (int l = ng; l >= 1; --l) {
    if (tot + bytes(l) > kCoarseLds) break;
    tot += bytes(l);
    ctx->Lc = l;
  }
  if (ctx->Lc > ng && ng > 1) return ctx->fail(LEPIC_ERR_ARG, "coarsest MG level does not fit in LDS");
  LEPIC_HIP(hipMalloc(&ctx->d_ctrl, (kMaxLevels + 1) * sizeof(LevelCtrl)));
  LEPIC_HIP(hipMemset(ctx->d_ctrl, 0, (kMaxLevels + 1) * sizeof(LevelCtrl)));
  int64_t maxblocks = (int64_t)((ctx->g.n1 + 1 + 255) / 256 + 1) * (ctx->g.n2 + 1);
  LEPIC_HIP(hipMalloc(&ctx->d_partials, maxblocks * 8));
  ctx->partials_cap = maxblocks;
  LEPIC_HIP(hipMalloc(&ctx->d_scalar, 64 * 8 + kVisitLog * sizeof(int)));
  LEPIC_HIP(hipMemset(ctx->d_scalar, 0, 64 * 8 + kVisitLog * sizeof(int)));
  return 0;
}

int poisson_solve(lepic_ctx* ctx, double eps, int ncycl, int start, lepic_solve_info* info,
                  int32_t* log, int32_t log_cap) {
  const int ng = ctx->ng;
  const bool mixed = ctx->g.precision == LEPIC_MIXED;
  const int64_t n = ctx->nodes;
  const dim3 g1((unsigned)((n + 255) / 256));
  if (start == LEPIC_START_COLD)
    hipLaunchKernelGGL(k_cold_start, g1, dim3(256), 0, ctx->stream, ctx->phi, (const int8_t*)ctx->bc, n);
  else if (start == LEPIC_START_GIVEN)
    hipLaunchKernelGGL(k_copy_interior, g1, dim3(256), 0, ctx->stream, ctx->phi,
                       (const double*)ctx->phi_given, (const int8_t*)ctx->bc, n);
  LEPIC_HIP(hipGetLastError());
  const double alpha = alpha_of(ctx);
  int* d_visit = (int*)(ctx->d_scalar + 64);
  double* d_vres = ctx->d_scalar + 16;
  int visit[kVisitLog];
  double vres = 0.0, res = 0.0, res_s = 0.0, ktot = 0.0;
  info->nlog = 0;
  info->start_used = start;
  double r0 = 0.0;
  // residual of the starting guess; in mixed mode it also fills the FP32
  // right-hand side of the first correction equation
  if (int rc = resid64(ctx, mixed && ng > 1, &r0)) return rc;
  info->res0 = r0 / alpha;
  int k = 1;
  info->status = 0;
  if (ng == 1) {
    // red-black SOR alone (mg.f90:197-198 with ng=1): the first call
    // iterates to eps (up to 10^6 sweeps), later calls do 10 sweeps.
    for (;;) {
      hipLaunchKernelGGL(k_ctrl_reset, dim3(1), dim3(1), 0, ctx->stream, ctx->d_ctrl + 1);
      const int nr = (k == 1) ? 1000000 : 10;
      int done_sweeps = 0;
      if (mixed) return ctx->fail(LEPIC_ERR_UNSUPPORTED, "ng=1 runs in FP64 only");
      for (;;) {
        const int batch = std::min(nr - done_sweeps, 64);
        if (int rc = launch_sweeps<double>(ctx, 1, false, 0, alpha * eps, nr, 1, batch, d_visit, d_vres)) return rc;
        done_sweeps += batch;
        LevelCtrl c;
        LEPIC_HIP(hipMemcpyAsync(ctx->h_pinned, ctx->d_ctrl + 1, sizeof(LevelCtrl), hipMemcpyDeviceToHost, ctx->stream));
        LEPIC_HIP(hipStreamSynchronize(ctx->stream));
        std::memcpy(&c, ctx->h_pinned, sizeof(c));
        if (c.done || done_sweeps >= nr) { vres = c.res; visit[1] = c.ksor; break; }
      }
      ktot += visit[1];
      if (log && info->nlog < log_cap) log[info->nlog] = visit[1];
      info->nlog++;
      res = vres / alpha;
      if (k == 1) res_s = res;
      else if (std::fabs(res / res_s) > 10.0) { info->vcycles = k; return ctx->fail(LEPIC_ERR_DIVERGED, "PDE solver is diverging"); }
      if (res <= eps || k == ncycl) break;
      ++k;
    }
    info->vcycles = k;
    info->status = (res <= eps) ? 0 : 1;
    info->res = res;
    info->ktot = ktot;
    return 0;
  }
  for (;;) {
    if (!mixed) {
      if (int rc = vcycle<double>(ctx, false, k, alpha * eps, d_visit, d_vres)) return rc;
      if (int rc = read_back(ctx, visit, &vres, d_visit, d_vres)) return rc;
    } else {
      // FP32 correction V-cycle on the FP64 residual, then phi += e; the new
      // FP64 residual is both the convergence test and the next right-hand side
      if (int rc = vcycle<float>(ctx, true, k, alpha * eps, d_visit, d_vres)) return rc;
      hipLaunchKernelGGL(k_add_correction, g1, dim3(256), 0, ctx->stream, ctx->phi,
                         (const float*)ctx->lv[1].uf, (const int8_t*)ctx->bc, n);
      LEPIC_HIP(hipGetLastError());
      if (int rc = resid64(ctx, true, &vres)) return rc;
      double dummy;
      if (int rc = read_back(ctx, visit, &dummy, d_visit, d_vres)) return rc;
    }
    for (int ig = 1; ig <= 2 * ng - 1; ++ig) {
      int l = ig < ng ? ig : (ig > ng ? 2 * ng - ig : ng);
      ktot += (double)visit[ig] / std::pow(2.0, 2 * (l - 1));
      if (log && info->nlog < log_cap) log[info->nlog] = visit[ig];
      info->nlog++;
    }
    res = vres / alpha;
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
  return 0;
}

}  // namespace lepic
