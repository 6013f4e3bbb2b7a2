// field.hip -- rhs assembly (calc_rho.f90:37-77), electric field
// (Efield.f90:1-128) and density reduction (calc_rho.f90:134-199).
// Node-parallel stencil kernels; every node is one thread, arrays in the
// reference's column-major layout (ix fastest, coalesced along x).
#include <hip/hip_runtime.h>

#include <algorithm>

#include <cstring>

#include "lepic_internal.h"

using namespace lepic;

namespace {

struct Charges { double q[kMaxSpecies]; const double* np[kMaxSpecies]; int n; };

// rhs(ix,iy) = 0 - sum_ptype charge*np, in species order (calc_rho.f90:58-67)
__global__ void k_rhs(Charges c, double* __restrict__ rhs, int n1, int n2) {
  int ix = blockIdx.x * blockDim.x + threadIdx.x + 1;
  int iy = blockIdx.y + 1;
  if (ix > n1 + 1) return;
  const int64_t nd = node_index(ix, iy, n1);
  double r = 0.0;
  for (int p = 0; p < c.n; ++p) r = r - c.q[p] * c.np[p][nd];
  rhs[(int64_t)(iy - 1) * (n1 + 1) + (ix - 1)] = r;
}

// Efield.f90:19-34: interior nodes (bcnd <= 0)
__global__ void k_efield_inner(const double* __restrict__ phi, const int8_t* __restrict__ bc,
                               double* __restrict__ E, int n1, int n2, double h1, double h2) {
  int ix = blockIdx.x * blockDim.x + threadIdx.x + 1;
  int iy = blockIdx.y + 1;
  if (ix > n1 + 1) return;
  const int64_t nd = node_index(ix, iy, n1);
  if (bc[nd] > 0) return;
  const int64_t W = n1 + 3;
  double ex = -(phi[nd + 1] - phi[nd - 1]) / (2.0 * h1);
  double ey = -(phi[nd + W] - phi[nd - W]) / (2.0 * h2);
  *(double2*)(E + 2 * nd) = make_double2(ex, ey);
}

// Efield.f90:87-118: wall nodes, zero then linear extrapolation from the
// open side (the right/top branches run after and may override).
__device__ inline void wall_node(const int8_t* bc, double* E, int ix, int iy, int n1) {
  const int64_t nd = node_index(ix, iy, n1), W = n1 + 3;
  double ex = 0.0, ey = 0.0;
  if (bc[nd + 1] <= 0) ex = 2.0 * E[2 * (nd + 1)] - E[2 * (nd + 2)];
  if (bc[nd - 1] <= 0) ex = 2.0 * E[2 * (nd - 1)] - E[2 * (nd - 2)];
  if (bc[nd + W] <= 0) ey = 2.0 * E[2 * (nd + W) + 1] - E[2 * (nd + 2 * W) + 1];
  if (bc[nd - W] <= 0) ey = 2.0 * E[2 * (nd - W) + 1] - E[2 * (nd - 2 * W) + 1];
  E[2 * nd] = ex;
  E[2 * nd + 1] = ey;
}

__global__ void k_efield_walls(const int8_t* __restrict__ bc, double* __restrict__ E, int n1, int n2) {
  int ix = blockIdx.x * blockDim.x + threadIdx.x + 1;
  int iy = blockIdx.y + 1;
  if (ix > n1 + 1) return;
  if (bc[node_index(ix, iy, n1)] < 1) return;
  wall_node(bc, E, ix, iy, n1);
}

// The same pass in the reference's loop order, for geometries where a wall
// node's extrapolation reads another wall node (one lane).
__global__ void k_efield_walls_serial(const int8_t* __restrict__ bc, double* __restrict__ E, int n1,
                                      int n2) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  for (int iy = 1; iy <= n2 + 1; ++iy)
    for (int ix = 1; ix <= n1 + 1; ++ix)
      if (bc[node_index(ix, iy, n1)] >= 1) wall_node(bc, E, ix, iy, n1);
}

__global__ void k_to_f32(const double* __restrict__ a, float* __restrict__ b, int64_t n) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) b[i] = (float)a[i];
}

// dens_red for one partition: np_red = 0 + np on nodes 1..n+1 (exact copy),
// ghosts zero; np_mx = max over interior nodes (values are non-negative, so
// the IEEE bit patterns order like the values).  Block b walks rows b,
// b + gridDim.x, ...; the maximum is taken over the block first and one
// atomic per block goes to the shared address (atomics on one address
// serialise; a maximum is exact in any order).
__global__ __launch_bounds__(256) void k_dens_red(const double* __restrict__ np, double* __restrict__ red,
                                                  const int8_t* __restrict__ bc, int n1, int n2,
                                                  unsigned long long* __restrict__ mx) {
  __shared__ unsigned long long ws[4];
  unsigned long long bits = 0ull;
  for (int iy = blockIdx.x; iy <= n2 + 2; iy += gridDim.x)
    for (int ix = threadIdx.x; ix <= n1 + 2; ix += blockDim.x) {
      const int64_t nd = node_index(ix, iy, n1);
      const bool in = ix >= 1 && ix <= n1 + 1 && iy >= 1 && iy <= n2 + 1;
      const double v = in ? 0.0 + np[nd] : 0.0;
      red[nd] = v;
      if (in && bc[nd] == -1 && v > 0.0) {
        const unsigned long long b = (unsigned long long)__double_as_longlong(v);
        bits = b > bits ? b : bits;
      }
    }
  for (int o = 32; o > 0; o >>= 1) {
    const unsigned long long t = __shfl_xor(bits, o, 64);
    bits = t > bits ? t : bits;
  }
  if ((threadIdx.x & 63) == 0) ws[threadIdx.x >> 6] = bits;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < 4; ++w) bits = ws[w] > bits ? ws[w] : bits;
    if (bits) atomicMax(mx, bits);
  }
}

}  // namespace

namespace lepic {

int rhs_launch(lepic_ctx* ctx) {
  Charges c{};
  c.n = ctx->nsp;
  for (int p = 0; p < ctx->nsp; ++p) { c.q[p] = ctx->sp[p].charge; c.np[p] = ctx->sp[p].np; }
  const int n1 = ctx->g.n1, n2 = ctx->g.n2;
  dim3 grid((n1 + 1 + 127) / 128, n2 + 1);
  hipLaunchKernelGGL(k_rhs, grid, dim3(128), 0, ctx->stream, c, ctx->rhs, n1, n2);
  LEPIC_HIP(hipGetLastError());
  return 0;
}

int efield_launch(lepic_ctx* ctx) {
  const int n1 = ctx->g.n1, n2 = ctx->g.n2;
  dim3 grid((n1 + 1 + 127) / 128, n2 + 1);
  hipLaunchKernelGGL(k_efield_inner, grid, dim3(128), 0, ctx->stream, ctx->phi, ctx->bc, ctx->E, n1,
                     n2, ctx->g.h1, ctx->g.h2);
  if (ctx->efield_serial)
    hipLaunchKernelGGL(k_efield_walls_serial, dim3(1), dim3(64), 0, ctx->stream, ctx->bc, ctx->E, n1, n2);
  else
    hipLaunchKernelGGL(k_efield_walls, grid, dim3(128), 0, ctx->stream, ctx->bc, ctx->E, n1, n2);
  if (ctx->Ef) {
    int64_t n = 2 * ctx->nodes;
    hipLaunchKernelGGL(k_to_f32, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, ctx->stream,
                       ctx->E, ctx->Ef, n);
  }
  LEPIC_HIP(hipGetLastError());
  return 0;
}

int dens_red_launch(lepic_ctx* ctx, double* np_mx_host) {
  const int n1 = ctx->g.n1, n2 = ctx->g.n2;
  unsigned long long* mx = (unsigned long long*)ctx->d_mcc_i64 + 32;  // 8 slots
  if (int rc = zero_async(ctx, mx, kMaxSpecies * 8)) return rc;
  const dim3 grid((unsigned)std::min(n2 + 3, 256));
  for (int p = 0; p < ctx->nsp; ++p)
    hipLaunchKernelGGL(k_dens_red, grid, dim3(256), 0, ctx->stream, ctx->sp[p].np, ctx->sp[p].np_red,
                       ctx->bc, n1, n2, mx + p);
  LEPIC_HIP(hipGetLastError());
  LEPIC_HIP(hipMemcpyAsync(ctx->h_pinned, mx, kMaxSpecies * 8, hipMemcpyDeviceToHost, ctx->stream));
  LEPIC_HIP(hipStreamSynchronize(ctx->stream));
  for (int p = 0; p < ctx->nsp; ++p) {
    double v;
    std::memcpy(&v, ctx->h_pinned + p, 8);
    if (np_mx_host) np_mx_host[p] = v;
    if (ctx->rx.set && (int)ctx->rx.np_mx.size() > p + 1) ctx->rx.np_mx[p + 1] = v;
  }
  return 0;
}

}  // namespace lepic
