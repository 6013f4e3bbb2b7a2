// mcc.hip -- null-collision Monte-Carlo collisions (collisions.f90:1-207,
// collision_OMP :209-1040) against neutral background targets.
//
// Check mode replays the reference exactly: one GPU lane consumes the
// reference's ran2 stream in the reference's order (candidate pick, target
// Maxwellian with the cached second z draw, bisection + scan of the union
// energy grid, stable insertion sort of the frequencies, the event draws
// including the two used only by the discarded SCATTER rotation), so it
// yields the same events as the reference given the same particle arrays.
//
// Production mode gives every candidate its own lane and its own Philox
// stream keyed by (seed, it); candidates are drawn without replacement
// through a claim bitmap, so no two lanes touch one particle.  Block-level
// LDS tallies feed a few global atomics; particles created by ionization are
// staged per candidate and appended in candidate order after a scan, so the
// result does not depend on scheduling.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>

#include "lepic_internal.h"

using namespace lepic;

namespace {

constexpr int kMaxPcol = 16;     // reactions per incident species
constexpr int kMaxCreate = 3;    // particles created per event
constexpr size_t kMccLdsTable = 48 * 1024;   // largest LDS-staged cross-section table

struct MccTab {
  int ncol, npt, ntype_all, n_rg, iyl, iyr, ixl[5], ixr[5];
  int col_info[kMaxReac][10];
  double Eth[kMaxReac];
  double charge[kMaxTypes + 1], mass[kMaxTypes + 1], Nm[kMaxTypes + 1], vt0[kMaxTypes + 1],
      np_mx[kMaxTypes + 1];
};

struct SpeciesMcc {
  int ptype;                     // 1-based incident species
  int p_ncol;
  int reac[kMaxPcol];            // sig_list(ptype, 1:p_ncol)
  int ttype[kMaxPcol];
  double nu_max;
  double nu_omp;                 // nu_max_OMP (collisions.f90:172)
  int64_t Nc;
  double rg_inv[5];              // 1/np_tot_rg(ptype, rg)
};

struct PartPtrs {
  double* d[kMaxSpecies][6];     // FP64 SoA of the current buffer
  int32_t* cell[kMaxSpecies];    // mixed
  float* f[kMaxSpecies][6];
  int8_t* fd[kMaxSpecies];
  int64_t cap[kMaxSpecies];
  double* src[kMaxSpecies];      // ss2D(1), ss2D(2)
  double* sink[kMaxSpecies];
};

PartPtrs part_ptrs(lepic_ctx* ctx) {
  PartPtrs P{};
  for (int p = 0; p < ctx->nsp; ++p) {
    Species& s = ctx->sp[p];
    for (int k = 0; k < 6; ++k) { P.d[p][k] = s.d[s.cur][k]; P.f[p][k] = s.f[s.cur][k]; }
    P.cell[p] = s.cell[s.cur];
    P.fd[p] = s.flag_dead;
    P.src[p] = s.src;
    P.sink[p] = s.sink;
    P.cap[p] = s.cap;
  }
  return P;
}

// ---------------------------------------------------------------------------
// check mode: one lane, FP64 particle state
// ---------------------------------------------------------------------------
struct CheckArgs {
  const MccTab* tab;
  const double* sig_Er;
  const double* sig;
  PartPtrs P;
  int64_t* counts;
  const double* np_red[kMaxSpecies];
  int nsp, n1;
  double h1, h2, zmax, dt, qe, pi;
  int ns_coll, n_cell;
  SpeciesMcc sp[kMaxSpecies];    // ptype 1..nsp; p_ncol==0 when no reactions
  int32_t* iseed;
  double* cnt_col;
  double* ploss3;
  unsigned long long* events;
  unsigned long long* created;
  int64_t* Nc_out;
  int32_t* status;
};

__global__ void k_mcc_check(CheckArgs A) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  const MccTab& T = *A.tab;
  int32_t seed = *A.iseed;
  const double qe = A.qe, pi = A.pi;
  const int npt = T.npt;
  int64_t Nc[kMaxSpecies] = {}, np_add[kMaxTypes + 1] = {};
  double rg_inv[kMaxSpecies][5];
  for (int q = 0; q < A.nsp; ++q) {  // collisions.f90:92-157
    const SpeciesMcc& S = A.sp[q];
    if (S.p_ncol == 0 || T.mass[q + 1] <= 0) continue;
    const int64_t N = A.counts[q];
    double Pmax = S.nu_max * (double)(float)A.ns_coll * A.dt;
    if (Pmax > 1.0) { *A.status = -11; return; }
    double dNc = (double)N * Pmax / (double)(float)1;
    int64_t Nct = (int64_t)dNc;
    double rnd = ran2(seed);
    if (rnd <= dNc - (double)Nct) Nct += 1;
    Nc[q] = Nct;
    if (A.n_cell > 0 && A.counts[0] / A.n_cell >= 5 && q == 0 && Nct <= 1) { *A.status = -12; return; }
    double kk = A.h1 * A.h2 * A.zmax / T.Nm[q + 1];
    for (int rg = 0; rg < T.n_rg; ++rg) {
      double sum = 0.0;
      for (int iy = T.iyl; iy <= T.iyr; ++iy)
        for (int ix = T.ixl[rg]; ix <= T.ixr[rg]; ++ix) sum += A.np_red[q][node_index(ix, iy, A.n1)];
      rg_inv[q][rg] = 1.0 / (double)(int64_t)(kk * sum);
    }
  }
  for (int q = 0; q < A.nsp; ++q) {  // lines 167-179
    const SpeciesMcc& S = A.sp[q];
    const int pt = q + 1;
    const int64_t N0 = A.counts[q];
    if (S.p_ncol == 0 || T.mass[pt] <= 0 || Nc[q] == 0 || N0 == 0) continue;
    A.Nc_out[q] = Nc[q];
    float fnum = (float)Nc[q] / (float)N0;
    const double nu_omp = (double)fnum / ((double)(float)A.ns_coll * A.dt);
    double vz_sav = 0.0;
    double* X = A.P.d[q][0]; double* Y = A.P.d[q][1];
    double* VX = A.P.d[q][3]; double* VY = A.P.d[q][4]; double* VZ = A.P.d[q][5];
    for (int64_t ic = 1; ic <= Nc[q]; ++ic) {
      double r1 = ran2(seed);
      int64_t ip = (int64_t)((double)N0 * r1) + 1;
      if (ip == N0 + 1) { *A.status = -13; return; }
      const int64_t i0 = ip - 1;
      double xp = X[i0], yp = Y[i0];
      int ix = (int)(xp / A.h1) + 1, iy = (int)(yp / A.h2) + 1;
      double vx1 = VX[i0], vy1 = VY[i0], vz1 = VZ[i0];
      double vx2 = 0, vy2 = 0, vz2 = 0;
      int tsel[kMaxTypes + 1];
      for (int k = 0; k <= T.ntype_all; ++k) tsel[k] = 0;
      double Ekr[kMaxTypes + 1], vr[kMaxTypes + 1], mu[kMaxTypes + 1];
      double vxcm[kMaxTypes + 1], vycm[kMaxTypes + 1], vzcm[kMaxTypes + 1];
      double nu[kMaxPcol + 1], sarr[kMaxPcol + 1];
      for (int k = 0; k <= kMaxPcol; ++k) { nu[k] = 0.0; sarr[k] = 0.0; }
      double sum_nu = 0.0;
      for (int icol = 1; icol <= S.p_ncol; ++icol) {
        const int c = S.reac[icol - 1];
        const int tt = S.ttype[icol - 1];
        if (!tsel[tt]) {
          if (T.charge[tt] != 0.0) { *A.status = -14; return; }
          double ra = ran2(seed), rb = ran2(seed);
          load_gauss(vx2, vy2, T.vt0[tt], ra, rb, pi);
          if (vz_sav == 0.0) {
            ra = ran2(seed); rb = ran2(seed);
            load_gauss(vz2, vz_sav, T.vt0[tt], ra, rb, pi);
          } else {
            vz2 = vz_sav;
            vz_sav = 0.0;
          }
          double mp = fabs(T.mass[pt]), mt = fabs(T.mass[tt]);
          mu[tt] = mp * mt / (mp + mt);
          vr[tt] = sqrt((vx1 - vx2) * (vx1 - vx2) + (vy1 - vy2) * (vy1 - vy2) + (vz1 - vz2) * (vz1 - vz2));
          Ekr[tt] = (0.5 * mu[tt] * (vr[tt] * vr[tt])) / qe;
          double sm = mp + mt;
          vxcm[tt] = (mp * vx1 + mt * vx2) / sm;
          vycm[tt] = (mp * vy1 + mt * vy2) / sm;
          vzcm[tt] = (mp * vz1 + mt * vz2) / sm;
          if (Ekr[tt] > A.sig_Er[npt - 1]) { *A.status = -15; return; }
          tsel[tt] = 1;
        }
        int L = 1, R = npt;  // dichotomy, lines 485-497
        for (int q5 = 1; q5 <= 5; ++q5) {
          int M = (int)((double)(float)(R - L) / 2.0) + L;
          double EkL = A.sig_Er[L - 1], EkR = A.sig_Er[R - 1], EkM = A.sig_Er[M - 1];
          if (Ekr[tt] >= EkL && Ekr[tt] <= EkM) R = M;
          if (Ekr[tt] > EkM && Ekr[tt] <= EkR) L = M;
        }
        for (int ipt = L + 1; ipt <= R; ++ipt) {  // scan, lines 503-611
          double EkL = A.sig_Er[ipt - 2], EkR = A.sig_Er[ipt - 1];
          if (Ekr[tt] > EkL && Ekr[tt] <= EkR) {
            double sL = A.sig[(int64_t)(c - 1) * npt + ipt - 2], sR = A.sig[(int64_t)(c - 1) * npt + ipt - 1];
            double sig_p = sL + (Ekr[tt] - EkL) * (sR - sL) / (EkR - EkL);
            nu[icol] = T.np_mx[tt] * sig_p * vr[tt] / nu_omp;
            sum_nu = sum_nu + nu[icol];
            for (int rg = 0; rg < T.n_rg; ++rg)
              if (ix >= T.ixl[rg] && ix <= T.ixr[rg] && iy >= T.iyl && iy <= T.iyr) {
                A.cnt_col[(0 * (int64_t)T.ncol + (c - 1)) * T.n_rg + rg] += nu[icol] * nu_omp;
                A.cnt_col[(1 * (int64_t)T.ncol + (c - 1)) * T.n_rg + rg] += 1.0;
                break;
              }
            sarr[icol] = nu[icol];
            break;
          }
        }
      }
      int indx[kMaxPcol + 1];  // indexx.f90:20-34
      for (int j = 1; j <= S.p_ncol; ++j) indx[j] = j;
      for (int j = 2; j <= S.p_ncol; ++j) {
        int it = indx[j];
        double a = sarr[it];
        int i;
        for (i = j - 1; i >= 1; --i) {
          if (sarr[indx[i]] <= a) break;
          indx[i + 1] = indx[i];
        }
        indx[i + 1] = it;
      }
      double rnd = ran2(seed);
      int c_ind = 0;
      if (rnd <= sum_nu) {
        double cs = 0.0;
        for (int icol = 1; icol <= S.p_ncol; ++icol) {
          cs = cs + nu[indx[icol]];
          if (rnd <= cs) { c_ind = S.reac[indx[icol] - 1]; break; }
        }
      }
      if (!c_ind) continue;
      const int* ci = T.col_info[c_ind - 1];
      const int n_re = ci[0], n_by = ci[1], rt = ci[2 + n_re + n_by];
      int tt = 0;
      for (int k = 0; k < S.p_ncol; ++k) if (S.reac[k] == c_ind) tt = S.ttype[k];
      atomicAdd(&A.events[c_ind - 1], 1ull);
      if (rt != 4) {
        double Eth = T.Eth[c_ind - 1];
        double Ee = 0.5 * mu[tt] * (vr[tt] * vr[tt]) - Eth * qe;
        double sum_mass = 0.0;
        for (int k = 1; k <= n_by; ++k) sum_mass = sum_mass + 1.0 / fabs(T.mass[ci[1 + n_re + k]]);
        double ra = ran2(seed), rb = ran2(seed);
        double th_add = 2.0 * pi / (double)(float)n_by;
        double th = acos(1.0 - 2.0 * ra);
        double phi = 2.0 * pi * rb;
        double cos_ph = cos(phi), sin_ph = sin(phi);
        (void)ran2(seed);  // cos_th_s, phi_s of the discarded SCATTER call
        (void)ran2(seed);
        if (Ee < 0.0) { *A.status = -16; return; }
        int flag_add1 = -1;
        for (int iby = 1; iby <= n_by; ++iby) {
          const int bt = ci[1 + n_re + iby];
          if (T.mass[bt] <= 0) continue;
          int opt_add = 1;
          if (iby == 1 && bt == pt) { opt_add = 0; flag_add1 = 0; }
          if (rt == 1 || rt == 3) {
            if (bt == tt) { *A.status = -14; return; }
            opt_add = 0;
            flag_add1 = 0;
          }
          if (bt < 1 || bt > A.nsp) { *A.status = -14; return; }
          const int bq = bt - 1;
          int64_t ib;
          double v2old;
          if (opt_add == 1) {
            np_add[bt] += 1;
            ib = A.counts[bq] + np_add[bt] - 1;
            if (ib >= A.P.cap[bq]) { *A.status = -17; return; }
            A.P.d[bq][0][ib] = X[i0]; A.P.d[bq][1][ib] = Y[i0]; A.P.d[bq][2][ib] = A.P.d[q][2][i0];
            A.P.fd[bq][ib] = 0;
            A.P.src[bq][node_index(ix, iy, A.n1)] += 1.0;   // collisions.f90:809
            v2old = 0.0;
          } else {
            ib = i0;
            v2old = A.P.d[bq][3][ib] * A.P.d[bq][3][ib] + A.P.d[bq][4][ib] * A.P.d[bq][4][ib] +
                    A.P.d[bq][5][ib] * A.P.d[bq][5][ib];
          }
          double c_t = cos(th), s_t = sin(th);
          th = th + th_add;
          double ex = c_t, ey = s_t * sin_ph, ez = s_t * cos_ph;
          double mb = fabs(T.mass[bt]);
          double vp = sqrt(2.0 * (Ee / sum_mass) / (mb * mb));
          double bx = vxcm[tt] + vp * ex, by = vycm[tt] + vp * ey, bz = vzcm[tt] + vp * ez;
          A.P.d[bq][3][ib] = bx; A.P.d[bq][4][ib] = by; A.P.d[bq][5][ib] = bz;
          A.ploss3[bt] = A.ploss3[bt] + 0.5 * T.Nm[bt] * T.mass[bt] * (bx * bx + by * by + bz * bz - v2old);
        }
        for (int ire = 1; ire <= n_re; ++ire) {  // reactants, lines 908-935
          const int ct = ci[1 + ire];
          if (T.mass[ct] <= 0) continue;
          int fa = (ire == 1) ? flag_add1 : ((rt == 1 || rt == 3) ? 0 : -1);
          if (fa == -1) {
            if (ct == tt || ct != pt) { *A.status = -14; return; }
            A.P.fd[q][i0] = 1;
            A.P.sink[q][node_index(ix, iy, A.n1)] += 1.0;   // collisions.f90:931
          }
        }
      } else {  // charge exchange, lines 942-1000
        for (int iby = 1; iby <= n_by; ++iby) {
          const int bt = ci[1 + n_re + iby];
          if (T.mass[bt] <= 0) continue;
          if (bt != pt) { *A.status = -14; return; }
          double v2old = VX[i0] * VX[i0] + VY[i0] * VY[i0] + VZ[i0] * VZ[i0];
          VX[i0] = vx2; VY[i0] = vy2; VZ[i0] = vz2;
          A.ploss3[bt] = A.ploss3[bt] + 0.5 * T.Nm[bt] * T.mass[bt] * (vx2 * vx2 + vy2 * vy2 + vz2 * vz2 - v2old);
        }
      }
      for (int rg = 0; rg < T.n_rg; ++rg)
        if (ix >= T.ixl[rg] && ix <= T.ixr[rg] && iy >= T.iyl && iy <= T.iyr) {
          A.cnt_col[(2 * (int64_t)T.ncol + (c_ind - 1)) * T.n_rg + rg] += rg_inv[q][rg];
          break;
        }
    }
  }
  for (int q = 0; q < A.nsp; ++q) {
    A.counts[q] += np_add[q + 1];
    A.created[q] += (unsigned long long)np_add[q + 1];
  }
  *A.iseed = seed;
  *A.status = 0;
}

// ---------------------------------------------------------------------------
// production mode
// ---------------------------------------------------------------------------
struct ProdArgs {
  const MccTab* tab;
  const double* sig_Er;
  const double* sig;
  PartPtrs P;
  const int64_t* counts;
  int64_t N0;                    // incident species' count before this call's appends
  int mixed, nsp, n1, q;         // q: incident species (0-based)
  double h1, h2, qe, pi;
  SpeciesMcc S;
  uint32_t k0, k1, it;
  uint32_t* claim;
  int8_t* st_type;               // [Nc*kMaxCreate] byproduct species (1-based) or 0
  int64_t* st_ip;                // [Nc] incident particle of the candidate
  double* st_v;                  // [Nc*kMaxCreate*3]
  uint32_t* ncreate;             // [kMaxSpecies][Nc]
  int64_t Nc;
  double* cnt_col;
  double* ploss3;
  unsigned long long* events;
  unsigned long long* missed;
  int32_t* status;
};

struct Rng {
  uint32_t k0, k1, c0, c1, c2;
  uint32_t n;  // call counter
  U4 buf;
  int used;
  __device__ double next() {
    if (used >= 2) {
      buf = philox4x32_10(U4{c0, c1, c2, n++}, k0, k1);
      used = 0;
    }
    double u = used == 0 ? u01(buf.x, buf.y) : u01(buf.z, buf.w);
    ++used;
    return u;
  }
};

template <bool kMixed>
__device__ inline void read_v(const ProdArgs& A, int q, int64_t i, double& vx, double& vy, double& vz) {
  if (kMixed) { vx = A.P.f[q][3][i]; vy = A.P.f[q][4][i]; vz = A.P.f[q][5][i]; }
  else { vx = A.P.d[q][3][i]; vy = A.P.d[q][4][i]; vz = A.P.d[q][5][i]; }
}

template <bool kMixed>
__device__ inline void write_v(const ProdArgs& A, int q, int64_t i, double vx, double vy, double vz) {
  if (kMixed) { A.P.f[q][3][i] = (float)vx; A.P.f[q][4][i] = (float)vy; A.P.f[q][5][i] = (float)vz; }
  else { A.P.d[q][3][i] = vx; A.P.d[q][4][i] = vy; A.P.d[q][5][i] = vz; }
}

// kLds: the union energy grid and the incident species' cross-section rows
// are staged in LDS (dynamic, (1 + p_ncol) * npt doubles) before the
// candidates search them; otherwise they are read from global memory.
template <bool kMixed, bool kLds>
__global__ __launch_bounds__(256) void k_mcc_prod(ProdArgs A) {
  __shared__ double s_cnt[3][kMaxPcol][5];
  __shared__ unsigned long long s_ev[kMaxPcol];
  __shared__ double s_pl[kMaxSpecies + 1];
  extern __shared__ double s_tab[];
  const MccTab& T = *A.tab;
  const SpeciesMcc& S = A.S;
  const int t = threadIdx.x;
  const int npt = T.npt;
  for (int e = t; e < 3 * kMaxPcol * 5; e += 256) (&s_cnt[0][0][0])[e] = 0.0;
  if (t < kMaxPcol) s_ev[t] = 0;
  if (t <= kMaxSpecies) s_pl[t] = 0.0;
  if (kLds) {
    for (int e = t; e < npt; e += 256) s_tab[e] = A.sig_Er[e];
    for (int icol = 0; icol < S.p_ncol; ++icol) {
      const double* row = A.sig + (int64_t)(S.reac[icol] - 1) * npt;
      for (int e = t; e < npt; e += 256) s_tab[(int64_t)(1 + icol) * npt + e] = row[e];
    }
  }
  __syncthreads();
  const double* Er = kLds ? s_tab : A.sig_Er;
  const int64_t c = (int64_t)blockIdx.x * 256 + t;
  const int q = A.q, pt = q + 1;
  // candidates come from the particles present when the call began: products
  // appended to this species by an earlier species' collisions in the same
  // call do not collide (collisions.f90:184, 275), and the claim bitmap is
  // sized for this count
  const int64_t N = A.N0;
  bool active = c < A.Nc && N > 0;
  Rng rng{A.k0, A.k1, (uint32_t)c, (uint32_t)(c >> 32) ^ ((uint32_t)pt << 24), A.it, 0u, U4{0, 0, 0, 0}, 2};
  int64_t ip = -1;
  if (active) {
    for (int tr = 0; tr < 8 && ip < 0; ++tr) {
      int64_t cand = (int64_t)((double)N * rng.next());
      if (cand >= N) cand = N - 1;
      uint32_t bit = 1u << (cand & 31);
      if (!(atomicOr(&A.claim[cand >> 5], bit) & bit)) ip = cand;
    }
    if (ip < 0) { atomicAdd(A.missed, 1ull); active = false; }
  }
  int c_ind = 0, rg_hit = -1, ix = 0, iy = 0;
  if (active) {
    if (kMixed) { int cc = A.P.cell[q][ip]; ix = cc & 0xFFFF; iy = cc >> 16; }
    else { ix = (int)(A.P.d[q][0][ip] / A.h1) + 1; iy = (int)(A.P.d[q][1][ip] / A.h2) + 1; }
    for (int rg = 0; rg < T.n_rg; ++rg)
      if (ix >= T.ixl[rg] && ix <= T.ixr[rg] && iy >= T.iyl && iy <= T.iyr) { rg_hit = rg; break; }
    double vx1, vy1, vz1;
    read_v<kMixed>(A, q, ip, vx1, vy1, vz1);
    // one Maxwellian target per distinct target species (collisions.f90:425-437)
    int tts[4] = {0, 0, 0, 0};
    double tv[4][3], tmu[4], tvr[4], tE[4], tcm[4][3];
    int nts = 0;
    double nu[kMaxPcol], sum_nu = 0.0;
    for (int icol = 0; icol < S.p_ncol; ++icol) {
      const int tt = S.ttype[icol];
      int k = 0;
      while (k < nts && tts[k] != tt) ++k;
      if (k == nts && nts == 4) { atomicOr(A.status, 16); nu[icol] = 0.0; continue; }
      if (k == nts) {
        tts[k] = tt;
        ++nts;
        double a, b, z0, z1;
        double r1 = rng.next(), r2 = rng.next(), r3 = rng.next(), r4 = rng.next();
        load_gauss(a, b, T.vt0[tt], r1, r2, A.pi);
        load_gauss(z0, z1, T.vt0[tt], r3, r4, A.pi);
        tv[k][0] = a; tv[k][1] = b; tv[k][2] = z0;
        double mp = fabs(T.mass[pt]), mt = fabs(T.mass[tt]);
        tmu[k] = mp * mt / (mp + mt);
        double dx = vx1 - a, dy = vy1 - b, dz = vz1 - z0;
        tvr[k] = sqrt(dx * dx + dy * dy + dz * dz);
        tE[k] = 0.5 * tmu[k] * tvr[k] * tvr[k] / A.qe;
        tcm[k][0] = (mp * vx1 + mt * a) / (mp + mt);
        tcm[k][1] = (mp * vy1 + mt * b) / (mp + mt);
        tcm[k][2] = (mp * vz1 + mt * z0) / (mp + mt);
      }
      nu[icol] = 0.0;
      const double E = tE[k];
      if (E > Er[0] && E <= Er[npt - 1]) {
        int lo = 0, hi = npt - 1;  // Er[lo] < E <= Er[hi]
        while (hi - lo > 1) {
          int mid = (lo + hi) >> 1;
          if (Er[mid] < E) lo = mid; else hi = mid;
        }
        const double* row = kLds ? s_tab + (int64_t)(1 + icol) * npt : A.sig + (int64_t)(S.reac[icol] - 1) * npt;
        double sL = row[lo], sR = row[hi];
        double EkL = Er[lo], EkR = Er[hi];
        double sig_p = sL + (E - EkL) * (sR - sL) / (EkR - EkL);
        nu[icol] = T.np_mx[tt] * sig_p * tvr[k] / S.nu_omp;
        sum_nu += nu[icol];
        if (rg_hit >= 0) {
          atomicAdd(&s_cnt[0][icol][rg_hit], nu[icol] * S.nu_omp);
          atomicAdd(&s_cnt[1][icol][rg_hit], 1.0);
        }
      } else if (E > Er[npt - 1]) {
        atomicOr(A.status, 2);  // energy above the table (the reference asks on stdin)
      }
    }
    int order[kMaxPcol];
    for (int j = 0; j < S.p_ncol; ++j) order[j] = j;
    for (int j = 1; j < S.p_ncol; ++j) {  // ascending, stable
      int v = order[j];
      int i = j - 1;
      while (i >= 0 && nu[order[i]] > nu[v]) { order[i + 1] = order[i]; --i; }
      order[i + 1] = v;
    }
    const double u = rng.next();
    int icol_hit = -1;
    if (u <= sum_nu) {
      double cs = 0.0;
      for (int j = 0; j < S.p_ncol; ++j) {
        cs += nu[order[j]];
        if (u <= cs) { icol_hit = order[j]; break; }
      }
    }
    if (icol_hit >= 0) {
      c_ind = S.reac[icol_hit];
      const int tt = S.ttype[icol_hit];
      int k = 0;
      while (k < nts - 1 && tts[k] != tt) ++k;
      const int* ci = T.col_info[c_ind - 1];
      const int n_re = ci[0], n_by = ci[1], rt = ci[2 + n_re + n_by];
      atomicAdd(&s_ev[icol_hit], 1ull);
      if (rg_hit >= 0) atomicAdd(&s_cnt[2][icol_hit][rg_hit], S.rg_inv[rg_hit]);
      int slot = 0;
      A.st_ip[c] = ip;
      if (rt != 4) {
        const double Ee = 0.5 * tmu[k] * tvr[k] * tvr[k] - T.Eth[c_ind - 1] * A.qe;
        if (Ee < 0.0) atomicOr(A.status, 4);
        double sum_mass = 0.0;
        for (int j = 1; j <= n_by; ++j) sum_mass += 1.0 / fabs(T.mass[ci[1 + n_re + j]]);
        const double th_add = 2.0 * A.pi / (double)n_by;
        double th = acos(1.0 - 2.0 * rng.next());
        const double ph = 2.0 * A.pi * rng.next();
        const double cph = cos(ph), sph = sin(ph);
        bool incident_kept = false;
        for (int iby = 1; iby <= n_by; ++iby) {
          const int bt = ci[1 + n_re + iby];
          if (T.mass[bt] <= 0) continue;
          const double ct = cos(th), st = sin(th);
          th += th_add;
          const double mb = fabs(T.mass[bt]);
          const double vp = sqrt(2.0 * (fmax(Ee, 0.0) / sum_mass) / (mb * mb));
          const double bx = tcm[k][0] + vp * ct, by = tcm[k][1] + vp * st * sph, bz = tcm[k][2] + vp * st * cph;
          const bool update_incident = (iby == 1 && bt == pt) || rt == 1 || rt == 3;
          if (update_incident) {
            incident_kept = true;
            double o0, o1, o2;
            read_v<kMixed>(A, q, ip, o0, o1, o2);
            write_v<kMixed>(A, q, ip, bx, by, bz);
            atomicAdd(&s_pl[pt], 0.5 * T.Nm[pt] * T.mass[pt] * (bx * bx + by * by + bz * bz - (o0 * o0 + o1 * o1 + o2 * o2)));
          } else if (slot < kMaxCreate && bt >= 1 && bt <= A.nsp) {
            const int64_t so = c * kMaxCreate + slot;
            A.st_type[so] = (int8_t)bt;
            A.st_v[3 * so] = bx; A.st_v[3 * so + 1] = by; A.st_v[3 * so + 2] = bz;
            A.ncreate[(int64_t)(bt - 1) * A.Nc + c] += 1;
            ++slot;
            atomic_add_f64(&A.P.src[bt - 1][node_index(ix, iy, A.n1)], 1.0);   // collisions.f90:809
            atomicAdd(&s_pl[bt], 0.5 * T.Nm[bt] * T.mass[bt] * (bx * bx + by * by + bz * bz));
          }
        }
        if (!incident_kept && T.mass[pt] > 0) {  // reactant killed
          A.P.fd[q][ip] = 1;
          atomic_add_f64(&A.P.sink[q][node_index(ix, iy, A.n1)], 1.0);   // collisions.f90:931
        }
      } else {
        double o0, o1, o2;
        read_v<kMixed>(A, q, ip, o0, o1, o2);
        write_v<kMixed>(A, q, ip, tv[k][0], tv[k][1], tv[k][2]);
        atomicAdd(&s_pl[pt], 0.5 * T.Nm[pt] * T.mass[pt] *
                                 (tv[k][0] * tv[k][0] + tv[k][1] * tv[k][1] + tv[k][2] * tv[k][2] -
                                  (o0 * o0 + o1 * o1 + o2 * o2)));
      }
    }
  }
  __syncthreads();
  for (int e = t; e < 3 * kMaxPcol * 5; e += 256) {
    const int kk = e / (kMaxPcol * 5), icol = (e / 5) % kMaxPcol, rg = e % 5;
    const double v = s_cnt[kk][icol][rg];
    if (v != 0.0 && icol < S.p_ncol && rg < T.n_rg)
      atomic_add_f64(&A.cnt_col[((int64_t)kk * T.ncol + (S.reac[icol] - 1)) * T.n_rg + rg], v);
  }
  if (t < S.p_ncol && s_ev[t]) atomicAdd(&A.events[S.reac[t] - 1], s_ev[t]);
  if (t <= kMaxSpecies && t >= 1 && s_pl[t] != 0.0) atomic_add_f64(&A.ploss3[t], s_pl[t]);
}

// append staged particles at N_b + exclusive(ncreate_b)[c], in candidate
// order, at the position of the candidate's incident particle
// (collisions.f90:804-816)
template <bool kMixed>
__global__ void k_mcc_append(ProdArgs A, const uint32_t* __restrict__ offs, int b) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= A.Nc) return;
  const int bt = b + 1, q = A.q;
  const int64_t base = A.counts[b] + offs[c];
  int k = 0;
  for (int s = 0; s < kMaxCreate; ++s) {
    const int64_t so = c * kMaxCreate + s;
    if (A.st_type[so] != bt) continue;
    const int64_t dst = base + k++;
    const int64_t ip = A.st_ip[c];
    if (dst >= A.P.cap[b]) { atomicOr(A.status, 8); continue; }
    if (kMixed) {
      A.P.cell[b][dst] = A.P.cell[q][ip];
      A.P.f[b][0][dst] = A.P.f[q][0][ip];
      A.P.f[b][1][dst] = A.P.f[q][1][ip];
      if (A.P.f[b][2]) A.P.f[b][2][dst] = A.P.f[q][2][ip];
      A.P.f[b][3][dst] = (float)A.st_v[3 * so];
      A.P.f[b][4][dst] = (float)A.st_v[3 * so + 1];
      A.P.f[b][5][dst] = (float)A.st_v[3 * so + 2];
    } else {
      A.P.d[b][0][dst] = A.P.d[q][0][ip];
      A.P.d[b][1][dst] = A.P.d[q][1][ip];
      A.P.d[b][2][dst] = A.P.d[q][2][ip];
      A.P.d[b][3][dst] = A.st_v[3 * so];
      A.P.d[b][4][dst] = A.st_v[3 * so + 1];
      A.P.d[b][5][dst] = A.st_v[3 * so + 2];
    }
    A.P.fd[b][dst] = 0;
  }
}

__global__ void k_add_count(int64_t* counts, int b, const uint32_t* offs, const uint32_t* cnt,
                            int64_t Nc, unsigned long long* created) {
  if (threadIdx.x || blockIdx.x || Nc <= 0) return;
  const uint64_t tot = (uint64_t)offs[Nc - 1] + cnt[Nc - 1];
  counts[b] += (int64_t)tot;
  created[b] += tot;
}

// np_tot_rg of collisions.f90:144-154: INT(k * SUM(np_red over the region)),
// summed in the reference's column-major order by one lane per region.
// Particles of one region (collisions.f90:144-154) for production mode: a
// fixed-order block reduction (thread partials over a strided walk, then the
// 16 wave sums in order), so the result is reproducible run to run; check
// mode keeps the reference's sequential sum inside k_mcc_check.
// 1/np_tot_rg of a region (collisions.f90): the sum of np_red over the
// region's nodes in two stages, each in a fixed order: block b sums rows
// iyl + b, iyl + b + kRegionBlocks, ... (lanes along the row), then one wave
// sums the block partials.
__global__ __launch_bounds__(256) void k_region_part(const double* __restrict__ np_red, int n1, int ixl, int ixr,
                                                     int iyl, int iyr, double* __restrict__ part) {
  __shared__ double ws[4];
  double s = 0.0;
  for (int iy = iyl + (int)blockIdx.x; iy <= iyr; iy += gridDim.x)
    for (int ix = ixl + (int)threadIdx.x; ix <= ixr; ix += blockDim.x) s += np_red[node_index(ix, iy, n1)];
  s = wave_sum(s);
  if ((threadIdx.x & 63) == 0) ws[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) part[blockIdx.x] = ((ws[0] + ws[1]) + ws[2]) + ws[3];
}

__global__ __launch_bounds__(64) void k_region_final(const double* __restrict__ part, int np, double kk,
                                                     double* __restrict__ out) {
  double s = 0.0;
  for (int q = threadIdx.x; q < np; q += 64) s += part[q];
  s = wave_sum(s);
  if (threadIdx.x == 0) *out = 1.0 / (double)(int64_t)(kk * s);
}

}  // namespace

// ===========================================================================
// host side
// ===========================================================================
namespace {

int build_species(lepic_ctx* ctx, SpeciesMcc* out) {
  const Reactions& R = ctx->rx;
  for (int q = 0; q < ctx->nsp; ++q) {
    SpeciesMcc& S = out[q];
    S = SpeciesMcc{};
    S.ptype = q + 1;
    const int pt = q + 1;
    S.p_ncol = (pt < (int)R.p_ncol.size()) ? R.p_ncol[pt] : 0;
    if (S.p_ncol > kMaxPcol) return ctx->fail(LEPIC_ERR_UNSUPPORTED, "more than 16 reactions for one species");
    double nu = 0.0;
    for (int icol = 1; icol <= S.p_ncol; ++icol) {
      const int c = R.sig_list[(size_t)pt * R.ncol + (icol - 1)];
      const int n_re = R.col_info[(size_t)(c - 1) * 10 + 0];
      const int tt = R.col_info[(size_t)(c - 1) * 10 + 1 + n_re];  // col_info(c, ind_nby+n_re)
      S.reac[icol - 1] = c;
      S.ttype[icol - 1] = tt;
      if (R.charge[tt] != 0.0) return ctx->fail(LEPIC_ERR_UNSUPPORTED, "charged (DSMC) collision targets are not supported by this library");
      nu = nu + R.np_mx[tt] * R.sigv_mx[(size_t)pt * R.ncol + (c - 1)];  // collisions.f90:103
    }
    S.nu_max = std::min(nu, R.nu_uplim[pt]);  // line 106
  }
  return 0;
}

int upload_tab(lepic_ctx* ctx, MccTab** dtab) {
  const Reactions& R = ctx->rx;
  MccTab T{};
  T.ncol = R.ncol; T.npt = R.npt; T.ntype_all = R.ntype_all; T.n_rg = R.n_rg;
  T.iyl = R.iyl_rg; T.iyr = R.iyr_rg;
  for (int k = 0; k < 5; ++k) { T.ixl[k] = R.ixl_rg[k]; T.ixr[k] = R.ixr_rg[k]; }
  for (int c = 0; c < R.ncol; ++c) {
    for (int k = 0; k < 10; ++k) T.col_info[c][k] = R.col_info[(size_t)c * 10 + k];
    T.Eth[c] = R.Eth[c];
  }
  for (int p = 0; p <= R.ntype_all; ++p) {
    T.charge[p] = R.charge[p]; T.mass[p] = R.mass[p]; T.Nm[p] = R.Nm[p];
    T.vt0[p] = R.vt0[p]; T.np_mx[p] = R.np_mx[p];
  }
  static_assert(sizeof(MccTab) < 16384, "table too large");
  if (!*dtab) LEPIC_HIP(hipMalloc(dtab, sizeof(MccTab)));
  LEPIC_HIP(hipMemcpyAsync(*dtab, &T, sizeof(MccTab), hipMemcpyHostToDevice, ctx->stream));
  return 0;
}

}  // namespace

namespace lepic {

int mcc_run(lepic_ctx* ctx, int it, int mode, uint64_t seed, int32_t* iseed, lepic_mcc_info* info) {
  Reactions& R = ctx->rx;
  if (!R.set) return ctx->fail(LEPIC_ERR_STATE, "lepic_set_reactions not called");
  if (R.flag_ionrep != 0) return ctx->fail(LEPIC_ERR_UNSUPPORTED, "flag_ionrep >= 1 is not supported by this library");
  SpeciesMcc S[kMaxSpecies];
  if (int rc = build_species(ctx, S)) return rc;
  for (int p = 0; p < ctx->nsp; ++p) ctx->sp[p].dead_maybe = true;   // collisions flag their reactants
  MccTab* dtab = (MccTab*)ctx->d_mcc_tab;
  if (int rc = upload_tab(ctx, &dtab)) return rc;
  ctx->d_mcc_tab = dtab;
  int32_t* d_status = ctx->d_error + 4;
  unsigned long long* d_events = (unsigned long long*)ctx->d_mcc_i64;        // [32]
  unsigned long long* d_created = (unsigned long long*)ctx->d_mcc_i64 + 40;  // [8]
  unsigned long long* d_missed = (unsigned long long*)ctx->d_mcc_i64 + 48;
  if (R.ncol > 32) return ctx->fail(LEPIC_ERR_UNSUPPORTED, "more than 32 reactions");
  if (int rc = zero_async(ctx, ctx->d_mcc_i64, 32 * 8)) return rc;
  if (int rc = zero_async(ctx, d_created, 16 * 8)) return rc;
  if (int rc = zero_async(ctx, d_status, 4)) return rc;
  double* d_ploss3 = ctx->d_mcc_f64;  // [ptype], accumulated until lepic_reset_tallies
  // candidates are drawn by particle index: dense arrays
  for (int q = 0; q < ctx->nsp; ++q)
    if (int rc = ensure_packed(ctx, q)) return rc;
  // counts on the host
  int64_t N[kMaxSpecies] = {};
  LEPIC_HIP(hipMemcpyAsync(ctx->h_pinned, ctx->d_counts, kMaxSpecies * 8, hipMemcpyDeviceToHost, ctx->stream));
  LEPIC_HIP(hipStreamSynchronize(ctx->stream));
  std::memcpy(N, ctx->h_pinned, kMaxSpecies * 8);
  PartPtrs P = part_ptrs(ctx);
  const double pi = 4.0 * std::atan(1.0);
  if (info) std::memset(info, 0, sizeof(*info));
  if (mode == LEPIC_MCC_CHECK) {
    if (ctx->g.precision != LEPIC_FP64) return ctx->fail(LEPIC_ERR_UNSUPPORTED, "check mode needs the FP64 particle state");
    if (!iseed) return LEPIC_ERR_ARG;
    CheckArgs A{};
    A.tab = dtab; A.sig_Er = R.d_sig_Er; A.sig = R.d_sig; A.P = P; A.counts = ctx->d_counts;
    for (int q = 0; q < ctx->nsp; ++q) A.np_red[q] = ctx->sp[q].np_red;
    A.nsp = ctx->nsp; A.n1 = ctx->g.n1; A.h1 = ctx->g.h1; A.h2 = ctx->g.h2; A.zmax = ctx->g.zmax;
    A.dt = ctx->g.dt; A.qe = ctx->g.qe; A.pi = pi; A.ns_coll = R.ns_coll; A.n_cell = R.n_cell;
    for (int q = 0; q < ctx->nsp; ++q) A.sp[q] = S[q];
    int32_t* d_seed = ctx->d_error + 8;
    LEPIC_HIP(hipMemcpyAsync(d_seed, iseed, 4, hipMemcpyHostToDevice, ctx->stream));
    A.iseed = d_seed;
    A.cnt_col = R.d_cnt_col; A.ploss3 = d_ploss3; A.events = d_events; A.created = d_created;
    int64_t* d_nc = (int64_t*)((unsigned long long*)ctx->d_mcc_i64 + 56);
    if (int rc = zero_async(ctx, d_nc, 8 * 8)) return rc;
    A.Nc_out = d_nc;
    A.status = d_status;
    hipLaunchKernelGGL(k_mcc_check, dim3(1), dim3(64), 0, ctx->stream, A);
    LEPIC_HIP(hipGetLastError());
    int32_t st = 0, sd = 0;
    LEPIC_HIP(hipMemcpyAsync(ctx->h_pinned, d_status, 4, hipMemcpyDeviceToHost, ctx->stream));
    LEPIC_HIP(hipMemcpyAsync((char*)ctx->h_pinned + 4, d_seed, 4, hipMemcpyDeviceToHost, ctx->stream));
    LEPIC_HIP(hipMemcpyAsync((char*)ctx->h_pinned + 8, d_nc, 64, hipMemcpyDeviceToHost, ctx->stream));
    LEPIC_HIP(hipStreamSynchronize(ctx->stream));
    std::memcpy(&st, ctx->h_pinned, 4);
    std::memcpy(&sd, (char*)ctx->h_pinned + 4, 4);
    if (info) std::memcpy(info->Nc, (char*)ctx->h_pinned + 8, 64);
    if (st) return ctx->fail(LEPIC_ERR_MCC, "collisions aborted (code " + std::to_string(st) + ")");
    *iseed = sd;
    if (info) info->seed = sd;
  } else {
    const uint32_t k0 = (uint32_t)seed, k1 = (uint32_t)(seed >> 32);
    for (int q = 0; q < ctx->nsp; ++q) {
      SpeciesMcc& Sq = S[q];
      const int pt = q + 1;
      if (Sq.p_ncol == 0 || R.mass[pt] <= 0 || N[q] == 0) continue;
      // number of candidates (collisions.f90:116-132); the rounding draw
      // comes from the Philox stream (it, species, ~0)
      const double Pmax = Sq.nu_max * (double)(float)R.ns_coll * ctx->g.dt;
      if (Pmax > 1.0) return ctx->fail(LEPIC_ERR_MCC, "collision probability greater than 1");
      const double dNc = (double)N[q] * Pmax;
      int64_t Nc = (int64_t)dNc;
      U4 r = philox4x32_10(U4{0xFFFFFFFFu, (uint32_t)pt, (uint32_t)it, 0u}, k0, k1);
      if (u01(r.x, r.y) <= dNc - (double)Nc) Nc += 1;
      if (Nc == 0) continue;
      const float fnum = (float)Nc / (float)N[q];
      Sq.nu_omp = (double)fnum / ((double)(float)R.ns_coll * ctx->g.dt);
      Sq.Nc = Nc;
      if (info) { info->Nc[q] = Nc; info->nu_max[q] = Sq.nu_max; }
      // 1/np_tot_rg per region
      double* d_rg = ctx->d_mcc_f64 + 40;
      const double kk = ctx->g.h1 * ctx->g.h2 * ctx->g.zmax / R.Nm[pt];
      for (int rg = 0; rg < R.n_rg; ++rg) {
        double* part = ctx->d_mcc_f64 + kMccF64Head + (size_t)rg * kRegionBlocks;
        hipLaunchKernelGGL(k_region_part, dim3(kRegionBlocks), dim3(256), 0, ctx->stream,
                           (const double*)ctx->sp[q].np_red, ctx->g.n1, R.ixl_rg[rg], R.ixr_rg[rg], R.iyl_rg,
                           R.iyr_rg, part);
        hipLaunchKernelGGL(k_region_final, dim3(1), dim3(64), 0, ctx->stream, (const double*)part, kRegionBlocks,
                           kk, d_rg + rg);
      }
      LEPIC_HIP(hipMemcpyAsync(ctx->h_pinned, d_rg, 5 * 8, hipMemcpyDeviceToHost, ctx->stream));
      // scratch: claim bitmap, staging
      const int64_t words = (N[q] + 31) / 32;
      if (words > ctx->claim_cap) {
        if (ctx->d_claim) LEPIC_HIP(hipFree(ctx->d_claim));
        LEPIC_HIP(hipMalloc(&ctx->d_claim, words * 4));
        ctx->claim_cap = words;
      }
      // staging: velocities, incident index, per-species create counts, the
      // scanned copy of one species' counts, byproduct types
      const size_t stage_bytes = (size_t)Nc * (kMaxCreate * (1 + 24) + 8 + 4 * kMaxSpecies + 4) + 1024;
      if ((int64_t)stage_bytes > ctx->stage_cap) {
        if (ctx->d_mcc_stage) LEPIC_HIP(hipFree(ctx->d_mcc_stage));
        LEPIC_HIP(hipMalloc(&ctx->d_mcc_stage, stage_bytes));
        ctx->stage_cap = (int64_t)stage_bytes;
      }
      char* st = (char*)ctx->d_mcc_stage;
      double* st_v = (double*)st;                                   st += (size_t)Nc * kMaxCreate * 24;
      int64_t* st_ip = (int64_t*)st;                                st += (size_t)Nc * 8;
      uint32_t* ncreate = (uint32_t*)st;                            st += (size_t)Nc * 4 * kMaxSpecies;
      uint32_t* offs = (uint32_t*)st;                               st += (size_t)Nc * 4;
      int8_t* st_type = (int8_t*)st;
      if (int rc = zero_async(ctx, ctx->d_claim, words * 4)) return rc;
      if (int rc = zero_async(ctx, ncreate, (size_t)Nc * 4 * kMaxSpecies)) return rc;
      if (int rc = zero_async(ctx, st_type, (size_t)Nc * kMaxCreate)) return rc;
      LEPIC_HIP(hipStreamSynchronize(ctx->stream));
      for (int rg = 0; rg < R.n_rg; ++rg) std::memcpy(&Sq.rg_inv[rg], ctx->h_pinned + rg, 8);
      ProdArgs A{};
      A.tab = dtab; A.sig_Er = R.d_sig_Er; A.sig = R.d_sig; A.P = P; A.counts = ctx->d_counts;
      A.N0 = N[q];
      A.mixed = ctx->g.precision == LEPIC_MIXED; A.nsp = ctx->nsp; A.n1 = ctx->g.n1; A.q = q;
      A.h1 = ctx->g.h1; A.h2 = ctx->g.h2; A.qe = ctx->g.qe; A.pi = pi; A.S = Sq;
      A.k0 = k0; A.k1 = k1; A.it = (uint32_t)it;
      A.claim = ctx->d_claim; A.st_type = st_type; A.st_ip = st_ip; A.st_v = st_v;
      A.ncreate = ncreate; A.Nc = Nc;
      A.cnt_col = R.d_cnt_col; A.ploss3 = d_ploss3; A.events = d_events; A.missed = d_missed;
      A.status = d_status;
      const dim3 g((unsigned)((Nc + 255) / 256));
      // cross-section rows of this species in LDS when they fit in 48 KB
      const size_t tab_bytes = (size_t)(1 + Sq.p_ncol) * R.npt * 8;
      const bool lds = tab_bytes <= kMccLdsTable;
      if (A.mixed && lds) hipLaunchKernelGGL((k_mcc_prod<true, true>), g, dim3(256), tab_bytes, ctx->stream, A);
      else if (A.mixed) hipLaunchKernelGGL((k_mcc_prod<true, false>), g, dim3(256), 0, ctx->stream, A);
      else if (lds) hipLaunchKernelGGL((k_mcc_prod<false, true>), g, dim3(256), tab_bytes, ctx->stream, A);
      else hipLaunchKernelGGL((k_mcc_prod<false, false>), g, dim3(256), 0, ctx->stream, A);
      LEPIC_HIP(hipGetLastError());
      for (int b = 0; b < ctx->nsp; ++b) {
        uint32_t* cnt = ncreate + (size_t)b * Nc;
        // keep the raw counts for the total: scan a copy
        LEPIC_HIP(hipMemcpyAsync(offs, cnt, (size_t)Nc * 4, hipMemcpyDeviceToDevice, ctx->stream));
        if (int rc = scan_exclusive_u32(ctx, offs, Nc)) return rc;
        if (A.mixed) hipLaunchKernelGGL(k_mcc_append<true>, g, dim3(256), 0, ctx->stream, A, (const uint32_t*)offs, b);
        else hipLaunchKernelGGL(k_mcc_append<false>, g, dim3(256), 0, ctx->stream, A, (const uint32_t*)offs, b);
        hipLaunchKernelGGL(k_add_count, dim3(1), dim3(64), 0, ctx->stream, ctx->d_counts, b,
                           (const uint32_t*)offs, (const uint32_t*)cnt, Nc, d_created);
        LEPIC_HIP(hipGetLastError());
      }
    }
    int32_t st = 0;
    LEPIC_HIP(hipMemcpyAsync(ctx->h_pinned, d_status, 4, hipMemcpyDeviceToHost, ctx->stream));
    LEPIC_HIP(hipStreamSynchronize(ctx->stream));
    std::memcpy(&st, ctx->h_pinned, 4);
    if (st & 8) return ctx->fail(LEPIC_ERR_CAPACITY, "particle capacity exceeded by collisions");
    if (st & 4) return ctx->fail(LEPIC_ERR_MCC, "negative collision energy");
    if (st & 2) return ctx->fail(LEPIC_ERR_MCC, "relative energy above the cross-section table");
    if (st & 16) return ctx->fail(LEPIC_ERR_UNSUPPORTED, "more than 4 target species for one incident species");
  }
  // counts and tallies back to the host
  unsigned long long ev[32], cr[8], ms = 0;
  LEPIC_HIP(hipMemcpyAsync(ctx->h_pinned, d_events, 32 * 8, hipMemcpyDeviceToHost, ctx->stream));
  LEPIC_HIP(hipStreamSynchronize(ctx->stream));
  std::memcpy(ev, ctx->h_pinned, 32 * 8);
  LEPIC_HIP(hipMemcpyAsync(ctx->h_pinned, d_created, 8 * 8, hipMemcpyDeviceToHost, ctx->stream));
  LEPIC_HIP(hipMemcpyAsync(ctx->h_pinned + 8, d_missed, 8, hipMemcpyDeviceToHost, ctx->stream));
  LEPIC_HIP(hipStreamSynchronize(ctx->stream));
  std::memcpy(cr, ctx->h_pinned, 8 * 8);
  std::memcpy(&ms, ctx->h_pinned + 8, 8);
  for (int q = 0; q < ctx->nsp; ++q) {
    ctx->sp[q].ub = N[q] + (int64_t)cr[q];
    if (int rc = set_packed_counts(ctx, q)) return rc;  // appended particles extend the last segments
  }
  if (info) {
    for (int c = 0; c < R.ncol && c < 32; ++c) info->events[c] = (int64_t)ev[c];
    for (int q = 0; q < ctx->nsp; ++q) info->created[q] = (int64_t)cr[q];
    info->missed = (int64_t)ms;
    for (int q = 0; q < ctx->nsp; ++q) info->nu_max[q] = S[q].nu_max;
  }
  for (int q = 0; q < ctx->nsp; ++q)
    if (ctx->sp[q].ub > ctx->sp[q].cap) return ctx->fail(LEPIC_ERR_CAPACITY, "nmax is too small (collisions.f90:187-191)");
  return 0;
}

}  // namespace lepic

extern "C" {

int lepic_set_reactions(lepic_ctx* ctx, const lepic_reactions* r) {
  if (!ctx || !r || r->ncol < 0 || r->ncol > kMaxReac || r->npt < 2 || r->ntype_all < 1 ||
      r->ntype_all > kMaxTypes || r->n_rg < 0 || r->n_rg > kMaxRegions)
    return LEPIC_ERR_ARG;
  if (r->ntype_all < ctx->nsp) return ctx->fail(LEPIC_ERR_ARG, "ntype_all < nspecies");
  Reactions& R = ctx->rx;
  R.ntype_all = r->ntype_all; R.ncol = r->ncol; R.npt = r->npt; R.ns_coll = r->ns_coll;
  R.n_cell = r->n_cell; R.n_rg = r->n_rg; R.iyl_rg = r->iyl_rg; R.iyr_rg = r->iyr_rg;
  for (int k = 0; k < 5; ++k) { R.ixl_rg[k] = r->ixl_rg[k]; R.ixr_rg[k] = r->ixr_rg[k]; }
  R.flag_ionrep = r->flag_ionrep;
  const int nt = r->ntype_all + 1;
  R.col_info.assign(r->col_info, r->col_info + (size_t)r->ncol * 10);
  R.p_ncol.assign(r->p_ncol, r->p_ncol + nt);
  R.sig_list.assign(r->sig_list, r->sig_list + (size_t)nt * r->ncol);
  R.sigv_mx.assign(r->sigv_mx, r->sigv_mx + (size_t)nt * r->ncol);
  R.Eth.assign(r->Eth, r->Eth + r->ncol);
  R.charge.assign(r->charge, r->charge + nt);
  R.mass.assign(r->mass, r->mass + nt);
  R.Nm.assign(r->Nm, r->Nm + nt);
  R.vt0.assign(r->vt0, r->vt0 + nt);
  R.np_mx.assign(r->np_mx, r->np_mx + nt);
  R.nu_uplim.assign(r->nu_uplim, r->nu_uplim + nt);
  LEPIC_HIP(hipStreamSynchronize(ctx->stream));
  if (R.d_sig_Er) LEPIC_HIP(hipFree(R.d_sig_Er));
  if (R.d_sig) LEPIC_HIP(hipFree(R.d_sig));
  if (R.d_cnt_col) LEPIC_HIP(hipFree(R.d_cnt_col));
  LEPIC_HIP(hipMalloc(&R.d_sig_Er, (size_t)r->npt * 8));
  LEPIC_HIP(hipMemcpyAsync(R.d_sig_Er, r->sig_Er, (size_t)r->npt * 8, hipMemcpyHostToDevice, ctx->stream));
  LEPIC_HIP(hipStreamSynchronize(ctx->stream));
  LEPIC_HIP(hipMalloc(&R.d_sig, (size_t)r->npt * std::max(r->ncol, 1) * 8));
  LEPIC_HIP(hipMemcpyAsync(R.d_sig, r->sig, (size_t)r->npt * r->ncol * 8, hipMemcpyHostToDevice, ctx->stream));
  LEPIC_HIP(hipStreamSynchronize(ctx->stream));
  const size_t ncc = 3 * (size_t)std::max(r->ncol, 1) * std::max(r->n_rg, 1);
  LEPIC_HIP(hipMalloc(&R.d_cnt_col, ncc * 8));
  LEPIC_HIP(hipMemsetAsync(R.d_cnt_col, 0, ncc * 8, ctx->stream));
  R.set = true;
  return 0;
}

int lepic_collisions(lepic_ctx* ctx, int32_t it, int32_t mode, uint64_t seed, int32_t* iseed,
                     lepic_mcc_info* info) {
  if (!ctx || (mode != LEPIC_MCC_CHECK && mode != LEPIC_MCC_PROD)) return LEPIC_ERR_ARG;
  return mcc_run(ctx, it, mode, seed, iseed, info);
}

int lepic_get_cnt_col(lepic_ctx* ctx, double* cnt_col) {
  if (!ctx || !cnt_col) return LEPIC_ERR_ARG;
  if (!ctx->rx.set) return ctx->fail(LEPIC_ERR_STATE, "no reactions");
  LEPIC_HIP(hipStreamSynchronize(ctx->stream));
  LEPIC_HIP(hipMemcpyAsync(cnt_col, ctx->rx.d_cnt_col, 3 * (size_t)ctx->rx.ncol * ctx->rx.n_rg * 8, hipMemcpyDeviceToHost, ctx->stream));
  LEPIC_HIP(hipStreamSynchronize(ctx->stream));
  return 0;
}

int lepic_get_mcc_ploss(lepic_ctx* ctx, double* P_loss3) {
  if (!ctx || !P_loss3) return LEPIC_ERR_ARG;
  LEPIC_HIP(hipStreamSynchronize(ctx->stream));
  LEPIC_HIP(hipMemcpyAsync(P_loss3, ctx->d_mcc_f64, (kMaxSpecies + 1) * 8, hipMemcpyDeviceToHost, ctx->stream));
  LEPIC_HIP(hipStreamSynchronize(ctx->stream));
  return 0;
}

}  // extern "C"
