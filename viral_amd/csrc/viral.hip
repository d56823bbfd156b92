// viral.hip -- MI355X (gfx950) implementation of the VIRAL viscous RK step
// behind the C ABI of include/viral.h.
//
// Device layout per event: structure of arrays, one ghosted FP64 array per
// field in the reference's HydroGrid[i][j][k] order (eta fastest), two copies
// of the state (the stage reads one and writes the other), RK registers
// Var0/Var1/Var2, a Result accumulator, face fluxes and face speeds of the
// current direction.
//
// One RK stage (NEWEST10/time.h:338-392 for RK2) is the launch sequence
//   for each direction d:  flux<d>   (face speeds + WENO + Rusanov face fluxes)
//   update   (flux divergence + sources + RK combination + T/pi from Var +
//             Newton root + velocities)
//   k0_broadcast (q2)    ghost_fill
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <mutex>
#include <algorithm>
#include <type_traits>
#include <string>
#include <vector>

#include "../../include/viral.h"
#include "adjoint.h"
#include "eos_table.h"
#include "physics.h"

using namespace viral;

namespace {

constexpr int NF = VIRAL_NFIELDS;
constexpr int SV = 15;
static_assert(SV % 2 == 1, "the SHASTA tile pairs rows and ends on an even one");
constexpr int BLK = 256;

struct Geom {
  int gx, gy, gz;
  int il, ir, jl, jr, kl, kr, k0;
  long n;
  long sx, sy;  // strides of i and j; k stride is 1
};

__device__ __forceinline__ long cidx(const Geom& g, int i, int j, int k) { return ((long)i * g.gy + j) * g.gz + k; }

// The device arrays a kernel finds through the event table are global memory,
// but a pointer loaded from memory is generic to the compiler, which then
// emits flat loads and stores: those count against the LDS wait counter as
// well, so every LDS wait also waits for them, and they take 64-bit vector
// addresses.  A global-address-space pointer gives global loads and stores
// (scalar base plus vector offset) whose waits are separate from LDS waits.
template <class T>
__device__ __forceinline__ __attribute__((address_space(1))) T* as_global(T* p)
{
  return (__attribute__((address_space(1))) T*)p;
}

__device__ __forceinline__ bool interior(const Geom& g, int i, int j, int k)
{
  return i >= g.il && i < g.ir && j >= g.jl && j < g.jr && k >= g.kl && k < g.kr;
}

// Box of cells [i0,i1) x [j0,j1) x [k0,k1), thread t -> (i,j,k) with k fastest.
struct Box {
  int i0, j0, k0, ni, nj, nk;
};

// Index decodes run in 32-bit unsigned arithmetic (a 64-bit division is a
// long software sequence; the light cell kernels spent a large part of their
// instructions on it).  Every count decoded this way -- the cells of one
// event, its ghost cells, lines and tiles -- is below 2^32: an event's state
// takes 29 doubles per ghosted cell in each of its two copies, so 2^32 cells
// would need 2 TB of HBM.
__device__ __forceinline__ unsigned udivmod(unsigned a, unsigned b, unsigned& rem)
{
  const unsigned q = a / b;
  rem = a - q * b;
  return q;
}
__device__ __forceinline__ bool box_cell(const Box& b, long t, int& i, int& j, int& k)
{
  if (t >= (long)b.ni * b.nj * b.nk) return false;
  unsigned rk, rj;
  const unsigned q = udivmod((unsigned)t, (unsigned)b.nk, rk);
  const unsigned qi = udivmod(q, (unsigned)b.nj, rj);
  k = b.k0 + (int)rk;
  j = b.j0 + (int)rj;
  i = b.i0 + (int)qi;
  return true;
}

// The same box in slab order: slabs of kSlabRows j-rows; inside a slab i is
// slowest, then j, then k.  The cell kernels read u at i +- 1, 2: in plain
// order those planes are a whole i-plane of traffic apart and leave L2 before
// their reuse; in slab order consecutive blocks advance through i within a
// slab of kSlabRows x nk cells per plane, so the five planes a block reads are
// the ones its in-flight neighbours read.  Only the slab's two edge rows on
// each side are read twice.
constexpr int kSlabRows = 16;
__device__ __forceinline__ bool slab_cell(const Box& b, long t, int& i, int& j, int& k)
{
  if (t >= (long)b.ni * b.nj * b.nk) return false;
  const unsigned nk = (unsigned)b.nk;
  const unsigned long long per64 = (unsigned long long)b.ni * kSlabRows * nk;  // above the box when nj < kSlabRows
  const unsigned per = per64 > 0xffffffffull ? 0xffffffffu : (unsigned)per64;
  unsigned r, rk, rj;
  const int s = (int)udivmod((unsigned)t, per, r);
  const int rows = min(kSlabRows, b.nj - s * kSlabRows);
  const unsigned q = udivmod(r, nk, rk);
  const unsigned qi = udivmod(q, (unsigned)rows, rj);
  k = b.k0 + (int)rk;
  j = b.j0 + s * kSlabRows + (int)rj;
  i = b.i0 + (int)qi;
  return true;
}

// faces c+1/2 along d for c in [lo-1, hi-1], interior in the other directions
__host__ __device__ inline Box face_box(const Geom& g, int d)
{
  Box b{g.il, g.jl, g.kl, g.ir - g.il, g.jr - g.jl, g.kr - g.kl};
  if (d == 0) { b.i0 -= 1; b.ni += 1; }
  if (d == 1) { b.j0 -= 1; b.nj += 1; }
  if (d == 2) { b.k0 -= 1; b.nk += 1; }
  return b;
}
__host__ __device__ inline Box interior_box(const Geom& g) { return Box{g.il, g.jl, g.kl, g.ir - g.il, g.jr - g.jl, g.kr - g.kl}; }

__device__ __forceinline__ long dstride(const Geom& g, int d) { return d == 0 ? g.sx : d == 1 ? g.sy : 1; }

// ------------------------------------------------- stage clock and event table

struct StageArgs {
  double tau_src;   // time of the sources / fluxes of this stage
  double ts_src;    // step used in d_tau u = (u - prevu)/ts_src
  double tau_new;   // time the RK combination lands on
  double ts;        // full step
  double taustep;   // SHASTA sub-step: ts, ts/2, ts/3 for RK1/2/3 (NEWEST10/time.h:314,352,411)
  int stage;        // 0,1,2
  int kind;         // 0: Var1 = Var0 + ts L; 1: RK2 final; 2: RK3 middle; 3: RK3 final
};

// Device-resident step clock.  Kernels read tau and dtau from here, so one
// captured graph serves every step; the last kernel of a step advances it
// (tau += dtau, as the reference's time loop, NEWEST10/main.cpp:156).
struct Clock {
  double tau, ts;
  long long step;  // steps taken since the clock was set (first-failure bookkeeping)
};

// Stage times of stage `stage` of an RK`rk` step from (tau, ts):
// NEWEST10/time.h:307-325 (RK1), 345-363 (RK2), 402-426 (RK3).
__host__ __device__ inline StageArgs stage_args(double tau, double ts, int rk, int stage)
{
  StageArgs a{};
  a.ts = ts;
  a.taustep = rk == 3 ? ts / 3.0 : rk == 2 ? ts / 2 : ts;
  a.stage = stage;
  if (stage == 0) { a.tau_src = tau; a.ts_src = ts; a.tau_new = tau + ts; a.kind = 0; }
  else if (rk == 2) { a.tau_src = tau + ts; a.ts_src = ts; a.tau_new = tau + ts; a.kind = 1; }
  else if (stage == 1) { a.tau_src = tau + ts; a.ts_src = ts; a.tau_new = tau + ts / 2.0; a.kind = 2; }
  else { a.tau_src = tau + ts / 2.0; a.ts_src = ts / 2.0; a.tau_new = tau + ts; a.kind = 3; }
  return a;
}

// One row of the device event table: geometry, physics and buffers of one
// event of the ensemble, plus its flux tile shape per direction.
struct EvDev {
  Geom g;
  Params p;
  double* st[2];
  double* var0;
  double* var1;
  double* var2;
  double* flux[3];
  double* src;
  int32_t* iters;
  uint8_t* flags;
  unsigned long long* bad;
  int tp[3], tk[3];
};

// Work of one launch: `total` tiles (blocks' worth of work) spread over the
// events ev0..ev0+nev-1; event e owns tiles [start[e-ev0], start[e-ev0+1]).
// A launch either gives every tile its own block (gridDim == total) or runs
// persistent blocks that stride over the tiles.  With xcd set, tile ids are
// permuted so the tiles one XCD (blocks b with b % 8 equal) receives are
// contiguous, keeping neighbouring tiles' halos in that XCD's L2.
struct Sched {
  const long* start;
  long total;
  int ev0, nev, xcd;
};

__host__ __device__ inline long xcd_tile(long w, long total)
{
  const long per = total / 8, rem = total % 8;
  const long x = w % 8, r = w / 8;
  return x * per + (x < rem ? x : rem) + r;
}

__device__ __forceinline__ int sched_find(const Sched& s, long w, long& local)
{
  const long t = s.xcd ? xcd_tile(w, s.total) : w;
  if (s.nev == 1) {
    local = t;
    return s.ev0;
  }
  int lo = 0, hi = s.nev - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (s.start[mid] <= t) lo = mid;
    else hi = mid - 1;
  }
  local = t - s.start[lo];
  return s.ev0 + lo;
}

// PERSIST: blocks stride over the tiles; otherwise one tile per block (the
// loop runs once, and the compiler keeps the event row in scalar registers)
#define TILE_LOOP(s) for (long w_ = blockIdx.x; w_ < (s).total; w_ += PERSIST ? (long)gridDim.x : (s).total)

// ---------------------------------------------------------------- kernels

// One thread per face c+1/2 of direction d: the WENO of the eigenvalue u^d/u^0
// gives the face speed a = max(|A_L(c)|, |A_R(c+1)|) (NEWEST10/hydroKT.h:
// 176-258, 280-329; under q1 the eigenvalue is zero outside the interior and
// the face before the first interior cell gets a = 0), then for each of the 15
// variables the WENO of the centred flux (hydroKT.h:16-98) and of Var
// (hydroKT.h:364-381) give the Rusanov flux
//   f = (F_L + F_R)/2 - a (V_R - V_L)/2      (hydroKT.h:444-506).
// The stencil cells' common terms are loaded once and kept in registers.
// CART: the Cartesian (TXYZ5/hydroKT.h:16-62) form -- no tau factors, the
// five evolved shear rows 4..8 and PI in row 14, and the T rows built from
// the full pi^{mu nu} that cart_pi10 derives per stencil cell.
template <bool FAST, bool CART>
__device__ __forceinline__ void flux_face(const Geom& g, const Params& p, int d, double tau,
                                          const double* __restrict__ st_, double* __restrict__ flux_, long t)
{
  const auto* __restrict__ st = as_global(st_);
  auto* __restrict__ flux = as_global(flux_);
  int i, j, k;
  Box b = face_box(g, d);
  if (!box_cell(b, t, i, j, k)) return;
  const long n = g.n, s = dstride(g, d), x = cidx(g, i, j, k);
  const int milne = CART ? 0 : p.milne;
  const bool q1 = (p.quirks & VIRAL_Q_GHOST_EIGEN) != 0;
  const int row = d + 1;
  double u0[6], ud[6], w[6], pp[6];
#pragma unroll
  for (int o = 0; o < 6; o++) {
    const long y = x + (o - 2) * s;
    const double e = st[VIRAL_F_EN * n + y], P = st[VIRAL_F_P * n + y], PI = st[VIRAL_F_BULK * n + y];
    u0[o] = st[VIRAL_F_U0 * n + y];
    ud[o] = st[(VIRAL_F_U1 + d) * n + y];
    w[o] = e + P - PI;
    pp[o] = (P - PI);
    if (d == 2 && milne) pp[o] = FAST ? pp[o] * (1.0 / (tau * tau)) : ((P - PI) / (tau * tau));
  }
  double vd[6];
  if (FAST) {
#pragma unroll
    for (int o = 0; o < 6; o++) vd[o] = ud[o] * frcp(u0[o]);
  }
  // face speed
  double amax;
  {
    double A[6];
    const int pos = d == 0 ? i : d == 1 ? j : k;
#pragma unroll
    for (int o = 0; o < 6; o++) {
      const int ii = d == 0 ? i + o - 2 : i, jj = d == 1 ? j + o - 2 : j, kk = d == 2 ? k + o - 2 : k;
      const bool in = interior(g, ii, jj, kk);
      A[o] = (in || !q1) ? (FAST ? vd[o] : ud[o] / u0[o]) : 0.0;
    }
    const double aL = wL<FAST>(A[0], A[1], A[2], A[3], A[4]);
    const double aR = wR<FAST>(A[1], A[2], A[3], A[4], A[5]);
    const int lo = d == 0 ? g.il : d == 1 ? g.jl : g.kl;
    double ml = 0, mr = 0;
    if (!(q1 && pos < lo)) {
      if (fabs(aL) > ml) ml = fabs(aL);
      if (fabs(aR) > mr) mr = fabs(aR);
    }
    amax = fmax(ml, mr);
  }
  const double tf = milne ? tau : 1.0;
  decltype(st) base[6];
#pragma unroll
  for (int o = 0; o < 6; o++) base[o] = st + x + (o - 2) * s;
  double piT[6][4];  // Cartesian pi^{row,m} of each stencil cell
  if (CART) {
#pragma unroll
    for (int o = 0; o < 6; o++) {
      const double uu[4] = {u0[o], base[o][(long)VIRAL_F_U1 * n], base[o][(long)VIRAL_F_U2 * n],
                            base[o][(long)VIRAL_F_U3 * n]};
      double pt[10];
      cart_pi10(uu, base[o][(long)VIRAL_F_PI0 * n], base[o][(long)(VIRAL_F_PI0 + 1) * n],
                base[o][(long)(VIRAL_F_PI0 + 2) * n], base[o][(long)(VIRAL_F_PI0 + 3) * n],
                base[o][(long)(VIRAL_F_PI0 + 4) * n], pt);
#pragma unroll
      for (int m = 0; m < 4; m++) piT[o][m] = pt[pidx(row, m)];
    }
  }
#pragma unroll
  for (int l = 0; l < SV; l++) {
    if (CART && l >= 9 && l < 14) continue;
    double F[6], V[6];
    // field feeding the flux row: pi^{row,l} for T rows, pi[l-4] / PI otherwise
    const int ff = l < 4 ? VIRAL_F_PI0 + pidx(row, l) : (l < 14 ? VIRAL_F_PI0 + l - 4 : VIRAL_F_BULK);
#pragma unroll
    for (int o = 0; o < 6; o++) {
      const double fv = (CART && l < 4) ? piT[o][l < 4 ? l : 0] : base[o][(long)ff * n];
      const double vv = base[o][(long)l * n];  // Var field index == state field index (T00..T30, pi0..pi9, PI)
      if (l < 4) {
        const double um = base[o][(long)(VIRAL_F_U0 + l) * n];
        double v;
        if (l == row)
          v = fv + pp[o] + ud[o] * um * w[o];
        else
          v = fv + ud[o] * um * w[o];
        F[o] = milne ? tau * (v) : v;
        V[o] = milne ? tf * vv : vv;
      } else {
        F[o] = FAST ? fv * vd[o] : (fv * ud[o]) / u0[o];
        V[o] = vv;
      }
    }
    const double FL = wL<FAST>(F[0], F[1], F[2], F[3], F[4]);
    const double FR = wR<FAST>(F[1], F[2], F[3], F[4], F[5]);
    const double VL = wL<FAST>(V[0], V[1], V[2], V[3], V[4]);
    const double VR = wR<FAST>(V[1], V[2], V[3], V[4], V[5]);
    flux[l * n + x] = FAST ? fma(-0.5 * amax, VR - VL, 0.5 * (FL + FR)) : 0.5 * (FL + FR) - 0.5 * (amax) * (VR - VL);
  }
}

template <bool FAST, bool PERSIST, bool CART>
__global__ void __launch_bounds__(BLK) k_flux_dir(const EvDev* __restrict__ tab, Sched s, const Clock* __restrict__ clk,
                                                  int stage, int cur, int d)
{
  TILE_LOOP(s)
  {
    long lb;
    const EvDev& E = tab[sched_find(s, w_, lb)];
    const StageArgs a = stage_args(clk->tau, clk->ts, E.p.rk, stage);
    flux_face<FAST, CART>(E.g, E.p, d, a.tau_src, E.st[cur], E.flux[d], lb * BLK + threadIdx.x);
  }
}

// Fast-mode KT flux of direction D on an LDS tile (same physics as
// k_flux_dir).  A block takes TK lines along D (interior in the other two
// directions) and TP consecutive cells of each; every cell's centred flux
// and Var of the current variable are computed once and staged in LDS, every
// cell reconstructs its left and right face values from one set of WENO
// smoothness indicators, and a face c+1/2 takes L from cell c and R from cell
// c+1 through a second, double-buffered LDS array.  TP-5 faces per line per
// tile, staged in LDS; each of the TP-6 cells whose two faces both lie in the
// tile forms its partial result -(1/h)(f(c+1/2) - f(c-1/2)), so consecutive
// tiles of a line overlap by one face.  The three directions share one
// 15-row array: x writes its partial results, y and eta add theirs (the
// phases run in that order on one stream), and the update reads the sum as it
// is -- 15 doubles per cell instead of 45.  The tile shape is a launch
// parameter.  Slots map to cells with the contiguous (eta) index fastest so
// loads coalesce in every direction.
//
// NT threads per block; each thread owns up to MAXR slots (TP*TK <= MAXR*NT,
// at most 1024 slots = 56 KiB of LDS).
// LDS arrays of a flux tile, one double per slot each: cell F and V, the
// double-buffered right face values of F and V, the face fluxes of the row
constexpr int kTileArrays = 7;

template <int NT>
struct TileSlots {
  static constexpr int MAXR = NT >= 256 ? 2 : 4;
  static constexpr int MAXS = MAXR * NT < 1024 ? MAXR * NT : 1024;
};

template <int D, int NT, bool CART>
__device__ __forceinline__ void flux_tile_body(const Geom& g, const Params& p, double tau,
                                               const double* __restrict__ st_, double* __restrict__ flux_, int TP,
                                               int TK, long tile, double* lds)
{
  constexpr int MAXR = TileSlots<NT>::MAXR;
  constexpr int BLK = NT;
  const auto* __restrict__ st = as_global(st_);
  auto* __restrict__ flux = as_global(flux_);
  const int S = TP * TK;
  double* sF = lds;
  double* sV = lds + S;
  double* rF0 = lds + 2 * S;
  double* rV0 = lds + 3 * S;
  double* rF1 = lds + 4 * S;
  double* rV1 = lds + 5 * S;
  double* Bf = lds + 6 * S;  // face fluxes of the current row
  const int na = D == 0 ? g.jr - g.jl : g.ir - g.il;  // slow line coordinate
  const int nb = D == 2 ? g.jr - g.jl : g.kr - g.kl;  // fast line coordinate
  const long nlines = (long)na * nb;
  const int lo = D == 0 ? g.il : D == 1 ? g.jl : g.kl;
  const int hi = D == 0 ? g.ir : D == 1 ? g.jr : g.kr;
  const int GD = D == 0 ? g.gx : D == 1 ? g.gy : g.gz;
  const int fpt = TP - 5;            // faces per tile
  const int cpt = fpt - 1;           // cells per tile: both of their faces are in the tile
  const int nseg = (hi - lo + cpt - 1) / cpt;
  unsigned sgu;
  const long lgrp = udivmod((unsigned)tile, (unsigned)nseg, sgu);
  const int sg = (int)sgu;
  const int c0 = lo - 1 + sg * cpt;  // first face cell of this tile
  const int p0 = c0 - 2;             // cell at tile position 0
  const long n = g.n;
  const int milne = CART ? 0 : p.milne;
  const bool q1 = (p.quirks & VIRAL_Q_GHOST_EIGEN) != 0;
  const long sd = D == 0 ? g.sx : D == 1 ? g.sy : 1;
  const int row = D + 1;
  const double tf = milne ? tau : 1.0;
  const double hD = D == 0 ? p.dx : D == 1 ? p.dy : p.deta;

  long xs[MAXR];
  int pos[MAXR];
  bool cell_ok[MAXR], face_ok[MAXR], own_ok[MAXR], in_[MAXR];
  double u0v[MAXR], udv[MAXR], vdv[MAXR], wv[MAXR], ppv[MAXR], amax[MAXR], LF[MAXR], LV[MAXR];
  double pT[MAXR][4];  // Cartesian pi^{row,m} of each owned cell
#pragma unroll
  for (int r = 0; r < MAXR; r++) {
    const int sl = threadIdx.x + r * BLK;
    cell_ok[r] = false;
    face_ok[r] = false;
    own_ok[r] = false;
    in_[r] = false;
    xs[r] = 0;
    pos[r] = -100;
    if (sl >= S) continue;
    const int ps = D < 2 ? sl / TK : sl % TP;
    const int li = D < 2 ? sl % TK : sl / TP;
    const long line = lgrp * TK + li;
    pos[r] = ps;
    const int cd = p0 + ps;
    if (line >= nlines || cd < 0 || cd >= GD) continue;
    unsigned bu;
    const int a = (int)udivmod((unsigned)line, (unsigned)nb, bu), b = (int)bu;
    int i, j, k;
    if (D == 0) { i = cd; j = g.jl + a; k = g.kl + b; }
    else if (D == 1) { i = g.il + a; j = cd; k = g.kl + b; }
    else { i = g.il + a; j = g.jl + b; k = cd; }
    xs[r] = cidx(g, i, j, k);
    cell_ok[r] = true;
    in_[r] = cd >= lo && cd < hi;
    face_ok[r] = ps >= 2 && ps < 2 + fpt && cd <= hi - 1;
    own_ok[r] = ps >= 3 && ps < 2 + fpt && cd >= lo && cd < hi;
  }
  // common terms of each owned cell
#pragma unroll
  for (int r = 0; r < MAXR; r++) {
    u0v[r] = 1.0;
    udv[r] = vdv[r] = wv[r] = ppv[r] = 0.0;
    amax[r] = LF[r] = LV[r] = 0.0;
    if (!cell_ok[r]) continue;
    const long y = xs[r];
    const double e = st[VIRAL_F_EN * n + y], P = st[VIRAL_F_P * n + y], PI = st[VIRAL_F_BULK * n + y];
    u0v[r] = st[VIRAL_F_U0 * n + y];
    udv[r] = st[(VIRAL_F_U1 + D) * n + y];
    vdv[r] = udv[r] * frcp(u0v[r]);
    wv[r] = e + P - PI;
    ppv[r] = (D == 2 && milne) ? (P - PI) * (1.0 / (tau * tau)) : (P - PI);
    if (CART) {
      const double uu[4] = {u0v[r], st[VIRAL_F_U1 * n + y], st[VIRAL_F_U2 * n + y], st[VIRAL_F_U3 * n + y]};
      double pt[10];
      cart_pi10(uu, st[VIRAL_F_PI0 * n + y], st[(VIRAL_F_PI0 + 1) * n + y], st[(VIRAL_F_PI0 + 2) * n + y],
                st[(VIRAL_F_PI0 + 3) * n + y], st[(VIRAL_F_PI0 + 4) * n + y], pt);
#pragma unroll
      for (int m = 0; m < 4; m++) pT[r][m] = pt[pidx(row, m)];
    }
  }
  // The state values row q reads: field q (T^{tau q}, pi[q-4] or PI) and, for
  // q < 4, pi^{row q} and u^q.  (Loading them at the top of the previous row,
  // with its reconstructions still to come, costs a wave per SIMD in
  // registers and is slower.)
  double ldA[MAXR], ldB[MAXR], ldC[MAXR];
  auto row_loads = [&](int q) {
#pragma unroll
    for (int r = 0; r < MAXR; r++) {
      ldA[r] = ldB[r] = ldC[r] = 0.0;
      if (!cell_ok[r]) continue;
      const long y = xs[r];
      ldA[r] = st[(long)q * n + y];
      if (q < 4) {
        if (!CART) ldB[r] = st[(long)(VIRAL_F_PI0 + pidx(row, q)) * n + y];
        ldC[r] = st[(long)(VIRAL_F_U0 + q) * n + y];
      }
    }
  };
  // quantity -1: eigenvalue -> face speeds; 0..14: flux rows
  for (int q = -1; q < SV; q++) {
    if (CART && q >= 9 && q < 14) continue;  // block-uniform: no barrier is skipped by part of the block
    double* rF = (q & 1) ? rF1 : rF0;
    double* rV = (q & 1) ? rV1 : rV0;
    // x and eta load each row's state after the previous row's
    // reconstructions (below), so the loads wait behind that row's face
    // fluxes and partial results; y, whose registers hold its early
    // partial-result load, loads here
    if (D == 1 ? q >= 0 : q == -1) row_loads(D == 1 ? q : 0);
    // y and eta add to the partial result the earlier directions wrote: its
    // load is issued here, three barriers ahead of the add (each cell's
    // partial result is written by one thread of one tile only).
    double prev[MAXR];
#pragma unroll
    for (int r = 0; r < MAXR; r++) prev[r] = (D > 0 && q >= 0 && own_ok[r]) ? flux[(long)q * n + xs[r]] : 0.0;
#pragma unroll
    for (int r = 0; r < MAXR; r++) {
      const int sl = threadIdx.x + r * BLK;
      if (sl >= S) continue;
      double F = 0.0, V = 0.0;
      if (cell_ok[r]) {
        if (q < 0) {
          V = (in_[r] || !q1) ? vdv[r] : 0.0;
        } else if (q < 4) {
          const double piv = CART ? (q == 0 ? pT[r][0] : q == 1 ? pT[r][1] : q == 2 ? pT[r][2] : pT[r][3]) : ldB[r];
          const double um = ldC[r];
          const double v = (q == row) ? piv + ppv[r] + udv[r] * um * wv[r] : piv + udv[r] * um * wv[r];
          F = milne ? tau * v : v;
          V = milne ? tf * ldA[r] : ldA[r];
        } else {
          V = ldA[r];  // pi[q-4] or PI: Var row q is state field q
          F = V * vdv[r];
        }
      }
      sF[sl] = F;
      sV[sl] = V;
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < MAXR; r++) {
      const int sl = threadIdx.x + r * BLK;
      const int ps = pos[r];
      if (sl >= S || ps < 2 || ps > TP - 3) continue;
      const int st1 = D < 2 ? TK : 1;  // slot stride along the line
      double Lq, Rq;
      weno_LR_fast(sV[sl - 2 * st1], sV[sl - st1], sV[sl], sV[sl + st1], sV[sl + 2 * st1], Lq, Rq);
      LV[r] = Lq;
      rV[sl] = Rq;
      if (q >= 0) {
        weno_LR_fast(sF[sl - 2 * st1], sF[sl - st1], sF[sl], sF[sl + st1], sF[sl + 2 * st1], Lq, Rq);
        LF[r] = Lq;
        rF[sl] = Rq;
      }
    }
    __syncthreads();
    if (D != 1 && q >= 0) {  // the next row's state (this row's F and V are in LDS)
      const int qn = (CART && q == 8) ? 14 : q + 1;
      if (qn < SV) row_loads(qn);
    }
#pragma unroll
    for (int r = 0; r < MAXR; r++) {
      if (!face_ok[r]) continue;
      const int sl = threadIdx.x + r * BLK;
      const int st1 = D < 2 ? TK : 1;
      const double VR = rV[sl + st1];
      if (q < 0) {
        const int cd = p0 + pos[r];
        double ml = 0, mr = 0;
        if (!(q1 && cd < lo)) {
          if (fabs(LV[r]) > ml) ml = fabs(LV[r]);
          if (fabs(VR) > mr) mr = fabs(VR);
        }
        amax[r] = fmax(ml, mr);
      } else {
        const double FR = rF[sl + st1];
        Bf[sl] = fma(-0.5 * amax[r], VR - LV[r], 0.5 * (LF[r] + FR));
      }
    }
    if (q < 0) continue;  // block-uniform
    __syncthreads();
    // the cell's partial result -(1/h)(f(c+1/2) - f(c-1/2)), as the update
    // would form it from the face fluxes, added to the earlier directions'
    // (x, then y, then eta: the update's order of summation)
#pragma unroll
    for (int r = 0; r < MAXR; r++) {
      if (!own_ok[r]) continue;
      const int sl = threadIdx.x + r * BLK;
      const int st1 = D < 2 ? TK : 1;
      const double df = Bf[sl] - Bf[sl - st1];
      const double pr = -df * (1.0 / hD);
      const long o = (long)q * n + xs[r];
      flux[o] = D == 0 ? pr : prev[r] + pr;
    }
  }
}

template <int D, int NT, bool PERSIST, bool CART>
// Milne tiles of 256 threads, one tile per block: registers capped for five
// waves per SIMD, the wave the early loads (the y tile's partial result, the
// x and eta tiles' next row) would take; the y tile spills 20 bytes per
// lane.  The Cartesian and persistent forms keep their occupancy uncapped.
__global__ void __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(NT == 256 && !PERSIST && !CART ? 5 : 1))) k_flux_tile(const EvDev* __restrict__ tab, Sched s,
                                                  const Clock* __restrict__ clk, int stage, int cur)
{
  extern __shared__ double lds[];
  TILE_LOOP(s)
  {
    long lb;
    const EvDev& E = tab[sched_find(s, w_, lb)];
    const StageArgs a = stage_args(clk->tau, clk->ts, E.p.rk, stage);
    flux_tile_body<D, NT, CART>(E.g, E.p, a.tau_src, E.st[cur], E.flux[0], E.tp[D], E.tk[D], lb, lds);
    __syncthreads();  // the next tile may lay its LDS arrays out differently
  }
}

// SHASTA flux-corrected transport along direction D (NEWEST10/hydroshasta.h:
// 8-32, 101-284): transport velocity v (row 0: T^{D,tau}/T^{tau tau}, other
// rows u^D/u^0), Q+- factors, low-order transported-diffused UTD, limited
// antidiffusion A, and per cell the partial result
//   (-U + UTD_c - A_c + A_{c-1}) / taustep,
// evaluated in the reference's operation order (exact mode is bitwise equal
// to the CPU restatement).  On an LDS line tile (the KT tiles' layout: TK
// lines x TP cells, slots with the contiguous eta index fastest, so loads
// coalesce in every direction) every slot loads its own cell once per
// quantity and the stencil neighbours come from LDS: the cell's v0 and vr,
// its Q+/Q- (positions 1..TP-2), per row its U and UTD (positions 1..TP-2),
// and the owned cells (positions 3..TP-4, which have the whole 7-cell
// stencil in the tile) form the antidiffusion and write the partial result.
// The Q factors of the rows that share u^D/u^0 are computed once.
template <int D, int NT, bool FAST>
__device__ __forceinline__ void shasta_tile_body(const Geom& g, const Params& p, double tau, double taustep,
                                                 const double* __restrict__ st_, double* __restrict__ pr_, int TP,
                                                 int TK, long tile, double* lds)
{
  constexpr int MAXR = TileSlots<NT>::MAXR;
  constexpr int BLK = NT;
  const auto* __restrict__ st = as_global(st_);
  auto* __restrict__ pr = as_global(pr_);
  const int S = TP * TK;
  // U and utd of even rows, then of odd rows (the v0 / vr stage uses the odd
  // pair), so one row's reads never meet the next row's writes
  double* U0 = lds;
  double* U1 = lds + S;
  double* T0 = lds + 2 * S;
  double* T1 = lds + 3 * S;
  const int na = D == 0 ? g.jr - g.jl : g.ir - g.il;  // slow line coordinate
  const int nb = D == 2 ? g.jr - g.jl : g.kr - g.kl;  // fast line coordinate
  const long nlines = (long)na * nb;
  const int lo = D == 0 ? g.il : D == 1 ? g.jl : g.kl;
  const int hi = D == 0 ? g.ir : D == 1 ? g.jr : g.kr;
  const int GD = D == 0 ? g.gx : D == 1 ? g.gy : g.gz;
  const int cpt = TP - 6;  // owned cells per tile
  const int nseg = (hi - lo + cpt - 1) / cpt;
  unsigned sgu;
  const long lgrp = udivmod((unsigned)tile, (unsigned)nseg, sgu);
  const int sg = (int)sgu;
  const int p0 = lo - 3 + sg * cpt;  // cell at tile position 0
  const long n = g.n;
  const int milne = p.milne;
  const double hh = D == 0 ? p.dx : D == 1 ? p.dy : p.deta;
  const double lam = taustep / hh;
  const double tf = milne ? tau : 1.0;
  const int st1 = D < 2 ? TK : 1;  // slot stride along the line

  long xs[MAXR];
  int pos[MAXR];
  bool cell_ok[MAXR], q_ok[MAXR], own_ok[MAXR];
#pragma unroll
  for (int r = 0; r < MAXR; r++) {
    const int sl = threadIdx.x + r * BLK;
    cell_ok[r] = q_ok[r] = own_ok[r] = false;
    xs[r] = 0;
    pos[r] = -100;
    if (sl >= S) continue;
    const int ps = D < 2 ? sl / TK : sl % TP;
    const int li = D < 2 ? sl % TK : sl / TP;
    const long line = lgrp * TK + li;
    pos[r] = ps;
    const int cd = p0 + ps;
    if (line >= nlines || cd < 0 || cd >= GD) continue;
    unsigned bu;
    const int a = (int)udivmod((unsigned)line, (unsigned)nb, bu), b = (int)bu;
    int i, j, k;
    if (D == 0) { i = cd; j = g.jl + a; k = g.kl + b; }
    else if (D == 1) { i = g.il + a; j = cd; k = g.kl + b; }
    else { i = g.il + a; j = g.jl + b; k = cd; }
    xs[r] = cidx(g, i, j, k);
    cell_ok[r] = true;
    q_ok[r] = ps >= 1 && ps <= TP - 2;
    own_ok[r] = ps >= 3 && ps < TP - 3 && cd >= lo && cd < hi;
  }
  // v0 and vr of every cell
#pragma unroll
  for (int r = 0; r < MAXR; r++) {
    const int sl = threadIdx.x + r * BLK;
    if (sl >= S) continue;
    double v0 = 0.0, vr = 0.0;
    if (cell_ok[r]) {
      const long y = xs[r];
      const double e = st[VIRAL_F_EN * n + y], P = st[VIRAL_F_P * n + y], PI = st[VIRAL_F_BULK * n + y];
      const double u0 = st[VIRAL_F_U0 * n + y], ud = st[(VIRAL_F_U1 + D) * n + y];
      const double pd = st[(VIRAL_F_PI0 + 1 + D) * n + y];
      v0 = (pd + ud * u0 * (e + P - PI)) / st[VIRAL_F_T00 * n + y];
      vr = ud / u0;
    }
    U1[sl] = v0;
    T1[sl] = vr;
  }
  __syncthreads();
  // Q+ and Q- of every cell with both neighbours in the tile
  double qp0[MAXR], qm0[MAXR], qpr[MAXR], qmr[MAXR];
#pragma unroll
  for (int r = 0; r < MAXR; r++) {
    qp0[r] = qm0[r] = qpr[r] = qmr[r] = 0.0;
    if (!q_ok[r]) continue;
    const int sl = threadIdx.x + r * BLK;
    const double* v0 = U1;
    const double* vr = T1;
    qp0[r] = (0.5 - v0[sl] * lam) / (1.0 + (v0[sl + st1] - v0[sl]) * lam);
    qm0[r] = (0.5 + v0[sl] * lam) / (1.0 - (v0[sl - st1] - v0[sl]) * lam);
    qpr[r] = (0.5 - vr[sl] * lam) / (1.0 + (vr[sl + st1] - vr[sl]) * lam);
    qmr[r] = (0.5 + vr[sl] * lam) / (1.0 - (vr[sl - st1] - vr[sl]) * lam);
  }
  // one row on the LDS pair of its parity (a compile-time choice, so every
  // LDS access is a ds instruction)
  auto row = [&](int l, double* __restrict__ U, double* __restrict__ T) {
#pragma unroll
    for (int r = 0; r < MAXR; r++) {
      const int sl = threadIdx.x + r * BLK;
      if (sl >= S) continue;
      double u = 0.0;
      if (cell_ok[r]) {
        const double vv = st[(long)l * n + xs[r]];
        u = (l < 4 && milne) ? tf * vv : vv;
      }
      U[sl] = u;
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < MAXR; r++) {
      if (!q_ok[r]) continue;
      const int sl = threadIdx.x + r * BLK;
      const double Qp = l == 0 ? qp0[r] : qpr[r], Qm = l == 0 ? qm0[r] : qmr[r];
      T[sl] = 0.5 * ((Qp * Qp * (U[sl + st1] - U[sl])) + (Qm * Qm * (U[sl - st1] - U[sl]))) + (Qp + Qm) * U[sl];
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < MAXR; r++) {
      if (!own_ok[r]) continue;
      const int sl = threadIdx.x + r * BLK;
      double utd[5];
#pragma unroll
      for (int m = 0; m < 5; m++) utd[m] = T[sl + (m - 2) * st1];
      // antidiffusion at c-1 (utd[0..3]) and c (utd[1..4])
      double A[2];
#pragma unroll
      for (int qq = 0; qq < 2; qq++) {
        double signa = 1;
        const double temp = 1 * 0.125 * (utd[qq + 2] - utd[qq + 1]);
        if (temp < 0) signa = -1;
        const double term1 = signa * (utd[qq + 3] - utd[qq + 2]);
        const double term3 = signa * (utd[qq + 1] - utd[qq]);
        const double mn = fmin(fmin(term1, fabs(temp)), term3);
        A[qq] = signa * fmax(0, mn);
      }
      const double num = -U[sl] + utd[2] - A[1] + A[0];
      pr[l * n + xs[r]] = FAST ? num * (1.0 / taustep) : num / taustep;
    }
  };
  for (int l = 0; l < SV - 1; l += 2) {
    row(l, U0, T0);
    row(l + 1, U1, T1);
  }
  row(SV - 1, U0, T0);  // SV is odd
}

template <int D, int NT, bool PERSIST, bool FAST>
__global__ void __launch_bounds__(NT) k_shasta_tile(const EvDev* __restrict__ tab, Sched s,
                                                    const Clock* __restrict__ clk, int stage, int cur)
{
  extern __shared__ double lds[];
  TILE_LOOP(s)
  {
    long lb;
    const EvDev& E = tab[sched_find(s, w_, lb)];
    const StageArgs a = stage_args(clk->tau, clk->ts, E.p.rk, stage);
    const double taustep = (E.p.quirks & VIRAL_Q_SHASTA_FULLSTEP) ? a.ts : a.taustep;
    shasta_tile_body<D, NT, FAST>(E.g, E.p, a.tau_src, taustep, E.st[cur], E.flux[D], E.tp[D], E.tk[D], lb, lds);
    __syncthreads();  // the next tile may lay its LDS arrays out differently
  }
}

// Geometric, CON-gradient and Israel-Stewart sources of one interior cell
// (NEWEST10/Source.h:3-101,195-536) from the stage-start state, into src.
// CART: TXYZ5/Source.h -- no geometric terms, the shear source of the five
// evolved components (rows 4..8) from the full tensor, and under q9 only in
// ghosted eta slices k < nz (the reference's loop k = 0..ZCMA-1,
// TXYZ5/Source.h:272; other slices keep zero sources).
template <bool FAST, bool CART, class Ptr>
__device__ __forceinline__ void source_terms(const Geom& g, const Params& p, const StageArgs& a,
                                             Ptr __restrict__ si, long x, double* S)
{
  const long n = g.n;
  const int milne = CART ? 0 : p.milne;
  const double tau = a.tau_src;
  const double e = si[VIRAL_F_EN * n + x], P = si[VIRAL_F_P * n + x], PI = si[VIRAL_F_BULK * n + x];
  double u[4], pu[4], pi[10];
#pragma unroll
  for (int l = 0; l < 4; l++) {
    u[l] = si[(VIRAL_F_U0 + l) * n + x];
    pu[l] = si[(VIRAL_F_PU0 + l) * n + x];
  }
  if (CART) {
    cart_pi10(u, si[VIRAL_F_PI0 * n + x], si[(VIRAL_F_PI0 + 1) * n + x], si[(VIRAL_F_PI0 + 2) * n + x],
              si[(VIRAL_F_PI0 + 3) * n + x], si[(VIRAL_F_PI0 + 4) * n + x], pi);
  } else {
#pragma unroll
    for (int l = 0; l < 10; l++) pi[l] = si[(VIRAL_F_PI0 + l) * n + x];
  }
  if (milne) {
    double tau2 = tau * tau;
    S[0] = -(P - PI) - tau2 * (pi[9] + (e + P - PI) * u[3] * u[3]);
    S[3] = (-2.0 * (pi[3] + (e + P - PI) * u[0] * u[3]));
  } else {
    S[0] = 0;
    S[3] = 0;
  }
  S[1] = 0;
  S[2] = 0;
  S[14] = 0;
  if (p.scheme == VIRAL_SHASTA) {
    // CON gradient sources, NEWEST10/Source.h:223-336
    const long sd[3] = {g.sx, g.sy, 1};
    const double hh[3] = {p.dx, p.dy, p.deta};
    const int ndir = p.lbi ? 2 : 3;
    const double tt = milne ? tau : 1.0;
#pragma unroll
    for (int dd = 0; dd < 3; dd++) {
      if (dd >= ndir) break;
      double q[4][5];
#pragma unroll
      for (int o = 0; o < 5; o++) {
        const long y = x + (o - 2) * sd[dd];
        const double Py = si[VIRAL_F_P * n + y], PIy = si[VIRAL_F_BULK * n + y];
        const double u0y = si[VIRAL_F_U0 * n + y], udy = si[(VIRAL_F_U1 + dd) * n + y];
        const double p2 = si[(VIRAL_F_PI0 + 1) * n + y], p3 = si[(VIRAL_F_PI0 + 2) * n + y],
                     p4 = si[(VIRAL_F_PI0 + 3) * n + y];
        q[0][o] = -tt * (0);
        if (dd == 0) {
          const double p5 = si[(VIRAL_F_PI0 + 4) * n + y], p6 = si[(VIRAL_F_PI0 + 5) * n + y],
                       p7 = si[(VIRAL_F_PI0 + 6) * n + y];
          q[1][o] = -tt * (Py - PIy - (udy / u0y) * p2 + p5);
          q[2][o] = -tt * (-(udy / u0y) * p3 + p6);
          q[3][o] = -tt * (-(udy / u0y) * p4 + p7);
        } else if (dd == 1) {
          const double p6 = si[(VIRAL_F_PI0 + 5) * n + y], p8 = si[(VIRAL_F_PI0 + 7) * n + y],
                       p9 = si[(VIRAL_F_PI0 + 8) * n + y];
          q[1][o] = -tt * (-(udy / u0y) * p2 + p6);
          q[2][o] = -tt * (Py - PIy - (udy / u0y) * p3 + p8);
          q[3][o] = -tt * (-(udy / u0y) * p4 + p9);
        } else {
          const double p7 = si[(VIRAL_F_PI0 + 6) * n + y], p9 = si[(VIRAL_F_PI0 + 8) * n + y],
                       p10 = si[(VIRAL_F_PI0 + 9) * n + y];
          q[1][o] = -tt * (-(udy / u0y) * p2 + p7);
          q[2][o] = -tt * (-(udy / u0y) * p3 + p9);
          q[3][o] = milne ? -tt * ((Py - PIy) / (tau * tau) - (udy / u0y) * p4 + p10)
                          : -tt * ((Py - PIy) - (udy / u0y) * p4 + p10);
        }
      }
#pragma unroll
      for (int m = 0; m < 4; m++)
        S[m] += wder<FAST>(q[m][0], q[m][1], q[m][2], q[m][3], q[m][4], hh[dd], 1.0 / hh[dd]);
    }
  }
  double du[4][4];
  const double hx[3] = {p.dx, p.dy, p.deta};
#pragma unroll
  for (int l = 0; l < 4; l++) {
    const auto ul = si + (VIRAL_F_U0 + l) * n;
    du[l][0] = FAST ? (u[l] - pu[l]) * (1.0 / a.ts_src) : (u[l] - pu[l]) / a.ts_src;
    du[l][1] = wder<FAST>(ul[x - 2 * g.sx], ul[x - g.sx], ul[x], ul[x + g.sx], ul[x + 2 * g.sx], hx[0], 1.0 / hx[0]);
    du[l][2] = wder<FAST>(ul[x - 2 * g.sy], ul[x - g.sy], ul[x], ul[x + g.sy], ul[x + 2 * g.sy], hx[1], 1.0 / hx[1]);
    if (!p.lbi)
      du[l][3] = wder<FAST>(ul[x - 2], ul[x - 1], ul[x], ul[x + 1], ul[x + 2], hx[2], 1.0 / hx[2]);
    else
      du[l][3] = 0;
  }
  {
    double T, s;
    if (FAST) {
      eos_Ts_fast(p, e, P, T, s);
    } else {
      T = eos_T(p, e);
      s = eos_s(p, e, P, T);
    }
    double SH = eta_of(p, s);
    double tpi = tpi_of(p, SH, P, e);
    double out[10];
    shear_source<FAST>(milne, tau, u, pi, du, SH, tpi, out);
    double sb = 0.0;
    if (p.bulk) {
      const double BU = zeta_of(p, s);
      sb = bulk_source<FAST>(milne, tau, u, PI, du, BU, tpib_of(p, BU, P));
    }
    if (CART) {
      // TXYZ5 evaluates the bulk row in the same slice loop (Source.h:272,327)
      const bool live = !(p.quirks & VIRAL_Q_SHEAR_SLICES) || (int)((unsigned)x % (unsigned)g.gz) < g.kr - g.kl;
#pragma unroll
      for (int l = 0; l < 5; l++) S[4 + l] = live ? out[cart5(l)] : 0.0;
#pragma unroll
      for (int l = 5; l < 10; l++) S[4 + l] = 0.0;
      S[14] = live ? sb : 0.0;
    } else {
#pragma unroll
      for (int l = 0; l < 10; l++) S[4 + l] = out[l];
      S[14] = sb;
    }
  }
}

template <bool FAST, bool PERSIST, bool CART>
// fast mode: registers capped for three waves per SIMD (a few bytes spill)
__global__ void __launch_bounds__(BLK) __attribute__((amdgpu_waves_per_eu(FAST ? 3 : 1))) k_source(const EvDev* __restrict__ tab, Sched s, const Clock* __restrict__ clk,
                                                int stage, int cur)
{
  TILE_LOOP(s)
  {
    long lb;
    const EvDev& E = tab[sched_find(s, w_, lb)];
    int i, j, k;
    if (!slab_cell(interior_box(E.g), lb * BLK + threadIdx.x, i, j, k)) continue;
    const StageArgs a = stage_args(clk->tau, clk->ts, E.p.rk, stage);
    const long x = cidx(E.g, i, j, k), n = E.g.n;
    double S[SV];
    source_terms<FAST, CART>(E.g, E.p, a, as_global((const double*)E.st[cur]), x, S);
    auto* __restrict__ src = as_global(E.src);
#pragma unroll
    for (int l = 0; l < SV; l++)
      if (!CART || l < 9 || l == 14) src[l * n + x] = S[l];
  }
}

// Flux divergence, L = Result + Source, the RK combination, T/pi from Var,
// the Newton root and the velocities for one interior cell
// (NEWEST10/hydroKT.h:393-403,458-504; time.h:2-180; root.h:301-424).  Reads
// the stage-start state `si`, the face fluxes of the three directions and the
// sources, writes the new state `so`.
//
// FUSED evaluates the sources in place (source_terms) instead of reading
// them from the source kernel's output: one launch and 15 doubles per cell
// less traffic, at the price of the union of both kernels' registers.
//
// CART: the Cartesian update (TXYZ5/time.h:100-165, TXYZ5/multiroot.h:322-479):
// T^{t mu} = Var, five pi rows, the 4-D gnewton root from (e, u^x, u^y, u^z),
// u^t = sqrt(1 + u.u), no k0 broadcast.
// Root and velocities of one Milne cell (NEWEST10/root.h:301-424) from its
// T^{tau mu}, pi, PI at tn, starting Newton from guess: e (with the
// VIRAL_Q_ROOT_SAFEGUARD re-solve), P, Temp, v, u^0 = gamma and the cell's
// iteration count and flags.  Shared by the stage update and viral_root.
template <bool FAST>
__device__ __forceinline__ void solve_cell(const Params& p, int milne, double tn, const double T0[4], const double pi[10],
                                           double PIn, double e, double& en, double& Pn, double& Temp, double& vx,
                                           double& vy, double& ve, double& gm, int& it, unsigned& fl)
{
  // ---- root (NEWEST10/root.h:301-364)
  double M2 = milne ? (T0[1] - pi[1]) * (T0[1] - pi[1]) + (T0[2] - pi[2]) * (T0[2] - pi[2]) +
                          tn * tn * (T0[3] - pi[3]) * (T0[3] - pi[3])
                    : (T0[1] - pi[1]) * (T0[1] - pi[1]) + (T0[2] - pi[2]) * (T0[2] - pi[2]) +
                          (T0[3] - pi[3]) * (T0[3] - pi[3]);
  double tv[4];
  bool have_tv = false;
  if (FAST) en = newton_cell_fast(p, M2, T0[0], pi[0], PIn, e, it, fl, tv, have_tv);
  else en = newton_cell(p, M2, T0[0], pi[0], PIn, e, it, fl);
  if ((p.quirks & VIRAL_Q_ROOT_SAFEGUARD) && ((fl & VIRAL_FLAG_SOLVER) || !(en > 0))) {
    double er;
    int ir;
    if (root_bracketed(p, M2, T0[0] - pi[0], PIn, er, ir)) {
      en = er;
      it = it + ir;
      fl = (fl & ~(VIRAL_FLAG_SOLVER | VIRAL_FLAG_MAXITER)) | VIRAL_FLAG_RESCUED;
      have_tv = false;
    }
  }
  if (FAST && have_tv) {
    // the table values of the root solve's last evaluation, at en
    Pn = tv[0];
    Temp = tv[2];
  } else if (FAST) {
    double a_, s_;
    eos_pa_fast(p, en, Pn, a_);
    eos_Ts_fast(p, en, Pn, Temp, s_);
  } else {
    Temp = eos_T(p, en);
    Pn = eos_p(p, en);
  }
  // ---- velocities (NEWEST10/root.h:372-424)
  double t2 = milne ? tn * tn : 1.0;
  if (FAST) {
    const double r = 1.0 / (T0[0] - pi[0] + Pn - PIn);
    vx = (T0[1] - pi[1]) * r;
    vy = (T0[2] - pi[2]) * r;
    ve = (T0[3] - pi[3]) * r;
  } else {
    vx = (T0[1] - pi[1]) / (T0[0] - pi[0] + Pn - PIn);
    vy = (T0[2] - pi[2]) / (T0[0] - pi[0] + Pn - PIn);
    ve = (T0[3] - pi[3]) / (T0[0] - pi[0] + Pn - PIn);
  }
  if (vx * vx + vy * vy + t2 * ve * ve >= 1) fl |= VIRAL_FLAG_ACAUSAL;
  if (en <= 0) fl |= VIRAL_FLAG_ENEG;
  gm = 1.0 / sqrt(1 - vx * vx - vy * vy - t2 * ve * ve);
  if (p.t_freeze > 0 && Temp < p.t_freeze) fl |= VIRAL_FLAG_FREEZEOUT;
}

// The eta ghost layers a cell feeds (NEWEST10/boundary.h:1-275, the eta
// part of the zero-gradient fill): slices 0..2 from kl, kr..kr+2 from kr-1.
// The kernels that write a boundary cell's field also write its eta ghost
// copies, while the cache lines are theirs; the ghost-fill kernel keeps the
// x and y faces.  Not in boost-invariant grids (no eta ghosts).
struct EtaGhosts {
  long lo = -1, hi = -1;  // first ghost of each side, or -1
  __device__ __forceinline__ EtaGhosts(const Geom& g, int lbi, int i, int j, int k)
  {
    if (lbi) return;
    if (k == g.kl) lo = cidx(g, i, j, 0);
    if (k == g.kr - 1) hi = cidx(g, i, j, g.kr);
  }
  __device__ __forceinline__ bool any() const { return lo >= 0 || hi >= 0; }
  template <class Ptr>
  __device__ __forceinline__ void put(Ptr st, long n, int f, double v) const
  {
    if (lo >= 0) {
      st[f * n + lo] = v;
      st[f * n + lo + 1] = v;
      st[f * n + lo + 2] = v;
    }
    if (hi >= 0) {
      st[f * n + hi] = v;
      st[f * n + hi + 1] = v;
      st[f * n + hi + 2] = v;
    }
  }
};

template <bool FAST, bool FUSED, bool CART>
__device__ __forceinline__ void update_cell(const Geom& g, const Params& p, const StageArgs& a, long long step,
                                            const double* __restrict__ si_, double* __restrict__ so_,
                                            const double* __restrict__ fx_, const double* __restrict__ fy_,
                                            const double* __restrict__ fz_, const double* __restrict__ src_,
                                            double* __restrict__ var0_, double* __restrict__ var1_,
                                            double* __restrict__ var2_, int32_t* __restrict__ iters_,
                                            uint8_t* __restrict__ flags_, unsigned long long* __restrict__ bad, long t)
{
  // global-address-space views of the event's arrays (the table hands them
  // over as generic pointers, whose flat accesses also count against the
  // scalar-load wait counter)
  const auto* __restrict__ si = as_global(si_);
  auto* __restrict__ so = as_global(so_);
  const auto* __restrict__ fx = as_global(fx_);
  const auto* __restrict__ fy = as_global(fy_);
  const auto* __restrict__ fz = as_global(fz_);
  const auto* __restrict__ src = as_global(src_);
  auto* __restrict__ var0 = as_global(var0_);
  auto* __restrict__ var1 = as_global(var1_);
  auto* __restrict__ var2 = as_global(var2_);
  auto* __restrict__ iters = as_global(iters_);
  auto* __restrict__ flags = as_global(flags_);
  int i, j, k;
  if (!slab_cell(interior_box(g), t, i, j, k)) return;
  const long n = g.n, x = cidx(g, i, j, k);
  const int milne = CART ? 0 : p.milne;
  const double tau = a.tau_src;
  // ---- sources first: the fused source evaluation's registers are free again
  // before the 15 sums of the flux divergence are live
  double S[SV];
  if (FUSED) {
    source_terms<FAST, CART>(g, p, a, si, x, S);
  } else {
#pragma unroll
    for (int l = 0; l < SV; l++) S[l] = (CART && l >= 9 && l < 14) ? 0.0 : src[l * n + x];
  }
  // ---- Result = sum over directions of -(1/h)(f(c+1/2) - f(c-1/2)), x then y then eta
  // (the fast KT tiles have already summed the three directions into fx)
  double Res[SV];
  if (FAST && p.scheme != VIRAL_SHASTA) {
#pragma unroll
    for (int l = 0; l < SV; l++) Res[l] = (CART && l >= 9 && l < 14) ? 0.0 : fx[l * n + x];
  } else {
    decltype(fx) fl[3] = {fx, fy, fz};
    const long sd[3] = {g.sx, g.sy, 1};
    const double hh[3] = {p.dx, p.dy, p.deta};
    const int ndir = p.lbi ? 2 : 3;
#pragma unroll
    for (int l = 0; l < SV; l++) Res[l] = 0.0;
#pragma unroll
    for (int dd = 0; dd < 3; dd++) {
      if (dd >= ndir) break;
      const auto f = fl[dd];
#pragma unroll
      for (int l = 0; l < SV; l++) {
        if (CART && l >= 9 && l < 14) continue;
        double pr;
        if (p.scheme == VIRAL_SHASTA) {
          pr = f[l * n + x];  // SHASTA stores the partial result itself
        } else {
          const double df = f[l * n + x] - f[l * n + x - sd[dd]];
          pr = -(1.0 / hh[dd]) * (df);
        }
        Res[l] = Res[l] + pr;
      }
    }
  }
  const double e = si[VIRAL_F_EN * n + x];
  double u[4], pu[4], pi[10];
#pragma unroll
  for (int l = 0; l < 4; l++) {
    u[l] = si[(VIRAL_F_U0 + l) * n + x];
    pu[l] = si[(VIRAL_F_PU0 + l) * n + x];
  }
  // ---- RK combination (NEWEST10/time.h:2-98)
  // Fast mode keeps only Var0: the Var of the previous stage (Var1 before the
  // RK2 final and the RK3 middle stage, Var2 before the RK3 final one) is the
  // stage-start state's own T^{tau mu} tau_src, pi and PI -- the stage before
  // stored T = Var / tau_new with tau_new = this stage's tau_src -- to within
  // two roundings of the T rows, instead of a 15-double copy written and read
  // back.
  double V[SV];
#pragma unroll
  for (int l = 0; l < SV; l++) {
    if (CART && l >= 9 && l < 14) {
      V[l] = 0.0;
      continue;
    }
    double L = (Res[l] + S[l]);
    // the stage-start state's Var (Var0 at stage 0)
    const double vs = l < 4 ? (milne ? tau * si[(VIRAL_F_T00 + l) * n + x] : si[(VIRAL_F_T00 + l) * n + x])
                            : si[(VIRAL_F_T00 + l) * n + x];
    if (a.kind == 0) {
      double v1 = FAST ? fma(a.ts, L, vs) : vs + a.ts * (L);
      var0[l * n + x] = vs;
      if (!FAST) var1[l * n + x] = v1;
      V[l] = v1;
    } else if (a.kind == 1) {
      V[l] = 0.5 * var0[l * n + x] + 0.5 * (FAST ? vs : var1[l * n + x]) + 0.5 * a.ts * L;
    } else if (a.kind == 2) {
      double v2 = 0.75 * var0[l * n + x] + 0.25 * (FAST ? vs : var1[l * n + x]) + 0.25 * a.ts * L;
      if (!FAST) var2[l * n + x] = v2;
      V[l] = v2;
    } else {
      V[l] = (1.0 / 3.0) * var0[l * n + x] + (2.0 / 3.0) * (FAST ? vs : var2[l * n + x]) + (2.0 / 3.0) * a.ts * L;
    }
  }
  if (CART) {
    // ---- TXYZ5/time.h:100-165 and TXYZ5/multiroot.h:322-479
    const double prm[10] = {V[0], V[1], V[2], V[3], V[4], V[5], V[6], V[7], V[8], V[14]};
    double r[4] = {e, u[1], u[2], u[3]};
    unsigned fl;
    const int it = cart_gnewton(p, prm, r, fl);
    const double En = r[0];
    const double Temp = eos_T(p, En);
    const double c0 = sqrt(1 + r[1] * r[1] + r[2] * r[2] + r[3] * r[3]);
    const double vx = r[1] / c0, vy = r[2] / c0, ve = r[3] / c0;
    if (sqrt(vx * vx + vy * vy + ve * ve) >= 1) fl |= VIRAL_FLAG_ACAUSAL;
    if (En <= 0) fl |= VIRAL_FLAG_ENEG;
    if (p.t_freeze > 0 && Temp < p.t_freeze) fl |= VIRAL_FLAG_FREEZEOUT;
#pragma unroll
    for (int l = 0; l < 4; l++) so[(VIRAL_F_T00 + l) * n + x] = V[l];
#pragma unroll
    for (int l = 0; l < 5; l++) so[(VIRAL_F_PI0 + l) * n + x] = V[4 + l];
    so[VIRAL_F_BULK * n + x] = V[14];
    so[VIRAL_F_EN * n + x] = En;
    so[VIRAL_F_P * n + x] = eos_p(p, En);
    so[VIRAL_F_TEMP * n + x] = Temp;
    so[VIRAL_F_VX * n + x] = vx;
    so[VIRAL_F_VY * n + x] = vy;
    so[VIRAL_F_VE * n + x] = ve;
    so[VIRAL_F_U0 * n + x] = c0;
    so[VIRAL_F_U1 * n + x] = r[1];
    so[VIRAL_F_U2 * n + x] = r[2];
    so[VIRAL_F_U3 * n + x] = r[3];
#pragma unroll
    for (int l = 0; l < 4; l++) so[(VIRAL_F_PU0 + l) * n + x] = a.stage == 0 ? u[l] : pu[l];
    const EtaGhosts eg(g, p.lbi, i, j, k);
    if (eg.any()) {
#pragma unroll
      for (int l = 0; l < 4; l++) eg.put(so, n, VIRAL_F_T00 + l, V[l]);
#pragma unroll
      for (int l = 0; l < 5; l++) eg.put(so, n, VIRAL_F_PI0 + l, V[4 + l]);
      // pi slots 5..9 are never written: their ghosts copy the cell's kept value
#pragma unroll
      for (int l = 5; l < 10; l++) eg.put(so, n, VIRAL_F_PI0 + l, si[(VIRAL_F_PI0 + l) * n + x]);
      eg.put(so, n, VIRAL_F_BULK, V[14]);
      eg.put(so, n, VIRAL_F_EN, En);
      eg.put(so, n, VIRAL_F_P, eos_p(p, En));
      eg.put(so, n, VIRAL_F_VX, vx);
      eg.put(so, n, VIRAL_F_VY, vy);
      eg.put(so, n, VIRAL_F_VE, ve);
      eg.put(so, n, VIRAL_F_U0, c0);
      eg.put(so, n, VIRAL_F_U1, r[1]);
      eg.put(so, n, VIRAL_F_U2, r[2]);
      eg.put(so, n, VIRAL_F_U3, r[3]);
#pragma unroll
      for (int l = 0; l < 4; l++) eg.put(so, n, VIRAL_F_PU0 + l, a.stage == 0 ? u[l] : pu[l]);
    }
    iters[x] = it;
    const uint8_t prev = a.stage == 0 ? 0 : flags[x];
    flags[x] = (uint8_t)((prev & ~VIRAL_FLAG_FREEZEOUT) | fl);
    if (fl & (VIRAL_FLAG_ACAUSAL | VIRAL_FLAG_ENEG | VIRAL_FLAG_SOLVER))
      atomicMin(bad, ((unsigned long long)step << 42) | ((unsigned long long)a.stage << 40) |
                         ((unsigned long long)x << 8) | (unsigned long long)fl);
    return;
  }
  // ---- primary variables at tau_new (NEWEST10/time.h:100-162)
  const double tn = a.tau_new;
  double T0[4];
#pragma unroll
  for (int l = 0; l < 4; l++) T0[l] = milne ? (FAST ? V[l] * (1.0 / tn) : (V[l]) / (tn)) : V[l];
#pragma unroll
  for (int l = 0; l < 10; l++) pi[l] = (V[4 + l]);
  const double PIn = (V[14]);
  // ---- store T/pi/PI before the root: pi[4..9] need no register across the
  // Newton loop
#pragma unroll
  for (int l = 0; l < 4; l++) so[(VIRAL_F_T00 + l) * n + x] = T0[l];
#pragma unroll
  for (int l = 0; l < 10; l++) so[(VIRAL_F_PI0 + l) * n + x] = pi[l];
  so[VIRAL_F_BULK * n + x] = PIn;
  const EtaGhosts eg(g, p.lbi, i, j, k);
  if (eg.any()) {
#pragma unroll
    for (int l = 0; l < 4; l++) eg.put(so, n, VIRAL_F_T00 + l, T0[l]);
#pragma unroll
    for (int l = 0; l < 10; l++) eg.put(so, n, VIRAL_F_PI0 + l, pi[l]);
    eg.put(so, n, VIRAL_F_BULK, PIn);
  }
  // ---- root and velocities (NEWEST10/root.h:301-424)
  int it;
  unsigned fl;
  double en, Temp, Pn, vx, vy, ve, gm;
  solve_cell<FAST>(p, milne, tn, T0, pi, PIn, e, en, Pn, Temp, vx, vy, ve, gm, it, fl);
  // ---- store (with the k0 broadcast on, the stage's broadcast kernel writes
  // v and u of every slice but k0 right after this kernel: those stores
  // would be dead)
  so[VIRAL_F_EN * n + x] = en;
  so[VIRAL_F_P * n + x] = Pn;
  so[VIRAL_F_TEMP * n + x] = Temp;
  const bool vu = !((p.quirks & VIRAL_Q_K0_BROADCAST) && !p.lbi && k != g.k0);
  if (vu) {
    so[VIRAL_F_VX * n + x] = vx;
    so[VIRAL_F_VY * n + x] = vy;
    so[VIRAL_F_VE * n + x] = ve;
    so[VIRAL_F_U0 * n + x] = gm;
    so[VIRAL_F_U1 * n + x] = gm * vx;
    so[VIRAL_F_U2 * n + x] = gm * vy;
    so[VIRAL_F_U3 * n + x] = gm * ve;
  }
#pragma unroll
  for (int l = 0; l < 4; l++) so[(VIRAL_F_PU0 + l) * n + x] = a.stage == 0 ? u[l] : pu[l];
  if (eg.any()) {
    eg.put(so, n, VIRAL_F_EN, en);
    eg.put(so, n, VIRAL_F_P, Pn);
    if (vu) {
      eg.put(so, n, VIRAL_F_VX, vx);
      eg.put(so, n, VIRAL_F_VY, vy);
      eg.put(so, n, VIRAL_F_VE, ve);
      eg.put(so, n, VIRAL_F_U0, gm);
      eg.put(so, n, VIRAL_F_U1, gm * vx);
      eg.put(so, n, VIRAL_F_U2, gm * vy);
      eg.put(so, n, VIRAL_F_U3, gm * ve);
    }
#pragma unroll
    for (int l = 0; l < 4; l++) eg.put(so, n, VIRAL_F_PU0 + l, a.stage == 0 ? u[l] : pu[l]);
  }
  iters[x] = it;
  uint8_t prev = a.stage == 0 ? 0 : flags[x];
  flags[x] = (uint8_t)((prev & ~VIRAL_FLAG_FREEZEOUT) | fl);
  if (fl & (VIRAL_FLAG_ACAUSAL | VIRAL_FLAG_ENEG | VIRAL_FLAG_SOLVER)) {
    // ordered by step, stage, then reference loop-order index (x < 2^32);
    // this stage's flags in the low byte
    atomicMin(bad, ((unsigned long long)step << 42) | ((unsigned long long)a.stage << 40) |
                       ((unsigned long long)x << 8) | (unsigned long long)fl);
  }
}

template <bool FAST, bool FUSED, bool PERSIST, bool CART>
__global__ void __launch_bounds__(BLK) k_update(const EvDev* __restrict__ tab, Sched s, const Clock* __restrict__ clk,
                                                int stage, int cur)
{
  TILE_LOOP(s)
  {
    long lb;
    const EvDev& E = tab[sched_find(s, w_, lb)];
    const StageArgs a = stage_args(clk->tau, clk->ts, E.p.rk, stage);
    update_cell<FAST, FUSED, CART>(E.g, E.p, a, clk->step, E.st[cur], E.st[cur ^ 1], E.flux[0], E.flux[1], E.flux[2],
                             E.src, E.var0, E.var1, E.var2, E.iters + (long)stage * E.g.n, E.flags, E.bad,
                             lb * BLK + threadIdx.x);
  }
}

// NEWEST10/root.h:426-449: every eta slice takes the velocity of slice k0
// GHOSTS: also the eta ghost copies of the cells it writes (the stage's
// broadcast, whose ghost fill covers only the x and y faces)
template <bool GHOSTS>
__device__ __forceinline__ void k0_cell(const Geom& g, const Params& p, double tau, double* __restrict__ st_, long t)
{
  auto* __restrict__ st = as_global(st_);
  int i, j, k;
  if (!box_cell(interior_box(g), t, i, j, k)) return;
  if (k == g.k0) return;
  const long n = g.n, x = cidx(g, i, j, k), x0 = cidx(g, i, j, g.k0);
  double t2 = p.milne ? tau * tau : 1.0;
  double vx = st[VIRAL_F_VX * n + x0], vy = st[VIRAL_F_VY * n + x0], ve = st[VIRAL_F_VE * n + x0];
  st[VIRAL_F_VE * n + x] = ve;
  st[VIRAL_F_VX * n + x] = vx;
  st[VIRAL_F_VY * n + x] = vy;
  double u0 = 1.0 / sqrt(1.0 - vx * vx - vy * vy - t2 * ve * ve);
  st[VIRAL_F_U0 * n + x] = u0;
  st[VIRAL_F_U1 * n + x] = u0 * vx;
  st[VIRAL_F_U2 * n + x] = u0 * vy;
  st[VIRAL_F_U3 * n + x] = u0 * ve;
  const EtaGhosts eg(g, p.lbi, i, j, k);
  if (GHOSTS && eg.any()) {
    eg.put(st, n, VIRAL_F_VX, vx);
    eg.put(st, n, VIRAL_F_VY, vy);
    eg.put(st, n, VIRAL_F_VE, ve);
    eg.put(st, n, VIRAL_F_U0, u0);
    eg.put(st, n, VIRAL_F_U1, u0 * vx);
    eg.put(st, n, VIRAL_F_U2, u0 * vy);
    eg.put(st, n, VIRAL_F_U3, u0 * ve);
  }
}

template <bool PERSIST>
__global__ void __launch_bounds__(BLK) k_k0_broadcast(const EvDev* __restrict__ tab, Sched s,
                                                      const Clock* __restrict__ clk, int stage, int cur)
{
  TILE_LOOP(s)
  {
    long lb;
    const EvDev& E = tab[sched_find(s, w_, lb)];
    const StageArgs a = stage_args(clk->tau, clk->ts, E.p.rk, stage);
    k0_cell<true>(E.g, E.p, a.tau_new, E.st[cur], lb * BLK + threadIdx.x);
  }
}

// fields copied into ghosts, NEWEST10/boundary.h:13-46 (all but Temp)
__constant__ int c_bfields[28] = {VIRAL_F_T00, VIRAL_F_T10, VIRAL_F_T20, VIRAL_F_T30, VIRAL_F_EN, VIRAL_F_VX,
                                  VIRAL_F_VY, VIRAL_F_VE, VIRAL_F_P, VIRAL_F_U0, VIRAL_F_U1, VIRAL_F_U2, VIRAL_F_U3,
                                  VIRAL_F_PU0, VIRAL_F_PU1, VIRAL_F_PU2, VIRAL_F_PU3, VIRAL_F_PI0, VIRAL_F_PI0 + 1,
                                  VIRAL_F_PI0 + 2, VIRAL_F_PI0 + 3, VIRAL_F_PI0 + 4, VIRAL_F_PI0 + 5,
                                  VIRAL_F_PI0 + 6, VIRAL_F_PI0 + 7, VIRAL_F_PI0 + 8, VIRAL_F_PI0 + 9, VIRAL_F_BULK};

// Zero-gradient ghost fill of the x and y layers of interior rows,
// NEWEST10/boundary.h:1-275 (the eta layers: EtaGhosts).  Thread = (ghost
// cell, field); the ghost slabs are disjoint and read only interior cells.
__host__ __device__ inline long ghost_count(const Geom& g, int lbi)
{
  // x and y faces; the eta faces are written by the update and k0 kernels
  // (EtaGhosts)
  (void)lbi;
  const long nx = g.ir - g.il, ny = g.jr - g.jl, nz = g.kr - g.kl;
  return 2L * VIRAL_BORDER * (ny * nz + nx * nz);
}

__device__ __forceinline__ void ghost_cell(const Geom& g, const Params& p, double* __restrict__ st_, int q, long t)
{
  auto* __restrict__ st = as_global(st_);
  const int f = c_bfields[q];
  const long n = g.n;
  const int nx = g.ir - g.il, ny = g.jr - g.jl, nz = g.kr - g.kl;
  const long cx = 2L * VIRAL_BORDER * ny * nz;  // x slabs
  const long cy = 2L * VIRAL_BORDER * nx * nz;  // y slabs
  const bool mirror = (p.quirks & VIRAL_Q_GUBSER_MIRROR) != 0 && p.milne && (f == VIRAL_F_U1 || f == VIRAL_F_U2);
  long dst, src;
  if (t < cx) {
    unsigned rk, rj;
    const unsigned r = udivmod((unsigned)t, (unsigned)nz, rk);
    int k = g.kl + (int)rk;
    int gg = (int)udivmod(r, (unsigned)ny, rj);  // 0..2B-1
    int j = g.jl + (int)rj;
    int side = gg / VIRAL_BORDER, h = gg % VIRAL_BORDER;
    int i = side == 0 ? h : g.ir + h;
    int is = side == 0 ? g.il : g.ir - 1;
    if (mirror) is = side == 0 ? g.il + VIRAL_BORDER - h - 1 : g.ir - h - 1;
    dst = cidx(g, i, j, k);
    src = cidx(g, is, j, k);
  } else if (t < cx + cy) {
    t -= cx;
    unsigned rk, ri;
    const unsigned r = udivmod((unsigned)t, (unsigned)nz, rk);
    int k = g.kl + (int)rk;
    int gg = (int)udivmod(r, (unsigned)nx, ri);
    int i = g.il + (int)ri;
    int side = gg / VIRAL_BORDER, h = gg % VIRAL_BORDER;
    int j = side == 0 ? h : g.jr + h;
    int js = side == 0 ? g.jl : g.jr - 1;
    if (mirror) js = side == 0 ? g.jl + VIRAL_BORDER - h - 1 : g.jr - h - 1;
    dst = cidx(g, i, j, k);
    src = cidx(g, i, js, k);
  } else {
    return;
  }
  st[f * n + dst] = st[f * n + src];
}

// tiles of one event: 28 fields x ceil(ghost cells / BLK)
template <bool PERSIST>
__global__ void __launch_bounds__(BLK) k_ghost_fill(const EvDev* __restrict__ tab, Sched s, int cur)
{
  TILE_LOOP(s)
  {
    long lb;
    const EvDev& E = tab[sched_find(s, w_, lb)];
    const long nb = (ghost_count(E.g, E.p.lbi) + BLK - 1) / BLK;
    unsigned rb;
    const int q = (int)udivmod((unsigned)lb, (unsigned)nb, rb);
    ghost_cell(E.g, E.p, E.st[cur], q, (long)rb * BLK + threadIdx.x);
  }
}

// first kernel of a step sequence: no failure recorded yet
__global__ void k_clear_bad(unsigned long long* __restrict__ bad, int nev)
{
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e < nev) bad[e] = ~0ull;
}

// last kernel of a step: tau += dtau (NEWEST10/main.cpp:156)
// stream-ordered resets used by the adjoint's recompute: the clock, and the
// two state buffers an event's table row points at
__global__ void k_set_clock(Clock* __restrict__ clk, double tau, double ts)
{
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    clk->tau = tau;
    clk->ts = ts;
    clk->step = 0;
  }
}

__global__ void k_set_state_ptrs(EvDev* __restrict__ tab, int ev, double* a, double* b)
{
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    tab[ev].st[0] = a;
    tab[ev].st[1] = b;
  }
}

// Copies what a stage writes from one state copy to another: every field of
// the interior cells, and the ghost-filled fields (all but Temp) of the face
// ghost cells (one index outside the interior).  The edge and corner ghosts,
// which no stage writes, keep the destination's values.
__global__ void __launch_bounds__(BLK) k_copy_written(Geom g, int lbi, const double* __restrict__ src,
                                                      double* __restrict__ dst)
{
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long)NF * g.n) return;
  const int f = (int)(t / g.n);
  const long x = t - (long)f * g.n;
  const int k = (int)(x % g.gz), j = (int)((x / g.gz) % g.gy), i = (int)(x / ((long)g.gz * g.gy));
  const int out = (i < g.il || i >= g.ir) + (j < g.jl || j >= g.jr) + (!lbi && (k < g.kl || k >= g.kr));
  if (out == 0 || (out == 1 && f != VIRAL_F_TEMP)) dst[t] = src[t];
}

// The FIX build's end of step (VIRAL_Q_FIX_VE; NEWEST10/time.h:280-290,
// FixTwo): the stored v_eta of every interior cell is zeroed when the last
// BMax exceeds 0.98; then DebugMSG's BMax (debug.h:30-74), the largest
// sqrt(vx^2 + vy^2 + tau^2 ve^2) over the interior at the step's start tau,
// is taken anew.  BMax is kept as the bits of a non-negative double, whose
// integer order is the value order, so one atomicMax per block combines it.
__global__ void __launch_bounds__(BLK) k_fix_ve(Geom g, double* __restrict__ st,
                                                const unsigned long long* __restrict__ bmax)
{
  int i, j, k;
  if (!box_cell(interior_box(g), (long)blockIdx.x * blockDim.x + threadIdx.x, i, j, k)) return;
  if (!(__longlong_as_double((long long)*bmax) > 0.98)) return;
  st[VIRAL_F_VE * g.n + cidx(g, i, j, k)] = 0.0;
}

__global__ void k_bmax_clear(unsigned long long* __restrict__ bmax)
{
  if (threadIdx.x == 0 && blockIdx.x == 0) *bmax = 0ull;
}

__global__ void __launch_bounds__(BLK) k_bmax(Geom g, const double* __restrict__ st, const Clock* __restrict__ clk,
                                              unsigned long long* __restrict__ bmax)
{
  int i, j, k;
  double m = 0.0;
  if (box_cell(interior_box(g), (long)blockIdx.x * blockDim.x + threadIdx.x, i, j, k)) {
    const long n = g.n, x = cidx(g, i, j, k);
    const double tau = clk->tau;
    const double vx = st[VIRAL_F_VX * n + x], vy = st[VIRAL_F_VY * n + x], ve = st[VIRAL_F_VE * n + x];
    const double b = sqrt(vx * vx + vy * vy + tau * tau * ve * ve);
    if (b > m) m = b;  // NaN never counts
  }
  for (int o = 32; o > 0; o >>= 1) m = fmax(m, __shfl_down(m, o));
  __shared__ double red[BLK / 64];
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < BLK / 64; w++) m = fmax(m, red[w]);
    if (m > 0.0) atomicMax(bmax, (unsigned long long)__double_as_longlong(m));
  }
}

__global__ void k_clock_advance(Clock* __restrict__ clk)
{
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    clk->tau = clk->tau + clk->ts;
    clk->step = clk->step + 1;
  }
}

__global__ void k_max_temp(Geom g, Params p, const double* __restrict__ st, unsigned long long* out)
{
  long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  int i, j, k;
  double v = 0;
  if (box_cell(interior_box(g), t, i, j, k)) {
    double tt = GEVFM * eos_T(p, st[VIRAL_F_EN * g.n + cidx(g, i, j, k)]);
    if (tt > v) v = tt;
  }
  // wave reduction then one atomic per wave; non-negative doubles order like their bits
  for (int o = 32; o > 0; o >>= 1) {
    double w = __shfl_down(v, o);
    if (w > v) v = w;
  }
  if ((threadIdx.x & 63) == 0 && v > 0) atomicMax(out, (unsigned long long)__double_as_longlong(v));
}

// RootSearchForEnVelUsingDerivatives (NEWEST10/root.h:212-424) on a state in
// place: every interior cell's e (Newton from its current e), P, Temp, v, u
// from its T^{tau mu}, pi, PI at tau; iteration counts, flags and the first
// bad cell as stage 0 of a step.  The k0 broadcast follows in k_root_k0.
template <bool FAST>
__global__ void __launch_bounds__(BLK) k_root(Geom g, Params p, double tau, double* __restrict__ st,
                                              int32_t* __restrict__ iters, uint8_t* __restrict__ flags,
                                              unsigned long long* __restrict__ bad)
{
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  int i, j, k;
  if (!box_cell(interior_box(g), t, i, j, k)) return;
  const long n = g.n, x = cidx(g, i, j, k);
  double T0[4], pi[10];
  for (int l = 0; l < 4; l++) T0[l] = st[(VIRAL_F_T00 + l) * n + x];
  for (int l = 0; l < 10; l++) pi[l] = st[(VIRAL_F_PI0 + l) * n + x];
  int it;
  unsigned fl;
  double en, Pn, Temp, vx, vy, ve, gm;
  solve_cell<FAST>(p, 1, tau, T0, pi, st[VIRAL_F_BULK * n + x], st[VIRAL_F_EN * n + x], en, Pn, Temp, vx, vy, ve, gm,
                   it, fl);
  st[VIRAL_F_EN * n + x] = en;
  st[VIRAL_F_P * n + x] = Pn;
  st[VIRAL_F_TEMP * n + x] = Temp;
  st[VIRAL_F_VX * n + x] = vx;
  st[VIRAL_F_VY * n + x] = vy;
  st[VIRAL_F_VE * n + x] = ve;
  st[VIRAL_F_U0 * n + x] = gm;
  st[VIRAL_F_U1 * n + x] = gm * vx;
  st[VIRAL_F_U2 * n + x] = gm * vy;
  st[VIRAL_F_U3 * n + x] = gm * ve;
  iters[x] = it;
  flags[x] = (uint8_t)fl;
  if (fl & (VIRAL_FLAG_ACAUSAL | VIRAL_FLAG_ENEG | VIRAL_FLAG_SOLVER))
    atomicMin(bad, ((unsigned long long)x << 8) | (unsigned long long)fl);
}

__global__ void __launch_bounds__(BLK) k_root_k0(Geom g, Params p, double tau, double* __restrict__ st)
{
  k0_cell<false>(g, p, tau, st, (long)blockIdx.x * blockDim.x + threadIdx.x);
}

// ---------------------------------------------------------------- Navier-Stokes initialisation
// The velocity gradients CalcDer4Vel (NEWEST10/Source.h:3-101) leaves on a
// freshly allocated (zeroed) grid, at ghosted cell (i,j,k): the time
// derivative (u - prevu)/ts on the interior; the x derivative on interior y-z
// rows for 2 <= i < gx-2 and the y derivative on interior x-z rows for
// 2 <= j < gy-2; the eta derivative on interior x-y columns for every k, the
// stencil clamped at both ends; zero everywhere else.
__device__ __forceinline__ void ns_du(const Geom& g, const Params& p, const double* __restrict__ st, int i, int j,
                                      int k, double ts, double du[4][4])
{
  const bool iin = i >= g.il && i < g.ir, jin = j >= g.jl && j < g.jr, kin = k >= g.kl && k < g.kr;
  const long n = g.n, x = cidx(g, i, j, k);
  for (int l = 0; l < 4; l++) {
    const double* ul = st + (VIRAL_F_U0 + l) * n;
    du[l][0] = (iin && jin && kin) ? (ul[x] - st[(VIRAL_F_PU0 + l) * n + x]) / ts : 0.0;
    du[l][1] = (jin && kin && i >= 2 && i < g.gx - 2)
                   ? weno_der(ul[x - 2 * g.sx], ul[x - g.sx], ul[x], ul[x + g.sx], ul[x + 2 * g.sx], p.dx)
                   : 0.0;
    du[l][2] = (iin && kin && j >= 2 && j < g.gy - 2)
                   ? weno_der(ul[x - 2 * g.sy], ul[x - g.sy], ul[x], ul[x + g.sy], ul[x + 2 * g.sy], p.dy)
                   : 0.0;
    du[l][3] = 0.0;
    if (!p.lbi && iin && jin) {
      const int Z = g.gz;
      const double qc = ul[x];
      double qm2, qm1, qp1, qp2;
      if (k > 1 && k < Z - 2) {
        qm1 = ul[x - 1]; qm2 = ul[x - 2]; qp1 = ul[x + 1]; qp2 = ul[x + 2];
      } else if (k == 0) {
        qm2 = qm1 = qc; qp1 = ul[x + 1]; qp2 = ul[x + 2];
      } else if (k == 1) {
        qm2 = qm1 = ul[x - 1]; qp1 = ul[x + 1]; qp2 = ul[x + 2];
      } else if (k == Z - 2) {
        qm1 = ul[x - 1]; qm2 = ul[x - 2]; qp2 = qp1 = ul[x + 1];
      } else {
        qm1 = ul[x - 1]; qm2 = ul[x - 2]; qp2 = qp1 = qc;
      }
      du[l][3] = weno_der(qm2, qm1, qc, qp1, qp2, p.deta);
    }
  }
}

// CalcNS / NSInit / ZeroInit (NEWEST10/Source.h:542-776) over every ghosted
// cell, one thread per cell.  MODE 0 writes the Navier-Stokes pi to
// out[10][n] (and, with bulk, PI = -zeta theta to out[10]) and leaves the
// state alone; MODE 1 sets pi and PI to them, MODE 2 both to zero, and both
// set T^{tau mu} = (e + P - PI) u^tau u^mu - (P - PI) g^{tau mu} + pi^{tau mu}
// in place (a cell writes only its own pi, PI and T; the gradients read u and
// prevu).  Exact arithmetic and the closed-form EOS in every mode, as the CPU
// restatement.
template <int MODE>
__global__ void __launch_bounds__(BLK) k_ns_init(Geom g, Params p, double tau, double ts, double* __restrict__ st,
                                                 double* __restrict__ out)
{
  const long x = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (x >= g.n) return;
  const long n = g.n;
  const int i = (int)(x / g.sx), j = (int)((x / g.sy) % g.gy), k = (int)(x % g.gz);
  double pn[10], nsPI = 0.0;
  for (int l = 0; l < 10; l++) pn[l] = 0.0;
  if (MODE != 2) {
    double du[4][4], u[4];
    ns_du(g, p, st, i, j, k, ts, du);
    for (int l = 0; l < 4; l++) u[l] = st[(VIRAL_F_U0 + l) * n + x];
    const double e = st[VIRAL_F_EN * n + x], P = st[VIRAL_F_P * n + x];
    const double T = eos_T(p, e);
    const double s = eos_s(p, e, P, T);
    ns_shear(tau, u, du, eta_of(p, s), pn);
    if (p.bulk) nsPI = ns_bulk(tau, u, du, zeta_of(p, s));
  }
  if (MODE == 0) {
    for (int l = 0; l < 10; l++) out[l * n + x] = pn[l];
    if (p.bulk) out[10 * n + x] = nsPI;
    return;
  }
  for (int l = 0; l < 10; l++) st[(VIRAL_F_PI0 + l) * n + x] = pn[l];
  st[VIRAL_F_BULK * n + x] = nsPI;
  const double e = st[VIRAL_F_EN * n + x], P = st[VIRAL_F_P * n + x], PI = nsPI;
  const double u0 = st[VIRAL_F_U0 * n + x];
  st[VIRAL_F_T00 * n + x] = -P + PI + (e + P - PI) * (u0 * u0) + pn[0];
  for (int l = 1; l < 4; l++) st[(VIRAL_F_T00 + l) * n + x] = (e + P - PI) * u0 * st[(VIRAL_F_U0 + l) * n + x] + pn[l];
}

// ---------------------------------------------------------------- Landau matching
// The per-cell part of ginit (NEWEST10/ginit.h:876-990) on a T^{mu nu} field
// (tmunu[10][n], pi index map order, contravariant at tau), one thread per
// ghosted cell: the Landau frame from landau_cell (FindVar's selection,
// ginit.h:521-629), then e, Temp, P, v and u; AT_START false (the call at
// tau - ts) stores prevu = u; AT_START true sets pi = T - T_ideal - PI
// (g^{-1} - u u) with PI = P - e/3, then PI = 0 (kept with bulk) and T^{tau mu}.  A
// cell without exactly one passing eigenvector ("EIGEN VALUE SEARCH FAILED",
// where the reference exits) keeps its state, gets VIRAL_FLAG_SOLVER and
// competes for the first bad cell.  Exact arithmetic and the closed-form EOS.
template <bool AT_START>
__global__ void __launch_bounds__(BLK) k_landau(Geom g, Params p, double tau, const double* __restrict__ tmunu,
                                                double* __restrict__ st, uint8_t* __restrict__ flags,
                                                unsigned long long* __restrict__ bad)
{
  const long x = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (x >= g.n) return;
  const long n = g.n;
  double T[10], E = 0.0, v[3] = {0.0, 0.0, 0.0};
  for (int l = 0; l < 10; l++) T[l] = tmunu[l * n + x];
  if (landau_cell(tau, T, &E, v) != 1) {
    flags[x] = (uint8_t)VIRAL_FLAG_SOLVER;
    atomicMin(bad, ((unsigned long long)x << 8) | (unsigned long long)VIRAL_FLAG_SOLVER);
    return;
  }
  flags[x] = 0;
  const double tau2 = tau * tau;
  const double guu[4] = {1.0, -1.0, -1.0, -1.0 / tau2};
  const double e = E, P = eos_p(p, E);
  st[VIRAL_F_EN * n + x] = e;
  st[VIRAL_F_TEMP * n + x] = eos_T(p, e);
  st[VIRAL_F_P * n + x] = P;
  st[VIRAL_F_VX * n + x] = v[0];
  st[VIRAL_F_VY * n + x] = v[1];
  st[VIRAL_F_VE * n + x] = v[2];
  const double u0 = 1.0 / sqrt(1.0 - v[0] * v[0] - v[1] * v[1] - tau2 * v[2] * v[2]);
  const double u[4] = {u0, u0 * v[0], u0 * v[1], u0 * v[2]};
  for (int l = 0; l < 4; l++) st[(VIRAL_F_U0 + l) * n + x] = u[l];
  if (!AT_START) {
    for (int l = 0; l < 4; l++) st[(VIRAL_F_PU0 + l) * n + x] = u[l];
    return;
  }
  double tid[4][4];
  for (int a = 0; a < 4; a++)
    for (int b = 0; b < 4; b++) tid[a][b] = (e + P) * u[a] * u[b] - P * (a == b ? guu[a] : 0.0);
  const double temp2 = tid[0][0] - tid[1][1] - tid[2][2] - tau * tau * tid[3][3];
  const double PIt = -temp2 / 3.0;
  double pn[10];
  for (int a = 0; a < 4; a++)
    for (int b = a; b < 4; b++) {
      const double PImat = PIt * ((a == b ? guu[a] : 0.0) - u[a] * u[b]);
      pn[pidx(a, b)] = T[pidx(a, b)] - tid[a][b] - PImat;
    }
  for (int l = 0; l < 10; l++) st[(VIRAL_F_PI0 + l) * n + x] = pn[l];
  const double PI = p.bulk ? PIt : 0.0;  // ginit.h:996-999 zeroes it without BULK
  st[VIRAL_F_BULK * n + x] = PI;
  st[VIRAL_F_T00 * n + x] = -P + PI + (e + P - PI) * (u0 * u0) + pn[0];
  for (int l = 1; l < 4; l++) st[(VIRAL_F_T00 + l) * n + x] = (e + P - PI) * u0 * u[l] + pn[l];
}

// ---------------------------------------------------------------- per-step diagnostics
// The driver loop's reductions over an event's interior: CheckRoot (NEWEST10/
// root.h:1-199; TXYZ5/multiroot.h:7-150), the DebugMSG extrema (debug.h:
// 3-100) and CheckPhysics (utils.h:231-258).  A fixed grid of DIAG_BLOCKS
// blocks walks the interior in reference loop order (grid stride); each
// block reduces its cells to one DiagPart per statistic and the host combines
// the blocks in block order, so sums are the same from run to run.
// Statistic rules (which cell a statistic reports):
//   MAX_FIRST  largest value, first cell in loop order reaching it; cells
//              with value <= 0 never count (the reference's `x > m` from 0)
//   MAX_LAST   largest value >= 0, last cell reaching it (`x >= m` from 0)
//   KEY_FIRST / KEY_LAST  first / last cell where a predicate holds, with
//              the value of that cell
// NaN values never count, as with the reference's comparisons.
constexpr int DIAG_BLOCKS = 512;
constexpr int DIAG_NSTAT = 9;
enum DiagRule { R_MAX_FIRST, R_MAX_LAST, R_KEY_FIRST, R_KEY_LAST };
enum DiagKind { D_ROOT, D_EXTREMA, D_PHYSICS };

struct DiagPart {
  double v;
  long long at;  // ghosted linear index, -1: no cell
  double sum;
};

__host__ __device__ inline DiagPart diag_combine(DiagPart a, DiagPart b, int rule)
{
  DiagPart r;
  r.sum = a.sum + b.sum;
  bool take_b;
  if (b.at < 0) take_b = false;
  else if (a.at < 0) take_b = true;
  else if ((rule == R_MAX_FIRST || rule == R_MAX_LAST) && b.v != a.v) take_b = b.v > a.v;
  else take_b = (rule == R_MAX_FIRST || rule == R_KEY_FIRST) ? b.at < a.at : b.at > a.at;
  r.v = take_b ? b.v : a.v;
  r.at = take_b ? b.at : a.at;
  return r;
}

// rules and count of the statistics of each kind
__host__ __device__ constexpr int diag_nstat(int kind, int milne)
{
  return kind == D_ROOT ? (milne ? 6 : 9) : kind == D_EXTREMA ? 3 : 2;
}
__host__ __device__ inline int diag_rule(int kind, int milne, int s)
{
  if (kind == D_ROOT) return (!milne && s == 8) ? R_KEY_LAST : R_MAX_FIRST;
  if (kind == D_EXTREMA) return s == 0 ? R_MAX_LAST : R_MAX_FIRST;
  return R_KEY_FIRST;
}

// |v|, NEWEST10/debug.h:35 and utils.h:242 (TXYZ5: no tau)
__device__ __forceinline__ double diag_speed(const Params& p, const double* st, long n, long x, double tau)
{
  const double vx = st[VIRAL_F_VX * n + x], vy = st[(VIRAL_F_VX + 1) * n + x], ve = st[(VIRAL_F_VX + 2) * n + x];
  if (p.milne) return sqrt(vx * vx + vy * vy + tau * tau * ve * ve);
  return sqrt(vx * vx + vy * vy + ve * ve);
}

// Per-cell values of the statistics (value, counts toward the statistic,
// added to the sum).  Exact arithmetic in the reference's operation order
// (this file is compiled without FP contraction), closed-form EOS.
// CART (TXYZ5) fixes the statistic layout at compile time, so the per-cell
// arrays stay in registers.
template <int KIND, bool CART>
__device__ __forceinline__ void diag_cell(const Params& p, const double* __restrict__ st, long n, long x, double tau,
                                          double val[DIAG_NSTAT], bool cand[DIAG_NSTAT], double add[DIAG_NSTAT])
{
  auto F = [&](int f) { return st[(long)f * n + x]; };
  if (KIND == D_EXTREMA) {
    const double b = diag_speed(p, st, n, x, tau), e = F(VIRAL_F_EN);
    val[0] = b; cand[0] = b >= 0;           // b >= vmax from 0: the last cell reaching the maximum
    val[1] = e; cand[1] = e > 0;            // e > emax from 0
    val[2] = -e; cand[2] = -e > 0;          // e < emin from 0
    add[0] = add[1] = add[2] = 0;
    return;
  }
  if (KIND == D_PHYSICS) {
    const double b = diag_speed(p, st, n, x, tau), e = F(VIRAL_F_EN);
    cand[0] = b >= 1; val[0] = b; add[0] = cand[0] ? 1.0 : 0.0;
    cand[1] = e <= 0; val[1] = e; add[1] = cand[1] ? 1.0 : 0.0;
    return;
  }
  const double e = F(VIRAL_F_EN), PI = F(VIRAL_F_BULK), P = eos_p(p, e);
  const double W = F(VIRAL_F_T00), X = F(VIRAL_F_T00 + 1), Y = F(VIRAL_F_T00 + 2), Z = F(VIRAL_F_T00 + 3);
  double u[4], pi[10], rows[4], trace;
  for (int m = 0; m < 4; m++) u[m] = F(VIRAL_F_U0 + m);
  constexpr int r0 = CART ? 4 : 1;
  if (!CART) {
    // root.h:29-44
    const double pi1 = F(VIRAL_F_PI0), pi2 = F(VIRAL_F_PI0 + 1), pi3 = F(VIRAL_F_PI0 + 2), pi4 = F(VIRAL_F_PI0 + 3);
    const double M2 = (X - pi2) * (X - pi2) + (Y - pi3) * (Y - pi3) + tau * tau * (Z - pi4) * (Z - pi4);
    const double res = e - (W - pi1) + (M2 / ((W - pi1) + P - PI));
    val[0] = fabs(res);
    for (int l = 0; l < 10; l++) pi[l] = F(VIRAL_F_PI0 + l);
    // root.h:72-90
    rows[0] = pi[0] * u[0] - pi[1] * u[1] - pi[2] * u[2] - pi[3] * u[3] * tau * tau;
    rows[1] = pi[1] * u[0] - pi[4] * u[1] - pi[5] * u[2] - pi[6] * u[3] * tau * tau;
    rows[2] = pi[2] * u[0] - pi[5] * u[1] - pi[7] * u[2] - pi[8] * u[3] * tau * tau;
    rows[3] = pi[3] * u[0] - pi[6] * u[1] - pi[8] * u[2] - pi[9] * u[3] * tau * tau;
    trace = pi[0] - pi[4] - pi[7] - pi[9] * tau * tau;
  } else {
    // multiroot.h:45-48 with pow(x, 0.5) as sqrt and pow(u, 2) as u*u
    const double u1 = u[1], u2 = u[2], u3 = u[3];
    const double p1 = F(VIRAL_F_PI0), p2 = F(VIRAL_F_PI0 + 1), p3 = F(VIRAL_F_PI0 + 2), p4 = F(VIRAL_F_PI0 + 3),
                 p5 = F(VIRAL_F_PI0 + 4);
    const double Q = 1 + u1 * u1 + u2 * u2 + u3 * u3, S3 = u1 * u1 + u2 * u2 + u3 * u3, rq = sqrt(Q);
    val[0] = fabs(-(u1 * X) - u2 * Y - u3 * Z + (-e + W) * rq);
    val[1] = fabs(-(p3 * u2) - p4 * u3 - P * u1 * Q - u1 * (p1 + e * Q - PI * Q) + X * rq);
    val[2] = fabs(-(p3 * u1) - (e + p2 - PI) * u2 - p5 * u3 - (e - PI) * u2 * S3 - P * u2 * Q + Y * rq);
    val[3] = fabs(-(p4 * u1) - p5 * u2 - P * u3 + u3 * (p1 + p2 + PI - W) + Z * rq);
    cart_pi10(u, p1, p2, p3, p4, p5, pi);
    // multiroot.h:90-104
    rows[0] = pi[0] * u[0] - pi[1] * u[1] - pi[2] * u[2] - pi[3] * u[3];
    rows[1] = pi[1] * u[0] - pi[4] * u[1] - pi[5] * u[2] - pi[6] * u[3];
    rows[2] = pi[2] * u[0] - pi[5] * u[1] - pi[7] * u[2] - pi[8] * u[3];
    rows[3] = pi[3] * u[0] - pi[6] * u[1] - pi[8] * u[2] - pi[9] * u[3];
    trace = pi[0] - pi[4] - pi[7] - pi[9];
  }
#pragma unroll
  for (int l = 0; l < 4; l++) val[r0 + l] = fabs(rows[l]);
  val[r0 + 4] = fabs(trace);
#pragma unroll
  for (int s = 0; s < r0 + 4; s++) { cand[s] = val[s] > 0; add[s] = val[s]; }
  // Cartesian trace: `fabs(trace) > trace` keeps the last negative trace
  cand[r0 + 4] = !CART ? val[r0 + 4] > 0 : fabs(trace) > trace;
  add[r0 + 4] = val[r0 + 4];
}

template <int KIND, bool CART>
__global__ void __launch_bounds__(BLK) k_diag(Geom g, Params p, const double* __restrict__ st, double tau,
                                              DiagPart* __restrict__ out)
{
  __shared__ DiagPart lds[BLK];
  constexpr int ns = diag_nstat(KIND, !CART);
  DiagPart acc[DIAG_NSTAT];
#pragma unroll
  for (int s = 0; s < DIAG_NSTAT; s++) acc[s] = DiagPart{0.0, -1, 0.0};
  const Box b = interior_box(g);
  const long total = (long)b.ni * b.nj * b.nk;
  for (long t = (long)blockIdx.x * BLK + threadIdx.x; t < total; t += (long)gridDim.x * BLK) {
    int i, j, k;
    box_cell(b, t, i, j, k);
    const long x = cidx(g, i, j, k);
    double val[DIAG_NSTAT], add[DIAG_NSTAT];
    bool cand[DIAG_NSTAT];
    diag_cell<KIND, CART>(p, st, g.n, x, tau, val, cand, add);
#pragma unroll
    for (int s = 0; s < ns; s++)
      acc[s] = diag_combine(acc[s], DiagPart{val[s], cand[s] ? (long long)x : -1, add[s]}, diag_rule(KIND, !CART, s));
  }
#pragma unroll
  for (int s = 0; s < ns; s++) {
    lds[threadIdx.x] = acc[s];
    __syncthreads();
    for (int h = BLK / 2; h > 0; h >>= 1) {
      if ((int)threadIdx.x < h)
        lds[threadIdx.x] = diag_combine(lds[threadIdx.x], lds[threadIdx.x + h], diag_rule(KIND, !CART, s));
      __syncthreads();
    }
    if (threadIdx.x == 0) out[(long)blockIdx.x * DIAG_NSTAT + s] = lds[0];
    __syncthreads();
  }
}

// Combines the block partials of k_diag in block order, one thread per
// statistic (the same fold as the host side of run_diag, so the results are
// equal bit for bit), into out[0..ns).
template <int KIND>
__global__ void k_diag_combine(const DiagPart* __restrict__ part, unsigned nb, int milne, DiagPart* __restrict__ out)
{
  const int s = (int)threadIdx.x;
  if (s >= diag_nstat(KIND, milne)) return;
  DiagPart acc{0.0, -1, 0.0};
  for (unsigned b = 0; b < nb; b++) acc = diag_combine(acc, part[(size_t)b * DIAG_NSTAT + s], diag_rule(KIND, milne, s));
  out[s] = acc;
}

// Device scratch of the allocation-free CheckRoot (viral_check_root_noalloc):
// the block partials of k_diag (at most DIAG_BLOCKS blocks) followed by the
// DIAG_NSTAT combined statistics, 110.8 KB per GPU for the life of the
// process.  Host code never passes this name to a kernel: it passes the
// address hipGetSymbolAddress returned on that device (scratch_ptr below).
constexpr int SCRATCH_PARTS = DIAG_BLOCKS * DIAG_NSTAT + DIAG_NSTAT;
static __device__ __align__(256) DiagPart scratch[SCRATCH_PARTS];

// write.h output records (NEWEST10/write.h:87-103): one thread per double of
// the record array; record r = ((oi * nyo) + oj) * nzo + ok holds
// {e, T, P, vx, vy, ve, pi[0..9], PI} of cell (il + oi*freq, jl + oj*freq,
// kl + ok*freqz), or of the k0 slice when xy.
__global__ void __launch_bounds__(BLK) k_output(Geom g, const double* __restrict__ st, int freq, int freqz, int xy,
                                                long nyo, long nzo, long total, double* __restrict__ out)
{
  const long t = (long)blockIdx.x * BLK + threadIdx.x;
  if (t >= total) return;
  const long r = t / 17;
  const int f = (int)(t - r * 17);
  const long ok = r % nzo, oj = (r / nzo) % nyo, oi = r / (nzo * nyo);
  const int i = g.il + (int)oi * freq, j = g.jl + (int)oj * freq;
  const int k = xy ? g.k0 : g.kl + (int)ok * freqz;
  const int field = f == 0 ? VIRAL_F_EN : f == 1 ? VIRAL_F_TEMP : f == 2 ? VIRAL_F_P : f < 6 ? VIRAL_F_VX + (f - 3)
                    : f < 16 ? VIRAL_F_PI0 + (f - 6) : VIRAL_F_BULK;
  out[t] = st[(long)field * g.n + cidx(g, i, j, k)];
}

// ---------------------------------------------------------------- adjoint kernels
// (FMA contraction on: the reverse pass is checked against finite differences,
// not against a bitwise operation order)
#pragma clang fp contract(fast)
// Reverse of one Milne KT stage for one event (conformal or s95p EOS, the
// latter in closed form or, in fast mode, from the table).  lam_out is
// the adjoint of the stage's output state (29 fields), lam_in receives the
// adjoint of its input state.  See adjoint.h for the per-cell pieces.

// Ghost fill in reverse: every interior cell takes the adjoint of the ghosts
// that copied from it (NEWEST10/boundary.h, including the q7 mirror).
template <int D>
__global__ void __launch_bounds__(BLK) k_adj_ghost(Geom g, Params p, double* __restrict__ lam)
{
  // one thread per (field, line across the D faces, side); lines are interior
  // in the other two directions, the fast one (eta, or y for D = 2) fastest
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const int na = D == 0 ? g.jr - g.jl : g.ir - g.il;
  const int nb = D == 2 ? g.jr - g.jl : g.kr - g.kl;
  const long per = (long)na * nb;
  if (t >= per * 2 * 28) return;
  unsigned bu, au, su;
  const unsigned tl = udivmod((unsigned)t, (unsigned)nb, bu);  // line index a + na * (side + 2 q)
  const unsigned ts = udivmod(tl, (unsigned)na, au);
  const int q = (int)udivmod(ts, 2u, su);
  const int b = (int)bu, a = (int)au, side = (int)su;
  const int f = c_bfields[q];
  const bool mirror = D < 2 && (p.quirks & VIRAL_Q_GUBSER_MIRROR) != 0 && p.milne && (f == VIRAL_F_U1 || f == VIRAL_F_U2);
  const int lo = D == 0 ? g.il : D == 1 ? g.jl : g.kl, hi = D == 0 ? g.ir : D == 1 ? g.jr : g.kr;
  auto at = [&](int cd) -> long {
    if (D == 0) return cidx(g, cd, g.jl + a, g.kl + b);
    if (D == 1) return cidx(g, g.il + a, cd, g.kl + b);
    return cidx(g, g.il + a, g.jl + b, cd);
  };
  const long n = g.n;
  double* lf = lam + (long)f * n;
  if (mirror) {
    // ghost h takes its mirror cell (NEWEST10/boundary.h:32-37): one ghost per source cell
    for (int h = 0; h < VIRAL_BORDER; h++) {
      const long src = side == 0 ? at(lo + VIRAL_BORDER - h - 1) : at(hi - h - 1);
      const long gh = side == 0 ? at(h) : at(hi + h);
      lf[src] += lf[gh];
    }
  } else {
    const long src = side == 0 ? at(lo) : at(hi - 1);
    double acc = lf[src];
    for (int h = 0; h < VIRAL_BORDER; h++) acc += lf[side == 0 ? at(h) : at(hi + h)];
    lf[src] = acc;
  }
}

// Cells a stage never writes keep their input values: corner ghosts (all
// fields; the fill covers only the x, y and eta slabs of interior rows) and
// Temp in the ghost slabs.  Their adjoint passes straight through.
__global__ void __launch_bounds__(BLK) k_adj_pass(Geom g, const double* __restrict__ lo, double* __restrict__ li)
{
  long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= g.n) return;
  unsigned ku, ju;
  const int i = (int)udivmod(udivmod((unsigned)t, (unsigned)g.gz, ku), (unsigned)g.gy, ju);
  const int k = (int)ku, j = (int)ju;
  const int out = (i < g.il || i >= g.ir) + (j < g.jl || j >= g.jr) + (k < g.kl || k >= g.kr);
  const long n = g.n;
  if (out == 1) {
    li[VIRAL_F_TEMP * n + t] += lo[VIRAL_F_TEMP * n + t];
  } else if (out > 1) {
    for (int f = 0; f < NF; f++) li[f * n + t] += lo[f * n + t];
  }
}

// k0 broadcast in reverse (NEWEST10/root.h:426-449): one wavefront per (i,j)
// line; its lanes take the eta slices, each collects what its slices took
// from slice k0, and a fixed-order wave reduction adds the sum to slice k0.
__global__ void __launch_bounds__(BLK) k_adj_k0(Geom g, Params p, double tn, const double* __restrict__ so,
                                                double* __restrict__ lam)
{
  const long wv = ((long)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int lane = threadIdx.x & 63;
  const int ni = g.ir - g.il, nj = g.jr - g.jl;
  if (wv >= (long)ni * nj) return;  // whole wavefronts leave together
  unsigned ju;
  const int i = g.il + (int)udivmod((unsigned)wv, (unsigned)nj, ju), j = g.jl + (int)ju;
  const long n = g.n, x0 = cidx(g, i, j, g.k0);
  const double v0[3] = {so[VIRAL_F_VX * n + x0], so[VIRAL_F_VY * n + x0], so[VIRAL_F_VE * n + x0]};
  const double t2 = tn * tn;
  double vb[3] = {0, 0, 0};
  for (int k = g.kl + lane; k < g.kr; k += 64) {
    if (k == g.k0) continue;
    const long x = cidx(g, i, j, k);
    const double ub[4] = {lam[VIRAL_F_U0 * n + x], lam[VIRAL_F_U1 * n + x], lam[VIRAL_F_U2 * n + x],
                          lam[VIRAL_F_U3 * n + x]};
    for (int c = 0; c < 3; c++) vb[c] += lam[(VIRAL_F_VX + c) * n + x];
    gamma_vjp(v0, t2, ub, vb);
    for (int c = 0; c < 3; c++) lam[(VIRAL_F_VX + c) * n + x] = 0.0;
    for (int c = 0; c < 4; c++) lam[(VIRAL_F_U0 + c) * n + x] = 0.0;
  }
  for (int o = 32; o > 0; o >>= 1)
    for (int c = 0; c < 3; c++) vb[c] += __shfl_down(vb[c], o);
  if (lane == 0)
    for (int c = 0; c < 3; c++) lam[(VIRAL_F_VX + c) * n + x0] += vb[c];
}

// Update and root of one interior cell in reverse.  so / lo: the stage's
// output state and its adjoint; vb0..vb2: adjoints of the RK registers
// Var0..Var2; writes resb, the adjoint of L = Result + Source (15 rows), for
// k_adj_src and the flux pass.  CART: the Cartesian update (Var rows
// {0..8, 14}, T = Var, the 4-D root; cart_update_vjp).
template <bool CART>
__global__ void __launch_bounds__(BLK) k_adj_root(Geom g, Params p, StageArgs a, const double* __restrict__ so,
                                                  const double* __restrict__ lo, double* __restrict__ li,
                                                  double* __restrict__ vb0, double* __restrict__ vb1,
                                                  double* __restrict__ vb2, double* __restrict__ resb)
{
  long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  int i, j, k;
  if (!box_cell(interior_box(g), t, i, j, k)) return;
  const long n = g.n, x = cidx(g, i, j, k);
  // prevu = stage 0 ? u : prevu
  for (int l = 0; l < 4; l++) {
    const double pb = lo[(VIRAL_F_PU0 + l) * n + x];
    if (a.stage == 0) li[(VIRAL_F_U0 + l) * n + x] += pb;
    else li[(VIRAL_F_PU0 + l) * n + x] += pb;
  }
  // the Cartesian step never writes pi slots 5..9 (it evolves five
  // components): they keep their values, and their adjoint passes through
  if (CART)
    for (int l = 5; l < 10; l++) li[(VIRAL_F_PI0 + l) * n + x] += lo[(VIRAL_F_PI0 + l) * n + x];
  // velocities and root
  double T0[4], pio[10], T0b[4], pibo[10], vbo[3], ubo[4];
  for (int l = 0; l < 4; l++) {
    T0[l] = so[(VIRAL_F_T00 + l) * n + x];
    T0b[l] = lo[(VIRAL_F_T00 + l) * n + x];
    ubo[l] = lo[(VIRAL_F_U0 + l) * n + x];
  }
  for (int l = 0; l < (CART ? 5 : 10); l++) {
    pio[l] = so[(VIRAL_F_PI0 + l) * n + x];
    pibo[l] = lo[(VIRAL_F_PI0 + l) * n + x];
  }
  for (int c = 0; c < 3; c++) vbo[c] = lo[(VIRAL_F_VX + c) * n + x];
  double Vb[SV];
  if (CART) {
    const double prm[10] = {T0[0], T0[1], T0[2], T0[3], pio[0], pio[1], pio[2], pio[3], pio[4],
                            so[VIRAL_F_BULK * n + x]};
    const double xr[4] = {so[VIRAL_F_EN * n + x], so[VIRAL_F_U1 * n + x], so[VIRAL_F_U2 * n + x],
                          so[VIRAL_F_U3 * n + x]};
    cart_update_vjp(p, prm, xr, T0b, pibo, lo[VIRAL_F_BULK * n + x], lo[VIRAL_F_EN * n + x], lo[VIRAL_F_P * n + x],
                    lo[VIRAL_F_TEMP * n + x], vbo, ubo, Vb);
  } else {
    update_vjp(p, a.tau_new, T0, pio, so[VIRAL_F_BULK * n + x], so[VIRAL_F_EN * n + x], lo[VIRAL_F_EN * n + x],
               lo[VIRAL_F_P * n + x], lo[VIRAL_F_TEMP * n + x], vbo, ubo, T0b, pibo, lo[VIRAL_F_BULK * n + x], Vb);
  }
  // RK combination (NEWEST10/time.h:2-98) in reverse
  for (int l = 0; l < SV; l++) {
    const long y = l * n + x;
    if (CART && l >= 9 && l < 14) {  // rows the Cartesian step does not evolve
      resb[y] = 0.0;
      continue;
    }
    double Lb;
    if (a.kind == 0) {
      const double Vt = Vb[l] + vb1[y];
      const double v0b = Vt + vb0[y];
      Lb = a.ts * Vt;
      const double sc = (l < 4 && !CART) ? a.tau_src : 1.0;
      li[(VIRAL_F_T00 + l) * n + x] += sc * v0b;  // Var row l is state field l
      vb0[y] = 0.0;
      vb1[y] = 0.0;
    } else if (a.kind == 1) {
      Lb = 0.5 * a.ts * Vb[l];
      vb0[y] += 0.5 * Vb[l];
      vb1[y] += 0.5 * Vb[l];
    } else if (a.kind == 2) {
      const double Vt = Vb[l] + vb2[y];
      Lb = 0.25 * a.ts * Vt;
      vb0[y] += 0.75 * Vt;
      vb1[y] += 0.25 * Vt;
      vb2[y] = 0.0;
    } else {
      Lb = (2.0 / 3.0) * a.ts * Vb[l];
      vb0[y] += (1.0 / 3.0) * Vb[l];
      vb2[y] += (2.0 / 3.0) * Vb[l];
    }
    resb[y] = Lb;
  }
}

// Sources of one interior cell in reverse (inputs from the stage's input
// state si; the source adjoint is resb, as L = Result + Source): adds to li,
// writes dubuf, the adjoint of the spatial velocity gradients du[l][1..3]
// (12 rows, taken back through their WENO derivatives by the flux tiles of
// each direction), and adds per-block sums of the adjoints of eta_coef and
// tpi_coef to partial.
// two waves per SIMD: without the cap the compiler takes 256 VGPRs plus 16
// AGPRs (one wave); with it, 88 bytes per lane spill and the kernel runs 9 %
// faster
// CART: the Cartesian sources (no geometric rows, the shear tensor built from
// the five evolved components; source_vjp_cart).  BULK: the bulk source
// (bulk_vjp), a separate instantiation so contexts without it keep the
// shear-only kernel's registers.
template <bool CART, bool BULK>
__global__ void __launch_bounds__(BLK) __attribute__((amdgpu_waves_per_eu(2))) k_adj_src(Geom g, Params p, StageArgs a, const double* __restrict__ si,
                                                 const double* __restrict__ resb, double* __restrict__ li,
                                                 double* __restrict__ dubuf, double* __restrict__ partial)
{
  long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  int i, j, k;
  double ceb = 0, ctb = 0;
  if (box_cell(interior_box(g), t, i, j, k)) {
    const long n = g.n, x = cidx(g, i, j, k);
    double Lb[SV];
    for (int l = 0; l < SV; l++) Lb[l] = resb[l * n + x];
    const double e = si[VIRAL_F_EN * n + x], P = si[VIRAL_F_P * n + x], PI = si[VIRAL_F_BULK * n + x];
    double u[4], pu[4], pi[10], du[4][4];
    for (int l = 0; l < 4; l++) {
      u[l] = si[(VIRAL_F_U0 + l) * n + x];
      pu[l] = si[(VIRAL_F_PU0 + l) * n + x];
    }
    for (int l = 0; l < 10; l++) pi[l] = si[(VIRAL_F_PI0 + l) * n + x];
    const double hx[3] = {p.dx, p.dy, p.deta};
    const long sd[3] = {g.sx, g.sy, 1};
    for (int l = 0; l < 4; l++) {
      const double* ul = si + (VIRAL_F_U0 + l) * n;
      du[l][0] = (u[l] - pu[l]) / a.ts_src;
      for (int d = 0; d < 3; d++) {
        const long s = sd[d];
        du[l][d + 1] = (d == 2 && p.lbi) ? 0.0
                                         : wder<true>(ul[x - 2 * s], ul[x - s], ul[x], ul[x + s], ul[x + 2 * s], hx[d],
                                                      1.0 / hx[d]);
      }
    }
    double eb = 0, Pb = 0, PIb = 0, ub[4] = {0, 0, 0, 0}, pib[10] = {0}, dub[4][4];
    for (int q = 0; q < 16; q++) dub[q / 4][q % 4] = 0;
    if (CART) {
      // TXYZ5's shear-source loop stops short of the top interior slices
      const bool live = !(p.quirks & VIRAL_Q_SHEAR_SLICES) || (int)((unsigned)x % (unsigned)g.gz) < g.kr - g.kl;
      source_vjp_cart(p, e, P, u, pi, live, du, Lb, eb, Pb, ub, pib, dub, ceb, ctb);
      if (BULK && live) bulk_vjp(p, 0, 1.0, e, P, u, PI, du, Lb[14], eb, Pb, PIb, ub, dub);
    } else {
      source_vjp(p, a.tau_src, e, P, PI, u, pi, du, Lb, eb, Pb, PIb, ub, pib, dub, ceb, ctb);
      if (BULK) bulk_vjp(p, 1, a.tau_src, e, P, u, PI, du, Lb[14], eb, Pb, PIb, ub, dub);
    }
    li[VIRAL_F_EN * n + x] += eb;
    li[VIRAL_F_P * n + x] += Pb;
    li[VIRAL_F_BULK * n + x] += PIb;
    for (int l = 0; l < (CART ? 5 : 10); l++) li[(VIRAL_F_PI0 + l) * n + x] += pib[l];
    for (int l = 0; l < 4; l++) {
      li[(VIRAL_F_U0 + l) * n + x] += ub[l] + dub[l][0] / a.ts_src;
      li[(VIRAL_F_PU0 + l) * n + x] -= dub[l][0] / a.ts_src;
      for (int d = 0; d < 3; d++) dubuf[(l * 3 + d) * n + x] = (d == 2 && p.lbi) ? 0.0 : dub[l][d + 1];
    }
  }
  // fixed-order block sums of the coefficient adjoints
  __shared__ double red[2][BLK / 64];
  for (int o = 32; o > 0; o >>= 1) {
    ceb += __shfl_down(ceb, o);
    ctb += __shfl_down(ctb, o);
  }
  if ((threadIdx.x & 63) == 0) {
    red[0][threadIdx.x >> 6] = ceb;
    red[1][threadIdx.x >> 6] = ctb;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double s0 = 0, s1 = 0;
    for (int w = 0; w < BLK / 64; w++) {
      s0 += red[0][w];
      s1 += red[1][w];
    }
    partial[2 * blockIdx.x] += s0;
    partial[2 * blockIdx.x + 1] += s1;
  }
}

// KT flux of direction D in reverse on an LDS line tile, with no global
// intermediate.  A block takes TK lines along D and a window of W
// consecutive cells of each (slot wi holds cell c0 - 5 + wi), W and TK chosen
// per direction on the host so the segments of W - 10 own cells cover a line
// with little waste.  Per row, every
// window cell stages its centred flux and Var in LDS; every cell with a full
// stencil in the window (slots 2..W-3) forms both WENO reconstructions of F
// and of Var (L at its right face, R at its left face) from one set of
// smoothness indicators and turns the adjoints of its two faces' Rusanov
// fluxes into contributions to its five stencil cells (weno_lr_vjp); every
// own cell (slots 5..W-6: its five contributors, and the eigenvalue
// reconstructions their face speeds need, lie in the window) sums them in a
// fixed order.  The face speeds a = max(|A_L|, |A_R|) come from the same
// per-cell reconstructions of the eigenvalue, and their adjoint (collected
// per face over the rows) goes back through it at the end.  Own cells cover
// the whole line, ghosts included, and tiles own disjoint cells, so each
// cell's state adjoint is updated by one thread.  Slots map to cells with the
// contiguous (eta) index fastest.
constexpr int NTA = 256;  // threads per block; the window shape is chosen per direction on the host

//
// CART: the Cartesian flux (tau = 1 from the host, no eta pressure factor),
// rows 9..13 skipped, and the T rows' pi^{d+1, l} built per window cell from
// u and the five evolved components (cart_pi10); its adjoint is collected per
// own cell and taken back through cart_pi10_vjp at the end.
template <int D, bool CART>
__global__ void __launch_bounds__(NTA) __attribute__((amdgpu_waves_per_eu(3)))
k_adj_tile(Geom g, Params p, double tau, const double* __restrict__ si, const double* __restrict__ resb,
           const double* __restrict__ dubuf, double* __restrict__ li, int W, int TK, int nseg)
{
  // one LDS slot per thread: thread t holds window slot wi of tile line kk,
  // the window index fastest along eta and the line index fastest otherwise,
  // so neighbouring window slots are st1 threads apart
  __shared__ double sF[NTA];
  __shared__ double sV[NTA];
  __shared__ double xL[NTA];  // per cell: A_L, then the face-adjoint parts
  __shared__ double xR[NTA];  // per cell: A_R
  __shared__ double cF[NTA][5];
  __shared__ double cV[NTA][5];
  __shared__ double sR[NTA];  // the row's partial-result adjoint, zero outside the interior
  const int t = threadIdx.x;
  const int st1 = D == 2 ? 1 : TK;
  const bool act = t < TK * W;               // threads past the last line idle (barriers only)
  const int kk = D == 2 ? t / W : t % TK;    // line of the tile
  const int wi = D == 2 ? t % W : t / TK;    // window slot
  const int GD = D == 0 ? g.gx : D == 1 ? g.gy : g.gz;
  const int lo = D == 0 ? g.il : D == 1 ? g.jl : g.kl;
  const int hi = D == 0 ? g.ir : D == 1 ? g.jr : g.kr;
  const int own = W - 10;                    // own cells per line and segment
  const int seg = (int)(blockIdx.x % nseg);
  const long lgrp = blockIdx.x / nseg;
  const int c0 = seg * own;
  const int nb = D == 2 ? g.jr - g.jl : g.kr - g.kl;
  const int na = D == 0 ? g.jr - g.jl : g.ir - g.il;
  const long ln = lgrp * TK + kk;
  const bool line_ok = act && ln < (long)na * nb;
  unsigned lbu = 0;
  const int la = line_ok ? (int)udivmod((unsigned)ln, (unsigned)nb, lbu) : 0, lb = (int)lbu;
  auto at = [&](int cd) -> long {
    if (D == 0) return cidx(g, cd, g.jl + la, g.kl + lb);
    if (D == 1) return cidx(g, g.il + la, cd, g.kl + lb);
    return cidx(g, g.il + la, g.jl + lb, cd);
  };
  const long n = g.n;
  const bool q1 = (p.quirks & VIRAL_Q_GHOST_EIGEN) != 0;
  const int row = D + 1;
  const double h = D == 0 ? p.dx : D == 1 ? p.dy : p.deta, ih = 1.0 / h;

  // window cell of this thread and its common terms
  const int c = c0 - 5 + wi;
  const bool w_ok = line_ok && c >= 0 && c < GD;
  const long xw = w_ok ? at(c) : 0;
  double u0 = 1.0, ud = 0.0, ww = 0.0, pp = 0.0;
  double pT[4] = {0.0, 0.0, 0.0, 0.0};  // Cartesian pi^{d+1, m} of the window cell
  if (w_ok) {
    const double e = si[VIRAL_F_EN * n + xw], P = si[VIRAL_F_P * n + xw], PI = si[VIRAL_F_BULK * n + xw];
    u0 = si[VIRAL_F_U0 * n + xw];
    ud = si[(VIRAL_F_U1 + D) * n + xw];
    ww = e + P - PI;
    pp = (D == 2 && !CART) ? (P - PI) / (tau * tau) : (P - PI);
    if (CART) {
      const double uu[4] = {u0, si[VIRAL_F_U1 * n + xw], si[VIRAL_F_U2 * n + xw], si[VIRAL_F_U3 * n + xw]};
      double pt[10];
      cart_pi10(uu, si[VIRAL_F_PI0 * n + xw], si[(VIRAL_F_PI0 + 1) * n + xw], si[(VIRAL_F_PI0 + 2) * n + xw],
                si[(VIRAL_F_PI0 + 3) * n + xw], si[(VIRAL_F_PI0 + 4) * n + xw], pt);
      for (int m = 0; m < 4; m++) pT[m] = pt[pidx(row, m)];
    }
  }
  const bool rc = act && wi >= 2 && wi < W - 2;  // full stencil in the window
  // faces of this cell: right c+1/2 (face index c), left c-1/2 (face index c-1)
  const bool fR = line_ok && rc && c >= lo - 1 && c <= hi - 1;
  const bool fL = line_ok && rc && c - 1 >= lo - 1 && c - 1 <= hi - 1;
  // the eigenvalue's reconstructions: A_L(c), A_R(c)
  sF[t] = w_ok && ((c >= lo && c < hi) || !q1) ? ud / u0 : 0.0;
  __syncthreads();
  double As[5], AL = 0.0, AR = 0.0;
  if (rc) {
    for (int o = 0; o < 5; o++) As[o] = sF[t + (o - 2) * st1];
    AL = weno_L(As[0], As[1], As[2], As[3], As[4]);
    AR = weno_R(As[0], As[1], As[2], As[3], As[4]);
  }
  xL[t] = AL;
  xR[t] = AR;
  __syncthreads();
  // face speeds and branches: face c+1/2 takes |A_L(c)| and |A_R(c+1)|, zero
  // under q1 before the first interior cell
  double aFR = 0.0, aFL = 0.0;
  bool brR = true, brL = true;  // true: the A_L side is the maximum
  if (fR && !(q1 && c < lo)) {
    const double ml = fabs(AL), mr = fabs(xR[t + st1]);
    aFR = fmax(ml, mr);
    brR = ml >= mr;
  }
  if (fL && !(q1 && c - 1 < lo)) {
    const double ml = fabs(xL[t - st1]), mr = fabs(AR);
    aFL = fmax(ml, mr);
    brL = ml >= mr;
  }
  // own cell of this thread
  const bool o_ok = line_ok && wi >= 5 && wi < W - 5 && c < GD;
  // The own cell's centred flux and Var in reverse (flux_cell_vjp, one row at
  // a time): each row's field adjoints go straight to li (only this thread
  // touches the cell); the adjoints of w = e + P - PI, the pressure term, u^d
  // and u^mu accumulate here.
  double wb = 0.0, ppb = 0.0, udb = 0.0, ub0 = 0.0, ub1 = 0.0, ub2 = 0.0, ub3 = 0.0;
  double pTb[4] = {0.0, 0.0, 0.0, 0.0};  // CART: adjoint of pi^{d+1, m}
  const double iu0 = frcp(u0);
  double abR = 0.0, abL = 0.0;  // this cell's parts of the speed adjoints of its two faces
  // a row's global loads, issued one row ahead so they overlap the previous
  // row's reconstructions: state field l, pi^{d+1,l} and u^l (l < 4), and the
  // partial-result adjoint of row l (zero outside the interior)
  double nA = 0.0, nB = 0.0, nM = 0.0, nR = 0.0;
  auto row_load = [&](int l) {
    if (!w_ok) return;
    nA = si[(long)l * n + xw];
    if (l < 4) {
      nB = CART ? (l == 0 ? pT[0] : l == 1 ? pT[1] : l == 2 ? pT[2] : pT[3])
                : si[(VIRAL_F_PI0 + pidx(row, l)) * n + xw];
      nM = si[(VIRAL_F_U0 + l) * n + xw];
    }
    nR = (c >= lo && c < hi) ? resb[l * n + xw] : 0.0;
  };
  row_load(0);
  for (int l = 0; l < SV; l++) {
    if (CART && l >= 9 && l < 14) {  // block-uniform: rows the Cartesian step does not evolve
      if (l == 13) row_load(14);
      continue;
    }
    // centred flux and Var of row l (cell_FVA)
    const double ldA = nA, ldB = nB, ldR = nR;
    double Fw = 0.0, Vw = 0.0, um = 0.0, fv = 0.0;
    if (w_ok) {
      if (l < 4) {
        um = nM;
        const double v = l == row ? ldB + pp + ud * um * ww : ldB + ud * um * ww;
        Fw = tau * v;
        Vw = tau * ldA;
      } else {
        fv = ldA;
        Fw = (fv * ud) / u0;
        Vw = fv;
      }
    }
    // (the previous row's sF, sV, sR and cF reads all precede the barrier below)
    sF[t] = Fw;
    sV[t] = Vw;
    sR[t] = ldR;
    if (l + 1 < SV && !(CART && l == 8)) row_load(l + 1);
    __syncthreads();
    {
      double gF[5] = {0, 0, 0, 0, 0}, gV[5] = {0, 0, 0, 0, 0};
      // adjoints of the face fluxes: partial result of cell x is
      // -(f(x+1/2) - f(x-1/2))/h, interior cells only
      const double fbR = fR ? sR[t + st1] * ih - sR[t] * ih : 0.0;
      const double fbL = fL ? sR[t] * ih - sR[t - st1] * ih : 0.0;
      if (fbR != 0.0 || fbL != 0.0) {
        double Vs[5], Fs[5], VL, VR, FL, FR;
        for (int o = 0; o < 5; o++) {
          Vs[o] = sV[t + (o - 2) * st1];
          Fs[o] = sF[t + (o - 2) * st1];
        }
        // f = (F_L + F_R)/2 - a (V_R - V_L)/2
        weno_lr_vjp(Vs, 0.5 * aFR * fbR, -0.5 * aFL * fbL, VL, VR, gV);
        weno_lr_vjp(Fs, 0.5 * fbR, 0.5 * fbL, FL, FR, gF);
        abR += 0.5 * fbR * VL;
        abL += -0.5 * fbL * VR;
      }
      for (int o = 0; o < 5; o++) {
        cF[t][o] = gF[o];
        cV[t][o] = gV[o];
      }
    }
    // the own cell's adjoints this row adds to (li of pi^{d+1,l} and T^{tau l}
    // for l < 4, of field l otherwise), loaded before the barrier so that the
    // load waits behind it (only this thread writes them)
    double lL1 = 0.0, lL2 = 0.0;
    if (o_ok) {
      if (l < 4) {
        if (!CART) lL1 = li[(VIRAL_F_PI0 + pidx(row, l)) * n + xw];
        lL2 = li[(VIRAL_F_T00 + l) * n + xw];
      } else {
        lL1 = li[(long)l * n + xw];
      }
    }
    __syncthreads();
    if (o_ok) {
      double sf = 0.0, sv = 0.0;
      for (int o = 0; o < 5; o++) {  // stencil cells c+2 .. c-2 of this cell's contributors
        sf += cF[t + (2 - o) * st1][o];
        sv += cV[t + (2 - o) * st1][o];
      }
      if (l < 4) {
        // F_l = tau (pi^{d+1,l} + [l = d+1] pp + u^{d+1} u^l w), V_l = tau T^{tau l}
        const double vb = tau * sf;
        if (CART) {
          if (l == 0) pTb[0] += vb;
          else if (l == 1) pTb[1] += vb;
          else if (l == 2) pTb[2] += vb;
          else pTb[3] += vb;
        } else {
          li[(VIRAL_F_PI0 + pidx(row, l)) * n + xw] = lL1 + vb;
        }
        li[(VIRAL_F_T00 + l) * n + xw] = lL2 + tau * sv;
        if (l == row) ppb += vb;
        udb += vb * um * ww;
        const double ubl = vb * ud * ww;
        if (l == 0) ub0 += ubl;
        else if (l == 1) ub1 += ubl;
        else if (l == 2) ub2 += ubl;
        else ub3 += ubl;
        wb += vb * ud * um;
      } else {
        // F_l = f u^{d+1}/u^0, V_l = f for the state field f = pi[l-4] or PI
        li[(long)l * n + xw] = lL1 + (sf * ud * iu0 + sv);
        udb += sf * fv * iu0;
        ub0 -= sf * fv * ud * iu0 * iu0;
      }
    }
  }
  // the sources' velocity gradients du[l][D] in reverse: their WENO
  // derivative along D (NEWEST10/Source.h:3-101, d u/dx = (L - R)/h) at every
  // interior centre, the adjoint dubuf of k_adj_src turned into contributions
  // to the five stencil cells, summed per own cell as above
  const bool cin = rc && line_ok && c >= lo && c < hi;
  // u^l and the gradient adjoint of row l are loaded one row ahead
  double nU = w_ok ? si[VIRAL_F_U0 * n + xw] : 0.0, nW = cin ? dubuf[(long)D * n + xw] : 0.0;
  for (int l = 0; l < 4; l++) {
    const double cW = nW;
    sF[t] = nU;
    if (l < 3) {
      nU = w_ok ? si[(VIRAL_F_U0 + l + 1) * n + xw] : 0.0;
      nW = cin ? dubuf[(long)((l + 1) * 3 + D) * n + xw] : 0.0;
    }
    __syncthreads();
    double gq[5] = {0, 0, 0, 0, 0};
    if (cin) {
      const double w = cW * ih;
      if (w != 0.0) {
        double q[5], dl, dr;
        for (int o = 0; o < 5; o++) q[o] = sF[t + (o - 2) * st1];
        weno_lr_vjp(q, w, -w, dl, dr, gq);
      }
    }
    for (int o = 0; o < 5; o++) cF[t][o] = gq[o];
    __syncthreads();
    if (o_ok) {
      double acc = 0.0;
      for (int o = 0; o < 5; o++) acc += cF[t + (2 - o) * st1][o];
      li[(VIRAL_F_U0 + l) * n + xw] += acc;
    }
  }
  // the own cell's adjoints of En, P, PI and u, which the end adds to,
  // loaded ahead of the face-speed phase's barriers
  double eL[7] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  if (o_ok) {
    eL[0] = li[VIRAL_F_EN * n + xw];
    eL[1] = li[VIRAL_F_P * n + xw];
    eL[2] = li[VIRAL_F_BULK * n + xw];
    for (int m = 0; m < 4; m++) eL[3 + m] = li[(VIRAL_F_U0 + m) * n + xw];
  }
  // the face speeds in reverse: a = max(|A_L|, |A_R|) on the branch taken;
  // abar(c+1/2) = abR(c) + abL(c+1)
  __syncthreads();
  xL[t] = abL;
  xR[t] = abR;
  __syncthreads();
  {
    double gA[5] = {0, 0, 0, 0, 0};
    double sAL = 0.0, sAR = 0.0;
    if (fR && !(q1 && c < lo) && brR) sAL = (abR + xL[t + st1]) * (AL > 0 ? 1.0 : AL < 0 ? -1.0 : 0.0);
    if (fL && !(q1 && c - 1 < lo) && !brL) sAR = (xR[t - st1] + abL) * (AR > 0 ? 1.0 : AR < 0 ? -1.0 : 0.0);
    if (sAL != 0.0 || sAR != 0.0) {
      double dl, dr;
      weno_lr_vjp(As, sAL, sAR, dl, dr, gA);
    }
    for (int o = 0; o < 5; o++) cF[t][o] = gA[o];
  }
  __syncthreads();
  if (!o_ok) return;
  double Ab = 0.0;
  for (int o = 0; o < 5; o++) Ab += cF[t + (2 - o) * st1][o];
  if (!((c >= lo && c < hi) || !q1)) Ab = 0.0;
  // A = u^{d+1}/u^0
  udb += Ab * iu0;
  ub0 -= Ab * ud * iu0 * iu0;
  if (row == 1) ub1 += udb;
  else if (row == 2) ub2 += udb;
  else ub3 += udb;
  if (CART) {
    // pi^{d+1, m} = cart_pi10(u, p1..p5) in reverse
    const double uu[4] = {u0, si[VIRAL_F_U1 * n + xw], si[VIRAL_F_U2 * n + xw], si[VIRAL_F_U3 * n + xw]};
    const double pe[5] = {si[VIRAL_F_PI0 * n + xw], si[(VIRAL_F_PI0 + 1) * n + xw], si[(VIRAL_F_PI0 + 2) * n + xw],
                          si[(VIRAL_F_PI0 + 3) * n + xw], si[(VIRAL_F_PI0 + 4) * n + xw]};
    double ob[10] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    for (int m = 0; m < 4; m++) ob[pidx(row, m)] = pTb[m];
    double ubc[4] = {0, 0, 0, 0}, peb[5] = {0, 0, 0, 0, 0};
    cart_pi10_vjp(uu, pe, ob, ubc, peb);
    ub0 += ubc[0];
    ub1 += ubc[1];
    ub2 += ubc[2];
    ub3 += ubc[3];
    for (int m = 0; m < 5; m++) li[(VIRAL_F_PI0 + m) * n + xw] += peb[m];
  }
  const double k = (D == 2 && !CART) ? 1.0 / (tau * tau) : 1.0;
  li[VIRAL_F_EN * n + xw] = eL[0] + wb;
  li[VIRAL_F_P * n + xw] = eL[1] + (ppb * k + wb);
  li[VIRAL_F_BULK * n + xw] = eL[2] + (-(ppb * k) - wb);
  li[VIRAL_F_U0 * n + xw] = eL[3] + ub0;
  li[(VIRAL_F_U0 + 1) * n + xw] = eL[4] + ub1;
  li[(VIRAL_F_U0 + 2) * n + xw] = eL[5] + ub2;
  li[(VIRAL_F_U0 + 3) * n + xw] = eL[6] + ub3;
}

#pragma clang fp contract(off)

}  // namespace

// ---------------------------------------------------------------- host side

// kernel phases of one RK stage; each phase is one launch per event, or one
// launch over every event of the ensemble
enum Phase { PH_FLUX_X = 0, PH_FLUX_Y, PH_FLUX_Z, PH_CELLS, PH_K0, PH_GHOST, PH_COUNT };

struct Event {
  viral_desc d;
  Geom g;
  Params p;
  double* st[2] = {nullptr, nullptr};  // ping-pong state, NF*n each
  double* var0 = nullptr;
  double* var1 = nullptr;
  double* var2 = nullptr;                         // RK3 only
  double* flux[3] = {nullptr, nullptr, nullptr};  // face fluxes of x, y, eta: [15][n] each
  double* src = nullptr;                          // sources of the stage: [15][n]
  int32_t* iters = nullptr;  // 3*n
  uint8_t* flags = nullptr;
  unsigned long long bad_host = ~0ull;
  int tile_tp[3] = {32, 32, 64};  // LDS tile of the fast flux kernel: cells per line
  int tile_tk[3] = {16, 16, 4};   //   and lines per tile, per direction
  double* tab_v = nullptr;        // tabulated s95p EOS (fast mode)
  unsigned char* tab_closed = nullptr;
};

// K_ADJ: the device part of viral_adjoint (recompute and reverse kernels,
// between the seed's upload and the gradient's download)
enum KernelId { K_FLUX_X = 0, K_FLUX_Y, K_FLUX_Z, K_SOURCE, K_UPDATE, K_K0, K_GHOST, K_ADJ, K_COUNT };
static const char* const kKernelNames[K_COUNT] = {"flux_x", "flux_y", "flux_eta", "source", "update",
                                                  "k0_broadcast", "ghost_fill", "adjoint_device"};

struct viral_ctx {
  std::vector<Event> ev;
  hipStream_t stream = nullptr;
  hipStream_t stream2 = nullptr;  // the adjoint's recompute stream (default priority, made on first use)
  int dev = 0;    // the device current at viral_create (the context's device)
  std::string err;
  unsigned long long* dtmp = nullptr;
  int cur = 0;    // which state copy is current (all events step together)
  int ncu = 256;  // compute units of the device
  int rk = 2, scheme = VIRAL_KT, fast = 0, cart = 0;
  viral_launch cfg{};
  // tile table: device event rows and, per phase, the first tile of every event
  EvDev* d_tab = nullptr;
  long* d_start = nullptr;  // PH_COUNT x (nev+1)
  std::vector<long> cnt[PH_COUNT];
  long total[PH_COUNT] = {};
  size_t shm[3] = {};  // LDS bytes of the batched flux launches
  bool dirty = true;   // table or launch configuration changed since the last upload / capture
  // step clock and failure words (device + pinned host mirrors)
  Clock* d_clk = nullptr;
  Clock* h_clk = nullptr;
  unsigned long long* d_bad = nullptr;
  unsigned long long* h_bad = nullptr;
  unsigned long long* d_bmax = nullptr;  // per event: BMax of the FIX build (VIRAL_Q_FIX_VE), as double bits
  std::vector<char> bmax_ok;             // per event: d_bmax belongs to the current state
  double last_tau = 1.0, last_ts = 0.01;  // clock the autotuner times with
  // captured RK step per starting parity
  hipGraph_t graph[2] = {nullptr, nullptr};
  hipGraphExec_t gexec[2] = {nullptr, nullptr};
  // per-kernel HIP-event timing (viral_timers)
  bool timing = false;
  std::vector<hipEvent_t> tev[K_COUNT];  // start/stop pairs
  std::vector<hipEvent_t> pool;
  // per-step diagnostics: block partials; output records: device staging
  DiagPart* d_diag = nullptr;
  double* d_out = nullptr;
  size_t out_cap = 0;
  // adjoint recording: the (tau, dtau) of every step since viral_adjoint_begin
  // and device copies of the state every `every` steps
  // (every == 0: no checkpoints past the first; each stage's output state is
  // kept instead, in one slab per viral_steps call; every > 0: the stage
  // states of the current segment, the steps since the last checkpoint, are
  // kept in a ring of every * rk state copies)
  struct {
    bool on = false;
    int every = 0;
    std::vector<double> taus, tss;
    std::vector<double*> ckpt;
    std::vector<double*> slabs, stages;
    double* ring = nullptr;
  } adj;
};

static hipEvent_t pool_event(viral_ctx* c)
{
  if (!c->pool.empty()) {
    hipEvent_t e = c->pool.back();
    c->pool.pop_back();
    return e;
  }
  hipEvent_t e;
  hipEventCreate(&e);
  return e;
}

// brackets one kernel launch with events on the context stream when timing
struct KTimer {
  viral_ctx* c;
  int id;
  KTimer(viral_ctx* c_, int id_) : c(c_), id(id_)
  {
    if (c->timing) {
      hipEvent_t e = pool_event(c);
      hipEventRecord(e, c->stream);
      c->tev[id].push_back(e);
    }
  }
  ~KTimer()
  {
    if (c->timing) {
      hipEvent_t e = pool_event(c);
      hipEventRecord(e, c->stream);
      c->tev[id].push_back(e);
    }
  }
};

static int fail(viral_ctx* c, int code, const std::string& msg)
{
  if (c) c->err = msg;
  return code;
}

// Per device: the address of `scratch` on that device, resolved once with
// the device current (the first viral_create on it), and whether that
// succeeded.
constexpr int MAX_DEVICES = 64;
static DiagPart* scratch_ptr[MAX_DEVICES] = {};
static std::mutex scratch_mu;

static hipError_t resolve_scratch(int dev)
{
  if (dev < 0 || dev >= MAX_DEVICES) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lock(scratch_mu);
  if (scratch_ptr[dev]) return hipSuccess;
  DiagPart* ptr = nullptr;
  const hipError_t e = hipGetSymbolAddress((void**)&ptr, HIP_SYMBOL(scratch));
  if (e != hipSuccess) return e;
  if (!ptr) return hipErrorInvalidSymbol;
  scratch_ptr[dev] = ptr;
  return hipSuccess;
}

#define HIPCHK(call)                                                                          \
  do {                                                                                        \
    hipError_t _e = (call);                                                                   \
    if (_e != hipSuccess) return fail(ctx, VIRAL_ERR_HIP, std::string(#call ": ") + hipGetErrorString(_e)); \
  } while (0)

static unsigned nblocks(long cells) { return (unsigned)((cells + BLK - 1) / BLK); }
static long box_count(const Box& b) { return (long)b.ni * b.nj * b.nk; }

static int max_slots(int nt)
{
  switch (nt) {
    case 64: return TileSlots<64>::MAXS;
    case 128: return TileSlots<128>::MAXS;
    case 256: return TileSlots<256>::MAXS;
    case 512: return TileSlots<512>::MAXS;
    default: return 0;
  }
}

// a tile needs at least 3 faces per line; cells past the end of a line are masked
static bool tile_ok(int TP, int TK, int nt) { return TP >= 8 && TK >= 1 && (long)TP * TK <= max_slots(nt); }

static void fill_geom(const viral_desc* d, Geom& g)
{
  viral_ghost_dims(d, &g.gx, &g.gy, &g.gz);
  g.il = VIRAL_BORDER; g.ir = VIRAL_BORDER + d->nx;
  g.jl = VIRAL_BORDER; g.jr = VIRAL_BORDER + d->ny;
  if (d->lbi) { g.kl = 0; g.kr = 1; g.k0 = 0; }
  else { g.kl = VIRAL_BORDER; g.kr = VIRAL_BORDER + d->nz; g.k0 = d->k0 >= 0 ? d->k0 : (g.gz - 1) / 2; }
  g.n = (long)g.gx * g.gy * g.gz;
  g.sx = (long)g.gy * g.gz;
  g.sy = g.gz;
}

// tiles of one event in one phase (a tile is one block's worth of work)
static long tile_count(const viral_desc* d, int ph, int tp, int tk)
{
  Geom g;
  fill_geom(d, g);
  const long ni = box_count(interior_box(g));
  switch (ph) {
    case PH_FLUX_X:
    case PH_FLUX_Y:
    case PH_FLUX_Z: {
      const int dd = ph - PH_FLUX_X;
      if (dd == 2 && d->lbi) return 0;
      if (d->scheme == VIRAL_KT && !d->fast_math) return nblocks(box_count(face_box(g, dd)));
      // fast KT and SHASTA: LDS line tiles
      const int na = dd == 0 ? g.jr - g.jl : g.ir - g.il;
      const int nb = dd == 2 ? g.jr - g.jl : g.kr - g.kl;
      const int lo = dd == 0 ? g.il : dd == 1 ? g.jl : g.kl, hi = dd == 0 ? g.ir : dd == 1 ? g.jr : g.kr;
      const long nlines = (long)na * nb;
      const long cpt = tp - 6;  // cells per tile (tp - 5 faces)
      const long nseg = (hi - lo + cpt - 1) / cpt;
      return ((nlines + tk - 1) / tk) * nseg;
    }
    case PH_CELLS: return nblocks(ni);
    case PH_K0:
      return ((d->quirks & VIRAL_Q_K0_BROADCAST) && !d->lbi && d->coords == VIRAL_MILNE) ? nblocks(ni) : 0;
    case PH_GHOST: return 28L * (long)nblocks(ghost_count(g, d->lbi));
    default: return 0;
  }
}

static void drop_graphs(viral_ctx* c)
{
  for (int q = 0; q < 2; q++) {
    if (c->gexec[q]) hipGraphExecDestroy(c->gexec[q]);
    if (c->graph[q]) hipGraphDestroy(c->graph[q]);
    c->gexec[q] = nullptr;
    c->graph[q] = nullptr;
  }
}

// rebuild the tile table from the events' tile shapes and upload it with the
// event rows; drops captured graphs (their launches baked the old table in)
static int refresh_table(viral_ctx* ctx)
{
  if (!ctx->dirty) return VIRAL_OK;
  drop_graphs(ctx);
  const int nev = (int)ctx->ev.size();
  std::vector<EvDev> rows(nev);
  std::vector<long> start((size_t)PH_COUNT * (nev + 1));
  for (int ph = 0; ph < PH_COUNT; ph++) {
    ctx->cnt[ph].assign(nev, 0);
    long acc = 0;
    for (int e = 0; e < nev; e++) {
      const Event& E = ctx->ev[e];
      const int dd = ph <= PH_FLUX_Z ? ph : 0;
      const long n = tile_count(&E.d, ph, E.tile_tp[dd], E.tile_tk[dd]);
      start[(size_t)ph * (nev + 1) + e] = acc;
      ctx->cnt[ph][e] = n;
      acc += n;
    }
    start[(size_t)ph * (nev + 1) + nev] = acc;
    ctx->total[ph] = acc;
  }
  for (int d = 0; d < 3; d++) {
    size_t m = 0;
    for (const Event& E : ctx->ev) m = std::max(m, kTileArrays * (size_t)E.tile_tp[d] * E.tile_tk[d] * sizeof(double));
    ctx->shm[d] = m;
  }
  for (int e = 0; e < nev; e++) {
    const Event& E = ctx->ev[e];
    EvDev& r = rows[e];
    r.g = E.g;
    r.p = E.p;
    r.st[0] = E.st[0];
    r.st[1] = E.st[1];
    r.var0 = E.var0;
    r.var1 = E.var1;
    r.var2 = E.var2;
    for (int q = 0; q < 3; q++) r.flux[q] = E.flux[q];
    r.src = E.src;
    r.iters = E.iters;
    r.flags = E.flags;
    r.bad = ctx->d_bad + e;
    for (int q = 0; q < 3; q++) {
      r.tp[q] = E.tile_tp[q];
      r.tk[q] = E.tile_tk[q];
    }
  }
  HIPCHK(hipStreamSynchronize(ctx->stream));
  HIPCHK(hipMemcpy(ctx->d_tab, rows.data(), nev * sizeof(EvDev), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(ctx->d_start, start.data(), start.size() * sizeof(long), hipMemcpyHostToDevice));
  ctx->dirty = false;
  return VIRAL_OK;
}

static unsigned grid_for(const viral_ctx* c, long tiles)
{
  long g = tiles;
  if (c->cfg.blocks_per_cu > 0) g = std::min(g, (long)c->ncu * c->cfg.blocks_per_cu);
  return (unsigned)g;
}

// run f(sched, grid, event or -1) once over all events (batched) or once per event
template <class F>
static void for_phase(viral_ctx* c, int ph, F&& f)
{
  const int nev = (int)c->ev.size();
  if (c->cfg.batched) {
    if (c->total[ph] == 0) return;
    Sched s{c->d_start + (size_t)ph * (nev + 1), c->total[ph], 0, nev, c->cfg.xcd};
    f(s, grid_for(c, c->total[ph]), -1);
  } else {
    for (int e = 0; e < nev; e++) {
      const long n = c->cnt[ph][e];
      if (n == 0) continue;
      Sched s{nullptr, n, e, 1, c->cfg.xcd};
      f(s, grid_for(c, n), e);
    }
  }
}

template <class F>
static void with_bool(bool b, F&& f)
{
  if (b) f(std::true_type{});
  else f(std::false_type{});
}

template <bool P, bool CART>
static void launch_tile(viral_ctx* c, int d, int nt, const Sched& s, unsigned grid, size_t shm, int stage, int cur)
{
#define TILE_CASE(D, NT)                                                                                          \
  hipLaunchKernelGGL((k_flux_tile<D, NT, P, CART>), dim3(grid), dim3(NT), shm, c->stream, c->d_tab, s, c->d_clk, \
                     stage, cur)
#define TILE_NT(D)                       \
  switch (nt) {                          \
    case 64: TILE_CASE(D, 64); break;    \
    case 128: TILE_CASE(D, 128); break;  \
    case 512: TILE_CASE(D, 512); break;  \
    default: TILE_CASE(D, 256); break;   \
  }
  if (d == 0) { TILE_NT(0) }
  else if (d == 1) { TILE_NT(1) }
  else { TILE_NT(2) }
#undef TILE_NT
#undef TILE_CASE
}

template <bool P>
static void launch_shasta_tile(viral_ctx* c, int d, int nt, const Sched& s, unsigned grid, size_t shm, int stage,
                               int cur)
{
#define STILE_CASE(D, NT)                                                                                   \
  do {                                                                                                      \
    if (c->fast)                                                                                            \
      hipLaunchKernelGGL((k_shasta_tile<D, NT, P, true>), dim3(grid), dim3(NT), shm, c->stream, c->d_tab, s, \
                         c->d_clk, stage, cur);                                                             \
    else                                                                                                    \
      hipLaunchKernelGGL((k_shasta_tile<D, NT, P, false>), dim3(grid), dim3(NT), shm, c->stream, c->d_tab,   \
                         s, c->d_clk, stage, cur);                                                          \
  } while (0)
#define STILE_NT(D)                       \
  switch (nt) {                           \
    case 64: STILE_CASE(D, 64); break;    \
    case 128: STILE_CASE(D, 128); break;  \
    case 512: STILE_CASE(D, 512); break;  \
    default: STILE_CASE(D, 256); break;   \
  }
  if (d == 0) { STILE_NT(0) }
  else if (d == 1) { STILE_NT(1) }
  else { STILE_NT(2) }
#undef STILE_NT
#undef STILE_CASE
}

// the tiled flux kernel of direction d (fast KT or SHASTA) for schedule s
template <bool P, bool CART>
static void launch_dir_tiles(viral_ctx* c, int d, int nt, const Sched& s, unsigned grid, size_t shm, int stage,
                             int cur)
{
  if (c->scheme == VIRAL_SHASTA) launch_shasta_tile<P>(c, d, nt, s, grid, shm, stage, cur);
  else launch_tile<P, CART>(c, d, nt, s, grid, shm, stage, cur);
}

static void launch_flux(viral_ctx* c, int d, int stage, int cur)
{
  KTimer t(c, K_FLUX_X + d);
  with_bool(c->cfg.blocks_per_cu > 0, [&](auto PT) {
    constexpr bool P = decltype(PT)::value;
    with_bool(c->cart, [&](auto CT) {
      constexpr bool CART = decltype(CT)::value;
      for_phase(c, PH_FLUX_X + d, [&](const Sched& s, unsigned grid, int e) {
        if (c->fast || c->scheme == VIRAL_SHASTA) {
          const size_t shm =
              e < 0 ? c->shm[d] : kTileArrays * (size_t)c->ev[e].tile_tp[d] * c->ev[e].tile_tk[d] * sizeof(double);
          launch_dir_tiles<P, CART>(c, d, c->cfg.flux_nt[d], s, grid, shm, stage, cur);
        } else {
          hipLaunchKernelGGL((k_flux_dir<false, P, CART>), dim3(grid), dim3(BLK), 0, c->stream, c->d_tab, s,
                             c->d_clk, stage, cur, d);
        }
      });
    });
  });
}

static void launch_cells(viral_ctx* c, int stage, int cur)
{
  with_bool(c->cfg.blocks_per_cu > 0, [&](auto PT) {
    constexpr bool P = decltype(PT)::value;
    with_bool(c->fast, [&](auto FT) {
      constexpr bool FAST = decltype(FT)::value;
      with_bool(c->cart, [&](auto CT) {
        constexpr bool CART = decltype(CT)::value;
        if (!c->cfg.fusion) {
          KTimer t(c, K_SOURCE);
          for_phase(c, PH_CELLS, [&](const Sched& s, unsigned grid, int) {
            hipLaunchKernelGGL((k_source<FAST, P, CART>), dim3(grid), dim3(BLK), 0, c->stream, c->d_tab, s, c->d_clk,
                               stage, cur);
          });
        }
        KTimer t(c, K_UPDATE);
        for_phase(c, PH_CELLS, [&](const Sched& s, unsigned grid, int) {
          if (c->cfg.fusion)
            hipLaunchKernelGGL((k_update<FAST, true, P, CART>), dim3(grid), dim3(BLK), 0, c->stream, c->d_tab, s,
                               c->d_clk, stage, cur);
          else
            hipLaunchKernelGGL((k_update<FAST, false, P, CART>), dim3(grid), dim3(BLK), 0, c->stream, c->d_tab, s,
                               c->d_clk, stage, cur);
        });
      });
    });
  });
}

// one RK stage of every event: fluxes and sources from copy `cur`, the new
// state into copy cur^1, then the k0 broadcast and the ghost fill of it
static void launch_stage(viral_ctx* c, int stage, int cur)
{
  for (int d = 0; d < 3; d++) launch_flux(c, d, stage, cur);
  launch_cells(c, stage, cur);
  with_bool(c->cfg.blocks_per_cu > 0, [&](auto PT) {
    constexpr bool P = decltype(PT)::value;
    {
      KTimer t(c, K_K0);
      for_phase(c, PH_K0, [&](const Sched& s, unsigned grid, int) {
        hipLaunchKernelGGL((k_k0_broadcast<P>), dim3(grid), dim3(BLK), 0, c->stream, c->d_tab, s, c->d_clk, stage,
                           cur ^ 1);
      });
    }
    KTimer t(c, K_GHOST);
    for_phase(c, PH_GHOST, [&](const Sched& s, unsigned grid, int) {
      hipLaunchKernelGGL((k_ghost_fill<P>), dim3(grid), dim3(BLK), 0, c->stream, c->d_tab, s, cur ^ 1);
    });
  });
}

// FIX build: BMax of event e's state copy `cur` at the clock's tau
static void launch_bmax(viral_ctx* c, int e, int cur)
{
  const Event& E = c->ev[e];
  hipLaunchKernelGGL(k_bmax_clear, dim3(1), dim3(64), 0, c->stream, c->d_bmax + e);
  hipLaunchKernelGGL(k_bmax, dim3(nblocks(box_count(interior_box(E.g)))), dim3(BLK), 0, c->stream, E.g, E.st[cur],
                     c->d_clk, c->d_bmax + e);
}

static bool fix_ve(const Event& E) { return (E.p.quirks & VIRAL_Q_FIX_VE) && E.p.milne; }

// every stage of one step starting from copy cur0, then the FIX build's end
// of step and the clock advance
static void launch_step(viral_ctx* c, int cur0)
{
  int cur = cur0;
  for (int stage = 0; stage < c->rk; stage++) {
    launch_stage(c, stage, cur);
    cur ^= 1;
  }
  for (int e = 0; e < (int)c->ev.size(); e++) {
    const Event& E = c->ev[e];
    if (!fix_ve(E)) continue;
    hipLaunchKernelGGL(k_fix_ve, dim3(nblocks(box_count(interior_box(E.g)))), dim3(BLK), 0, c->stream, E.g,
                       E.st[cur], c->d_bmax + e);
    launch_bmax(c, e, cur);
  }
  hipLaunchKernelGGL(k_clock_advance, dim3(1), dim3(64), 0, c->stream, c->d_clk);
}

static int capture_step(viral_ctx* ctx, int cur0)
{
  if (ctx->gexec[cur0]) return VIRAL_OK;
  HIPCHK(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
  launch_step(ctx, cur0);
  hipGraph_t g = nullptr;
  HIPCHK(hipStreamEndCapture(ctx->stream, &g));
  ctx->graph[cur0] = g;
  HIPCHK(hipGraphInstantiate(&ctx->gexec[cur0], g, nullptr, nullptr, 0));
  return VIRAL_OK;
}

static EosTab host_tab()
{
  static std::vector<double> v;
  static std::vector<unsigned char> c;
  static bool built = false;
  if (!built) {
    build_eos_table(v, c);
    built = true;
  }
  return EosTab{v.data(), c.data(), kTabEmin, kTabNoct, kTabK};
}

// elapsed ms of reps runs of fn on the context stream (after one warm run)
template <class F>
static int time_ms(viral_ctx* ctx, int reps, float* ms, F&& fn)
{
  hipEvent_t t0, t1;
  HIPCHK(hipEventCreate(&t0));
  HIPCHK(hipEventCreate(&t1));
  fn();
  HIPCHK(hipEventRecord(t0, ctx->stream));
  for (int r = 0; r < reps; r++) fn();
  HIPCHK(hipEventRecord(t1, ctx->stream));
  HIPCHK(hipEventSynchronize(t1));
  HIPCHK(hipEventElapsedTime(ms, t0, t1));
  hipEventDestroy(t0);
  hipEventDestroy(t1);
  HIPCHK(hipGetLastError());
  return VIRAL_OK;
}

extern "C" {

void viral_desc_default(viral_desc* d, int nx, int ny, int nz)
{
  memset(d, 0, sizeof(*d));
  d->nx = nx;
  d->ny = ny;
  d->nz = nz;
  d->lbi = 0;
  d->coords = VIRAL_MILNE;
  d->scheme = VIRAL_KT;
  d->rk = 2;
  d->eos = VIRAL_EOS_CONFORMAL;
  d->dx = d->dy = d->deta = 0.1;
  d->dof = 42.25;
  d->eta_law = VIRAL_ETA_TIMES_S;
  d->eta_coef = 0.2;
  d->tpi_law = VIRAL_TPI_OVER_EP;
  d->tpi_coef = 5.0;
  d->quirks = VIRAL_Q_DEFAULT;
  d->k0 = -1;
  d->max_iter = 5000;
  d->t_freeze = 0.0;
  d->fast_math = 0;
  d->eos_table = 0;
  d->bulk = 0;
  d->zeta_law = VIRAL_ZETA_CONST;
  d->zeta_coef = 1.0;
  d->tpib_law = VIRAL_TPIB_CONST;
  d->tpib_coef = 0.1;
}

void viral_ghost_dims(const viral_desc* d, int* gx, int* gy, int* gz)
{
  *gx = d->nx + 2 * VIRAL_BORDER;
  *gy = d->ny + 2 * VIRAL_BORDER;
  *gz = d->lbi ? 1 : d->nz + 2 * VIRAL_BORDER;
}

size_t viral_cells(const viral_desc* d)
{
  int gx, gy, gz;
  viral_ghost_dims(d, &gx, &gy, &gz);
  return (size_t)gx * gy * gz;
}

long long viral_tile_count(const viral_desc* d, int phase, int tp, int tk)
{
  if (!d || phase < 0 || phase >= PH_COUNT) return -1;
  if (phase <= PH_FLUX_Z && (d->fast_math || d->scheme == VIRAL_SHASTA) && (tp < 8 || tk < 1)) return -1;
  return tile_count(d, phase, tp, tk);
}

long long viral_xcd_tile(long long w, long long total)
{
  if (total <= 0 || w < 0 || w >= total) return -1;
  return xcd_tile(w, total);
}

static int check_desc(viral_ctx* ctx, const viral_desc* d)
{
  if (d->nx < 1 || d->ny < 1 || d->nz < 1) return fail(ctx, VIRAL_ERR_ARG, "grid sizes must be positive");
  if (d->lbi && d->nz != 1) return fail(ctx, VIRAL_ERR_ARG, "lbi grids need nz == 1");
  if (d->rk < 1 || d->rk > 3) return fail(ctx, VIRAL_ERR_ARG, "rk must be 1, 2 or 3");
  if (d->scheme != VIRAL_KT && d->scheme != VIRAL_SHASTA) return fail(ctx, VIRAL_ERR_ARG, "unknown scheme");
  if (d->coords != VIRAL_MILNE && d->coords != VIRAL_CARTESIAN) return fail(ctx, VIRAL_ERR_ARG, "unknown coordinates");
  if (d->coords == VIRAL_CARTESIAN && d->scheme != VIRAL_KT)
    return fail(ctx, VIRAL_ERR_UNSUPPORTED,
                "Cartesian coordinates run the KT scheme only (the reference's TXYZ5 SHASTA sources do not compile)");
  if (d->coords == VIRAL_CARTESIAN && d->lbi)
    return fail(ctx, VIRAL_ERR_UNSUPPORTED, "boost invariance needs Milne coordinates");
  if (d->eos != VIRAL_EOS_CONFORMAL && d->eos != VIRAL_EOS_S95P) return fail(ctx, VIRAL_ERR_ARG, "unknown eos");
  if (d->max_iter < 1) return fail(ctx, VIRAL_ERR_ARG, "max_iter must be >= 1");
  // the mirrored ghost layer h of U1, U2 copies the (B-1-h)-th cell from the
  // face: on a grid narrower than the layers that is a ghost cell the same
  // fill writes, an order the parallel fill does not reproduce
  if ((d->quirks & VIRAL_Q_GUBSER_MIRROR) && d->coords == VIRAL_MILNE &&
      (d->nx < VIRAL_BORDER || d->ny < VIRAL_BORDER))
    return fail(ctx, VIRAL_ERR_ARG, "the mirror quirk needs at least 3 interior cells in x and y");
  if (d->bulk && (d->zeta_law < VIRAL_ZETA_CONST || d->zeta_law > VIRAL_ZETA_S_OVER_4PI ||
                  d->tpib_law < VIRAL_TPIB_CONST || d->tpib_law > VIRAL_TPIB_OVER_P))
    return fail(ctx, VIRAL_ERR_ARG, "unknown bulk viscosity or bulk relaxation law");
  if (d->eos_table && !d->fast_math)
    return fail(ctx, VIRAL_ERR_ARG, "the tabulated EOS is a fast-mode option (exact mode evaluates the closed form)");
  int gx, gy, gz;
  viral_ghost_dims(d, &gx, &gy, &gz);
  if (!d->lbi && d->k0 >= gz) return fail(ctx, VIRAL_ERR_ARG, "k0 outside the grid");
  if ((double)gx * gy * gz > 4.0e9) return fail(ctx, VIRAL_ERR_ARG, "grid too large");
  return VIRAL_OK;
}

int viral_create(const viral_desc* descs, int nev, viral_ctx** out)
{
  if (!out || !descs || nev < 1) return VIRAL_ERR_ARG;
  *out = nullptr;
  viral_ctx* ctx = new viral_ctx();
  for (int e = 0; e < nev; e++) {
    int r = check_desc(ctx, &descs[e]);
    if (!r && (descs[e].scheme != descs[0].scheme || descs[e].rk != descs[0].rk ||
               descs[e].fast_math != descs[0].fast_math || descs[e].coords != descs[0].coords))
      r = VIRAL_ERR_ARG;  // the events of one context step through one launch sequence
    if (r) {
      delete ctx;
      return r;
    }
  }
  ctx->rk = descs[0].rk;
  ctx->scheme = descs[0].scheme;
  ctx->fast = descs[0].fast_math != 0;
  ctx->cart = descs[0].coords == VIRAL_CARTESIAN;
  ctx->cfg.batched = nev > 1;
  ctx->cfg.graph = 0;
  ctx->cfg.blocks_per_cu = 0;
  ctx->cfg.fusion = 0;
  ctx->cfg.xcd = 0;
  for (int d = 0; d < 3; d++) ctx->cfg.flux_nt[d] = 256;
  {
    int dev = 0, ncu = 0;
    if (hipGetDevice(&dev) != hipSuccess || resolve_scratch(dev) != hipSuccess) {
      delete ctx;
      return VIRAL_ERR_HIP;
    }
    ctx->dev = dev;
    if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess && ncu > 0)
      ctx->ncu = ncu;
  }
  hipError_t he = hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking);
  if (he != hipSuccess) {
    delete ctx;
    return VIRAL_ERR_HIP;
  }
  bool ok = true;
  ok = ok && hipMalloc(&ctx->d_tab, nev * sizeof(EvDev)) == hipSuccess;
  ok = ok && hipMalloc(&ctx->d_start, (size_t)PH_COUNT * (nev + 1) * sizeof(long)) == hipSuccess;
  ok = ok && hipMalloc(&ctx->d_clk, sizeof(Clock)) == hipSuccess;
  ok = ok && hipHostMalloc(&ctx->h_clk, sizeof(Clock), hipHostMallocDefault) == hipSuccess;
  ok = ok && hipMalloc(&ctx->d_bad, nev * sizeof(unsigned long long)) == hipSuccess;
  ok = ok && hipMalloc(&ctx->d_bmax, nev * sizeof(unsigned long long)) == hipSuccess;
  ctx->bmax_ok.assign(nev, 0);
  ok = ok && hipHostMalloc(&ctx->h_bad, nev * sizeof(unsigned long long), hipHostMallocDefault) == hipSuccess;
  ok = ok && hipMalloc(&ctx->dtmp, sizeof(unsigned long long)) == hipSuccess;
  ctx->ev.resize(nev);
  for (int e = 0; e < nev && ok; e++) {
    Event& E = ctx->ev[e];
    const viral_desc* d = &descs[e];
    E.d = *d;
    Geom& g = E.g;
    fill_geom(d, g);
    Params& p = E.p;
    p.milne = d->coords == VIRAL_MILNE;
    p.lbi = d->lbi;
    p.eos = d->eos;
    p.eta_law = d->eta_law;
    p.tpi_law = d->tpi_law;
    p.max_iter = d->max_iter;
    p.rk = d->rk;
    p.scheme = d->scheme;
    p.quirks = d->quirks;
    p.dx = d->dx; p.dy = d->dy; p.deta = d->deta;
    p.eta_coef = d->eta_coef;
    p.tpi_coef = d->tpi_coef;
    p.bulk = d->bulk;
    p.zeta_law = d->zeta_law;
    p.tpib_law = d->tpib_law;
    p.zeta_coef = d->zeta_coef;
    p.tpib_coef = d->tpib_coef;
    p.factor = 3.0 * d->dof * PIE * PIE / 90.0;
    p.t_freeze = d->t_freeze;
    p.use_tab = 0;
    p.tab = EosTab{nullptr, nullptr, 0, 0, 0};
    size_t n = (size_t)g.n;
    ok = ok && hipMalloc(&E.st[0], NF * n * sizeof(double)) == hipSuccess;
    ok = ok && hipMalloc(&E.st[1], NF * n * sizeof(double)) == hipSuccess;
    // fast mode keeps Var0 only (the stage-start state stands in for Var1 and
    // Var2), and its KT tiles sum the three directions into flux[0]
    const bool fast = descs[0].fast_math != 0;
    ok = ok && hipMalloc(&E.var0, SV * n * sizeof(double)) == hipSuccess;
    if (!fast) ok = ok && hipMalloc(&E.var1, SV * n * sizeof(double)) == hipSuccess;
    if (!fast && d->rk == 3) ok = ok && hipMalloc(&E.var2, SV * n * sizeof(double)) == hipSuccess;
    const int nflux = (fast && d->scheme == VIRAL_KT) ? 1 : d->lbi ? 2 : 3;
    for (int q = 0; q < nflux; q++) ok = ok && hipMalloc(&E.flux[q], SV * n * sizeof(double)) == hipSuccess;
    ok = ok && hipMalloc(&E.src, SV * n * sizeof(double)) == hipSuccess;
    ok = ok && hipMalloc(&E.iters, 3 * n * sizeof(int32_t)) == hipSuccess;
    ok = ok && hipMalloc(&E.flags, n) == hipSuccess;
    if (!ok) break;
    {
      // default LDS tiles: 32-cell lines x 16 lines across x and y; along eta
      // the whole ghosted line when it fits (one tile per line), as many lines
      // as the default block size holds
      const int gz = g.gz;
      E.tile_tp[2] = gz <= 128 ? std::max(gz, 8) : 64;  // lbi (gz = 1): the eta phase is empty anyway
      E.tile_tk[2] = std::max(1, TileSlots<256>::MAXS / E.tile_tp[2]);
    }
    if (d->eos_table && d->eos == VIRAL_EOS_S95P) {
      std::vector<double> tv;
      std::vector<unsigned char> tc;
      build_eos_table(tv, tc);
      ok = ok && hipMalloc(&E.tab_v, tv.size() * sizeof(double)) == hipSuccess;
      ok = ok && hipMalloc(&E.tab_closed, tc.size()) == hipSuccess;
      if (!ok) break;
      hipMemcpy(E.tab_v, tv.data(), tv.size() * sizeof(double), hipMemcpyHostToDevice);
      hipMemcpy(E.tab_closed, tc.data(), tc.size(), hipMemcpyHostToDevice);
      p.use_tab = 1;
      p.tab = EosTab{E.tab_v, E.tab_closed, kTabEmin, kTabNoct, kTabK};
    }
    hipMemset(E.st[0], 0, NF * n * sizeof(double));
    hipMemset(E.st[1], 0, NF * n * sizeof(double));
    hipMemset(E.iters, 0, 3 * n * sizeof(int32_t));
    hipMemset(E.flags, 0, n);
  }
  if (!ok || hipDeviceSynchronize() != hipSuccess) {
    viral_destroy(ctx);
    return VIRAL_ERR_HIP;
  }
  ctx->h_clk->tau = 1.0;
  ctx->h_clk->ts = 0.01;
  ctx->h_clk->step = 0;
  if (refresh_table(ctx) != VIRAL_OK) {
    viral_destroy(ctx);
    return VIRAL_ERR_HIP;
  }
  *out = ctx;
  return VIRAL_OK;
}

void viral_destroy(viral_ctx* ctx)
{
  if (!ctx) return;
  if (ctx->stream) hipStreamSynchronize(ctx->stream);
  drop_graphs(ctx);
  // the adjoint recording: checkpoints and, when every == 0, the stage-state slabs
  for (double* c : ctx->adj.ckpt) hipFree(c);
  for (double* c : ctx->adj.slabs) hipFree(c);
  if (ctx->adj.ring) hipFree(ctx->adj.ring);
  ctx->adj.ring = nullptr;
  ctx->adj.ckpt.clear();
  ctx->adj.slabs.clear();
  ctx->adj.stages.clear();
  for (Event& E : ctx->ev) {
    hipFree(E.st[0]); hipFree(E.st[1]); hipFree(E.var0);
    if (E.var1) hipFree(E.var1);
    if (E.var2) hipFree(E.var2);
    for (int q = 0; q < 3; q++)
      if (E.flux[q]) hipFree(E.flux[q]);
    hipFree(E.src);
    hipFree(E.iters); hipFree(E.flags);
    if (E.tab_v) hipFree(E.tab_v);
    if (E.tab_closed) hipFree(E.tab_closed);
  }
  if (ctx->d_tab) hipFree(ctx->d_tab);
  if (ctx->d_start) hipFree(ctx->d_start);
  if (ctx->d_clk) hipFree(ctx->d_clk);
  if (ctx->h_clk) hipHostFree(ctx->h_clk);
  if (ctx->d_bad) hipFree(ctx->d_bad);
  if (ctx->d_bmax) hipFree(ctx->d_bmax);
  if (ctx->h_bad) hipHostFree(ctx->h_bad);
  if (ctx->dtmp) hipFree(ctx->dtmp);
  if (ctx->d_diag) hipFree(ctx->d_diag);
  if (ctx->d_out) hipFree(ctx->d_out);
  for (int k = 0; k < K_COUNT; k++)
    for (hipEvent_t e : ctx->tev[k]) hipEventDestroy(e);
  for (hipEvent_t e : ctx->pool) hipEventDestroy(e);
  if (ctx->stream) hipStreamDestroy(ctx->stream);
  if (ctx->stream2) hipStreamDestroy(ctx->stream2);
  delete ctx;
}

const char* viral_last_error(viral_ctx* ctx) { return ctx ? ctx->err.c_str() : "no context"; }

static Event* get_ev(viral_ctx* ctx, int ev)
{
  if (!ctx || ev < 0 || ev >= (int)ctx->ev.size()) return nullptr;
  return &ctx->ev[ev];
}

int viral_upload_state(viral_ctx* ctx, int ev, const double* const* fields)
{
  Event* E = get_ev(ctx, ev);
  if (!E || !fields) return fail(ctx, VIRAL_ERR_ARG, "bad event or fields");
  size_t n = (size_t)E->g.n;
  for (int f = 0; f < NF; f++) {
    if (!fields[f]) return fail(ctx, VIRAL_ERR_ARG, "null field pointer");
    // both copies: cells the step never writes (corner ghosts, Temp in ghosts)
    // keep their uploaded values whichever copy is current
    HIPCHK(hipMemcpyAsync(E->st[0] + f * n, fields[f], n * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(hipMemcpyAsync(E->st[1] + f * n, fields[f], n * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
  }
  HIPCHK(hipStreamSynchronize(ctx->stream));
  ctx->bmax_ok[ev] = 0;
  return VIRAL_OK;
}

int viral_download_field(viral_ctx* ctx, int ev, int f, double* out)
{
  Event* E = get_ev(ctx, ev);
  if (!E || !out || f < 0 || f >= NF) return fail(ctx, VIRAL_ERR_ARG, "bad event, field or pointer");
  size_t n = (size_t)E->g.n;
  HIPCHK(hipMemcpyAsync(out, E->st[ctx->cur] + f * n, n * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  return VIRAL_OK;
}

int viral_download_state(viral_ctx* ctx, int ev, double* const* fields)
{
  Event* E = get_ev(ctx, ev);
  if (!E || !fields) return fail(ctx, VIRAL_ERR_ARG, "bad event or fields");
  size_t n = (size_t)E->g.n;
  for (int f = 0; f < NF; f++) {
    if (!fields[f]) return fail(ctx, VIRAL_ERR_ARG, "null field pointer");
    HIPCHK(hipMemcpyAsync(fields[f], E->st[ctx->cur] + f * n, n * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  }
  HIPCHK(hipStreamSynchronize(ctx->stream));
  return VIRAL_OK;
}

int viral_steps(viral_ctx* ctx, double tau, double ts, int nsteps, int* done)
{
  if (!ctx || nsteps < 0) return VIRAL_ERR_ARG;
  if (done) *done = 0;
  if (nsteps == 0) return VIRAL_OK;
  if (nsteps >= (1 << 22)) return fail(ctx, VIRAL_ERR_ARG, "at most 2^22 - 1 steps per call");
  if (int r = refresh_table(ctx)) return r;
  const int nev = (int)ctx->ev.size();
  HIPCHK(hipStreamSynchronize(ctx->stream));  // the pinned clock may still be in flight
  ctx->h_clk->tau = tau;
  ctx->h_clk->ts = ts;
  ctx->h_clk->step = 0;
  HIPCHK(hipMemcpyAsync(ctx->d_clk, ctx->h_clk, sizeof(Clock), hipMemcpyHostToDevice, ctx->stream));
  hipLaunchKernelGGL(k_clear_bad, dim3((nev + 63) / 64), dim3(64), 0, ctx->stream, ctx->d_bad, nev);
  for (int e = 0; e < nev; e++)
    if (fix_ve(ctx->ev[e]) && !ctx->bmax_ok[e]) {  // init()'s DebugMSG of the uploaded state
      launch_bmax(ctx, e, ctx->cur);
      ctx->bmax_ok[e] = 1;
    }
  // adjoint recording that keeps every stage state: the stages write straight
  // into a slab of stage states (the event's table row points there), and the
  // last one is copied back into the state at the end
  const bool store = ctx->adj.on && ctx->adj.every == 0;
  const size_t sbytes = store ? (size_t)NF * ctx->ev[0].g.n * sizeof(double) : 0;
  double* slab = nullptr;
  if (store) {
    if (hipMalloc(&slab, (size_t)nsteps * ctx->rk * sbytes) != hipSuccess)
      return fail(ctx, VIRAL_ERR_HIP, "out of device memory for the recorded stage states");
    ctx->adj.slabs.push_back(slab);
  }
  // adjoint recording with checkpoints: the stages of the current segment
  // write straight into the ring of stage states (stage q of the segment
  // reads q - 1, its first stage the segment's checkpoint), so viral_adjoint
  // reverses the last segment without recomputing it; a checkpoint copies the
  // segment's last stage state, and the last one is copied back into the
  // state at the end
  const bool ringrec = ctx->adj.on && ctx->adj.every > 0;
  const size_t rslot = ringrec ? (size_t)NF * ctx->ev[0].g.n : 0;
  const bool graph = ctx->cfg.graph && !ctx->timing && !store && !ringrec;
  double trec = tau;
  for (int s = 0; s < nsteps; s++) {
    if (ringrec) {
      const size_t q0 = (ctx->adj.taus.size() % ctx->adj.every) * ctx->rk;
      for (int st = 0; st < ctx->rk; st++) {
        const size_t q = q0 + st;
        double* in = q == 0 ? ctx->adj.ckpt.back() : ctx->adj.ring + (q - 1) * rslot;
        hipLaunchKernelGGL(k_set_state_ptrs, dim3(1), dim3(64), 0, ctx->stream, ctx->d_tab, 0, in,
                           ctx->adj.ring + q * rslot);
        launch_stage(ctx, st, 0);
      }
      hipLaunchKernelGGL(k_clock_advance, dim3(1), dim3(64), 0, ctx->stream, ctx->d_clk);
    } else if (store) {
      for (int st = 0; st < ctx->rk; st++) {
        // the call's first stage reads the state itself (equal, on every cell a
        // stage reads, to the last recorded stage state or the first checkpoint)
        double* in = (s == 0 && st == 0) ? ctx->ev[0].st[ctx->cur] : ctx->adj.stages.back();
        double* out = slab + ((size_t)s * ctx->rk + st) * (sbytes / sizeof(double));
        hipLaunchKernelGGL(k_set_state_ptrs, dim3(1), dim3(64), 0, ctx->stream, ctx->d_tab, 0, in, out);
        launch_stage(ctx, st, 0);
        ctx->adj.stages.push_back(out);
      }
      hipLaunchKernelGGL(k_clock_advance, dim3(1), dim3(64), 0, ctx->stream, ctx->d_clk);
    } else if (graph) {
      if (int r = capture_step(ctx, ctx->cur)) return r;
      HIPCHK(hipGraphLaunch(ctx->gexec[ctx->cur], ctx->stream));
    } else {
      launch_step(ctx, ctx->cur);
    }
    ctx->cur ^= (ctx->rk & 1);
    if (ctx->adj.on) {
      ctx->adj.taus.push_back(trec);
      ctx->adj.tss.push_back(ts);
      if (ringrec && ctx->adj.taus.size() % ctx->adj.every == 0) {
        const size_t bytes = rslot * sizeof(double);
        double* c = nullptr;
        HIPCHK(hipMalloc(&c, bytes));
        ctx->adj.ckpt.push_back(c);
        HIPCHK(hipMemcpyAsync(c, ctx->adj.ring + ((size_t)ctx->adj.every * ctx->rk - 1) * rslot, bytes,
                              hipMemcpyDeviceToDevice, ctx->stream));
      }
    }
    trec = trec + ts;
  }
  if (ringrec) {
    Event& E = ctx->ev[0];
    const size_t last = ((ctx->adj.taus.size() - 1) % ctx->adj.every) * ctx->rk + ctx->rk - 1;
    hipLaunchKernelGGL(k_set_state_ptrs, dim3(1), dim3(64), 0, ctx->stream, ctx->d_tab, 0, E.st[0], E.st[1]);
    hipLaunchKernelGGL(k_copy_written, dim3(nblocks((long)NF * E.g.n)), dim3(BLK), 0, ctx->stream, E.g, E.p.lbi,
                       ctx->adj.ring + last * rslot, E.st[ctx->cur]);
  }
  if (store) {
    Event& E = ctx->ev[0];
    hipLaunchKernelGGL(k_set_state_ptrs, dim3(1), dim3(64), 0, ctx->stream, ctx->d_tab, 0, E.st[0], E.st[1]);
    hipLaunchKernelGGL(k_copy_written, dim3(nblocks((long)NF * E.g.n)), dim3(BLK), 0, ctx->stream, E.g, E.p.lbi,
                       ctx->adj.stages.back(), E.st[ctx->cur]);
  }
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(ctx->h_bad, ctx->d_bad, nev * sizeof(unsigned long long), hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  double t = tau;
  for (int s = 0; s < nsteps; s++) t = t + ts;
  ctx->last_tau = t;
  ctx->last_ts = ts;
  long long first = -1;
  for (int e = 0; e < nev; e++) {
    ctx->ev[e].bad_host = ctx->h_bad[e];
    if (ctx->h_bad[e] != ~0ull) {
      const long long st = (long long)(ctx->h_bad[e] >> 42);
      if (first < 0 || st < first) first = st;
    }
  }
  if (first >= 0) {
    if (done) *done = (int)first;
    return fail(ctx, VIRAL_ERR_PHYSICS, "a cell became acausal, non-positive or unsolvable");
  }
  if (done) *done = nsteps;
  return VIRAL_OK;
}

int viral_step(viral_ctx* ctx, double tau, double ts) { return viral_steps(ctx, tau, ts, 1, nullptr); }

int viral_flags(viral_ctx* ctx, int ev, uint8_t* out)
{
  Event* E = get_ev(ctx, ev);
  if (!E || !out) return fail(ctx, VIRAL_ERR_ARG, "bad event or pointer");
  HIPCHK(hipMemcpyAsync(out, E->flags, (size_t)E->g.n, hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  return VIRAL_OK;
}

int viral_iterations(viral_ctx* ctx, int ev, int stage, int32_t* out)
{
  Event* E = get_ev(ctx, ev);
  if (!E || !out || stage < 0 || stage > 2) return fail(ctx, VIRAL_ERR_ARG, "bad event, stage or pointer");
  HIPCHK(hipMemcpyAsync(out, E->iters + (long)stage * E->g.n, (size_t)E->g.n * sizeof(int32_t),
                        hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  return VIRAL_OK;
}

int viral_first_bad_cell(viral_ctx* ctx, int ev, int* i, int* j, int* k, int* stage, unsigned* flags)
{
  Event* E = get_ev(ctx, ev);
  if (!E || !i || !j || !k || !stage || !flags) return fail(ctx, VIRAL_ERR_ARG, "bad event or pointer");
  if (E->bad_host == ~0ull) return -1;
  long x = (long)((E->bad_host >> 8) & 0xffffffffull);
  *stage = (int)((E->bad_host >> 40) & 3);
  *k = (int)(x % E->g.gz);
  long r = x / E->g.gz;
  *j = (int)(r % E->g.gy);
  *i = (int)(r / E->g.gy);
  *flags = (unsigned)(E->bad_host & 0xff);
  return 0;
}

int viral_root(viral_ctx* ctx, int ev, double tau)
{
  Event* E = get_ev(ctx, ev);
  if (!E) return fail(ctx, VIRAL_ERR_ARG, "bad event");
  if (!E->p.milne) return fail(ctx, VIRAL_ERR_UNSUPPORTED, "viral_root covers the Milne (1-D) root solve");
  ctx->bmax_ok[ev] = 0;
  const long ni = box_count(interior_box(E->g));
  hipLaunchKernelGGL(k_clear_bad, dim3(1), dim3(64), 0, ctx->stream, ctx->d_bad + ev, 1);
  double* st = E->st[ctx->cur];
  if (E->d.fast_math)
    hipLaunchKernelGGL(k_root<true>, dim3(nblocks(ni)), dim3(BLK), 0, ctx->stream, E->g, E->p, tau, st, E->iters,
                       E->flags, ctx->d_bad + ev);
  else
    hipLaunchKernelGGL(k_root<false>, dim3(nblocks(ni)), dim3(BLK), 0, ctx->stream, E->g, E->p, tau, st, E->iters,
                       E->flags, ctx->d_bad + ev);
  if ((E->p.quirks & VIRAL_Q_K0_BROADCAST) && !E->p.lbi)
    hipLaunchKernelGGL(k_root_k0, dim3(nblocks(ni)), dim3(BLK), 0, ctx->stream, E->g, E->p, tau, st);
  HIPCHK(hipGetLastError());
  unsigned long long w = ~0ull;
  HIPCHK(hipMemcpyAsync(&w, ctx->d_bad + ev, sizeof(w), hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  E->bad_host = w;
  return w == ~0ull ? VIRAL_OK : VIRAL_ERR_PHYSICS;
}

int viral_ns_init(viral_ctx* ctx, int ev, double tau, double ts, int mode, double* const* nspi)
{
  Event* E = get_ev(ctx, ev);
  if (!E) return fail(ctx, VIRAL_ERR_ARG, "bad event");
  if (!E->p.milne) return fail(ctx, VIRAL_ERR_UNSUPPORTED, "the Navier-Stokes initialisation is Milne only");
  if (mode < VIRAL_NS_OUTPUT || mode > VIRAL_ZERO_INIT || !(ts > 0) || !(tau > 0))
    return fail(ctx, VIRAL_ERR_ARG, "bad mode, tau or ts");
  const int nout = E->p.bulk ? 11 : 10;  // with bulk, nspi[10] receives the Navier-Stokes PI
  if (mode == VIRAL_NS_OUTPUT) {
    if (!nspi) return fail(ctx, VIRAL_ERR_ARG, "VIRAL_NS_OUTPUT needs 10 output arrays (11 with bulk)");
    for (int l = 0; l < nout; l++)
      if (!nspi[l]) return fail(ctx, VIRAL_ERR_ARG, "VIRAL_NS_OUTPUT needs 10 output arrays (11 with bulk)");
  }
  const long n = E->g.n;
  double* st = E->st[ctx->cur];
  const unsigned grid = nblocks(n);
  if (mode != VIRAL_NS_OUTPUT) ctx->bmax_ok[ev] = 0;
  if (mode == VIRAL_NS_OUTPUT) {
    const size_t m = (size_t)nout * n;
    if (m > ctx->out_cap) {
      HIPCHK(hipStreamSynchronize(ctx->stream));
      if (ctx->d_out) HIPCHK(hipFree(ctx->d_out));
      ctx->d_out = nullptr;
      ctx->out_cap = 0;
      HIPCHK(hipMalloc(&ctx->d_out, m * sizeof(double)));
      ctx->out_cap = m;
    }
    hipLaunchKernelGGL(k_ns_init<0>, dim3(grid), dim3(BLK), 0, ctx->stream, E->g, E->p, tau, ts, st, ctx->d_out);
    HIPCHK(hipGetLastError());
    for (int l = 0; l < nout; l++)
      HIPCHK(hipMemcpyAsync(nspi[l], ctx->d_out + (size_t)l * n, n * sizeof(double), hipMemcpyDeviceToHost,
                            ctx->stream));
  } else if (mode == VIRAL_NS_INIT) {
    hipLaunchKernelGGL(k_ns_init<1>, dim3(grid), dim3(BLK), 0, ctx->stream, E->g, E->p, tau, ts, st, nullptr);
  } else {
    hipLaunchKernelGGL(k_ns_init<2>, dim3(grid), dim3(BLK), 0, ctx->stream, E->g, E->p, tau, ts, st, nullptr);
  }
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(ctx->stream));
  return VIRAL_OK;
}

int viral_landau(viral_ctx* ctx, int ev, double tau, const double* const* tmunu, int at_start)
{
  Event* E = get_ev(ctx, ev);
  if (!E) return fail(ctx, VIRAL_ERR_ARG, "bad event");
  if (!E->p.milne) return fail(ctx, VIRAL_ERR_UNSUPPORTED, "Landau matching is Milne only (ginit)");
  if (!tmunu || !(tau > 0)) return fail(ctx, VIRAL_ERR_ARG, "bad T^{mu nu} arrays or tau");
  for (int l = 0; l < 10; l++)
    if (!tmunu[l]) return fail(ctx, VIRAL_ERR_ARG, "bad T^{mu nu} arrays or tau");
  const long n = E->g.n;
  const size_t m = 10 * (size_t)n;
  ctx->bmax_ok[ev] = 0;
  if (m > ctx->out_cap) {
    HIPCHK(hipStreamSynchronize(ctx->stream));
    if (ctx->d_out) HIPCHK(hipFree(ctx->d_out));
    ctx->d_out = nullptr;
    ctx->out_cap = 0;
    HIPCHK(hipMalloc(&ctx->d_out, m * sizeof(double)));
    ctx->out_cap = m;
  }
  for (int l = 0; l < 10; l++)
    HIPCHK(hipMemcpyAsync(ctx->d_out + (size_t)l * n, tmunu[l], n * sizeof(double), hipMemcpyHostToDevice,
                          ctx->stream));
  hipLaunchKernelGGL(k_clear_bad, dim3(1), dim3(64), 0, ctx->stream, ctx->d_bad + ev, 1);
  double* st = E->st[ctx->cur];
  if (at_start)
    hipLaunchKernelGGL(k_landau<true>, dim3(nblocks(n)), dim3(BLK), 0, ctx->stream, E->g, E->p, tau, ctx->d_out, st,
                       E->flags, ctx->d_bad + ev);
  else
    hipLaunchKernelGGL(k_landau<false>, dim3(nblocks(n)), dim3(BLK), 0, ctx->stream, E->g, E->p, tau, ctx->d_out,
                       st, E->flags, ctx->d_bad + ev);
  HIPCHK(hipGetLastError());
  unsigned long long w = ~0ull;
  HIPCHK(hipMemcpyAsync(&w, ctx->d_bad + ev, sizeof(w), hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  E->bad_host = w;
  return w == ~0ull ? VIRAL_OK : VIRAL_ERR_PHYSICS;
}

int viral_max_temp_gev(viral_ctx* ctx, int ev, double* out)
{
  Event* E = get_ev(ctx, ev);
  if (!E || !out) return fail(ctx, VIRAL_ERR_ARG, "bad event or pointer");
  HIPCHK(hipMemsetAsync(ctx->dtmp, 0, sizeof(unsigned long long), ctx->stream));
  long ni = box_count(interior_box(E->g));
  hipLaunchKernelGGL(k_max_temp, dim3(nblocks(ni)), dim3(BLK), 0, ctx->stream, E->g, E->p, E->st[ctx->cur], ctx->dtmp);
  unsigned long long bits = 0;
  HIPCHK(hipMemcpyAsync(&bits, ctx->dtmp, sizeof(bits), hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  double v;
  memcpy(&v, &bits, sizeof(v));
  *out = v;
  return VIRAL_OK;
}

// Runs one diagnostic reduction over event E's current state and combines
// the block partials in block order into res[0..nstat).
static void launch_diag(viral_ctx* ctx, const Event* E, int kind, unsigned nb, const double* st, double tau,
                        DiagPart* out)
{
  const bool cart = !E->p.milne;
  if (kind == D_ROOT && !cart)
    hipLaunchKernelGGL((k_diag<D_ROOT, false>), dim3(nb), dim3(BLK), 0, ctx->stream, E->g, E->p, st, tau, out);
  else if (kind == D_ROOT)
    hipLaunchKernelGGL((k_diag<D_ROOT, true>), dim3(nb), dim3(BLK), 0, ctx->stream, E->g, E->p, st, tau, out);
  else if (kind == D_EXTREMA)
    hipLaunchKernelGGL((k_diag<D_EXTREMA, false>), dim3(nb), dim3(BLK), 0, ctx->stream, E->g, E->p, st, tau, out);
  else
    hipLaunchKernelGGL((k_diag<D_PHYSICS, false>), dim3(nb), dim3(BLK), 0, ctx->stream, E->g, E->p, st, tau, out);
}

static int run_diag(viral_ctx* ctx, Event* E, int kind, double tau, DiagPart* res, int* nstat)
{
  if (!ctx->d_diag) HIPCHK(hipMalloc(&ctx->d_diag, sizeof(DiagPart) * DIAG_BLOCKS * DIAG_NSTAT));
  const long ni = box_count(interior_box(E->g));
  const unsigned nb = (unsigned)std::min<long>(DIAG_BLOCKS, std::max<long>(1, (ni + BLK - 1) / BLK));
  const double* st = E->st[ctx->cur];
  launch_diag(ctx, E, kind, nb, st, tau, ctx->d_diag);
  HIPCHK(hipGetLastError());
  std::vector<DiagPart> part((size_t)nb * DIAG_NSTAT);
  HIPCHK(hipMemcpyAsync(part.data(), ctx->d_diag, part.size() * sizeof(DiagPart), hipMemcpyDeviceToHost,
                        ctx->stream));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  const int ns = diag_nstat(kind, E->p.milne);
  for (int q = 0; q < ns; q++) {
    DiagPart acc{0.0, -1, 0.0};
    for (unsigned b = 0; b < nb; b++) acc = diag_combine(acc, part[(size_t)b * DIAG_NSTAT + q], diag_rule(kind, E->p.milne, q));
    res[q] = acc;
  }
  if (nstat) *nstat = ns;
  return VIRAL_OK;
}

static void cell_ijk(const Geom& g, long long x, int* ijk)
{
  if (x < 0) { ijk[0] = ijk[1] = ijk[2] = -1; return; }
  ijk[2] = (int)(x % g.gz);
  ijk[1] = (int)((x / g.gz) % g.gy);
  ijk[0] = (int)(x / ((long long)g.gz * g.gy));
}

int viral_check_root(viral_ctx* ctx, int ev, double tau, double* max, double* avg, int* ijk, int* nstat)
{
  Event* E = get_ev(ctx, ev);
  if (!E || !max || !avg) return fail(ctx, VIRAL_ERR_ARG, "bad event or pointer");
  DiagPart r[DIAG_NSTAT];
  int ns = 0, rc = run_diag(ctx, E, D_ROOT, tau, r, &ns);
  if (rc != VIRAL_OK) return rc;
  const double count = (double)box_count(interior_box(E->g));
  for (int q = 0; q < ns; q++) {
    max[q] = r[q].v;
    avg[q] = r[q].sum / count;
    if (ijk) cell_ijk(E->g, r[q].at, ijk + 3 * q);
  }
  if (nstat) *nstat = ns;
  return VIRAL_OK;
}

int viral_check_root_noalloc(viral_ctx* ctx, int ev, double tau, double* max, double* avg, int* ijk, int* nstat)
{
  Event* E = get_ev(ctx, ev);
  if (!E || !max || !avg) return fail(ctx, VIRAL_ERR_ARG, "bad event or pointer");
  DiagPart* sp = (ctx->dev >= 0 && ctx->dev < MAX_DEVICES) ? scratch_ptr[ctx->dev] : nullptr;
  if (!sp) return fail(ctx, VIRAL_ERR_HIP, "device scratch of viral_check_root_noalloc not resolved");
  const long ni = box_count(interior_box(E->g));
  const unsigned nb = (unsigned)std::min<long>(DIAG_BLOCKS, std::max<long>(1, (ni + BLK - 1) / BLK));
  const int ns = diag_nstat(D_ROOT, E->p.milne);
  // the partials of nb blocks and ns statistics must fit the scratch (k_diag
  // strides over the interior with at most DIAG_BLOCKS blocks, so every grid
  // fits; the check guards the layout)
  if ((size_t)nb * DIAG_NSTAT + ns > (size_t)SCRATCH_PARTS || ns > DIAG_NSTAT)
    return fail(ctx, VIRAL_ERR_ARG, "viral_check_root_noalloc: partials exceed the device scratch");
  DiagPart* part = sp;
  DiagPart* res = sp + (size_t)DIAG_BLOCKS * DIAG_NSTAT;
  launch_diag(ctx, E, D_ROOT, nb, E->st[ctx->cur], tau, part);
  hipLaunchKernelGGL(k_diag_combine<D_ROOT>, dim3(1), dim3(64), 0, ctx->stream, part, nb, E->p.milne, res);
  HIPCHK(hipGetLastError());
  DiagPart r[DIAG_NSTAT];
  HIPCHK(hipMemcpyAsync(r, res, (size_t)ns * sizeof(DiagPart), hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  const double count = (double)ni;
  for (int q = 0; q < ns; q++) {
    max[q] = r[q].v;
    avg[q] = r[q].sum / count;
    if (ijk) cell_ijk(E->g, r[q].at, ijk + 3 * q);
  }
  if (nstat) *nstat = ns;
  return VIRAL_OK;
}

int viral_debug_extrema(viral_ctx* ctx, int ev, double tau, double* out, int* ijk)
{
  Event* E = get_ev(ctx, ev);
  if (!E || !out) return fail(ctx, VIRAL_ERR_ARG, "bad event or pointer");
  DiagPart r[DIAG_NSTAT];
  int rc = run_diag(ctx, E, D_EXTREMA, tau, r, nullptr);
  if (rc != VIRAL_OK) return rc;
  out[0] = r[0].at >= 0 ? r[0].v : 0.0;
  for (int m = 0; m < 3; m++) {
    out[1 + m] = 0.0;
    if (r[0].at >= 0)
      HIPCHK(hipMemcpy(&out[1 + m], E->st[ctx->cur] + (long)(VIRAL_F_VX + m) * E->g.n + r[0].at, sizeof(double),
                       hipMemcpyDeviceToHost));
  }
  out[4] = r[1].at >= 0 ? r[1].v : 0.0;
  out[5] = r[2].at >= 0 ? -r[2].v : 0.0;
  if (ijk)
    for (int q = 0; q < 3; q++) cell_ijk(E->g, r[q].at, ijk + 3 * q);
  return VIRAL_OK;
}

int viral_check_physics(viral_ctx* ctx, int ev, double tau, long long* count, int* ijk)
{
  Event* E = get_ev(ctx, ev);
  if (!E || !count) return fail(ctx, VIRAL_ERR_ARG, "bad event or pointer");
  DiagPart r[DIAG_NSTAT];
  int rc = run_diag(ctx, E, D_PHYSICS, tau, r, nullptr);
  if (rc != VIRAL_OK) return rc;
  for (int q = 0; q < 2; q++) {
    count[q] = (long long)r[q].sum;
    if (ijk) cell_ijk(E->g, r[q].at, ijk + 3 * q);
  }
  return VIRAL_OK;
}

int viral_output_size(const viral_desc* d, int freq, int freqz, int xy, size_t* ndoubles)
{
  if (!d || !ndoubles) return VIRAL_ERR_ARG;
  if (freq < 1 || freqz < 1 || d->nx % freq || d->ny % freq || (!xy && !d->lbi && d->nz % freqz)) return VIRAL_ERR_ARG;
  const size_t nzo = (xy || d->lbi) ? 1 : (size_t)(d->nz / freqz);
  *ndoubles = (size_t)(d->nx / freq) * (size_t)(d->ny / freq) * nzo * 17;
  return VIRAL_OK;
}

int viral_output(viral_ctx* ctx, int ev, int freq, int freqz, int xy, double* out, int async)
{
  Event* E = get_ev(ctx, ev);
  if (!E || !out) return fail(ctx, VIRAL_ERR_ARG, "bad event or pointer");
  size_t m = 0;
  if (viral_output_size(&E->d, freq, freqz, xy, &m) != VIRAL_OK)
    return fail(ctx, VIRAL_ERR_ARG, "output strides must divide the grid");
  if (m > ctx->out_cap) {
    HIPCHK(hipStreamSynchronize(ctx->stream));
    if (ctx->d_out) HIPCHK(hipFree(ctx->d_out));
    ctx->d_out = nullptr;
    ctx->out_cap = 0;
    HIPCHK(hipMalloc(&ctx->d_out, m * sizeof(double)));
    ctx->out_cap = m;
  }
  const long nyo = E->d.ny / freq, nzo = (xy || E->d.lbi) ? 1 : E->d.nz / freqz;
  hipLaunchKernelGGL(k_output, dim3(nblocks((long)m)), dim3(BLK), 0, ctx->stream, E->g, E->st[ctx->cur], freq, freqz,
                     xy || E->d.lbi, nyo, nzo, (long)m, ctx->d_out);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(out, ctx->d_out, m * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  if (!async) HIPCHK(hipStreamSynchronize(ctx->stream));
  return VIRAL_OK;
}

int viral_sync(viral_ctx* ctx)
{
  if (!ctx) return VIRAL_ERR_ARG;
  HIPCHK(hipStreamSynchronize(ctx->stream));
  return VIRAL_OK;
}

int viral_host_alloc(size_t bytes, void** p)
{
  if (!p) return VIRAL_ERR_ARG;
  *p = nullptr;
  if (hipHostMalloc(p, bytes, hipHostMallocDefault) != hipSuccess) return VIRAL_ERR_HIP;
  return VIRAL_OK;
}

int viral_host_free(void* p)
{
  if (p && hipHostFree(p) != hipSuccess) return VIRAL_ERR_HIP;
  return VIRAL_OK;
}

int viral_eos_eval(const viral_desc* d, size_t n, const double* e, double* p, double* dpde, double* T, double* s)
{
  if (!d || !e || !p || !dpde || !T || !s) return VIRAL_ERR_ARG;
  const double factor = 3.0 * d->dof * PIE * PIE / 90.0;
  const bool tab = d->eos == VIRAL_EOS_S95P && d->eos_table;
  EosTab t = tab ? host_tab() : EosTab{nullptr, nullptr, 0, 0, 0};
  for (size_t q = 0; q < n; q++) {
    const double x = e[q];
    if (d->eos == VIRAL_EOS_S95P) {
      double out[4];
      if (!(tab && eos_tab_lookup(t, x, out))) s95p_all(x, out);
      p[q] = out[0];
      dpde[q] = out[1];
      T[q] = out[2];
      s[q] = out[3];
    } else {
      p[q] = x / 3.0;
      dpde[q] = 1.0 / 3.0;
      T[q] = sqrt(sqrt(x / factor));
      s[q] = (x + p[q]) * (1.0 / T[q]);
    }
  }
  return VIRAL_OK;
}

int viral_set_tile(viral_ctx* ctx, int ev, int dir, int tp, int tk)
{
  Event* E = get_ev(ctx, ev);
  if (!E || dir < 0 || dir > 2 || !tile_ok(tp, tk, ctx->cfg.flux_nt[dir])) return fail(ctx, VIRAL_ERR_ARG, "bad tile");
  E->tile_tp[dir] = tp;
  E->tile_tk[dir] = tk;
  ctx->dirty = true;
  return VIRAL_OK;
}

int viral_get_tile(viral_ctx* ctx, int ev, int dir, int* tp, int* tk)
{
  Event* E = get_ev(ctx, ev);
  if (!E || dir < 0 || dir > 2 || !tp || !tk) return fail(ctx, VIRAL_ERR_ARG, "bad tile query");
  *tp = E->tile_tp[dir];
  *tk = E->tile_tk[dir];
  return VIRAL_OK;
}

int viral_get_launch(viral_ctx* ctx, viral_launch* out)
{
  if (!ctx || !out) return VIRAL_ERR_ARG;
  *out = ctx->cfg;
  return VIRAL_OK;
}

int viral_set_launch(viral_ctx* ctx, const viral_launch* in)
{
  if (!ctx || !in) return VIRAL_ERR_ARG;
  if (in->blocks_per_cu < 0 || in->blocks_per_cu > 64) return fail(ctx, VIRAL_ERR_ARG, "blocks_per_cu in [0, 64]");
  for (int d = 0; d < 3; d++) {
    if (!max_slots(in->flux_nt[d])) return fail(ctx, VIRAL_ERR_ARG, "flux_nt must be 64, 128, 256 or 512");
    for (const Event& E : ctx->ev)
      if (!tile_ok(E.tile_tp[d], E.tile_tk[d], in->flux_nt[d]))
        return fail(ctx, VIRAL_ERR_ARG, "an event's flux tile does not fit flux_nt (set smaller tiles first)");
  }
  // SHASTA runs its LDS line tiles with 256 threads per block, one tile per
  // block, and the sources in their own kernel: the configuration its tests
  // cover.  An autotuner pass over the other options on a 256x256x96 SHASTA
  // context took a GPU memory fault whose cause was not found, so they stay
  // closed to SHASTA.
  if (ctx->scheme == VIRAL_SHASTA &&
      (in->blocks_per_cu != 0 || in->fusion != 0 || in->flux_nt[0] != 256 || in->flux_nt[1] != 256 ||
       in->flux_nt[2] != 256))
    return fail(ctx, VIRAL_ERR_UNSUPPORTED,
                "SHASTA runs with 256-thread flux blocks, no persistent blocks and no source fusion");
  viral_launch c = *in;
  c.batched = in->batched != 0;
  c.graph = in->graph != 0;
  c.fusion = in->fusion != 0;
  c.xcd = in->xcd != 0;
  HIPCHK(hipStreamSynchronize(ctx->stream));
  ctx->cfg = c;
  ctx->dirty = true;
  return VIRAL_OK;
}

int viral_tile_table(viral_ctx* ctx, int phase, long long* start)
{
  if (!ctx || !start || phase < 0 || phase >= PH_COUNT) return fail(ctx, VIRAL_ERR_ARG, "bad tile table query");
  if (int r = refresh_table(ctx)) return r;
  long long acc = 0;
  for (size_t e = 0; e < ctx->ev.size(); e++) {
    start[e] = acc;
    acc += ctx->cnt[phase][e];
  }
  start[ctx->ev.size()] = acc;
  return VIRAL_OK;
}

// Autotuner.  Times launches on the current state at the clock of the last
// step (stage 0 writes only the spare state copy, the RK registers and the
// per-step diagnostics, which the next step rewrites) and keeps the fastest:
//   ev >= 0: the fast KT flux tile shape of event ev per direction;
//   ev <  0: for the whole context, per direction the flux block size and one
//            tile shape for every event (or each event's whole eta line),
//            then the fusion depth, the persistent-block count and the XCD
//            tile order, each timed over a whole stage.
int viral_autotune(viral_ctx* ctx, int ev, int reps)
{
  if (!ctx || reps < 1 || ev >= (int)ctx->ev.size()) return fail(ctx, VIRAL_ERR_ARG, "bad autotune call");
  if (ctx->scheme == VIRAL_SHASTA)
    return fail(ctx, VIRAL_ERR_UNSUPPORTED, "SHASTA contexts keep the default launch (see viral_set_launch)");
  if (ctx->timing) return fail(ctx, VIRAL_ERR_ARG, "switch the timers off before autotuning");
  HIPCHK(hipStreamSynchronize(ctx->stream));
  ctx->h_clk->tau = ctx->last_tau;
  ctx->h_clk->ts = ctx->last_ts;
  ctx->h_clk->step = 0;
  HIPCHK(hipMemcpyAsync(ctx->d_clk, ctx->h_clk, sizeof(Clock), hipMemcpyHostToDevice, ctx->stream));
  static const int shapes[][2] = {{32, 8},  {32, 16}, {32, 32}, {64, 4},  {64, 8},   {64, 16},
                                  {128, 4}, {128, 8}, {256, 4}, {16, 32}, {16, 64}, {0, 0}};
  static const int nts[] = {64, 128, 256, 512};
  const int nev = (int)ctx->ev.size();
  const int cur = ctx->cur;
  if (ctx->fast) {
    for (int d = 0; d < 3; d++) {
      bool any = false;
      for (int e = 0; e < nev; e++) any = any || !(d == 2 && ctx->ev[e].d.lbi);
      if (!any) continue;
      int bnt = ctx->cfg.flux_nt[d];
      std::vector<int> btp(nev), btk(nev);
      for (int e = 0; e < nev; e++) {
        btp[e] = ctx->ev[e].tile_tp[d];
        btk[e] = ctx->ev[e].tile_tk[d];
      }
      // the configuration in place is the first candidate: tuning never
      // leaves a slower one behind
      float best = 1e30f;
      {
        ctx->dirty = true;
        if (int r = refresh_table(ctx)) return r;
        int r;
        if (ev >= 0) {
          Sched s{nullptr, ctx->cnt[PH_FLUX_X + d][ev], ev, 1, ctx->cfg.xcd};
          const size_t shm = kTileArrays * (size_t)ctx->ev[ev].tile_tp[d] * ctx->ev[ev].tile_tk[d] * sizeof(double);
          const unsigned grid = grid_for(ctx, s.total);
          r = VIRAL_OK;
          with_bool(ctx->cfg.blocks_per_cu > 0, [&](auto PT) {
            with_bool(ctx->cart, [&](auto CT) {
              r = time_ms(ctx, reps, &best, [&] {
                launch_dir_tiles<decltype(PT)::value, decltype(CT)::value>(ctx, d, bnt, s, grid, shm, 0, cur);
              });
            });
          });
        } else {
          r = time_ms(ctx, reps, &best, [&] { launch_flux(ctx, d, 0, cur); });
        }
        if (r) return r;
      }
      for (int nt : nts) {
        if (ev >= 0 && nt != ctx->cfg.flux_nt[d]) continue;  // one event: keep the context's block size
        for (auto& sh : shapes) {
          // candidate: every event (or just ev) gets shape sh; {0,0} = whole ghosted line along eta
          if (sh[0] == 0 && d != 2) continue;
          bool valid = true;
          for (int e = 0; e < nev; e++) {
            if (ev >= 0 && e != ev) continue;
            Event& E = ctx->ev[e];
            if (d == 2 && E.d.lbi) continue;  // no eta phase
            const int G = d == 0 ? E.g.gx : d == 1 ? E.g.gy : E.g.gz;
            int tp = sh[0], tk = sh[1];
            if (tp == 0) {  // as many whole lines as the block's slots hold
              tp = std::max(8, G);
              tk = std::max(1, max_slots(nt) / tp);
            }
            if (!tile_ok(tp, tk, nt) || (tp > G + 8 && sh[0] != 0)) {
              valid = false;
              break;
            }
            E.tile_tp[d] = tp;
            E.tile_tk[d] = tk;
          }
          if (!valid) continue;
          const int keep_nt = ctx->cfg.flux_nt[d];
          ctx->cfg.flux_nt[d] = nt;
          ctx->dirty = true;
          if (int r = refresh_table(ctx)) return r;
          float ms = 0;
          int r;
          if (ev >= 0) {
            Sched s{nullptr, ctx->cnt[PH_FLUX_X + d][ev], ev, 1, ctx->cfg.xcd};
            const size_t shm = kTileArrays * (size_t)ctx->ev[ev].tile_tp[d] * ctx->ev[ev].tile_tk[d] * sizeof(double);
            const unsigned grid = grid_for(ctx, s.total);
            r = VIRAL_OK;
            with_bool(ctx->cfg.blocks_per_cu > 0, [&](auto PT) {
              with_bool(ctx->cart, [&](auto CT) {
                r = time_ms(ctx, reps, &ms, [&] {
                  launch_dir_tiles<decltype(PT)::value, decltype(CT)::value>(ctx, d, nt, s, grid, shm, 0, cur);
                });
              });
            });
          } else {
            r = time_ms(ctx, reps, &ms, [&] { launch_flux(ctx, d, 0, cur); });
          }
          ctx->cfg.flux_nt[d] = keep_nt;
          if (r) return r;
          if (ms < best) {
            best = ms;
            bnt = nt;
            for (int e = 0; e < nev; e++) {
              btp[e] = ctx->ev[e].tile_tp[d];
              btk[e] = ctx->ev[e].tile_tk[d];
            }
          }
        }
      }
      ctx->cfg.flux_nt[d] = bnt;
      for (int e = 0; e < nev; e++) {
        ctx->ev[e].tile_tp[d] = btp[e];
        ctx->ev[e].tile_tk[d] = btk[e];
      }
      ctx->dirty = true;
    }
  }
  if (ev < 0) {
    auto stage_ms = [&](float* ms) -> int {
      ctx->dirty = true;
      if (int r = refresh_table(ctx)) return r;
      return time_ms(ctx, reps, ms, [&] { launch_stage(ctx, 0, cur); });
    };
    float best = 1e30f, ms = 0;
    int bf = ctx->cfg.fusion;
    for (int f = 0; f < 2; f++) {
      ctx->cfg.fusion = f;
      if (int r = stage_ms(&ms)) return r;
      if (ms < best) { best = ms; bf = f; }
    }
    ctx->cfg.fusion = bf;
    static const int bpcs[] = {0, 1, 2, 4, 8};
    best = 1e30f;
    int bb = ctx->cfg.blocks_per_cu;
    for (int b : bpcs) {
      ctx->cfg.blocks_per_cu = b;
      if (int r = stage_ms(&ms)) return r;
      if (ms < best) { best = ms; bb = b; }
    }
    ctx->cfg.blocks_per_cu = bb;
    best = 1e30f;
    int bx = ctx->cfg.xcd;
    for (int x = 0; x < 2; x++) {
      ctx->cfg.xcd = x;
      if (int r = stage_ms(&ms)) return r;
      if (ms < best) { best = ms; bx = x; }
    }
    ctx->cfg.xcd = bx;
  }
  ctx->dirty = true;
  if (int r = refresh_table(ctx)) return r;
  HIPCHK(hipStreamSynchronize(ctx->stream));
  return VIRAL_OK;
}

// ---------------------------------------------------------------- adjoint

int viral_adjoint_end(viral_ctx* ctx)
{
  if (!ctx) return VIRAL_ERR_ARG;
  HIPCHK(hipStreamSynchronize(ctx->stream));
  for (double* c : ctx->adj.ckpt) hipFree(c);
  for (double* c : ctx->adj.slabs) hipFree(c);
  if (ctx->adj.ring) hipFree(ctx->adj.ring);
  ctx->adj.ring = nullptr;
  ctx->adj.ckpt.clear();
  ctx->adj.slabs.clear();
  ctx->adj.stages.clear();
  ctx->adj.taus.clear();
  ctx->adj.tss.clear();
  ctx->adj.on = false;
  return VIRAL_OK;
}

int viral_adjoint_begin(viral_ctx* ctx, int every)
{
  if (!ctx || every < 0) return fail(ctx, VIRAL_ERR_ARG, "bad adjoint recording request");
  if (ctx->ev.size() != 1) return fail(ctx, VIRAL_ERR_UNSUPPORTED, "the adjoint runs on single-event contexts");
  const Event& E = ctx->ev[0];
  if (E.d.scheme != VIRAL_KT)
    return fail(ctx, VIRAL_ERR_UNSUPPORTED, "the adjoint covers the KT step (Milne and Cartesian)");
  if (E.d.quirks & VIRAL_Q_FIX_VE) return fail(ctx, VIRAL_ERR_UNSUPPORTED, "the adjoint does not cover the FIX build");
  if (E.d.coords == VIRAL_MILNE && (E.d.quirks & VIRAL_Q_K0_BROADCAST) && !E.d.lbi &&
      (E.g.k0 < E.g.kl || E.g.k0 >= E.g.kr))
    return fail(ctx, VIRAL_ERR_UNSUPPORTED, "the adjoint needs k0 inside the interior");
  if (int r = viral_adjoint_end(ctx)) return r;
  const size_t bytes = (size_t)NF * E.g.n * sizeof(double);
  double* c = nullptr;
  HIPCHK(hipMalloc(&c, bytes));
  ctx->adj.ckpt.push_back(c);
  HIPCHK(hipMemcpyAsync(c, E.st[ctx->cur], bytes, hipMemcpyDeviceToDevice, ctx->stream));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  if (every > 0 && hipMalloc(&ctx->adj.ring, (size_t)every * ctx->rk * bytes) != hipSuccess) {
    ctx->adj.ring = nullptr;
    viral_adjoint_end(ctx);
    return fail(ctx, VIRAL_ERR_HIP, "out of device memory for the recorded segment's stage states");
  }
  ctx->adj.every = every;
  ctx->adj.on = true;
  return VIRAL_OK;
}

int viral_adjoint(viral_ctx* ctx, const double* const* seed, double* const* grad, double* pgrad)
{
  if (!ctx || !seed || !grad || !pgrad) return fail(ctx, VIRAL_ERR_ARG, "bad adjoint arguments");
  if (!ctx->adj.on) return fail(ctx, VIRAL_ERR_ARG, "no recorded steps: call viral_adjoint_begin before stepping");
  if (int r = refresh_table(ctx)) return r;
  Event& E = ctx->ev[0];
  const Geom& g = E.g;
  const Params& p = E.p;
  const size_t n = (size_t)g.n;
  const size_t fbytes = (size_t)NF * n * sizeof(double);
  const int M = (int)ctx->adj.taus.size();
  const int every = ctx->adj.every;
  const int rk = ctx->rk;
  const long ni = box_count(interior_box(g));
  const unsigned nbi = nblocks(ni);
  // device buffers of this call; the guard frees them on every return path
  // (after draining the stream, so no kernel still reads them)
  const bool stored = ctx->adj.every == 0;  // every stage state kept: one segment, no recompute
  if (!stored && !ctx->stream2) HIPCHK(hipStreamCreateWithFlags(&ctx->stream2, hipStreamNonBlocking));
  struct Bufs {
    hipStream_t s, s2;
    std::vector<double*> v;
    std::vector<hipEvent_t> e;
    ~Bufs()
    {
      if (s2) hipStreamSynchronize(s2);
      hipStreamSynchronize(s);
      for (double* q : v) hipFree(q);
      for (hipEvent_t x : e) hipEventDestroy(x);
    }
  } bufs{ctx->stream, stored ? nullptr : ctx->stream2, {}, {}};
  auto dalloc = [&](size_t count) -> double* {
    double* q = nullptr;
    if (hipMalloc(&q, count * sizeof(double)) != hipSuccess) return nullptr;
    bufs.v.push_back(q);
    return q;
  };
  const int seglen = stored ? std::max(M, 1) : every;
  const int nsnap = stored ? 1 : std::min(every, std::max(M, 1)) * rk + 1;
  double* lamA = dalloc(NF * n);
  double* lamB = dalloc(NF * n);
  double* vb0 = dalloc(SV * n);
  double* vb1 = dalloc(SV * n);
  double* vb2 = dalloc(SV * n);
  double* resb = dalloc(SV * n);
  double* dubuf = dalloc(12 * n);
  double* partial = dalloc(2 * (size_t)nbi);
  double* saved = dalloc(NF * n);
  if (!lamA || !lamB || !vb0 || !vb1 || !vb2 || !resb || !dubuf || !partial || !saved)
    return fail(ctx, VIRAL_ERR_HIP, "out of device memory for the adjoint");
  for (int f = 0; f < NF; f++) {
    if (!seed[f] || !grad[f]) return fail(ctx, VIRAL_ERR_ARG, "null field pointer");
    HIPCHK(hipMemcpyAsync(lamA + f * n, seed[f], n * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
  }
  HIPCHK(hipMemsetAsync(partial, 0, 2 * (size_t)nbi * sizeof(double), ctx->stream));
  HIPCHK(hipMemsetAsync(vb0, 0, SV * n * sizeof(double), ctx->stream));
  HIPCHK(hipMemsetAsync(vb1, 0, SV * n * sizeof(double), ctx->stream));
  HIPCHK(hipMemsetAsync(vb2, 0, SV * n * sizeof(double), ctx->stream));
  HIPCHK(hipMemcpyAsync(saved, E.st[ctx->cur], fbytes, hipMemcpyDeviceToDevice, ctx->stream));
  const bool cart = ctx->cart != 0;
  const bool k0b = !cart && (p.quirks & VIRAL_Q_K0_BROADCAST) && !p.lbi;  // TXYZ5 has no k0 broadcast
  const int ndir = p.lbi ? 2 : 3;
  std::vector<const double*> sp;  // the segment's stage states: sp[idx] feeds stage idx
  const int nseg = M > 0 ? (M - 1) / seglen + 1 : 0;
  // Checkpointed mode: the recording's ring holds the stage states of the
  // last segment.  Every earlier segment j is recomputed on the second stream
  // into snapshot buffer j % 2 while the reverse sweep of segment j + 1 runs
  // on the context stream; ev_rec[b] marks buffer b's recompute done, ev_rev[b]
  // the reverse sweep that last read it.  The recompute alone uses the device
  // table and clock (the reverse kernels take their pointers and clock as
  // arguments), so the two streams share no written memory.
  const int nrec = stored ? 0 : std::max(nseg - 1, 0);  // segments to recompute
  double* sbuf[2] = {nullptr, nullptr};
  hipEvent_t ev_rec[2] = {}, ev_rev[2] = {}, ev_join = nullptr;
  if (nrec > 0) {
    for (int b = 0; b < std::min(nrec, 2); b++)
      if (!(sbuf[b] = dalloc((size_t)nsnap * NF * n)))
        return fail(ctx, VIRAL_ERR_HIP, "out of device memory for the adjoint");
    for (hipEvent_t* x : {&ev_rec[0], &ev_rec[1], &ev_rev[0], &ev_rev[1], &ev_join}) {
      HIPCHK(hipEventCreateWithFlags(x, hipEventDisableTiming));
      bufs.e.push_back(*x);
    }
  }
  // every stage reads snapshot idx and writes snapshot idx + 1 in place:
  // the table row points the stage's two state copies there (cells no stage
  // writes, the corner ghosts, are never read)
  auto recompute = [&](int seg) -> int {
    const int m0 = seg * seglen, m1 = std::min(m0 + seglen, M);
    double* buf = sbuf[seg & 1];
    struct Swap {
      viral_ctx* c;
      hipStream_t keep;
      ~Swap() { c->stream = keep; }
    } swap{ctx, ctx->stream};
    ctx->stream = ctx->stream2;  // launch_stage launches on the context's stream member
    HIPCHK(hipMemcpyAsync(buf, ctx->adj.ckpt[seg], fbytes, hipMemcpyDeviceToDevice, ctx->stream));
    int idx = 0;
    for (int m = m0; m < m1; m++) {
      hipLaunchKernelGGL(k_set_clock, dim3(1), dim3(64), 0, ctx->stream, ctx->d_clk, ctx->adj.taus[m],
                         ctx->adj.tss[m]);
      for (int st = 0; st < rk; st++, idx++) {
        hipLaunchKernelGGL(k_set_state_ptrs, dim3(1), dim3(64), 0, ctx->stream, ctx->d_tab, 0,
                           buf + (size_t)idx * NF * n, buf + (size_t)(idx + 1) * NF * n);
        launch_stage(ctx, st, 0);
      }
    }
    hipLaunchKernelGGL(k_set_state_ptrs, dim3(1), dim3(64), 0, ctx->stream, ctx->d_tab, 0, E.st[0], E.st[1]);
    HIPCHK(hipGetLastError());
    HIPCHK(hipEventRecord(ev_rec[seg & 1], ctx->stream));
    return VIRAL_OK;
  };
  auto adj_timer = std::make_unique<KTimer>(ctx, K_ADJ);  // stops before the gradient's download
  if (nrec > 0) {
    // the recompute starts after the seed's upload and the timer's start
    HIPCHK(hipEventRecord(ev_join, ctx->stream));
    HIPCHK(hipStreamWaitEvent(ctx->stream2, ev_join, 0));
  }
  for (int seg = nseg - 1; seg >= 0; seg--) {
    const int m0 = seg * seglen, m1 = std::min(m0 + seglen, M);
    sp.clear();
    if (stored) {
      sp.push_back(ctx->adj.ckpt[0]);
      sp.insert(sp.end(), ctx->adj.stages.begin(), ctx->adj.stages.end());
    } else if (seg == nseg - 1) {
      // the last segment: its checkpoint, then the ring's stage outputs
      sp.push_back(ctx->adj.ckpt[seg]);
      for (int q = 0; q < (m1 - m0) * rk; q++) sp.push_back(ctx->adj.ring + (size_t)q * NF * n);
    } else {
      for (int q = 0; q <= (m1 - m0) * rk; q++) sp.push_back(sbuf[seg & 1] + (size_t)q * NF * n);
      HIPCHK(hipStreamWaitEvent(ctx->stream, ev_rec[seg & 1], 0));
    }
    if (!stored && seg > 0) {
      // buffer (seg - 1) % 2 was last read by the reverse sweep of seg + 1,
      // if that segment was recomputed
      if (seg + 1 < nseg - 1) HIPCHK(hipStreamWaitEvent(ctx->stream2, ev_rev[(seg + 1) & 1], 0));
      if (int r = recompute(seg - 1)) return r;
    }
    // reverse sweep over the segment
    for (int m = m1 - 1; m >= m0; m--) {
      // vb0..vb2 are zero here: cleared once below before the first segment,
      // and k_adj_root's first-stage (and RK3 middle-stage) reverse zeroes
      // what it consumes, so each reversed step leaves them zero
      for (int st = rk - 1; st >= 0; st--) {
        const size_t base = (size_t)((m - m0) * rk + st);
        const double* si = sp[base];
        const double* so = sp[base + 1];
        const StageArgs a = stage_args(ctx->adj.taus[m], ctx->adj.tss[m], rk, st);
        // ghost fill in reverse, one face pair at a time (x, y, eta) so that cells
        // next to two faces sum in a fixed order
        {
          const long nx_ = g.ir - g.il, ny_ = g.jr - g.jl, nz_ = g.kr - g.kl;
          hipLaunchKernelGGL(k_adj_ghost<0>, dim3(nblocks(ny_ * nz_ * 2 * 28)), dim3(BLK), 0, ctx->stream, g, p, lamA);
          hipLaunchKernelGGL(k_adj_ghost<1>, dim3(nblocks(nx_ * nz_ * 2 * 28)), dim3(BLK), 0, ctx->stream, g, p, lamA);
          if (!p.lbi)
            hipLaunchKernelGGL(k_adj_ghost<2>, dim3(nblocks(nx_ * ny_ * 2 * 28)), dim3(BLK), 0, ctx->stream, g, p,
                               lamA);
        }
        if (k0b) {
          const long nl = (long)(g.ir - g.il) * (g.jr - g.jl);
          hipLaunchKernelGGL(k_adj_k0, dim3(nblocks(nl * 64)), dim3(BLK), 0, ctx->stream, g, p, a.tau_new, so, lamA);
        }
        HIPCHK(hipMemsetAsync(lamB, 0, fbytes, ctx->stream));
        hipLaunchKernelGGL(k_adj_pass, dim3(nblocks(g.n)), dim3(BLK), 0, ctx->stream, g, lamA, lamB);
        with_bool(cart, [&](auto CT) {
          constexpr bool CART = decltype(CT)::value;
          hipLaunchKernelGGL(k_adj_root<CART>, dim3(nbi), dim3(BLK), 0, ctx->stream, g, p, a, so, lamA, lamB, vb0,
                             vb1, vb2, resb);
          if (p.bulk)
            hipLaunchKernelGGL((k_adj_src<CART, true>), dim3(nbi), dim3(BLK), 0, ctx->stream, g, p, a, si, resb, lamB,
                               dubuf, partial);
          else
            hipLaunchKernelGGL((k_adj_src<CART, false>), dim3(nbi), dim3(BLK), 0, ctx->stream, g, p, a, si, resb, lamB,
                               dubuf, partial);
        });
        const double ttile = cart ? 1.0 : a.tau_src;  // the Cartesian flux has no tau factors
        for (int d = 0; d < ndir; d++) {
          // lines of direction d (interior in the other two), TK per block, each cut into nseg segments
          // of W - 10 own cells; (nseg, W, TK) minimises blocks per line, nseg / TK, with at least
          // 8 lines per block across eta (the lines' loads are TK consecutive eta cells)
          const long nlines = d == 0 ? (long)(g.jr - g.jl) * (g.kr - g.kl)
                              : d == 1 ? (long)(g.ir - g.il) * (g.kr - g.kl)
                                       : (long)(g.ir - g.il) * (g.jr - g.jl);
          const int GD = d == 0 ? g.gx : d == 1 ? g.gy : g.gz;
          int bs = 0, bW = 0, bT = 0;
          for (int ns = 1; ns <= GD; ns++) {
            const int W = (GD + ns - 1) / ns + 10;
            if (W > NTA) continue;
            const int TK = NTA / W;
            if (d < 2 && TK < 8 && W > 18) continue;
            if (!bs || (long)ns * bT < (long)bs * TK) {
              bs = ns;
              bW = W;
              bT = TK;
            }
            if (W <= 16) break;
          }
          const unsigned nt = (unsigned)(((nlines + bT - 1) / bT) * bs);
          with_bool(cart, [&](auto CT) {
            constexpr bool CART = decltype(CT)::value;
            if (d == 0)
              hipLaunchKernelGGL((k_adj_tile<0, CART>), dim3(nt), dim3(NTA), 0, ctx->stream, g, p, ttile, si, resb,
                                 dubuf, lamB, bW, bT, bs);
            if (d == 1)
              hipLaunchKernelGGL((k_adj_tile<1, CART>), dim3(nt), dim3(NTA), 0, ctx->stream, g, p, ttile, si, resb,
                                 dubuf, lamB, bW, bT, bs);
            if (d == 2)
              hipLaunchKernelGGL((k_adj_tile<2, CART>), dim3(nt), dim3(NTA), 0, ctx->stream, g, p, ttile, si, resb,
                                 dubuf, lamB, bW, bT, bs);
          });
        }
        std::swap(lamA, lamB);
      }
    }
    if (nrec > 0) HIPCHK(hipEventRecord(ev_rev[seg & 1], ctx->stream));
  }
  if (nrec > 0) {
    // join the recompute stream (its last work precedes the last reverse sweep already)
    HIPCHK(hipEventRecord(ev_join, ctx->stream2));
    HIPCHK(hipStreamWaitEvent(ctx->stream, ev_join, 0));
  }
  adj_timer.reset();
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(E.st[ctx->cur], saved, fbytes, hipMemcpyDeviceToDevice, ctx->stream));
  HIPCHK(hipMemcpyAsync(E.st[ctx->cur ^ 1], saved, fbytes, hipMemcpyDeviceToDevice, ctx->stream));
  for (int f = 0; f < NF; f++)
    HIPCHK(hipMemcpyAsync(grad[f], lamA + f * n, n * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  std::vector<double> part(2 * (size_t)nbi);
  HIPCHK(hipMemcpyAsync(part.data(), partial, part.size() * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  double s0 = 0, s1 = 0;
  for (unsigned b = 0; b < nbi; b++) {
    s0 += part[2 * b];
    s1 += part[2 * b + 1];
  }
  pgrad[0] = s0;
  pgrad[1] = s1;
  return VIRAL_OK;
}

int viral_kernel_count(void) { return K_COUNT; }
const char* viral_kernel_name(int kernel) { return (kernel >= 0 && kernel < K_COUNT) ? kKernelNames[kernel] : ""; }

int viral_timers(viral_ctx* ctx, int enable)
{
  if (!ctx) return VIRAL_ERR_ARG;
  HIPCHK(hipStreamSynchronize(ctx->stream));
  for (int k = 0; k < K_COUNT; k++) {
    for (hipEvent_t e : ctx->tev[k]) ctx->pool.push_back(e);
    ctx->tev[k].clear();
  }
  ctx->timing = enable != 0;
  return VIRAL_OK;
}

int viral_timer_read(viral_ctx* ctx, int kernel, double* total_ms, long long* launches)
{
  if (!ctx || kernel < 0 || kernel >= K_COUNT || !total_ms || !launches) return fail(ctx, VIRAL_ERR_ARG, "bad timer");
  HIPCHK(hipStreamSynchronize(ctx->stream));
  double t = 0;
  const std::vector<hipEvent_t>& v = ctx->tev[kernel];
  for (size_t q = 0; q + 1 < v.size(); q += 2) {
    float ms = 0;
    HIPCHK(hipEventElapsedTime(&ms, v[q], v[q + 1]));
    t += ms;
  }
  *total_ms = t;
  *launches = (long long)(v.size() / 2);
  return VIRAL_OK;
}

}  // extern "C"
