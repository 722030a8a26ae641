// poisson.hip -- MI355X (gfx950) pressure-Poisson solver: HIP kernels and
// the C ABI of include/cup2d_poisson.h.
//
// Method (reference source/Poisson/BiCGSTAB.cu:412-620): right-preconditioned
// BiCGSTAB with the dense 64x64 block-Jacobi inverse P_inv = -K^-1
// (ExpAMRSolver.cpp:25-95), restart on breakdown, best iterate x_opt, then
// removal of the h^2-weighted mean (ExpAMRSolver.cpp:383-415).
//
// MI355X design:
//  * one wavefront per 8x8 block in the operator kernels (lane = cell); the
//    5-point part of the AMR operator is applied from a per-block neighbour
//    table, only coarse-fine rows are read from a sparse list;
//  * the vector passes of one iteration are fused into five kernels; every
//    scalar (alpha, omega, beta, restart, convergence) lives on the device;
//    a reduction is finished by its consumer kernel, which re-reduces the
//    producer's per-workgroup partials (no device-wide fences or atomics);
//  * iterations are replayed in chunks from a hipGraph; kernels of a
//    finished solve exit at launch on the device-side `done` flag;
//  * mixed precision: FP64 iterative refinement (r = b - A x and x += d in
//    FP64) around an inner BiCGSTAB with FP32 vectors and SpMV, FP64
//    accumulation of every inner product, and the block preconditioner as a
//    BF16 MFMA contraction with FP32 accumulation.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/cup2d_poisson.h"
#include "assemble_device.h"
#include "assembly.h"
#include "kernels.h"
#include "flow.h"

using namespace cup2d;

namespace {

constexpr int MAXG = 512;  // workgroups of the block part of the vector kernels (grid-stride beyond)
constexpr int MAXGE = 64;  // workgroups of the explicit-row part
// workgroups of the FP64 preconditioner kernels (grid-stride over tiles of
// TILE blocks): a workgroup loads its P fragments and reduces the previous
// kernel's partials once for all of its tiles, which on meshes of 1k+ tiles
// pays more than one tile per workgroup (6-level 16k-block mesh: the p
// update 18 -> 13 us with 512 instead of 2048; 256 alike)
constexpr int MAXGP = 512;
constexpr int SPMV_U = 2;  // blocks per wavefront in flight in the operator kernels

// ----------------------------------------------------------------------------
// scalar recurrence
// ----------------------------------------------------------------------------

// Evaluates the state of this iteration in every workgroup (identically);
// workgroup 0 stores it to slot `par`.
// The previous slot is copied to LDS (a word per thread) while the partials
// are reduced, so the evaluation waits on no global load.
__device__ __forceinline__ void stage_scal(const Scal &src, Scal &dst) {
  static_assert(sizeof(Scal) % 8 == 0 && sizeof(Scal) / 8 <= WG, "Scal staging");
  if (threadIdx.x < sizeof(Scal) / 8)
    reinterpret_cast<unsigned long long *>(&dst)[threadIdx.x] =
        reinterpret_cast<const unsigned long long *>(&src)[threadIdx.x];
}

__device__ __forceinline__ void top_of_iteration(Scal *st, int par, const double *p5, int G5,
                                                 Scal &cur) {
  __shared__ Scal prev;
  stage_scal(st[1 - par], prev);
  double o[NQ5];
  load_totals<NQ5, 1u>(p5, G5, o);  // its barrier publishes prev
  if (threadIdx.x == 0) cur = evaluate(prev, o);
  __syncthreads();
  if (blockIdx.x == 0 && threadIdx.x == 0) st[par] = cur;
}

// ----------------------------------------------------------------------------
// initialisation
// ----------------------------------------------------------------------------

// Partial sums of h^2 x (the mean row of A x when the mean constraint is on).
template <class T>
__global__ __launch_bounds__(WG) void k_hsum(int nb, const T *__restrict__ x,
                                             const double *__restrict__ h2, double *pm) {
  const long n = (long)nb * 64;
  double v[1] = {0.0};
  for (long i = (long)blockIdx.x * WG + threadIdx.x; i < n; i += (long)gridDim.x * WG)
    v[0] += h2[i >> 6] * (double)x[i];
  store_partials<1, 0u>(v, pm);
}

// r = b - A x0 in FP64, stored as TR to r and rhat; p = nu = 0; x = x_opt =
// x0 (x may alias x0) (BiCGSTAB.cu:431-465); partials of the update totals
// of the stored r.  Workgroups below Gb take the blocks (U per wave), the
// others the explicit rows.
template <class TR>
__global__ __launch_bounds__(WG) void k_init(OpView A, int Gb, const double *__restrict__ b,
                                             const double *x0, double *x,
                                             double *__restrict__ xopt, TR *__restrict__ r,
                                             TR *__restrict__ rhat, TR *__restrict__ p,
                                             TR *__restrict__ nu, const double *pm, int Gm,
                                             double *p5) {
  double m0[1] = {0.0};
  if (A.mean_row >= 0) load_totals<1, 0u>(pm, Gm, m0);
  double v[NQ5] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  // mean row: b = 0 (ExpAMRSolver.cpp:340-345) and (A x)_row = sum h^2 x
  auto store_r = [&](long i, double y, double bi) {
    const double ri = (i == A.mean_row) ? 0.0 - m0[0] : bi - y;
    const TR rt = (TR)ri;
    r[i] = rt;
    rhat[i] = rt;
    const double rd = (double)rt;
    v[0] = nmax(v[0], fabs(rd));
    v[1] += rd * rd;
    v[4] += rd;
  };
  if ((int)blockIdx.x < Gb) {
    const int c = lane_id();
    for (int g = blockIdx.x * WAVES + wave_id(); g * SPMV_U < A.nb; g += Gb * WAVES) {
      int bl[SPMV_U];
      bool ok[SPMV_U];
#pragma unroll
      for (int u = 0; u < SPMV_U; u++) {
        ok[u] = g * SPMV_U + u < A.nb;
        bl[u] = ok[u] ? g * SPMV_U + u : g * SPMV_U;
      }
      BlockRows<SPMV_U, double> R;
      R.load(A, x0, b, bl, c);
#pragma unroll
      for (int u = 0; u < SPMV_U; u++) {
        if (!ok[u]) continue;
        bool plain;
        const double y = R.y(u, c, plain);
        const long i = R.i(u, c);
        p[i] = TR(0);
        nu[i] = TR(0);
        x[i] = R.xc(u);
        xopt[i] = R.xc(u);
        if (plain) store_r(i, y, R.a(u));
      }
    }
  } else {
    const int ge = gridDim.x - Gb;
    for (int e = (blockIdx.x - Gb) * WG + threadIdx.x; e < A.n_exc; e += ge * WG) {
      const int row = A.erow[e];
      store_r(row, explicit_row<double>(A, x0, e, ell_run_width(A, e)), b[row]);
    }
  }
  v[2] = v[1];
  v[3] = v[1];
  v[5] = v[4];
  store_partials<NQ5, 1u>(v, p5);
}

// ----------------------------------------------------------------------------
// iteration kernels, FP64 (MFMA preconditioner: a workgroup per tile of 16
// blocks, a wave per 16 output rows, see tile_prec64)
// ----------------------------------------------------------------------------

// The thread's four cells of tile t: first index and whether its block exists
// (a lane past the last block addresses the tile's first block).
__device__ __forceinline__ long tile_cells(int t, int nb, int w, int lane, bool &ok) {
  const int b = t * TILE + (lane & 15);
  ok = b < nb;
  return (long)(ok ? b : t * TILE) * 64 + cell4(w, lane);
}

// Top of iteration: p = r + beta (p - omega nu) (or p = rhat = r after a
// restart), z1 = P p (BiCGSTAB.cu:471-520).  The x_opt copy of the
// iteration's state is made by k_xrupdate, which reads x anyway.  The P
// fragments and the first tile's loads are issued before the state is
// evaluated.
__global__ __launch_bounds__(WG) void k_pupdate(int nb, const double *__restrict__ pf,
                                                const double *__restrict__ r,
                                                double *__restrict__ rhat, double *__restrict__ p,
                                                const double *__restrict__ nu,
                                                double *__restrict__ z1, Scal *st, int par,
                                                const double *p5, int G5) {
  __shared__ Scal cur;
  __shared__ double tile[TILE][TROW];
  const int lane = lane_id(), w = wave_id();
  const int ntiles = (nb + TILE - 1) / TILE;
  PFrag64 F;
  load_pfrag64(F, pf, w, lane);
  int t = blockIdx.x;
  bool ok = false;
  long i0 = 0;
  double rv[4], pv[4], nv[4];
  auto load = [&](int tt) {
    i0 = tile_cells(tt, nb, w, lane, ok);
    ld4(r, i0, rv);
    ld4(p, i0, pv);
    ld4(nu, i0, nv);
  };
  if (t < ntiles) load(t);
  top_of_iteration(st, par, p5, G5, cur);
  if (cur.done) return;
  const double beta = cur.beta, omega = cur.omega;
  const int restart = cur.restart;
  for (; t < ntiles; t += gridDim.x) {
    double pi[4];
#pragma unroll
    for (int k = 0; k < 4; k++) pi[k] = restart ? rv[k] : rv[k] + beta * (pv[k] - omega * nv[k]);
    if (ok) {
      if (restart) st4(rhat, i0, rv);
      st4(p, i0, pi);
    }
    tile_put(tile, w, lane, ok, pi);
    const bool oo = ok;
    const long io = i0;
    if (t + (int)gridDim.x < ntiles) load(t + gridDim.x);  // in flight during the product
    __syncthreads();
    const f64x4 z = tile_prec64(tile, F, lane);
    __syncthreads();  // the tile is rewritten by the next pass
    if (oo) st4(z1, io, z);
  }
}

// alpha = rho / ((rhat, nu) + eps); s = r - alpha nu (stored in r); z2 = P s
// (BiCGSTAB.cu:527-545).  Loads of P and of the first tile first.
__global__ __launch_bounds__(WG) void k_supdate(int nb, const double *__restrict__ pf,
                                                double *__restrict__ r, double *__restrict__ nu,
                                                const double *__restrict__ rhat,
                                                double *__restrict__ z2, Scal *st, int par,
                                                const double *p2, int G2, long mean_row) {
  __shared__ double tile[TILE][TROW];
  const int lane = lane_id(), w = wave_id();
  const int ntiles = (nb + TILE - 1) / TILE;
  PFrag64 F;
  load_pfrag64(F, pf, w, lane);
  int t = blockIdx.x;
  bool ok = false;
  long i0 = 0;
  double rv[4], nv[4];
  auto load = [&](int tt) {
    i0 = tile_cells(tt, nb, w, lane, ok);
    ld4(r, i0, rv);
    ld4(nu, i0, nv);
  };
  if (t < ntiles) load(t);
  const double rh_mr = mean_row >= 0 ? rhat[mean_row] : 0.0;
  const int done = st[par].done;  // state read before the reduction's barrier
  const double rho = st[par].rho, eps = st[par].eps;
  double o[2];
  load_totals<2, 0u>(p2, G2, o);
  if (done) return;
  const double rn = o[0] + (mean_row >= 0 ? rh_mr * o[1] : 0.0);
  const double alpha = rho / (rn + eps);
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    st[par].alpha = alpha;
    st[par].red[2] = rn;
  }
  for (; t < ntiles; t += gridDim.x) {
    double si[4];
#pragma unroll
    for (int k = 0; k < 4; k++) {
      if (i0 + k == mean_row) {
        nv[k] = o[1];
        if (ok) nu[mean_row] = nv[k];
      }
      si[k] = rv[k] - alpha * nv[k];
    }
    if (ok) st4(r, i0, si);
    tile_put(tile, w, lane, ok, si);
    const bool oo = ok;
    const long io = i0;
    if (t + (int)gridDim.x < ntiles) load(t + gridDim.x);  // in flight during the product
    __syncthreads();
    const f64x4 z = tile_prec64(tile, F, lane);
    __syncthreads();
    if (oo) st4(z2, io, z);
  }
}

// ----------------------------------------------------------------------------
// iteration kernels, FP32 inner solve (BF16 MFMA preconditioner)
// ----------------------------------------------------------------------------

// Top of iteration (FP32 storage): r -= mu (null-space projection),
// p = r + beta (p - omega nu) or p = rhat = r after a restart, z1 = P p on
// MFMA   (BiCGSTAB.cu:471-520).  The P fragments and the wave's first tile are
// loaded before the state is evaluated.
__global__ __launch_bounds__(WG) void k_pupdate_mf(int nb, const bf16x8 *__restrict__ pf,
                                                   float *__restrict__ r,
                                                   float *__restrict__ rhat, float *__restrict__ p,
                                                   const float *__restrict__ nu,
                                                   float *__restrict__ z1, Scal *st, int par,
                                                   const double *p5, int G5) {
  __shared__ Scal cur;
  const int lane = lane_id(), w = wave_id(), q = lane >> 4, j = lane & 15;
  const int ntiles = (nb + TILE - 1) / TILE, nw = gridDim.x * WAVES;
  PFrag P;
  load_pfrag(P, pf, lane);
  int tile = blockIdx.x * WAVES + w;
  bool ok = false;
  long base = 0;
  MfIn in;
  auto load_tile = [&](int t) {
    ok = t * TILE + j < nb;
    base = (long)(ok ? t * TILE + j : t * TILE) * 64;
    mf_load(r, p, nu, base, q, in);
  };
  if (tile < ntiles) load_tile(tile);
  top_of_iteration(st, par, p5, G5, cur);
  if (cur.done) return;
  const double beta = cur.beta, omega = cur.omega, mu = cur.mu;
  const int restart = cur.restart;
  for (; tile < ntiles; tile += nw) {
    float pin[16];
    mf_pupdate(in, r, rhat, p, base, q, ok, beta, omega, mu, restart, pin);
    const bool oo = ok;
    const long bo = base;
    if (tile + nw < ntiles) load_tile(tile + nw);  // in flight during the product
    float z[16];
    prec_mfma(P, pin, z);
    if (oo) st16_out(z1, bo, q, z);
  }
}

// alpha = rho / ((rhat, nu) + eps); s = r - alpha nu (stored in r); z2 = P s
// on MFMA (BiCGSTAB.cu:527-545).  Loads first.
__global__ __launch_bounds__(WG) void k_supdate_mf(int nb, const bf16x8 *__restrict__ pf,
                                                   float *__restrict__ r, float *__restrict__ nu,
                                                   const float *__restrict__ rhat,
                                                   float *__restrict__ z2, Scal *st, int par,
                                                   const double *p2, int G2, long mean_row) {
  const int lane = lane_id(), w = wave_id(), q = lane >> 4, j = lane & 15;
  const int ntiles = (nb + TILE - 1) / TILE, nw = gridDim.x * WAVES;
  PFrag P;
  load_pfrag(P, pf, lane);
  int tile = blockIdx.x * WAVES + w;
  bool ok = false;
  long base = 0;
  float rv[16], nv[16];
  auto load_tile = [&](int t) {
    ok = t * TILE + j < nb;
    base = (long)(ok ? t * TILE + j : t * TILE) * 64;
    ld16(r, base, q, rv);
    ld16(nu, base, q, nv);
  };
  if (tile < ntiles) load_tile(tile);
  const double rh_mr = mean_row >= 0 ? (double)rhat[mean_row] : 0.0;
  const int done = st[par].done;  // state read before the reduction's barrier
  const double rho = st[par].rho, eps = st[par].eps;
  double o[2];
  load_totals<2, 0u>(p2, G2, o);
  if (done) return;
  const double rn = o[0] + (mean_row >= 0 ? rh_mr * o[1] : 0.0);
  const double alpha = rho / (rn + eps);
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    st[par].alpha = alpha;
    st[par].red[2] = rn;
  }
  for (; tile < ntiles; tile += nw) {
    if (ok && mean_row >= base && mean_row < base + 64) {
#pragma unroll
      for (int k = 0; k < 16; k++)
        if (base + in_cell(q, k) == mean_row) {
          nv[k] = (float)o[1];
          nu[mean_row] = nv[k];
        }
    }
    float sv[16];
#pragma unroll
    for (int k = 0; k < 16; k++) sv[k] = ok ? (float)((double)rv[k] - alpha * (double)nv[k]) : 0.f;
    if (ok) st16_in(r, base, q, sv);
    const bool oo = ok;
    const long bo = base;
    if (tile + nw < ntiles) load_tile(tile + nw);  // in flight during the product
    float z[16];
    prec_mfma(P, sv, z);
    if (oo) st16_out(z2, bo, q, z);
  }
}

// ----------------------------------------------------------------------------
// iteration kernels shared by both precisions
// ----------------------------------------------------------------------------

// State evaluation alone (end of an iteration chunk, one workgroup), so the
// host sees convergence of the chunk's last iteration.
__global__ __launch_bounds__(WG) void k_eval(Scal *st, int par, const double *p5, int G5) {
  double o[NQ5];
  __shared__ Scal prev;
  stage_scal(st[1 - par], prev);
  load_totals<NQ5, 1u>(p5, G5, o);
  if (threadIdx.x == 0) st[par] = evaluate(prev, o);
}

// Runs f(R, u) for the blocks of the wave in the block part of an operator
// kernel (SPMV_U blocks per wave at a time, grid-stride over Gb
// workgroups); the next group's loads are issued before the current group
// is processed.  Returns false, having only issued the first loads, when
// `stop` holds.
template <class TV, class F>
__device__ __forceinline__ bool block_groups(const OpView &A, int Gb, const TV *__restrict__ x,
                                             const TV *__restrict__ aux, bool stop, F &&f) {
  const int c = lane_id(), stride = Gb * WAVES;
  int g = blockIdx.x * WAVES + wave_id();
  auto index = [&](int gg, int (&bl)[SPMV_U], bool (&ok)[SPMV_U]) {
#pragma unroll
    for (int u = 0; u < SPMV_U; u++) {
      ok[u] = gg * SPMV_U + u < A.nb;
      bl[u] = ok[u] ? gg * SPMV_U + u : gg * SPMV_U;
    }
  };
  if (g * SPMV_U >= A.nb) return !stop;
  int bl[SPMV_U];
  bool ok[SPMV_U];
  index(g, bl, ok);
  BlockRows<SPMV_U, TV> R;
  R.load(A, x, aux, bl, c);
  if (stop) return false;
  for (;;) {
    const int gn = g + stride;
    const bool more = gn * SPMV_U < A.nb;
    int bln[SPMV_U];
    bool okn[SPMV_U];
    BlockRows<SPMV_U, TV> Rn;
    if (more) {
      index(gn, bln, okn);
      Rn.load(A, x, aux, bln, c);
    }
#pragma unroll
    for (int u = 0; u < SPMV_U; u++)
      if (ok[u]) f(R, u);
    if (!more) break;
    R = Rn;
#pragma unroll
    for (int u = 0; u < SPMV_U; u++) ok[u] = okn[u];
    g = gn;
  }
  return true;
}

// nu = A z1; partials of (rhat, nu) and sum h^2 z1   (BiCGSTAB.cu:523-527)
template <class T>
__global__ __launch_bounds__(WG) void k_spmv_nu(OpView A, int Gb, const T *__restrict__ z1,
                                                T *__restrict__ nu, const T *__restrict__ rhat,
                                                const Scal *st, int par, double *p2) {
  const int done = st[par].done;  // checked once the first loads are issued
  double v[2] = {0.0, 0.0};
  auto store_y = [&](long i, double y, double rh) {
    if (i == A.mean_row) y = 0.0;  // replaced by the consumer
    const T yt = (T)y;
    nu[i] = yt;
    v[0] += rh * (double)yt;
  };
  if ((int)blockIdx.x < Gb) {
    const int c = lane_id();
    if (!block_groups<T>(A, Gb, z1, rhat, done, [&](const BlockRows<SPMV_U, T> &R, int u) {
          bool plain;
          const double y = R.y(u, c, plain);
          if (plain) store_y(R.i(u, c), y, R.a(u));
          v[1] += R.hh[u] * R.xc(u);
        }))
      return;
  } else {
    const int ge = gridDim.x - Gb;
    if (done) return;
    for (int e = (blockIdx.x - Gb) * WG + threadIdx.x; e < A.n_exc; e += ge * WG) {
      const int row = A.erow[e];
      store_y(row, explicit_row<T>(A, z1, e, ell_run_width(A, e)), (double)rhat[row]);
    }
  }
  if (done) return;
  store_partials<2, 0u>(v, p2);
}

// t = A z2; partials of (t, s), (t, t), sum h^2 z2 and, from workgroup 0,
// s at the mean row (BiCGSTAB.cu:548-553).  s is stable during this kernel,
// unlike in k_xrupdate, which overwrites it.
template <class T>
__global__ __launch_bounds__(WG) void k_spmv_t(OpView A, int Gb, const T *__restrict__ z2,
                                               T *__restrict__ t, const T *__restrict__ sv,
                                               const Scal *st, int par, double *p4) {
  const int done = st[par].done;  // checked once the first loads are issued
  double v[4] = {0.0, 0.0, 0.0, 0.0};
  if (A.mean_row >= 0 && blockIdx.x == 0 && threadIdx.x == 0) v[3] = (double)sv[A.mean_row];
  auto store_y = [&](long i, double y, double si) {
    if (i == A.mean_row) y = 0.0;
    const T yt = (T)y;
    t[i] = yt;
    const double yd = (double)yt;
    v[0] += yd * si;
    v[1] += yd * yd;
  };
  if ((int)blockIdx.x < Gb) {
    const int c = lane_id();
    if (!block_groups<T>(A, Gb, z2, sv, done, [&](const BlockRows<SPMV_U, T> &R, int u) {
          bool plain;
          const double y = R.y(u, c, plain);
          if (plain) store_y(R.i(u, c), y, R.a(u));
          v[2] += R.hh[u] * R.xc(u);
        }))
      return;
  } else {
    const int ge = gridDim.x - Gb;
    if (done) return;
    for (int e = (blockIdx.x - Gb) * WG + threadIdx.x; e < A.n_exc; e += ge * WG) {
      const int row = A.erow[e];
      store_y(row, explicit_row<T>(A, z2, e, ell_run_width(A, e)), (double)sv[row]);
    }
  }
  if (done) return;
  store_partials<4, 0u>(v, p4);
}

// omega = (t, s) / ((t, t) + eps); x_opt = x when the previous iterate was
// the best so far; x += alpha z1 + omega z2 (x in FP64); r = s - omega t;
// partials of the update totals   (BiCGSTAB.cu:501-503, 552-572)
template <class T>
__global__ __launch_bounds__(WG) void k_xrupdate(int nb, double *__restrict__ x,
                                                 double *__restrict__ xopt,
                                                 const T *__restrict__ z1,
                                                 const T *__restrict__ z2, T *__restrict__ r,
                                                 const T *__restrict__ t,
                                                 const T *__restrict__ rhat, Scal *st, int par,
                                                 const double *p4, int G4, double *p5,
                                                 long mean_row) {
  const long n = (long)nb * 64;
  // the first cell's loads are issued before the scalars are reduced
  long i = (long)blockIdx.x * WG + threadIdx.x;
  const long i0 = i < n ? i : 0;
  double xo = x[i0], a1 = (double)z1[i0], a2 = (double)z2[i0], rr = (double)r[i0],
         ti = (double)t[i0], rh = (double)rhat[i0];
  const int done = st[par].done;  // state read before the reduction's barrier
  const double eps = st[par].eps, alpha = st[par].alpha;
  const int improved = st[par].improved;
  double o[4];
  load_totals<4, 0u>(p4, G4, o);
  if (done) return;
  double ts = o[0], tt = o[1];
  if (mean_row >= 0) {
    ts += o[2] * o[3];  // t and s at the mean row
    tt += o[2] * o[2];
  }
  const double omega = ts / (tt + eps);
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    st[par].omega = omega;
    st[par].red[3] = ts;
    st[par].red[4] = tt;
  }
  double v[NQ5] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  const long stride = (long)gridDim.x * WG;
  for (; i < n; i += stride) {
    if (improved) xopt[i] = xo;
    x[i] = xo + alpha * a1 + omega * a2;
    if (i == mean_row) ti = o[2];
    const T rt = (T)(rr - omega * ti);
    r[i] = rt;
    const double rd = (double)rt;
    const double rhi = rh;
    if (i + stride < n) {
      const long in = i + stride;
      xo = x[in];
      a1 = (double)z1[in];
      a2 = (double)z2[in];
      rr = (double)r[in];
      ti = (double)t[in];
      rh = (double)rhat[in];
    }
    v[0] = nmax(v[0], fabs(rd));
    v[1] += rd * rd;
    v[2] += rhi * rd;
    v[3] += rhi * rhi;
    v[4] += rd;
    v[5] += rhi;
  }
  store_partials<NQ5, 1u>(v, p5);
}

// ----------------------------------------------------------------------------
// end of solve
// ----------------------------------------------------------------------------

// x = x_opt unless the last iteration was the best one (end of an inner
// solve of the mixed solver, before the FP64 residual check)
__global__ __launch_bounds__(WG) void k_select_best(int nb, double *__restrict__ x,
                                                    const double *__restrict__ xopt, const Scal *st,
                                                    int par) {
  if (st[par].improved) return;
  const long n = (long)nb * 64;
  for (long i = (long)blockIdx.x * WG + threadIdx.x; i < n; i += (long)gridDim.x * WG) x[i] = xopt[i];
}

// mean of the returned iterate (x if the last iteration improved or
// `always_x`, else x_opt)
__global__ __launch_bounds__(WG) void k_final_mean(int nb, const double *__restrict__ x,
                                                   const double *__restrict__ xopt,
                                                   const double *__restrict__ h2, const Scal *st,
                                                   int par, int always_x, double *pf) {
  const double *src = (always_x || st[par].improved) ? x : xopt;
  const long n = (long)nb * 64;
  double v[2] = {0.0, 0.0};
  for (long i = (long)blockIdx.x * WG + threadIdx.x; i < n; i += (long)gridDim.x * WG) {
    const double hh = h2[i >> 6];
    v[0] += src[i] * hh;
    v[1] += hh;
  }
  store_partials<2, 0u>(v, pf);
}

__global__ __launch_bounds__(WG) void k_final_write(int nb, const double *__restrict__ x,
                                                    const double *__restrict__ xopt,
                                                    double *__restrict__ out, Scal *st, int par,
                                                    int always_x, const double *pf, int Gf) {
  double o[2];
  load_totals<2, 0u>(pf, Gf, o);
  const double mean = o[0] / o[1];
  const double *src = (always_x || st[par].improved) ? x : xopt;
  if (blockIdx.x == 0 && threadIdx.x == 0) st[par].mean = mean;
  const long n = (long)nb * 64;
  for (long i = (long)blockIdx.x * WG + threadIdx.x; i < n; i += (long)gridDim.x * WG)
    out[i] = src[i] - mean;
}

// ----------------------------------------------------------------------------
// the pressure step around the solve (PressureSingle.cpp:842-879)
// ----------------------------------------------------------------------------

// b = rhs - L p_old with L the operator without the mean-row replacement:
// the old pressure's Laplacian taken off the RHS (updatePressureRHS1,
// PressureSingle.cpp:401-456, applied at :848-849).
__global__ __launch_bounds__(WG) void k_rhs_update(OpView A, int Gb, const double *__restrict__ p_old,
                                                   const double *__restrict__ rhs,
                                                   double *__restrict__ b) {
  if ((int)blockIdx.x < Gb) {
    const int c = lane_id();
    block_groups<double>(A, Gb, p_old, rhs, false, [&](const BlockRows<SPMV_U, double> &R, int u) {
      bool plain;
      const double yi = R.y(u, c, plain);
      if (plain) b[R.i(u, c)] = R.a(u) - yi;
    });
  } else {
    const int ge = gridDim.x - Gb;
    for (int e = (blockIdx.x - Gb) * WG + threadIdx.x; e < A.n_exc; e += ge * WG) {
      const int r = A.erow[e];
      b[r] = rhs[r] - explicit_row<double>(A, p_old, e, ell_run_width(A, e));
    }
  }
}

// Partials of sum h^2 dp and sum h^2 (PressureSingle.cpp:853-868).
__global__ __launch_bounds__(WG) void k_h2mean_part(int nb, const double *__restrict__ dp,
                                                    const double *__restrict__ h2, double *pf) {
  const long n = (long)nb * 64;
  double v[2] = {0.0, 0.0};
  for (long i = (long)blockIdx.x * WG + threadIdx.x; i < n; i += (long)gridDim.x * WG) {
    const double hh = h2[i >> 6];
    v[0] += dp[i] * hh;
    v[1] += hh;
  }
  store_partials<2, 0u>(v, pf);
}

// p = dp + (p_old - avg), avg the h^2-weighted mean of dp (PressureSingle.cpp:869-879).
__global__ __launch_bounds__(WG) void k_pressure_add(int nb, double *__restrict__ p,
                                                     const double *__restrict__ dp, const double *pf,
                                                     int Gf) {
  double o[2];
  load_totals<2, 0u>(pf, Gf, o);
  const double avg = o[0] / o[1];
  const long n = (long)nb * 64;
  for (long i = (long)blockIdx.x * WG + threadIdx.x; i < n; i += (long)gridDim.x * WG)
    p[i] = dp[i] + (p[i] - avg);
}


// ----------------------------------------------------------------------------
// operator pieces exposed by the C ABI
// ----------------------------------------------------------------------------

// y = A x (stencil + explicit rows); partials of sum h^2 x for the mean row
template <class T>
__global__ __launch_bounds__(WG) void k_spmv(OpView A, int Gb, const T *__restrict__ x,
                                             T *__restrict__ y, double *pm) {
  double v[1] = {0.0};
  if ((int)blockIdx.x < Gb) {
    const int c = lane_id();
    block_groups<T>(A, Gb, x, x, false, [&](const BlockRows<SPMV_U, T> &R, int u) {
      bool plain;
      const double yi = R.y(u, c, plain);
      if (plain) y[R.i(u, c)] = (T)yi;
      v[0] += R.hh[u] * R.xc(u);
    });
  } else {
    const int ge = gridDim.x - Gb;
    for (int e = (blockIdx.x - Gb) * WG + threadIdx.x; e < A.n_exc; e += ge * WG)
      y[A.erow[e]] = (T)explicit_row<T>(A, x, e, ell_run_width(A, e));
  }
  if (A.mean_row >= 0) store_partials<1, 0u>(v, pm);
}

// y = A x, plain CSR, one row per thread
__global__ __launch_bounds__(WG) void k_spmv_csr(long n, const int *__restrict__ rp,
                                                 const int *__restrict__ col,
                                                 const double *__restrict__ val,
                                                 const double *__restrict__ x,
                                                 double *__restrict__ y) {
  for (long i = (long)blockIdx.x * WG + threadIdx.x; i < n; i += (long)gridDim.x * WG) {
    double acc = 0.0;
    const int k1 = rp[i + 1];
    for (int k = rp[i]; k < k1; k++) acc += val[k] * x[col[k]];
    y[i] = acc;
  }
}

// y[mean_row] = sum h^2 x from the partials (one workgroup)
template <class T>
__global__ __launch_bounds__(WG) void k_meanrow_fix(const double *pm, int G, long mean_row,
                                                    T *__restrict__ y) {
  double o[1];
  load_totals<1, 0u>(pm, G, o);
  if (threadIdx.x == 0) y[mean_row] = (T)o[0];
}

// z = P r on every block, FP64 on MFMA (a workgroup per tile of 16 blocks)
__global__ __launch_bounds__(WG) void k_prec(int nb, const double *__restrict__ pf,
                                             const double *__restrict__ r,
                                             double *__restrict__ z) {
  __shared__ double tile[TILE][TROW];
  const int lane = lane_id(), w = wave_id();
  const int ntiles = (nb + TILE - 1) / TILE;
  PFrag64 F;
  load_pfrag64(F, pf, w, lane);
  for (int t = blockIdx.x; t < ntiles; t += gridDim.x) {
    bool ok;
    const long i0 = tile_cells(t, nb, w, lane, ok);
    double v[4];
    ld4(r, i0, v);
    tile_put(tile, w, lane, ok, v);
    __syncthreads();
    const f64x4 zz = tile_prec64(tile, F, lane);
    __syncthreads();
    if (ok) st4(z, i0, zz);
  }
}

// z = P r on every block, BF16 MFMA (FP32 in/out)
__global__ __launch_bounds__(WG) void k_prec_mf(int nb, const bf16x8 *__restrict__ pf,
                                                const float *__restrict__ r,
                                                float *__restrict__ z) {
  const int lane = lane_id(), w = wave_id(), q = lane >> 4, j = lane & 15;
  PFrag P;
  load_pfrag(P, pf, lane);
  const int ntiles = (nb + TILE - 1) / TILE, nw = gridDim.x * WAVES;
  for (int tile = blockIdx.x * WAVES + w; tile < ntiles; tile += nw) {
    const int bb = tile * TILE + j;
    const bool ok = bb < nb;
    const long base = (long)bb * 64;
    float v[16];
    if (ok) {
      ld16(r, base, q, v);
    } else {
#pragma unroll
      for (int k = 0; k < 16; k++) v[k] = 0.f;
    }
    float out[16];
    prec_mfma(P, v, out);
    if (ok) st16_out(z, base, q, out);
  }
}

// partials of |b - y|^2, |b - y|_inf, |b|^2, |b|_inf; y holds A x
__global__ __launch_bounds__(WG) void k_residual(long n, const double *__restrict__ b,
                                                 const double *__restrict__ y, long mean_row,
                                                 double *pr) {
  double v[4] = {0.0, 0.0, 0.0, 0.0};
  for (long i = (long)blockIdx.x * WG + threadIdx.x; i < n; i += (long)gridDim.x * WG) {
    const double bi = (i == mean_row) ? 0.0 : b[i];
    const double ri = bi - y[i];
    v[0] += ri * ri;
    v[1] = fmax(v[1], fabs(ri));
    v[2] += bi * bi;
    v[3] = fmax(v[3], fabs(bi));
  }
  store_partials<4, 0xAu>(v, pr);
}

// ----------------------------------------------------------------------------
// the unfused comparator: one operation per launch, the structure of the
// reference GPU loop (BiCGSTAB.cu:412-620), whose BLAS-1 calls, scalar
// kernels (BiCGSTAB.cu:255-294) and host synchronisations it keeps
// ----------------------------------------------------------------------------

// device scalars, in the order of BiCGSTABScalars (BiCGSTAB.cuh:12-22);
// `one` is the constant 1.0 the reference passes as d_eye_
struct UScal {
  double alpha, beta, omega, eps, rho_prev, rho_curr, buff_1, buff_2, one;
};

// partial sums of a.b (ddot) or of a.a (dnrm2, b == a), per workgroup
__global__ __launch_bounds__(WG) void ku_dot(long n, const double *__restrict__ a,
                                             const double *__restrict__ b, double *part) {
  double v[1] = {0.0};
  for (long i = (long)blockIdx.x * WG + threadIdx.x; i < n; i += (long)gridDim.x * WG) v[0] += a[i] * b[i];
  store_partials<1, 0u>(v, part);
}
// partial maxima of |a| (idamax)
__global__ __launch_bounds__(WG) void ku_amax(long n, const double *__restrict__ a, double *part) {
  double v[1] = {0.0};
  for (long i = (long)blockIdx.x * WG + threadIdx.x; i < n; i += (long)gridDim.x * WG)
    v[0] = nmax(v[0], fabs(a[i]));
  store_partials<1, 1u>(v, part);
}
// second stage of a reduction: *dst = total (mode 0: sum, 1: sqrt of the
// sum -- dnrm2 --, 2: max)
__global__ __launch_bounds__(WG) void ku_total(const double *part, int G, int mode, double *dst) {
  double o[1];
  if (mode == 2) load_totals<1, 1u>(part, G, o);
  else load_totals<1, 0u>(part, G, o);
  if (threadIdx.x == 0) *dst = mode == 1 ? sqrt(o[0]) : o[0];
}
// y += a x, a read from device memory (daxpy with a device scalar)
__global__ __launch_bounds__(WG) void ku_axpy(long n, const double *a, const double *__restrict__ x,
                                              double *__restrict__ y) {
  const double av = *a;
  for (long i = (long)blockIdx.x * WG + threadIdx.x; i < n; i += (long)gridDim.x * WG) y[i] += av * x[i];
}
// x *= a (dscal)
__global__ __launch_bounds__(WG) void ku_scal(long n, const double *a, double *__restrict__ x) {
  const double av = *a;
  for (long i = (long)blockIdx.x * WG + threadIdx.x; i < n; i += (long)gridDim.x * WG) x[i] *= av;
}
// the reference's one-thread scalar kernels
__global__ void ku_set_negative(double *dst, const double *src) { *dst = -*src; }
__global__ void ku_set_squared(double *v) { *v *= *v; }
__global__ void ku_set_zero(double *v) { *v = 0.0; }
__global__ void ku_set_beta(UScal *c) {
  c->beta = (c->rho_curr / (c->rho_prev + c->eps)) * (c->alpha / (c->omega + c->eps));
}
__global__ void ku_breakdown_update(UScal *c) {
  c->rho_prev = 1.;
  c->alpha = 1.;
  c->omega = 1.;
  c->beta = (c->rho_curr / (c->rho_prev + c->eps)) * (c->alpha / (c->omega + c->eps));
}
__global__ void ku_set_alpha(UScal *c) { c->alpha = c->rho_curr / (c->buff_1 + c->eps); }
__global__ void ku_set_omega(UScal *c) { c->omega = c->buff_1 / (c->buff_2 + c->eps); }
__global__ void ku_set_rho(UScal *c) { c->rho_prev = c->rho_curr; }

// ----------------------------------------------------------------------------
// host side
// ----------------------------------------------------------------------------

thread_local std::string g_create_err;  // failures without a handle, per calling thread

}  // namespace

struct DevBuf {
  void *p;
  size_t bytes;
};

struct cup2d_ps {
  cup2d_ps_config cfg{};
  mutable std::string err;  // also set by the const queries
  int device = 0;
  bool mixed = false;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  bool have_mesh = false;
  Operator op;
  // device operator
  int4 *d_nbr = nullptr;
  int2 *d_fid = nullptr;                 // [4 nb] coarse-fine neighbours of each face
  unsigned long long *d_mask = nullptr;
  int *d_erow = nullptr, *d_ecol = nullptr;
  uint16_t *d_ecode = nullptr;
  uint8_t *d_ewid = nullptr;
  double *d_etab = nullptr, *d_h2 = nullptr;
  double *d_h = nullptr;                 // cell size per block
  int ell_w = 0;
  long long n_exc = 0;                   // explicit rows
  bool has_csr = false;                  // plain CSR copy present (host-assembled meshes)
  int *d_csr_rp = nullptr, *d_csr_col = nullptr;
  double *d_csr_val = nullptr;
  double *d_P = nullptr;                 // FP64 MFMA fragments of P_inv (make_pfrag64)
  bf16x8 *d_pf = nullptr;                // BF16 MFMA fragments of P_inv
  // FP64 work vectors (x is the solution in both precisions)
  double *d_x = nullptr, *d_xopt = nullptr, *d_r = nullptr, *d_rhat = nullptr, *d_p = nullptr,
         *d_nu = nullptr, *d_z1 = nullptr, *d_z2 = nullptr, *d_t = nullptr, *d_b = nullptr,
         *d_xio = nullptr;
  // FP32-stored Krylov vectors of the mixed-precision solver (x stays FP64)
  float *f_r = nullptr, *f_rhat = nullptr, *f_p = nullptr, *f_nu = nullptr, *f_z1 = nullptr,
        *f_z2 = nullptr, *f_t = nullptr;
  Scal *d_s = nullptr;                   // two state slots (iteration parity)
  double *scratch = nullptr;             // this device's allocation-free work space (see below)
  int2 *d_bxy = nullptr;                 // [nb] block index (bx, by) at its level
  double h_min = 0.0;                    // smallest cell size of the mesh
  double *d_u = nullptr;                 // scalars of the unfused comparator (UScal + constants)
  double *d_p2 = nullptr, *d_p4 = nullptr, *d_p5 = nullptr, *d_pm = nullptr, *d_pfin = nullptr;
  Scal *h_s = nullptr;                   // pinned copy of the state slots
  int G = 1;                             // workgroups of the SpMV / vector kernels
  int Gb = 1;                            // of which the block part of the operator kernels
  int Gp = 1;                            // workgroups of the FP64 preconditioner kernels
  int Gm = 1;                            // workgroups of the BF16 preconditioner kernels
  // graph of one chunk of iterations
  hipGraphExec_t gexec = nullptr;
  int gchunk = 0;
  // device buffers of the current mesh, and those of earlier meshes kept
  // for reuse: a mesh change of similar size allocates (and frees) nothing
  std::vector<DevBuf> allocs, spare;
  // per-phase timers (cfg.profile): an event after every launch, tagged
  // with the launch's phase; the interval since the previous event is the
  // launch's time
  std::vector<hipEvent_t> pev;
  std::vector<int> pph;
  size_t pused = 0;
  double pms[5] = {0, 0, 0, 0, 0};
};

namespace {

int fail(cup2d_ps *h, int code, const std::string &msg) {
  if (h) h->err = msg;
  else g_create_err = msg;
  return code;
}

// Runs an entry point's body so that no C++ exception crosses the C ABI: a
// failed host allocation becomes CUP2D_PS_ENOMEM, any other exception EINVAL,
// with its text as the last error.
template <class F>
auto guarded(cup2d_ps *h, F &&body) -> decltype(body()) {
  try {
    return body();
  } catch (const std::bad_alloc &) {
    return fail(h, CUP2D_PS_ENOMEM, "host allocation failed");
  } catch (const std::exception &e) {
    return fail(h, CUP2D_PS_EINVAL, e.what());
  }
}

#define HIPCHK(h, expr)                                                                      \
  do {                                                                                       \
    hipError_t _e = (expr);                                                                  \
    if (_e != hipSuccess)                                                                    \
      return fail((h), _e == hipErrorOutOfMemory ? CUP2D_PS_ENOMEM : CUP2D_PS_EHIP,          \
                  std::string(#expr) + ": " + hipGetErrorString(_e));                        \
  } while (0)

// A buffer of at least `bytes`: the smallest spare one that is not more
// than twice as large, else a new one.
int take(cup2d_ps *h, size_t bytes, DevBuf &out) {
  size_t best = h->spare.size();
  for (size_t i = 0; i < h->spare.size(); i++) {
    const size_t b = h->spare[i].bytes;
    if (b >= bytes && b <= 2 * bytes + 4096 && (best == h->spare.size() || b < h->spare[best].bytes))
      best = i;
  }
  if (best < h->spare.size()) {
    out = h->spare[best];
    h->spare.erase(h->spare.begin() + best);
    return CUP2D_PS_OK;
  }
  out = DevBuf{nullptr, bytes};
  HIPCHK(h, hipMalloc(&out.p, bytes));
  return CUP2D_PS_OK;
}

template <class T>
int dalloc(cup2d_ps *h, T **p, size_t count) {
  *p = nullptr;
  DevBuf b;
  const int rc = take(h, std::max<size_t>(count, 1) * sizeof(T), b);
  if (rc) return rc;
  h->allocs.push_back(b);
  *p = static_cast<T *>(b.p);
  return CUP2D_PS_OK;
}

// Frees spare buffers, largest first, until they hold no more than the
// current mesh's buffers.
void trim_spare(cup2d_ps *h) {
  size_t live = 0, kept = 0;
  for (const DevBuf &b : h->allocs) live += b.bytes;
  for (const DevBuf &b : h->spare) kept += b.bytes;
  std::sort(h->spare.begin(), h->spare.end(),
            [](const DevBuf &a, const DevBuf &b) { return a.bytes > b.bytes; });
  size_t i = 0;
  for (; i < h->spare.size() && kept > live; i++) {
    (void)hipFree(h->spare[i].p);
    kept -= h->spare[i].bytes;
  }
  h->spare.erase(h->spare.begin(), h->spare.begin() + i);
}

template <class T>
int upload(cup2d_ps *h, T **p, const std::vector<T> &v) {
  int rc = dalloc(h, p, v.size());
  if (rc) return rc;
  if (!v.empty()) HIPCHK(h, hipMemcpy(*p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
  return CUP2D_PS_OK;
}

void free_mesh(cup2d_ps *h) {
  if (h->gexec) {
    (void)hipGraphExecDestroy(h->gexec);
    h->gexec = nullptr;
  }
  // the stream is idle (callers synchronise first): the buffers can be
  // handed out again in stream order
  h->spare.insert(h->spare.end(), h->allocs.begin(), h->allocs.end());
  h->allocs.clear();
  h->have_mesh = false;
}

OpView view(const cup2d_ps *h) {
  OpView A;
  A.nb = (int)h->op.nb;
  A.mean_row = (int)h->op.mean_row;
  A.nbr = h->d_nbr;
  A.mask = h->d_mask;
  A.n_exc = (int)h->n_exc;
  A.ell_w = h->ell_w;
  A.erow = h->d_erow;
  A.ecol = h->d_ecol;
  A.ecode = h->d_ecode;
  A.etab = h->d_etab;
  A.ewid = h->d_ewid;
  A.h2 = h->d_h2;
  return A;
}

// Profiling (cfg.profile): records an event after the launch just made,
// tagged with its phase (-1: a start mark).  Errors of the timing calls
// leave the phase times incomplete, never the solve.
void ptick(cup2d_ps *h, int ph) {
  if (!h->cfg.profile) return;
  if (h->pused == h->pev.size()) {
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return;
    h->pev.push_back(e);
    h->pph.push_back(-1);
  }
  if (hipEventRecord(h->pev[h->pused], h->stream) != hipSuccess) return;
  h->pph[h->pused++] = ph;
}
// After a synchronisation of the stream: adds every interval to its phase
// and keeps the last event as the start of the next.
void pcollect(cup2d_ps *h) {
  if (!h->cfg.profile) return;
  for (size_t i = 1; i < h->pused; i++) {
    float ms = 0;
    if (h->pph[i] >= 0 && hipEventElapsedTime(&ms, h->pev[i - 1], h->pev[i]) == hipSuccess)
      h->pms[h->pph[i]] += ms;
  }
  if (h->pused > 0) {
    std::swap(h->pev[0], h->pev[h->pused - 1]);
    h->pph[0] = -1;
    h->pused = 1;
  }
}
void pstart(cup2d_ps *h) {
  for (double &m : h->pms) m = 0.0;
  h->pused = 0;
  ptick(h, -1);
}
bool graphs_on(const cup2d_ps *h) { return h->cfg.use_graph && !h->cfg.profile; }

void launch_iter_fp64(cup2d_ps *h, int par) {
  const OpView A = view(h);
  const int G = h->G, Gp = h->Gp, nb = (int)h->op.nb;
  const long mr = h->op.mean_row;
  hipStream_t st = h->stream;
  k_pupdate<<<Gp, WG, 0, st>>>(nb, h->d_P, h->d_r, h->d_rhat, h->d_p, h->d_nu, h->d_z1,
                                   h->d_s, par, h->d_p5, G);
  ptick(h, CUP2D_PS_PHASE_PREC);
  k_spmv_nu<double><<<G, WG, 0, st>>>(A, h->Gb, h->d_z1, h->d_nu, h->d_rhat, h->d_s, par, h->d_p2);
  ptick(h, CUP2D_PS_PHASE_SPMV);
  k_supdate<<<Gp, WG, 0, st>>>(nb, h->d_P, h->d_r, h->d_nu, h->d_rhat, h->d_z2, h->d_s, par,
                                   h->d_p2, G, mr);
  ptick(h, CUP2D_PS_PHASE_PREC);
  k_spmv_t<double><<<G, WG, 0, st>>>(A, h->Gb, h->d_z2, h->d_t, h->d_r, h->d_s, par, h->d_p4);
  ptick(h, CUP2D_PS_PHASE_SPMV);
  k_xrupdate<double><<<G, WG, 0, st>>>(nb, h->d_x, h->d_xopt, h->d_z1, h->d_z2, h->d_r, h->d_t, h->d_rhat,
                                       h->d_s, par, h->d_p4, G, h->d_p5, mr);
  ptick(h, CUP2D_PS_PHASE_UPDATE);
}

void launch_iter_mixed(cup2d_ps *h, int par) {
  const OpView A = view(h);
  const int G = h->G, Gm = h->Gm, nb = (int)h->op.nb;
  const long mr = h->op.mean_row;
  hipStream_t st = h->stream;
  k_pupdate_mf<<<Gm, WG, 0, st>>>(nb, h->d_pf, h->f_r, h->f_rhat, h->f_p, h->f_nu, h->f_z1,
                                  h->d_s, par, h->d_p5, G);
  ptick(h, CUP2D_PS_PHASE_PREC);
  k_spmv_nu<float><<<G, WG, 0, st>>>(A, h->Gb, h->f_z1, h->f_nu, h->f_rhat, h->d_s, par, h->d_p2);
  ptick(h, CUP2D_PS_PHASE_SPMV);
  k_supdate_mf<<<Gm, WG, 0, st>>>(nb, h->d_pf, h->f_r, h->f_nu, h->f_rhat, h->f_z2, h->d_s, par,
                                  h->d_p2, G, mr);
  ptick(h, CUP2D_PS_PHASE_PREC);
  k_spmv_t<float><<<G, WG, 0, st>>>(A, h->Gb, h->f_z2, h->f_t, h->f_r, h->d_s, par, h->d_p4);
  ptick(h, CUP2D_PS_PHASE_SPMV);
  k_xrupdate<float><<<G, WG, 0, st>>>(nb, h->d_x, h->d_xopt, h->f_z1, h->f_z2, h->f_r, h->f_t, h->f_rhat,
                                      h->d_s, par, h->d_p4, G, h->d_p5, mr);
  ptick(h, CUP2D_PS_PHASE_UPDATE);
}

// Launches one BiCGSTAB iteration (parity `par` selects the state slot).
void launch_iteration(cup2d_ps *h, int par) {
  if (h->mixed) launch_iter_mixed(h, par);
  else launch_iter_fp64(h, par);
}

// Iterations per chunk: even, so that every chunk starts on state slot 0.
int chunk_of(const cup2d_ps *h) {
  int c = h->cfg.chunk > 0 ? h->cfg.chunk : 32;
  return c + (c & 1);
}

// One chunk: C iterations, then the evaluation of the next state into slot 0.
void launch_chunk(cup2d_ps *h) {
  const int C = chunk_of(h);
  for (int k = 0; k < C; k++) launch_iteration(h, k & 1);
  k_eval<<<1, WG, 0, h->stream>>>(h->d_s, 0, h->d_p5, h->G);
  ptick(h, CUP2D_PS_PHASE_SCALAR);
}

int ensure_graph(cup2d_ps *h) {
  const int C = chunk_of(h);
  if (h->gexec && h->gchunk == C) return CUP2D_PS_OK;
  if (h->gexec) {
    (void)hipGraphExecDestroy(h->gexec);
    h->gexec = nullptr;
  }
  hipGraph_t g = nullptr;
  HIPCHK(h, hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
  launch_chunk(h);
  HIPCHK(h, hipStreamEndCapture(h->stream, &g));
  hipError_t e = hipGraphInstantiate(&h->gexec, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  if (e != hipSuccess)
    return fail(h, CUP2D_PS_EHIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
  h->gchunk = C;
  return CUP2D_PS_OK;
}

// Runs iteration chunks until the device state reports done (or max_iter +
// one chunk have been launched).
int run_iterations(cup2d_ps *h, int max_iter) {
  const int C = chunk_of(h);
  if (graphs_on(h)) {
    int rc = ensure_graph(h);
    if (rc) return rc;
  }
  int launched = 0;
  for (;;) {
    if (graphs_on(h)) {
      HIPCHK(h, hipGraphLaunch(h->gexec, h->stream));
    } else {
      launch_chunk(h);
      HIPCHK(h, hipGetLastError());
    }
    launched += C;
    HIPCHK(h, hipMemcpyAsync(&h->h_s[0].done, &h->d_s[0].done, sizeof(int),
                             hipMemcpyDeviceToHost, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    pcollect(h);
    if (h->h_s[0].done || launched > max_iter + C) break;
  }
  return CUP2D_PS_OK;
}

Scal marker_state(const cup2d_ps *h, double tol_abs, double tol_rel, int max_restarts, int max_iter,
                  int stall = STALL_BASE) {
  Scal s{};
  s.n = (double)h->op.n;
  // the projection applies to the singular Neumann operator only: with the
  // mean constraint the replaced row makes A nonsingular and 1 is no longer
  // its left null vector
  s.project = (h->mixed && h->op.mean_row < 0) ? 1 : 0;
  s.stall = h->mixed ? stall : 0;
  s.tol_abs = tol_abs;
  s.tol_rel = tol_rel;
  s.eps = h->cfg.eps;
  s.bd_tol = h->cfg.breakdown_tol;
  s.max_iter = max_iter;
  s.max_restarts = max_restarts;
  s.norm = h->cfg.norm;
  s.iter = -1;
  return s;
}

int finish(cup2d_ps *h, double *d_out, int always_x) {
  const int G = h->G, nb = (int)h->op.nb;
  k_final_mean<<<G, WG, 0, h->stream>>>(nb, h->d_x, h->d_xopt, h->d_h2, h->d_s, 0, always_x,
                                        h->d_pfin);
  k_final_write<<<G, WG, 0, h->stream>>>(nb, h->d_x, h->d_xopt, d_out, h->d_s, 0, always_x,
                                         h->d_pfin, G);
  HIPCHK(h, hipGetLastError());
  return CUP2D_PS_OK;
}

int solve_fp64(cup2d_ps *h, const double *d_b, const double *d_x0, double *d_out, double tol_abs,
               double tol_rel, int max_restarts, cup2d_ps_stats *st) {
  const int G = h->G, nb = (int)h->op.nb;
  h->h_s[0] = h->h_s[1] = marker_state(h, tol_abs, tol_rel, max_restarts, h->cfg.max_iter);
  HIPCHK(h, hipEventRecord(h->ev0, h->stream));
  HIPCHK(h, hipMemcpyAsync(h->d_s, h->h_s, 2 * sizeof(Scal), hipMemcpyHostToDevice, h->stream));
  if (h->op.mean_row >= 0) k_hsum<double><<<G, WG, 0, h->stream>>>(nb, d_x0, h->d_h2, h->d_pm);
  k_init<double><<<G, WG, 0, h->stream>>>(view(h), h->Gb, d_b, d_x0, h->d_x, h->d_xopt, h->d_r,
                                          h->d_rhat, h->d_p, h->d_nu, h->d_pm, G, h->d_p5);
  HIPCHK(h, hipGetLastError());
  ptick(h, CUP2D_PS_PHASE_SETUP);
  int rc = run_iterations(h, h->cfg.max_iter);
  if (rc) return rc;
  if ((rc = finish(h, d_out, 0))) return rc;
  ptick(h, CUP2D_PS_PHASE_SETUP);
  HIPCHK(h, hipEventRecord(h->ev1, h->stream));
  HIPCHK(h, hipMemcpyAsync(h->h_s, h->d_s, 2 * sizeof(Scal), hipMemcpyDeviceToHost, h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  if (st) {
    const Scal &f = h->h_s[0];
    std::memset(st, 0, sizeof(*st));
    st->iterations = f.iter;
    st->restarts = f.restarts;
    st->converged = f.converged;
    st->err_init = f.err0;
    st->err_final = f.err_opt;
    float ms = 0;
    HIPCHK(h, hipEventElapsedTime(&ms, h->ev0, h->ev1));
    st->ms = ms;
  }
  return CUP2D_PS_OK;
}

// FP64 norms of r = b - A x from the k_init partials (host side).
int outer_residual(cup2d_ps *h, double &err) {
  std::vector<double> part(NQ5 * (size_t)h->G);
  HIPCHK(h, hipMemcpyAsync(part.data(), h->d_p5, part.size() * sizeof(double),
                           hipMemcpyDeviceToHost, h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  double rinf = 0, r2 = 0;
  for (int i = 0; i < h->G; i++) {
    const double a = part[i];
    rinf = (a != a || rinf != rinf) ? a : std::max(rinf, a);
    r2 += part[h->G + i];
  }
  err = h->cfg.norm ? std::sqrt(r2) : rinf;
  return CUP2D_PS_OK;
}

// Mixed precision: BiCGSTAB whose Krylov vectors (r, rhat, p, nu, z, t) are
// stored in FP32 and whose preconditioner runs on BF16 MFMA, with the
// solution x, every inner product and all arithmetic in FP64; r is projected
// onto the range of the Neumann operator every iteration.  FP64 iterative
// refinement wraps it: r = b - A x is recomputed in FP64 from x, the
// tolerance is checked on that true residual, and if the FP32 recurrence
// stopped short the iteration restarts from the current x.
int solve_mixed(cup2d_ps *h, const double *d_b, const double *d_x0, double *d_out, double tol_abs,
                double tol_rel, int max_restarts, cup2d_ps_stats *st) {
  const int G = h->G, nb = (int)h->op.nb;
  const size_t bytes = h->op.n * sizeof(double);
  HIPCHK(h, hipEventRecord(h->ev0, h->stream));
  HIPCHK(h, hipMemcpyAsync(h->d_x, d_x0, bytes, hipMemcpyDeviceToDevice, h->stream));
  const OpView A = view(h);
  int total_iters = 0, total_restarts = 0, outer = 0;
  double err0 = -1, err = 0;
  bool conv = false;
  const int max_outer = 64;
  int stall = STALL_BASE;  // doubled after a pass that gained less than 2x in FP64
  double err_prev = -1;
  for (;;) {
    if (h->op.mean_row >= 0) k_hsum<double><<<G, WG, 0, h->stream>>>(nb, h->d_x, h->d_h2, h->d_pm);
    k_init<float><<<G, WG, 0, h->stream>>>(A, h->Gb, d_b, h->d_x, h->d_x, h->d_xopt, h->f_r, h->f_rhat,
                                           h->f_p, h->f_nu, h->d_pm, G, h->d_p5);
    HIPCHK(h, hipGetLastError());
    ptick(h, CUP2D_PS_PHASE_SETUP);
    int rc = outer_residual(h, err);
    pcollect(h);
    if (rc) return rc;
    if (err != err) return fail(h, CUP2D_PS_EINVAL, "non-finite residual (NaN/Inf in b or x0)");
    if (err0 < 0) err0 = err;
    if (err == 0.0 || err <= tol_abs || err / err0 <= tol_rel) {
      conv = true;
      break;
    }
    if (outer >= max_outer || total_iters >= h->cfg.max_iter) break;
    if (err_prev > 0 && err > 0.5 * err_prev) stall = std::min(2 * stall, STALL_MAX);
    err_prev = err;
    // iterate towards the remaining reduction, at most inner_tol (0: 1e-2)
    // of this pass's initial residual: the FP32 recurrence is trusted that
    // far; the FP64 check above decides whether another pass is needed
    const double itol = h->cfg.inner_tol > 0 ? h->cfg.inner_tol : 1e-2;
    const double rel = std::max(tol_rel > 0 ? tol_rel * err0 / err : 0.0, itol);
    h->h_s[0] = h->h_s[1] =
        marker_state(h, tol_abs, rel, max_restarts, h->cfg.max_iter - total_iters, stall);
    HIPCHK(h, hipMemcpyAsync(h->d_s, h->h_s, 2 * sizeof(Scal), hipMemcpyHostToDevice, h->stream));
    if ((rc = run_iterations(h, h->cfg.max_iter - total_iters))) return rc;
    k_select_best<<<G, WG, 0, h->stream>>>(nb, h->d_x, h->d_xopt, h->d_s, 0);
    ptick(h, CUP2D_PS_PHASE_SETUP);
    HIPCHK(h, hipMemcpyAsync(h->h_s, h->d_s, sizeof(Scal), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    total_iters += h->h_s[0].iter;
    total_restarts += h->h_s[0].restarts;
    outer++;
    if (h->h_s[0].iter == 0) break;  // no progress possible
  }
  int rc = finish(h, d_out, 1);
  if (rc) return rc;
  ptick(h, CUP2D_PS_PHASE_SETUP);
  HIPCHK(h, hipEventRecord(h->ev1, h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  if (st) {
    std::memset(st, 0, sizeof(*st));
    st->iterations = total_iters;
    st->restarts = total_restarts;
    st->converged = conv;
    st->refinements = outer;
    st->err_init = err0;
    st->err_final = err;
    float ms = 0;
    HIPCHK(h, hipEventElapsedTime(&ms, h->ev0, h->ev1));
    st->ms = ms;
  }
  return CUP2D_PS_OK;
}

// The unfused FP64 comparator (cfg.schedule 1): BiCGSTAB.cu:412-620 step by
// step, one operation per launch (two for a reduction, as cuBLAS), the
// scalars on the device, and the reference's blocking host reads: the three
// reductions of the breakdown test, (rhat, nu), (t, s) with |t|^2, and the
// stopping norm, every iteration.  Same operator and preconditioner kernels
// as the fused solver; stopping norm, eps, breakdown tolerance and the
// iteration cap from the configuration (the reference: infinity norm,
// 1e-21, 1e-16, 1000).
int solve_unfused(cup2d_ps *h, const double *d_b, const double *d_x0, double *d_out,
                  double tol_abs, double tol_rel, int max_restarts, cup2d_ps_stats *st) {
  const long n = (long)h->op.n;
  const long mr = h->op.mean_row;
  const size_t bytes = n * sizeof(double);
  const int Gv = (int)std::min<long>(MAXG, (n + WG - 1) / WG), G = h->G;
  const OpView A = view(h);
  hipStream_t s = h->stream;
  UScal *c = reinterpret_cast<UScal *>(h->d_u);
  double *hr = reinterpret_cast<double *>(h->h_s);  // pinned staging of scalar reads
  double *x = h->d_x, *xopt = h->d_xopt, *r = h->d_r, *rhat = h->d_rhat, *p = h->d_p,
         *nu = h->d_nu, *z = h->d_z1, *t = h->d_t, *part = h->d_pm;
  auto spmv = [&](const double *in, double *out) {  // hd_cusparseSpMV, BiCGSTAB.cu:315-410
    k_spmv<double><<<G, WG, 0, s>>>(A, h->Gb, in, out, h->d_pm);
    if (mr >= 0) k_meanrow_fix<double><<<1, WG, 0, s>>>(h->d_pm, G, mr, out);
    ptick(h, CUP2D_PS_PHASE_SPMV);
  };
  auto prec = [&](const double *in, double *out) {  // cublasDgemm, BiCGSTAB.cu:520, :545
    k_prec<<<h->Gp, WG, 0, s>>>((int)h->op.nb, h->d_P, in, out);
    ptick(h, CUP2D_PS_PHASE_PREC);
  };
  auto reduce = [&](const double *a, const double *b, int mode, double *dst) {
    if (mode == 2) ku_amax<<<Gv, WG, 0, s>>>(n, a, part);
    else ku_dot<<<Gv, WG, 0, s>>>(n, a, b, part);
    ku_total<<<1, WG, 0, s>>>(part, Gv, mode, dst);
    ptick(h, CUP2D_PS_PHASE_UPDATE);
  };
  auto axpy = [&](const double *a, const double *xv, double *yv) {
    ku_axpy<<<Gv, WG, 0, s>>>(n, a, xv, yv);
    ptick(h, CUP2D_PS_PHASE_UPDATE);
  };
  auto scalar = [&]() { ptick(h, CUP2D_PS_PHASE_SCALAR); };
  auto read = [&](const double *src, int count) -> int {  // blocking read of device scalars
    HIPCHK(h, hipMemcpyAsync(hr, src, count * sizeof(double), hipMemcpyDeviceToHost, s));
    HIPCHK(h, hipStreamSynchronize(s));
    return CUP2D_PS_OK;
  };
  const int norm_mode = h->cfg.norm ? 1 : 2;  // dnrm2 or idamax of r
  HIPCHK(h, hipEventRecord(h->ev0, s));
  const UScal c0{1., 1., 1., h->cfg.eps, 1., 1., 0., 0., 1.};
  HIPCHK(h, hipMemcpyAsync(c, &c0, sizeof(UScal), hipMemcpyHostToDevice, s));
  // 1. r = b - A x0 (b's mean row is 0, ExpAMRSolver.cpp:340-345)
  HIPCHK(h, hipMemcpyAsync(x, d_x0, bytes, hipMemcpyDeviceToDevice, s));
  spmv(x, nu);
  HIPCHK(h, hipMemcpyAsync(r, d_b, bytes, hipMemcpyDeviceToDevice, s));
  if (mr >= 0) ku_set_zero<<<1, 1, 0, s>>>(r + mr);
  ku_set_negative<<<1, 1, 0, s>>>(&c->buff_2, &c->one);  // the reference's d_nye_ = -1
  axpy(&c->buff_2, nu, r);
  reduce(r, r, norm_mode, &c->buff_1);
  int rc = read(&c->buff_1, 1);
  if (rc) return rc;
  const double error_init = hr[0];
  double error = error_init, error_opt = error_init;
  if (error != error) return fail(h, CUP2D_PS_EINVAL, "non-finite residual (NaN/Inf in b or x0)");
  HIPCHK(h, hipMemcpyAsync(xopt, x, bytes, hipMemcpyDeviceToDevice, s));
  HIPCHK(h, hipMemcpyAsync(rhat, r, bytes, hipMemcpyDeviceToDevice, s));
  HIPCHK(h, hipMemsetAsync(nu, 0, bytes, s));
  HIPCHK(h, hipMemsetAsync(p, 0, bytes, s));
  ptick(h, CUP2D_PS_PHASE_SETUP);
  bool converged = error == 0.0;
  int restarts = 0, iters = 0;
  for (int k = 0; k < h->cfg.max_iter && !converged; k++) {
    // rho = (rhat, r); |r|^2 and |rhat|^2 for the breakdown test
    reduce(rhat, r, 0, &c->rho_curr);
    reduce(r, r, 1, &c->buff_1);
    reduce(rhat, rhat, 1, &c->buff_2);
    if ((rc = read(&c->rho_curr, 3))) return rc;  // rho_curr, buff_1, buff_2 are contiguous
    const double rho = hr[0], r2 = hr[1] * hr[1], rh2 = hr[2] * hr[2];
    const bool breakdown = rho * rho < h->cfg.breakdown_tol * r2 * rh2;
    ku_set_beta<<<1, 1, 0, s>>>(c);
    scalar();
    if (breakdown && max_restarts > 0) {
      restarts++;
      if (restarts >= max_restarts) break;
      HIPCHK(h, hipMemcpyAsync(rhat, r, bytes, hipMemcpyDeviceToDevice, s));
      reduce(rhat, rhat, 1, &c->rho_curr);
      ku_set_squared<<<1, 1, 0, s>>>(&c->rho_curr);
      HIPCHK(h, hipMemsetAsync(nu, 0, bytes, s));
      HIPCHK(h, hipMemsetAsync(p, 0, bytes, s));
      ku_breakdown_update<<<1, 1, 0, s>>>(c);
      scalar();
    }
    // p = r + beta (p - omega nu)
    ku_set_negative<<<1, 1, 0, s>>>(&c->buff_1, &c->omega);
    scalar();
    axpy(&c->buff_1, nu, p);
    ku_scal<<<Gv, WG, 0, s>>>(n, &c->beta, p);
    ptick(h, CUP2D_PS_PHASE_UPDATE);
    axpy(&c->one, r, p);
    prec(p, z);
    spmv(z, nu);
    // alpha = rho / (rhat, nu)
    reduce(rhat, nu, 0, &c->buff_1);
    if ((rc = read(&c->buff_1, 1))) return rc;
    ku_set_alpha<<<1, 1, 0, s>>>(c);
    scalar();
    axpy(&c->alpha, z, x);
    ku_set_negative<<<1, 1, 0, s>>>(&c->buff_1, &c->alpha);
    scalar();
    axpy(&c->buff_1, nu, r);  // s = r - alpha nu, in r
    prec(r, z);
    spmv(z, t);
    // omega = (t, s) / (t, t)
    reduce(t, r, 0, &c->buff_1);
    reduce(t, t, 1, &c->buff_2);
    ku_set_squared<<<1, 1, 0, s>>>(&c->buff_2);
    scalar();
    if ((rc = read(&c->buff_1, 2))) return rc;
    ku_set_omega<<<1, 1, 0, s>>>(c);
    scalar();
    axpy(&c->omega, z, x);
    ku_set_negative<<<1, 1, 0, s>>>(&c->buff_1, &c->omega);
    scalar();
    axpy(&c->buff_1, t, r);
    reduce(r, r, norm_mode, &c->buff_1);
    if ((rc = read(&c->buff_1, 1))) return rc;
    pcollect(h);
    error = hr[0];
    iters = k + 1;
    if (error < error_opt) {
      error_opt = error;
      HIPCHK(h, hipMemcpyAsync(xopt, x, bytes, hipMemcpyDeviceToDevice, s));
      if (error <= tol_abs || error / error_init <= tol_rel) converged = true;
    }
    ku_set_rho<<<1, 1, 0, s>>>(c);
    scalar();
  }
  HIPCHK(h, hipGetLastError());
  HIPCHK(h, hipMemcpyAsync(x, xopt, bytes, hipMemcpyDeviceToDevice, s));
  if ((rc = finish(h, d_out, 1))) return rc;
  ptick(h, CUP2D_PS_PHASE_SETUP);
  HIPCHK(h, hipEventRecord(h->ev1, s));
  HIPCHK(h, hipStreamSynchronize(s));
  if (st) {
    std::memset(st, 0, sizeof(*st));
    st->iterations = iters;
    st->restarts = restarts;
    st->converged = converged;
    st->err_init = error_init;
    st->err_final = error_opt;
    float ms = 0;
    HIPCHK(h, hipEventElapsedTime(&ms, h->ev0, h->ev1));
    st->ms = ms;
  }
  return CUP2D_PS_OK;
}

// |b - A x|_2, |b - A x|_inf, |b|_2, |b|_inf in FP64 on the device (b's
// mean row taken as 0); uses the work vector t and the final partials.
int residual_norms(cup2d_ps *h, const double *d_b, const double *d_x, double out[4]) {
  k_spmv<double><<<h->G, WG, 0, h->stream>>>(view(h), h->Gb, d_x, h->d_t, h->d_pm);
  if (h->op.mean_row >= 0)
    k_meanrow_fix<double><<<1, WG, 0, h->stream>>>(h->d_pm, h->G, h->op.mean_row, h->d_t);
  k_residual<<<h->G, WG, 0, h->stream>>>((long)h->op.n, d_b, h->d_t, h->op.mean_row, h->d_pfin);
  HIPCHK(h, hipGetLastError());
  std::vector<double> part(4 * (size_t)h->G);
  HIPCHK(h, hipMemcpyAsync(part.data(), h->d_pfin, part.size() * sizeof(double),
                           hipMemcpyDeviceToHost, h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  double r2 = 0, rinf = 0, b2 = 0, binf = 0;
  for (int i = 0; i < h->G; i++) {
    r2 += part[i];
    rinf = std::max(rinf, part[h->G + i]);
    b2 += part[2 * h->G + i];
    binf = std::max(binf, part[3 * h->G + i]);
  }
  out[0] = std::sqrt(r2);
  out[1] = rinf;
  out[2] = std::sqrt(b2);
  out[3] = binf;
  return CUP2D_PS_OK;
}

int solve_impl(cup2d_ps *h, const double *d_b, const double *d_x0, double *d_out, double tol_abs,
               double tol_rel, int max_restarts, cup2d_ps_stats *st) {
  if (!h->have_mesh) return fail(h, CUP2D_PS_ENOMESH, "solve before cup2d_ps_set_mesh");
  HIPCHK(h, hipSetDevice(h->device));
  return guarded(h, [&]() -> int {
    pstart(h);
    int rc = h->cfg.schedule == 1
                 ? solve_unfused(h, d_b, d_x0, d_out, tol_abs, tol_rel, max_restarts, st)
                 : h->mixed ? solve_mixed(h, d_b, d_x0, d_out, tol_abs, tol_rel, max_restarts, st)
                            : solve_fp64(h, d_b, d_x0, d_out, tol_abs, tol_rel, max_restarts, st);
    if (rc) return rc;
    // the returned solution's true residual, recomputed in FP64
    double nr[4];
    if ((rc = residual_norms(h, d_b, d_out, nr))) return rc;
    ptick(h, CUP2D_PS_PHASE_SETUP);
    HIPCHK(h, hipStreamSynchronize(h->stream));
    pcollect(h);
    if (st) {
      st->true_rel2 = nr[2] > 0 ? nr[0] / nr[2] : 0.0;
      st->true_relinf = nr[3] > 0 ? nr[1] / nr[3] : 0.0;
      for (int k = 0; k < 5; k++) st->ms_phase[k] = h->cfg.profile ? h->pms[k] : 0.0;
    }
    return CUP2D_PS_OK;
  });
}

// ----------------------------------------------------------------------------
// Allocation-free solve (cup2d_ps_solve_device_noalloc).  Its work space --
// the nine Krylov / solution vectors, the per-workgroup partial sums and the
// two solver-state slots -- is this file-scope device array, not the
// handle's buffers, so the call allocates nothing.  Each device holds its own
// copy until the process ends; host code takes its address from
// hipGetSymbolAddress (once per device, at handle creation) and never uses
// the array's name.  Every call on a device shares the device's array:
// calls that overlap on one device race.  The solve initializes every part
// of it it reads (r, rhat, p, nu, x, x_opt by the first kernel, the state
// slots from the host, the partials before each use).
// ----------------------------------------------------------------------------
constexpr long long NOALLOC_N = CUP2D_PS_NOALLOC_MAX_N;   // unknowns
constexpr size_t NOALLOC_GT = MAXG + MAXGE;                 // partial slots per quantity
constexpr size_t NOALLOC_SCAL = 2 * ((sizeof(Scal) + 255) / 256) * 32;  // doubles: 2 slots, 256-byte aligned
constexpr size_t NOALLOC_DOUBLES = NOALLOC_SCAL + 9 * (size_t)NOALLOC_N + 17 * NOALLOC_GT + 8;
static_assert(NOALLOC_N % 64 == 0, "the work vectors are whole blocks");

static __device__ __align__(256) double scratch[NOALLOC_DOUBLES];

std::mutex g_scratch_mu;
double *g_scratch_ptr[64] = {};  // per device ordinal, set once

// The device address of `scratch` on the current device (`dev`), looked up
// once per device.
int scratch_address(int dev, double *&out) {
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (dev < 0 || dev >= 64) return CUP2D_PS_EINVAL;
  if (!g_scratch_ptr[dev]) {
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) return CUP2D_PS_EHIP;
    g_scratch_ptr[dev] = static_cast<double *>(p);
  }
  out = g_scratch_ptr[dev];
  return CUP2D_PS_OK;
}

// The handle's work-buffer pointers (what the allocation-free solve swaps).
struct WorkBufs {
  double *x, *xopt, *r, *rhat, *p, *nu, *z1, *z2, *t, *p2, *p4, *p5, *pm, *pfin;
  Scal *s;
};
WorkBufs get_bufs(const cup2d_ps *h) {
  return WorkBufs{h->d_x, h->d_xopt, h->d_r, h->d_rhat, h->d_p, h->d_nu, h->d_z1, h->d_z2, h->d_t,
                  h->d_p2, h->d_p4, h->d_p5, h->d_pm, h->d_pfin, h->d_s};
}
void set_bufs(cup2d_ps *h, const WorkBufs &b) {
  h->d_x = b.x; h->d_xopt = b.xopt; h->d_r = b.r; h->d_rhat = b.rhat; h->d_p = b.p;
  h->d_nu = b.nu; h->d_z1 = b.z1; h->d_z2 = b.z2; h->d_t = b.t; h->d_p2 = b.p2; h->d_p4 = b.p4;
  h->d_p5 = b.p5; h->d_pm = b.pm; h->d_pfin = b.pfin; h->d_s = b.s;
}
// The same buffers carved out of the scratch array for n unknowns.
WorkBufs scratch_bufs(double *base, size_t n) {
  WorkBufs b;
  b.s = reinterpret_cast<Scal *>(base);
  double *q = base + NOALLOC_SCAL;
  double **vec[] = {&b.x, &b.xopt, &b.r, &b.rhat, &b.p, &b.nu, &b.z1, &b.z2, &b.t};
  for (double **v : vec) {
    *v = q;
    q += n;  // n is a multiple of 64: every vector starts 512-byte aligned
  }
  b.p2 = q;
  b.p4 = b.p2 + 2 * NOALLOC_GT;
  b.p5 = b.p4 + 4 * NOALLOC_GT;
  b.pm = b.p5 + NQ5 * NOALLOC_GT;
  b.pfin = b.pm + NOALLOC_GT;
  return b;
}

// The FP64 solve of cup2d_ps_solve_device with the scratch array as work
// space: the same kernels in the same order (plain stream launches: a graph
// would be instantiated, which allocates), then the true residual, reduced
// on the device into the scratch and read back through the handle's pinned
// state buffer.
int solve_noalloc_body(cup2d_ps *h, const double *d_b, double *d_x, double tol_abs, double tol_rel,
                       int max_restarts, cup2d_ps_stats *st) {
  if (!h || !d_b || !d_x) return fail(h, CUP2D_PS_EINVAL, "NULL argument");
  if (!(tol_abs >= 0) || !(tol_rel >= 0) || max_restarts < 0)
    return fail(h, CUP2D_PS_EINVAL, "negative tolerance or restart count");
  if (!h->have_mesh) return fail(h, CUP2D_PS_ENOMESH, "solve before cup2d_ps_set_mesh");
  if (h->mixed || h->cfg.schedule != 0)
    return fail(h, CUP2D_PS_EINVAL,
                "the allocation-free solve takes an FP64 handle with the fused schedule");
  if (h->op.n > NOALLOC_N)
    return fail(h, CUP2D_PS_EINVAL,
                "mesh has " + std::to_string(h->op.n) + " unknowns; the allocation-free solve takes at most " +
                    std::to_string(NOALLOC_N) + " (its work space is a fixed device array)");
  if (!h->scratch) return fail(h, CUP2D_PS_EHIP, "no work-space address for this device");
  if (d_b == d_x) return fail(h, CUP2D_PS_EINVAL, "b and x must not alias");
  HIPCHK(h, hipSetDevice(h->device));
  const WorkBufs own = get_bufs(h);
  const WorkBufs sb = scratch_bufs(h->scratch, (size_t)h->op.n);
  const int graph = h->cfg.use_graph, prof = h->cfg.profile;
  set_bufs(h, sb);
  h->cfg.use_graph = 0;
  h->cfg.profile = 0;
  int rc = guarded(h, [&]() -> int {
    int r = solve_fp64(h, d_b, d_x, d_x, tol_abs, tol_rel, max_restarts, st);
    if (r) return r;
    // true residual of the returned x: partials of |r|^2, |r|_inf, |b|^2,
    // |b|_inf, each reduced by one workgroup into scratch
    const int G = h->G;
    k_spmv<double><<<G, WG, 0, h->stream>>>(view(h), h->Gb, d_x, h->d_t, h->d_pm);
    if (h->op.mean_row >= 0)
      k_meanrow_fix<double><<<1, WG, 0, h->stream>>>(h->d_pm, G, h->op.mean_row, h->d_t);
    k_residual<<<G, WG, 0, h->stream>>>((long)h->op.n, d_b, h->d_t, h->op.mean_row, h->d_pfin);
    double *tot = h->d_p2;  // free again: four totals
    for (int q = 0; q < 4; q++)
      ku_total<<<1, WG, 0, h->stream>>>(h->d_pfin + (size_t)q * G, G, (q & 1) ? 2 : 0, tot + q);
    HIPCHK(h, hipGetLastError());
    double *pin = reinterpret_cast<double *>(h->h_s);
    HIPCHK(h, hipMemcpyAsync(pin, tot, 4 * sizeof(double), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    if (st) {
      st->true_rel2 = pin[2] > 0 ? std::sqrt(pin[0]) / std::sqrt(pin[2]) : 0.0;
      st->true_relinf = pin[3] > 0 ? pin[1] / pin[3] : 0.0;
    }
    return CUP2D_PS_OK;
  });
  set_bufs(h, own);
  h->cfg.use_graph = graph;
  h->cfg.profile = prof;
  return rc;
}

}  // namespace

extern "C" {

void cup2d_ps_config_default(cup2d_ps_config *cfg) {
  std::memset(cfg, 0, sizeof(*cfg));
  cfg->device = 0;
  cfg->precision = CUP2D_PS_FP64;
  cfg->mean_constraint = 0;
  cfg->norm = CUP2D_PS_NORM_INF;
  cfg->max_iter = 1000;
  cfg->chunk = 0;
  cfg->use_graph = 1;
  cfg->eps = 1e-100;
  cfg->breakdown_tol = 1e-16;
  cfg->inner_tol = 0.0;
}

static int create_body(const cup2d_ps_config *cfg, cup2d_ps **out) {
  if (!out) return fail(nullptr, CUP2D_PS_EINVAL, "out is NULL");
  *out = nullptr;
  cup2d_ps_config c;
  if (cfg) c = *cfg;
  else cup2d_ps_config_default(&c);
  if (c.max_iter <= 0 || c.chunk < 0 || c.eps < 0 || c.breakdown_tol < 0 || c.inner_tol < 0 ||
      c.inner_tol >= 1 || (c.precision != CUP2D_PS_FP64 && c.precision != CUP2D_PS_MIXED) ||
      (c.norm != CUP2D_PS_NORM_INF && c.norm != CUP2D_PS_NORM_L2) || (c.profile != 0 && c.profile != 1) ||
      c.item_blocks < 0 || (c.schedule != 0 && c.schedule != 1) ||
      (c.schedule == 1 && c.precision != CUP2D_PS_FP64))
    return fail(nullptr, CUP2D_PS_EINVAL, "invalid configuration");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0)
    return fail(nullptr, CUP2D_PS_ENODEV, "no HIP device");
  if (c.device < 0 || c.device >= ndev) return fail(nullptr, CUP2D_PS_EINVAL, "bad device ordinal");
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, c.device) != hipSuccess)
    return fail(nullptr, CUP2D_PS_ENODEV, "hipGetDeviceProperties failed");
  if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
    return fail(nullptr, CUP2D_PS_ENODEV, std::string("built for gfx950, device is ") + prop.gcnArchName);
  auto h = std::make_unique<cup2d_ps>();
  h->cfg = c;
  h->device = c.device;
  h->mixed = c.precision == CUP2D_PS_MIXED;
  if (hipSetDevice(c.device) != hipSuccess) return fail(nullptr, CUP2D_PS_EHIP, "hipSetDevice failed");
  // the allocation-free solve's work space on this device (looked up once)
  if (scratch_address(c.device, h->scratch) != CUP2D_PS_OK)
    return fail(nullptr, CUP2D_PS_EHIP, "hipGetSymbolAddress of the work space failed");
  if (hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreate(&h->ev0) != hipSuccess || hipEventCreate(&h->ev1) != hipSuccess ||
      hipHostMalloc((void **)&h->h_s, 2 * sizeof(Scal), hipHostMallocDefault) != hipSuccess) {
    cup2d_ps_destroy(h.release());  // releases whichever of them were made
    return fail(nullptr, CUP2D_PS_EHIP, "stream/event/pinned allocation failed");
  }
  *out = h.release();
  return CUP2D_PS_OK;
}

int cup2d_ps_destroy(cup2d_ps *h) {
  if (!h) return CUP2D_PS_OK;
  (void)hipSetDevice(h->device);
  if (h->stream) (void)hipStreamSynchronize(h->stream);
  free_mesh(h);
  for (const DevBuf &b : h->spare) (void)hipFree(b.p);
  h->spare.clear();
  if (h->h_s) (void)hipHostFree(h->h_s);
  for (hipEvent_t e : h->pev) (void)hipEventDestroy(e);
  if (h->ev0) (void)hipEventDestroy(h->ev0);
  if (h->ev1) (void)hipEventDestroy(h->ev1);
  if (h->stream) (void)hipStreamDestroy(h->stream);
  delete h;
  return CUP2D_PS_OK;
}

static int finish_mesh(cup2d_ps *h);

// Cell size per block, h0 / 2^level (the BlockInfo::h of the reference's
// operators), for the kernels either side of the solve.
static int upload_h(cup2d_ps *h, const cup2d_mesh_desc *m) {
  const double h0 = m->extent / (double(std::max(m->bpdx, m->bpdy)) * BS);
  std::vector<double> hb((size_t)m->n_blocks);
  std::vector<int2> bxy((size_t)m->n_blocks);
  for (long long b = 0; b < m->n_blocks; b++) {
    hb[b] = h0 / double(1LL << m->level[b]);
    bxy[b] = make_int2(m->bx[b], m->by[b]);
  }
  h->h_min = *std::min_element(hb.begin(), hb.end());
  int rc = upload(h, &h->d_h, hb);
  return rc ? rc : upload(h, &h->d_bxy, bxy);
}

static int set_mesh_body(cup2d_ps *h, const cup2d_mesh_desc *m) {
  if (!h || !m) return fail(h, CUP2D_PS_EINVAL, "NULL argument");
  HIPCHK(h, hipSetDevice(h->device));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  free_mesh(h);
  MeshDesc md{m->bpdx, m->bpdy, m->levels, m->periodic_x, m->periodic_y, m->extent, m->n_blocks,
              m->level, m->bx, m->by};
  std::string err;
  if (!build_operator(md, h->cfg.mean_constraint, h->op, err)) return fail(h, CUP2D_PS_EINVAL, err);
  const Operator &op = h->op;
  int rc = 0;
  std::vector<int4> nbr(op.nb);
  for (long long b = 0; b < op.nb; b++)
    nbr[b] = make_int4(op.nbr[4 * b], op.nbr[4 * b + 1], op.nbr[4 * b + 2], op.nbr[4 * b + 3]);
  std::vector<int2> fid(4 * op.nb);
  for (long long k = 0; k < 4 * op.nb; k++) fid[k] = make_int2(op.fid[2 * k], op.fid[2 * k + 1]);
  std::vector<unsigned long long> mask(op.exc_mask.begin(), op.exc_mask.end());
  h->ell_w = ell_width(op);
  if (h->ell_w > ELL_MAX)
    return fail(h, CUP2D_PS_EINVAL, "explicit row with more than " + std::to_string(ELL_MAX) + " entries");
  EllRows ell;
  ell.W = h->ell_w;
  std::string eerr;
  if (!ell.append(op, 0, eerr)) return fail(h, CUP2D_PS_EINVAL, eerr);
  const size_t n_exc = ell.erow.size();
  std::vector<uint8_t> ewid((n_exc + 63) / 64);
  for (size_t r = 0; r < ewid.size(); r++)
    ewid[r] = (uint8_t)ell.width((long long)(64 * r), (long long)std::min(n_exc, 64 * r + 64));
  if ((rc = upload(h, &h->d_nbr, nbr)) || (rc = upload(h, &h->d_mask, mask)) ||
      (rc = upload(h, &h->d_fid, fid)) || (rc = upload(h, &h->d_erow, ell.erow)) ||
      (rc = upload(h, &h->d_ecol, ell_transpose(ell.ecol_rm, h->ell_w))) ||
      (rc = upload(h, &h->d_ecode, ell_transpose(ell.ecode_rm, h->ell_w))) ||
      (rc = upload(h, &h->d_etab, ell.values)) || (rc = upload(h, &h->d_ewid, ewid)) ||
      (rc = upload(h, &h->d_h2, op.h2)) || (rc = upload(h, &h->d_csr_rp, op.rp)) ||
      (rc = upload(h, &h->d_csr_col, op.col)) || (rc = upload(h, &h->d_csr_val, op.val)))
    return rc;
  h->n_exc = (long long)n_exc;
  h->has_csr = true;
  if ((rc = upload_h(h, m))) return rc;
  return finish_mesh(h);
}

// The block preconditioner, the work vectors and the launch shapes: the part
// of a mesh change that does not depend on where the operator was assembled.
static int finish_mesh(cup2d_ps *h) {
  const Operator &op = h->op;
  int rc = 0;
  static const std::vector<double> P = make_pinv();  // the same for every mesh
  static const std::vector<uint16_t> pf = make_pfrag(P);
  static const std::vector<double> pf64 = make_pfrag64(P);
  if ((rc = upload(h, &h->d_P, pf64)) || (rc = dalloc(h, &h->d_pf, 8 * 64))) return rc;
  HIPCHK(h, hipMemcpy(h->d_pf, pf.data(), pf.size() * sizeof(uint16_t), hipMemcpyHostToDevice));
  const size_t n = op.n;
  double **vecs[] = {&h->d_x, &h->d_xopt, &h->d_r, &h->d_rhat, &h->d_p, &h->d_nu,
                     &h->d_z1, &h->d_z2, &h->d_t, &h->d_b, &h->d_xio};
  for (double **v : vecs)
    if ((rc = dalloc(h, v, n))) return rc;
  float **fvecs[] = {&h->f_r, &h->f_rhat, &h->f_p, &h->f_nu, &h->f_z1, &h->f_z2, &h->f_t};
  for (float **v : fvecs)
    if ((rc = dalloc(h, v, h->mixed ? n : 64))) return rc;
  const size_t GT = MAXG + MAXGE;
  if ((rc = dalloc(h, &h->d_s, 2)) || (rc = dalloc(h, &h->d_u, 16)) || (rc = dalloc(h, &h->d_p2, 2 * GT)) ||
      (rc = dalloc(h, &h->d_p4, 4 * GT)) || (rc = dalloc(h, &h->d_p5, NQ5 * GT)) ||
      (rc = dalloc(h, &h->d_pm, GT)) || (rc = dalloc(h, &h->d_pfin, 4 * GT)))
    return rc;
  HIPCHK(h, hipMemset(h->d_s, 0, 2 * sizeof(Scal)));
  const long long waves = (op.nb + SPMV_U - 1) / SPMV_U;
  h->Gb = (int)std::max<long long>(1, std::min<long long>(MAXG, (waves + WAVES - 1) / WAVES));
  const int Ge = h->n_exc > 0 ? (int)std::min<long long>(MAXGE, (h->n_exc + WG - 1) / WG) : 0;
  h->G = h->Gb + Ge;
  const long long tiles = (op.nb + TILE - 1) / TILE;
  h->Gp = (int)std::max<long long>(1, std::min<long long>(MAXGP, tiles));
  h->Gm = (int)std::max<long long>(1, std::min<long long>(MAXG, (tiles + WAVES - 1) / WAVES));
  trim_spare(h);
  h->have_mesh = true;
  return CUP2D_PS_OK;
}

// The operator assembled by kernels (assemble_device.hip) straight into the
// arrays the solver reads; no host assembly, no plain CSR copy.
static int set_mesh_on_device_body(cup2d_ps *h, const cup2d_mesh_desc *m) {
  if (!h || !m) return fail(h, CUP2D_PS_EINVAL, "NULL argument");
  HIPCHK(h, hipSetDevice(h->device));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  free_mesh(h);
  const MeshDesc md{m->bpdx, m->bpdy, m->levels, m->periodic_x, m->periodic_y, m->extent,
                    m->n_blocks, m->level, m->bx, m->by};
  DeviceOperator D;
  std::string err;
  const DeviceAlloc alloc = [h](size_t bytes) -> void * {
    uint8_t *p = nullptr;
    return dalloc(h, &p, bytes) ? nullptr : p;
  };
  std::vector<DevBuf> temps;  // the assembly's work arrays, spare again afterwards
  const DeviceAlloc temp = [h, &temps](size_t bytes) -> void * {
    DevBuf b;
    if (take(h, bytes, b)) return nullptr;
    temps.push_back(b);
    return b.p;
  };
  int bad = -1;
  const int rc = assemble_on_device(&md, 1, h->cfg.mean_constraint, h->stream, alloc, temp, D, bad, err);
  (void)hipStreamSynchronize(h->stream);
  h->spare.insert(h->spare.end(), temps.begin(), temps.end());
  if (rc) {
    free_mesh(h);
    return fail(h, rc, err);
  }
  h->op = Operator();
  h->op.nb = D.nb;
  h->op.n = D.n;
  h->op.nnz_rows = D.nnz;
  h->op.mean_row = D.mean_row;
  h->op.corner_block = D.corner_block;
  h->d_nbr = D.nbr;
  h->d_fid = D.fid;
  h->d_mask = D.mask;
  h->d_h2 = D.h2;
  h->d_erow = D.erow;
  h->d_ecol = D.ecol;
  h->d_ecode = D.ecode;
  h->d_etab = D.etab;
  h->d_ewid = D.ewid;
  h->ell_w = D.ell_w;
  h->n_exc = D.n_exc;
  h->has_csr = false;
  if (const int hr = upload_h(h, m)) return hr;
  h->d_csr_rp = h->d_csr_col = nullptr;
  h->d_csr_val = nullptr;
  return finish_mesh(h);
}

int cup2d_ps_solve_device(cup2d_ps *h, const double *d_b, double *d_x, double tol_abs,
                          double tol_rel, int max_restarts, cup2d_ps_stats *st) {
  if (!h || !d_b || !d_x) return fail(h, CUP2D_PS_EINVAL, "NULL argument");
  if (tol_abs < 0 || tol_rel < 0 || max_restarts < 0) return fail(h, CUP2D_PS_EINVAL, "negative tolerance");
  return solve_impl(h, d_b, d_x, d_x, tol_abs, tol_rel, max_restarts, st);
}

int cup2d_ps_solve_device_noalloc(cup2d_ps *h, const double *d_b, double *d_x, double tol_abs,
                                  double tol_rel, int max_restarts, cup2d_ps_stats *st) {
  return guarded(h, [&] { return solve_noalloc_body(h, d_b, d_x, tol_abs, tol_rel, max_restarts, st); });
}

int cup2d_ps_solve(cup2d_ps *h, const double *b, double *x, double tol_abs, double tol_rel,
                   int max_restarts, cup2d_ps_stats *st) {
  if (!h || !b || !x) return fail(h, CUP2D_PS_EINVAL, "NULL argument");
  if (!h->have_mesh) return fail(h, CUP2D_PS_ENOMESH, "solve before cup2d_ps_set_mesh");
  if (tol_abs < 0 || tol_rel < 0 || max_restarts < 0) return fail(h, CUP2D_PS_EINVAL, "negative tolerance");
  HIPCHK(h, hipSetDevice(h->device));
  const size_t bytes = h->op.n * sizeof(double);
  HIPCHK(h, hipMemcpyAsync(h->d_b, b, bytes, hipMemcpyHostToDevice, h->stream));
  HIPCHK(h, hipMemcpyAsync(h->d_xio, x, bytes, hipMemcpyHostToDevice, h->stream));
  int rc = solve_impl(h, h->d_b, h->d_xio, h->d_xio, tol_abs, tol_rel, max_restarts, st);
  if (rc) return rc;
  HIPCHK(h, hipMemcpyAsync(x, h->d_xio, bytes, hipMemcpyDeviceToHost, h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  return CUP2D_PS_OK;
}

// The pressure step of PressureSingle::operator() around the solve, on the
// device: b = rhs - L p_old, dp from x0 = 0, p = p_old + dp - <dp>_h2.
static int pressure_step_body(cup2d_ps *h, const double *d_rhs, double *d_p, double tol_abs,
                              double tol_rel, int max_restarts, cup2d_ps_stats *st) {
  if (!h || !d_rhs || !d_p) return fail(h, CUP2D_PS_EINVAL, "NULL argument");
  if (!h->have_mesh) return fail(h, CUP2D_PS_ENOMESH, "pressure step before cup2d_ps_set_mesh");
  if (!(tol_abs >= 0) || !(tol_rel >= 0) || max_restarts < 0)
    return fail(h, CUP2D_PS_EINVAL, "negative tolerance or restart count");
  if (d_rhs == d_p) return fail(h, CUP2D_PS_EINVAL, "rhs and p must not alias");
  HIPCHK(h, hipSetDevice(h->device));
  const int nb = (int)h->op.nb, G = h->G;
  k_rhs_update<<<G, WG, 0, h->stream>>>(view(h), h->Gb, d_p, d_rhs, h->d_b);
  HIPCHK(h, hipGetLastError());
  HIPCHK(h, hipMemsetAsync(h->d_xio, 0, h->op.n * sizeof(double), h->stream));  // x0 = 0 (:845)
  const int rc = solve_impl(h, h->d_b, h->d_xio, h->d_xio, tol_abs, tol_rel, max_restarts, st);
  if (rc) return rc;
  k_h2mean_part<<<G, WG, 0, h->stream>>>(nb, h->d_xio, h->d_h2, h->d_pfin);
  k_pressure_add<<<G, WG, 0, h->stream>>>(nb, d_p, h->d_xio, h->d_pfin, G);
  HIPCHK(h, hipGetLastError());
  HIPCHK(h, hipStreamSynchronize(h->stream));
  return CUP2D_PS_OK;
}

// One thread per cell (no partials to re-reduce, so the grid is not capped
// like the solver's), up to 16384 workgroups.
static int cell_grid(const cup2d_ps *h) {
  return (int)std::min<long long>((h->op.n + WG - 1) / WG, 16384);
}

static int projection_checks(cup2d_ps *h, double dt) {
  if (!h->have_mesh) return fail(h, CUP2D_PS_ENOMESH, "no mesh");
  if (!(dt > 0)) return fail(h, CUP2D_PS_EINVAL, "dt must be positive");
  return CUP2D_PS_OK;
}

static int pressure_rhs_body(cup2d_ps *h, const double *d_u, const double *d_v, const double *d_ud,
                             const double *d_vd, const double *d_chi, double dt, double *d_rhs) {
  if (!h || !d_u || !d_v || !d_ud || !d_vd || !d_chi || !d_rhs)
    return fail(h, CUP2D_PS_EINVAL, "NULL argument");
  // the kernel reads neighbours of every input while writing rhs
  if (d_rhs == d_u || d_rhs == d_v || d_rhs == d_ud || d_rhs == d_vd || d_rhs == d_chi)
    return fail(h, CUP2D_PS_EINVAL, "rhs must not alias an input");
  if (const int rc = projection_checks(h, dt)) return rc;
  HIPCHK(h, hipSetDevice(h->device));
  k_pressure_rhs<<<cell_grid(h), WG, 0, h->stream>>>((int)h->op.nb, h->d_nbr, h->d_mask, h->d_h, d_u,
                                                    d_v, d_ud, d_vd, d_chi, dt, d_rhs);
  if (h->n_exc > 0)
    k_pressure_rhs_cf<<<exc_grid((int)h->n_exc), WG, 0, h->stream>>>((int)h->n_exc, h->d_erow, h->d_nbr,
                                                                    h->d_fid, h->d_h, d_u, d_v, d_ud,
                                                                    d_vd, d_chi, dt, d_rhs);
  HIPCHK(h, hipGetLastError());
  HIPCHK(h, hipStreamSynchronize(h->stream));
  return CUP2D_PS_OK;
}

static int pressure_correct_body(cup2d_ps *h, const double *d_p, double dt, double *d_u, double *d_v) {
  if (!h || !d_p || !d_u || !d_v) return fail(h, CUP2D_PS_EINVAL, "NULL argument");
  if (d_u == d_v || d_p == d_u || d_p == d_v) return fail(h, CUP2D_PS_EINVAL, "p, u and v must not alias");
  if (const int rc = projection_checks(h, dt)) return rc;
  HIPCHK(h, hipSetDevice(h->device));
  k_pressure_correct<<<cell_grid(h), WG, 0, h->stream>>>((int)h->op.nb, h->d_nbr, h->d_mask, h->d_h, d_p,
                                                        dt, d_u, d_v);
  if (h->n_exc > 0)
    k_pressure_correct_cf<<<exc_grid((int)h->n_exc), WG, 0, h->stream>>>((int)h->n_exc, h->d_erow,
                                                                        h->d_nbr, h->d_fid, h->d_h, d_p,
                                                                        dt, d_u, d_v);
  HIPCHK(h, hipGetLastError());
  HIPCHK(h, hipStreamSynchronize(h->stream));
  return CUP2D_PS_OK;
}

int cup2d_ps_pressure_rhs_device(cup2d_ps *h, const double *d_u, const double *d_v,
                                 const double *d_udef_u, const double *d_udef_v, const double *d_chi,
                                 double dt, double *d_rhs) {
  return guarded(h, [&] { return pressure_rhs_body(h, d_u, d_v, d_udef_u, d_udef_v, d_chi, dt, d_rhs); });
}

int cup2d_ps_pressure_correct_device(cup2d_ps *h, const double *d_p, double dt, double *d_u,
                                     double *d_v) {
  return guarded(h, [&] { return pressure_correct_body(h, d_p, dt, d_u, d_v); });
}

int cup2d_ps_spmv_device(cup2d_ps *h, const double *d_x, double *d_y, int kind) {
  if (!h || !d_x || !d_y) return fail(h, CUP2D_PS_EINVAL, "NULL argument");
  if (!h->have_mesh) return fail(h, CUP2D_PS_ENOMESH, "no mesh");
  if (kind != 0 && kind != 1) return fail(h, CUP2D_PS_EINVAL, "kind must be 0 or 1");
  if (kind == 1 && !h->has_csr) return fail(h, CUP2D_PS_EINVAL, "no CSR copy: the mesh was assembled on the device");
  HIPCHK(h, hipSetDevice(h->device));
  const long n = (long)h->op.n;
  if (kind == 0) {
    k_spmv<double><<<h->G, WG, 0, h->stream>>>(view(h), h->Gb, d_x, d_y, h->d_pm);
  } else {
    const int G = (int)std::min<long>(4096, (n + WG - 1) / WG);
    k_spmv_csr<<<G, WG, 0, h->stream>>>(n, h->d_csr_rp, h->d_csr_col, h->d_csr_val, d_x, d_y);
    if (h->op.mean_row >= 0)
      k_hsum<double><<<h->G, WG, 0, h->stream>>>((int)h->op.nb, d_x, h->d_h2, h->d_pm);
  }
  if (h->op.mean_row >= 0)
    k_meanrow_fix<double><<<1, WG, 0, h->stream>>>(h->d_pm, h->G, h->op.mean_row, d_y);
  HIPCHK(h, hipGetLastError());
  HIPCHK(h, hipStreamSynchronize(h->stream));
  return CUP2D_PS_OK;
}

int cup2d_ps_spmv_time(cup2d_ps *h, const double *d_x, double *d_y, int kind, int reps,
                       double *ms_per_launch) {
  if (!h || !d_x || !d_y || !ms_per_launch || reps <= 0) return fail(h, CUP2D_PS_EINVAL, "bad argument");
  if (!h->have_mesh) return fail(h, CUP2D_PS_ENOMESH, "no mesh");
  if (kind != 0 && kind != 1) return fail(h, CUP2D_PS_EINVAL, "kind must be 0 or 1");
  if (kind == 1 && !h->has_csr) return fail(h, CUP2D_PS_EINVAL, "no CSR copy: the mesh was assembled on the device");
  HIPCHK(h, hipSetDevice(h->device));
  const long n = (long)h->op.n;
  const int Gc = (int)std::min<long>(4096, (n + WG - 1) / WG);
  const OpView A = view(h);
  for (int k = -1; k < reps; k++) {  // k = -1: warm-up launch, untimed
    if (k == 0) HIPCHK(h, hipEventRecord(h->ev0, h->stream));
    if (kind == 0) k_spmv<double><<<h->G, WG, 0, h->stream>>>(A, h->Gb, d_x, d_y, h->d_pm);
    else k_spmv_csr<<<Gc, WG, 0, h->stream>>>(n, h->d_csr_rp, h->d_csr_col, h->d_csr_val, d_x, d_y);
  }
  HIPCHK(h, hipEventRecord(h->ev1, h->stream));
  HIPCHK(h, hipEventSynchronize(h->ev1));
  HIPCHK(h, hipGetLastError());
  float ms = 0;
  HIPCHK(h, hipEventElapsedTime(&ms, h->ev0, h->ev1));
  *ms_per_launch = ms / reps;
  return CUP2D_PS_OK;
}

static int kernel_time_body(cup2d_ps *h, int which, int reps, double *ms_per_launch) {
  if (!h || !ms_per_launch || reps <= 0 || which < 0 || which > 4)
    return fail(h, CUP2D_PS_EINVAL, "bad argument");
  if (!h->have_mesh) return fail(h, CUP2D_PS_ENOMESH, "no mesh");
  HIPCHK(h, hipSetDevice(h->device));
  const size_t n = h->op.n;
  // inputs: x = 0, b = a fixed non-constant pattern (a constant b would be
  // projected away entirely); r0 = b - A x and the initial state in slot 0
  std::vector<double> bh(n);
  for (size_t i = 0; i < n; i++) bh[i] = double(int(i % 7) - 3);
  HIPCHK(h, hipMemsetAsync(h->d_xio, 0, n * sizeof(double), h->stream));
  HIPCHK(h, hipMemcpyAsync(h->d_b, bh.data(), n * sizeof(double), hipMemcpyHostToDevice, h->stream));
  h->h_s[0] = h->h_s[1] = marker_state(h, 0.0, 0.0, 1 << 20, 1 << 30);
  h->h_s[0].stall = h->h_s[1].stall = 0;
  HIPCHK(h, hipMemcpyAsync(h->d_s, h->h_s, 2 * sizeof(Scal), hipMemcpyHostToDevice, h->stream));
  const int G = h->G, nb = (int)h->op.nb;
  if (h->op.mean_row >= 0) k_hsum<double><<<G, WG, 0, h->stream>>>(nb, h->d_xio, h->d_h2, h->d_pm);
  if (h->mixed)
    k_init<float><<<G, WG, 0, h->stream>>>(view(h), h->Gb, h->d_b, h->d_xio, h->d_x, h->d_xopt,
                                           h->f_r, h->f_rhat, h->f_p, h->f_nu, h->d_pm, G, h->d_p5);
  else
    k_init<double><<<G, WG, 0, h->stream>>>(view(h), h->Gb, h->d_b, h->d_xio, h->d_x, h->d_xopt,
                                            h->d_r, h->d_rhat, h->d_p, h->d_nu, h->d_pm, G, h->d_p5);
  k_eval<<<1, WG, 0, h->stream>>>(h->d_s, 0, h->d_p5, G);
  // one full iteration (slot 1) fills every partial; then the kernel alone,
  // again in slot 1: each launch redoes the same work from the same state
  launch_iteration(h, 1);
  HIPCHK(h, hipGetLastError());
  const OpView A = view(h);
  const long mr = h->op.mean_row;
  hipStream_t st = h->stream;
  auto launch = [&]() {
    if (h->mixed) {
      switch (which) {
        case 0: k_pupdate_mf<<<h->Gm, WG, 0, st>>>(nb, h->d_pf, h->f_r, h->f_rhat, h->f_p, h->f_nu,
                                                   h->f_z1, h->d_s, 1, h->d_p5, G); break;
        case 1: k_spmv_nu<float><<<G, WG, 0, st>>>(A, h->Gb, h->f_z1, h->f_nu, h->f_rhat, h->d_s, 1,
                                                   h->d_p2); break;
        case 2: k_supdate_mf<<<h->Gm, WG, 0, st>>>(nb, h->d_pf, h->f_r, h->f_nu, h->f_rhat, h->f_z2,
                                                   h->d_s, 1, h->d_p2, G, mr); break;
        case 3: k_spmv_t<float><<<G, WG, 0, st>>>(A, h->Gb, h->f_z2, h->f_t, h->f_r, h->d_s, 1,
                                                  h->d_p4); break;
        default: k_xrupdate<float><<<G, WG, 0, st>>>(nb, h->d_x, h->d_xopt, h->f_z1, h->f_z2, h->f_r,
                                                     h->f_t, h->f_rhat, h->d_s, 1, h->d_p4, G,
                                                     h->d_p5, mr);
      }
    } else {
      switch (which) {
        case 0:
          k_pupdate<<<h->Gp, WG, 0, st>>>(nb, h->d_P, h->d_r, h->d_rhat, h->d_p, h->d_nu, h->d_z1, h->d_s, 1, h->d_p5, G);
          break;
        case 1: k_spmv_nu<double><<<G, WG, 0, st>>>(A, h->Gb, h->d_z1, h->d_nu, h->d_rhat, h->d_s, 1,
                                                    h->d_p2); break;
        case 2:
          k_supdate<<<h->Gp, WG, 0, st>>>(nb, h->d_P, h->d_r, h->d_nu, h->d_rhat, h->d_z2, h->d_s, 1, h->d_p2, G, mr);
          break;
        case 3: k_spmv_t<double><<<G, WG, 0, st>>>(A, h->Gb, h->d_z2, h->d_t, h->d_r, h->d_s, 1,
                                                   h->d_p4); break;
        default: k_xrupdate<double><<<G, WG, 0, st>>>(nb, h->d_x, h->d_xopt, h->d_z1, h->d_z2,
                                                      h->d_r, h->d_t, h->d_rhat, h->d_s, 1, h->d_p4,
                                                      G, h->d_p5, mr);
      }
    }
  };
  launch();  // warm-up
  HIPCHK(h, hipEventRecord(h->ev0, st));
  for (int k = 0; k < reps; k++) launch();
  HIPCHK(h, hipEventRecord(h->ev1, st));
  HIPCHK(h, hipEventSynchronize(h->ev1));
  HIPCHK(h, hipGetLastError());
  float ms = 0;
  HIPCHK(h, hipEventElapsedTime(&ms, h->ev0, h->ev1));
  *ms_per_launch = ms / reps;
  return CUP2D_PS_OK;
}

int cup2d_ps_precond_device(cup2d_ps *h, const double *d_r, double *d_z) {
  if (!h || !d_r || !d_z) return fail(h, CUP2D_PS_EINVAL, "NULL argument");
  if (!h->have_mesh) return fail(h, CUP2D_PS_ENOMESH, "no mesh");
  HIPCHK(h, hipSetDevice(h->device));
  const int nb = (int)h->op.nb;
  k_prec<<<h->Gp, WG, 0, h->stream>>>(nb, h->d_P, d_r, d_z);
  HIPCHK(h, hipGetLastError());
  HIPCHK(h, hipStreamSynchronize(h->stream));
  return CUP2D_PS_OK;
}

int cup2d_ps_precond_bf16_device(cup2d_ps *h, const float *d_r, float *d_z) {
  if (!h || !d_r || !d_z) return fail(h, CUP2D_PS_EINVAL, "NULL argument");
  if (!h->have_mesh) return fail(h, CUP2D_PS_ENOMESH, "no mesh");
  HIPCHK(h, hipSetDevice(h->device));
  k_prec_mf<<<h->Gm, WG, 0, h->stream>>>((int)h->op.nb, h->d_pf, d_r, d_z);
  HIPCHK(h, hipGetLastError());
  HIPCHK(h, hipStreamSynchronize(h->stream));
  return CUP2D_PS_OK;
}

static int residual_body(cup2d_ps *h, const double *d_b, const double *d_x, double out[4]) {
  if (!h || !d_b || !d_x || !out) return fail(h, CUP2D_PS_EINVAL, "NULL argument");
  if (!h->have_mesh) return fail(h, CUP2D_PS_ENOMESH, "no mesh");
  HIPCHK(h, hipSetDevice(h->device));
  return residual_norms(h, d_b, d_x, out);
}

int cup2d_ps_sizes(const cup2d_ps *h, long long *n, long long *nnz, long long *n_explicit,
                   long long *mean_row) {
  if (!h) return CUP2D_PS_EINVAL;
  if (!h->have_mesh) return CUP2D_PS_ENOMESH;
  if (n) *n = h->op.n;
  if (nnz) *nnz = h->op.nnz();
  if (n_explicit) *n_explicit = h->n_exc;
  if (mean_row) *mean_row = h->op.mean_row;
  return CUP2D_PS_OK;
}

int cup2d_ps_get_csr(const cup2d_ps *h, long long *row_ptr, int *col, double *val) {
  if (!h || !row_ptr || !col || !val) return CUP2D_PS_EINVAL;
  if (!h->have_mesh) return CUP2D_PS_ENOMESH;
  if (!h->has_csr) {
    h->err = "no CSR copy: the mesh was assembled on the device";
    return CUP2D_PS_EINVAL;
  }
  for (size_t i = 0; i < h->op.rp.size(); i++) row_ptr[i] = h->op.rp[i];
  std::memcpy(col, h->op.col.data(), h->op.col.size() * sizeof(int));
  std::memcpy(val, h->op.val.data(), h->op.val.size() * sizeof(double));
  return CUP2D_PS_OK;
}

static long long op_csr_body(const cup2d_mesh_desc *m, int mean_constraint, long long *row_ptr,
                             int *col, double *val, long long cap, long long *mean_row) {
  if (!m) {
    g_create_err = "NULL mesh";
    return -1;
  }
  MeshDesc md{m->bpdx, m->bpdy, m->levels, m->periodic_x, m->periodic_y, m->extent, m->n_blocks,
              m->level, m->bx, m->by};
  Operator op;
  std::string err;
  if (!build_operator(md, mean_constraint, op, err)) {
    g_create_err = err;
    return -1;
  }
  if (mean_row) *mean_row = op.mean_row;
  const long long nnz = op.nnz();
  if (nnz > cap || !row_ptr || !col || !val) return -nnz;
  for (size_t i = 0; i < op.rp.size(); i++) row_ptr[i] = op.rp[i];
  std::memcpy(col, op.col.data(), nnz * sizeof(int));
  std::memcpy(val, op.val.data(), nnz * sizeof(double));
  return nnz;
}

int cup2d_ps_dmalloc(cup2d_ps *h, size_t bytes, void **ptr) {
  if (!h || !ptr) return fail(h, CUP2D_PS_EINVAL, "NULL argument");
  HIPCHK(h, hipSetDevice(h->device));
  HIPCHK(h, hipMalloc(ptr, bytes ? bytes : 1));
  return CUP2D_PS_OK;
}
int cup2d_ps_dfree(cup2d_ps *h, void *ptr) {
  if (!h) return CUP2D_PS_EINVAL;
  HIPCHK(h, hipSetDevice(h->device));
  HIPCHK(h, hipFree(ptr));
  return CUP2D_PS_OK;
}
int cup2d_ps_h2d(cup2d_ps *h, void *dst, const void *src, size_t bytes) {
  if (!h || !dst || !src) return fail(h, CUP2D_PS_EINVAL, "NULL argument");
  HIPCHK(h, hipSetDevice(h->device));
  HIPCHK(h, hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
  return CUP2D_PS_OK;
}
int cup2d_ps_d2h(cup2d_ps *h, void *dst, const void *src, size_t bytes) {
  if (!h || !dst || !src) return fail(h, CUP2D_PS_EINVAL, "NULL argument");
  HIPCHK(h, hipSetDevice(h->device));
  HIPCHK(h, hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
  return CUP2D_PS_OK;
}

const char *cup2d_ps_last_error(const cup2d_ps *h) {
  return h ? h->err.c_str() : g_create_err.c_str();
}

int cup2d_ps_create(const cup2d_ps_config *cfg, cup2d_ps **out) {
  return guarded(nullptr, [&] { return create_body(cfg, out); });
}

int cup2d_ps_set_mesh(cup2d_ps *h, const cup2d_mesh_desc *m) {
  return guarded(h, [&] { return set_mesh_body(h, m); });
}

int cup2d_ps_set_mesh_on_device(cup2d_ps *h, const cup2d_mesh_desc *m) {
  return guarded(h, [&] { return set_mesh_on_device_body(h, m); });
}

int cup2d_ps_pressure_step_device(cup2d_ps *h, const double *d_rhs, double *d_p, double tol_abs,
                                  double tol_rel, int max_restarts, cup2d_ps_stats *stats) {
  return guarded(h, [&] {
    return pressure_step_body(h, d_rhs, d_p, tol_abs, tol_rel, max_restarts, stats);
  });
}

int cup2d_ps_kernel_time(cup2d_ps *h, int which, int reps, double *ms_per_launch) {
  return guarded(h, [&] { return kernel_time_body(h, which, reps, ms_per_launch); });
}

int cup2d_ps_residual_device(cup2d_ps *h, const double *d_b, const double *d_x, double out[4]) {
  return guarded(h, [&] { return residual_body(h, d_b, d_x, out); });
}

// ---- flow operators of the projection step (flow.h) --------------------------

static FlowMesh flow_mesh(const cup2d_ps *h) {
  return FlowMesh{(int)h->op.nb, h->d_nbr, h->d_fid, h->d_h, h->d_bxy};
}
static int flow_grid(const cup2d_ps *h) {
  return (int)std::min<long long>((h->op.nb + FWB - 1) / FWB, 8192);
}

static int advect_diffuse_body(cup2d_ps *h, double *d_u, double *d_v, double uinf_x, double uinf_y,
                               double nu, double dt) {
  if (!h || !d_u || !d_v) return fail(h, CUP2D_PS_EINVAL, "NULL argument");
  if (d_u == d_v) return fail(h, CUP2D_PS_EINVAL, "u and v must not alias");
  if (!h->have_mesh) return fail(h, CUP2D_PS_ENOMESH, "no mesh");
  if (!(dt > 0) || !(nu >= 0)) return fail(h, CUP2D_PS_EINVAL, "dt must be positive and nu non-negative");
  HIPCHK(h, hipSetDevice(h->device));
  const FlowMesh M = flow_mesh(h);
  // the solver's work vectors serve as the stage's work space: u1, v1 in
  // p, nu; the right-hand sides in z1, z2
  double *u1 = h->d_p, *v1 = h->d_nu, *tu = h->d_z1, *tv = h->d_z2;
  const int Gf = flow_grid(h), Gc = cell_grid(h);
  hipStream_t s = h->stream;
  kf_advdiff_rhs<<<Gf, WG, 0, s>>>(M, d_u, d_v, uinf_x, uinf_y, nu, dt, tu, tv);
  kf_stage<<<Gc, WG, 0, s>>>(M, d_u, d_v, tu, tv, 1, u1, v1);
  kf_advdiff_rhs<<<Gf, WG, 0, s>>>(M, u1, v1, uinf_x, uinf_y, nu, dt, tu, tv);
  kf_stage<<<Gc, WG, 0, s>>>(M, d_u, d_v, tu, tv, 0, d_u, d_v);
  HIPCHK(h, hipGetLastError());
  HIPCHK(h, hipStreamSynchronize(s));
  return CUP2D_PS_OK;
}

int cup2d_ps_advect_diffuse_device(cup2d_ps *h, double *d_u, double *d_v, double uinf_x,
                                   double uinf_y, double nu, double dt) {
  return guarded(h, [&] { return advect_diffuse_body(h, d_u, d_v, uinf_x, uinf_y, nu, dt); });
}

int cup2d_ps_penalize_device(cup2d_ps *h, const double *d_chi, double us, double vs, double lambda,
                             double dt, double *d_u, double *d_v) {
  return guarded(h, [&]() -> int {
    if (!h || !d_chi || !d_u || !d_v) return fail(h, CUP2D_PS_EINVAL, "NULL argument");
    if (d_u == d_v || d_chi == d_u || d_chi == d_v) return fail(h, CUP2D_PS_EINVAL, "chi, u and v must not alias");
    if (!h->have_mesh) return fail(h, CUP2D_PS_ENOMESH, "no mesh");
    if (!(dt > 0) || !(lambda >= 0)) return fail(h, CUP2D_PS_EINVAL, "dt must be positive and lambda non-negative");
    HIPCHK(h, hipSetDevice(h->device));
    kf_penalize<<<cell_grid(h), WG, 0, h->stream>>>((long)h->op.n, d_chi, us, vs, lambda, dt, d_u, d_v);
    HIPCHK(h, hipGetLastError());
    HIPCHK(h, hipStreamSynchronize(h->stream));
    return CUP2D_PS_OK;
  });
}

int cup2d_ps_chi_disk_device(cup2d_ps *h, double cx, double cy, double r, double *d_chi) {
  return guarded(h, [&]() -> int {
    if (!h || !d_chi) return fail(h, CUP2D_PS_EINVAL, "NULL argument");
    if (!h->have_mesh) return fail(h, CUP2D_PS_ENOMESH, "no mesh");
    if (!(r > 0)) return fail(h, CUP2D_PS_EINVAL, "radius must be positive");
    HIPCHK(h, hipSetDevice(h->device));
    kf_chi_disk<<<cell_grid(h), WG, 0, h->stream>>>(flow_mesh(h), cx, cy, r, d_chi);
    HIPCHK(h, hipGetLastError());
    HIPCHK(h, hipStreamSynchronize(h->stream));
    return CUP2D_PS_OK;
  });
}

int cup2d_ps_max_speed_device(cup2d_ps *h, const double *d_u, const double *d_v, double uinf_x,
                              double uinf_y, double *umax) {
  return guarded(h, [&]() -> int {
    if (!h || !d_u || !d_v || !umax) return fail(h, CUP2D_PS_EINVAL, "NULL argument");
    if (!h->have_mesh) return fail(h, CUP2D_PS_ENOMESH, "no mesh");
    HIPCHK(h, hipSetDevice(h->device));
    const int G = std::min<int>(MAXG, cell_grid(h));
    kf_max_speed<<<G, WG, 0, h->stream>>>((long)h->op.n, d_u, d_v, uinf_x, uinf_y, h->d_pfin);
    HIPCHK(h, hipGetLastError());
    double part[4 * MAXG];
    HIPCHK(h, hipMemcpyAsync(part, h->d_pfin, 4 * (size_t)G * sizeof(double), hipMemcpyDeviceToHost,
                             h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    double m = 0.0;
    for (int i = 0; i < 4 * G; i++) m = std::max(m, part[i]);
    *umax = m;
    return CUP2D_PS_OK;
  });
}

int cup2d_ps_block_score_device(cup2d_ps *h, const double *d_u, const double *d_v,
                                const double *d_chi, double rtol, double *score) {
  return guarded(h, [&]() -> int {
    if (!h || !d_u || !d_v || !d_chi || !score) return fail(h, CUP2D_PS_EINVAL, "NULL argument");
    if (!h->have_mesh) return fail(h, CUP2D_PS_ENOMESH, "no mesh");
    if (!(rtol > 0)) return fail(h, CUP2D_PS_EINVAL, "rtol must be positive");
    HIPCHK(h, hipSetDevice(h->device));
    double *d_sc = h->d_t;  // nb doubles of the solver's work vector (n = 64 nb)
    kf_block_score<<<flow_grid(h), WG, 0, h->stream>>>(flow_mesh(h), d_u, d_v, d_chi, rtol, h->h_min, d_sc);
    HIPCHK(h, hipGetLastError());
    HIPCHK(h, hipMemcpyAsync(score, d_sc, (size_t)h->op.nb * sizeof(double), hipMemcpyDeviceToHost,
                             h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    return CUP2D_PS_OK;
  });
}

static int transfer_body(cup2d_ps *h, const cup2d_mesh_desc *om, const cup2d_mesh_desc *nm,
                         const int *src_kind, const long long *src, int n_fields,
                         const double *const *d_old, double *const *d_new) {
  if (!h || !om || !nm || !src_kind || !src || n_fields < 0 || (n_fields > 0 && (!d_old || !d_new)))
    return fail(h, CUP2D_PS_EINVAL, "NULL argument");
  const long long onb = om->n_blocks, nnb = nm->n_blocks;
  if (onb <= 0 || nnb <= 0) return fail(h, CUP2D_PS_EINVAL, "empty mesh");
  for (long long b = 0; b < nnb; b++) {  // the plan's sources are old blocks of the right kind
    const int k = src_kind[b];
    if (k < 0 || k > 2) return fail(h, CUP2D_PS_EINVAL, "src_kind out of range");
    for (int q = 0; q < (k == 2 ? 4 : 1); q++)
      if (src[4 * b + q] < 0 || src[4 * b + q] >= onb) return fail(h, CUP2D_PS_EINVAL, "src out of range");
    bool ok = true;
    if (k == 0) {
      const long long a = src[4 * b];
      ok = nm->level[b] == om->level[a] && nm->bx[b] == om->bx[a] && nm->by[b] == om->by[a];
    } else if (k == 1) {
      const long long a = src[4 * b];
      const int d = nm->level[b] - om->level[a];
      ok = d >= 1 && d <= 15 && (nm->bx[b] >> d) == om->bx[a] && (nm->by[b] >> d) == om->by[a];
    } else {
      for (int q = 0; q < 4; q++) {
        const long long c = src[4 * b + q];
        ok = ok && om->level[c] == nm->level[b] + 1 && om->bx[c] == 2 * nm->bx[b] + (q & 1) &&
             om->by[c] == 2 * nm->by[b] + (q >> 1);
      }
    }
    if (!ok) return fail(h, CUP2D_PS_EINVAL, "block " + std::to_string(b) + " does not match its source");
  }
  for (int f = 0; f < n_fields; f++) {
    if (!d_old[f] || !d_new[f]) return fail(h, CUP2D_PS_EINVAL, "NULL field");
    for (int g = 0; g < n_fields; g++)
      if ((const double *)d_new[g] == d_old[f] || (g != f && d_new[g] == d_new[f]))
        return fail(h, CUP2D_PS_EINVAL, "fields must not alias");
  }
  HIPCHK(h, hipSetDevice(h->device));
  std::vector<int2> nij(nnb), oij(onb);
  for (long long b = 0; b < nnb; b++) nij[b] = make_int2(nm->bx[b], nm->by[b]);
  for (long long b = 0; b < onb; b++) oij[b] = make_int2(om->bx[b], om->by[b]);
  void *buf[6] = {};
  const size_t sz[6] = {nnb * sizeof(int), 4 * nnb * sizeof(long long), nnb * sizeof(int),
                        nnb * sizeof(int2), onb * sizeof(int), onb * sizeof(int2)};
  const void *host[6] = {src_kind, src, nm->level, nij.data(), om->level, oij.data()};
  int rc = CUP2D_PS_OK;
  for (int k = 0; k < 6 && !rc; k++) {
    if (hipMalloc(&buf[k], sz[k]) != hipSuccess) rc = fail(h, CUP2D_PS_ENOMEM, "device allocation failed");
    else if (hipMemcpyAsync(buf[k], host[k], sz[k], hipMemcpyHostToDevice, h->stream) != hipSuccess)
      rc = fail(h, CUP2D_PS_EHIP, "upload failed");
  }
  if (!rc) {
    const Transfer T{(int)nnb, (const int *)buf[0], (const long long *)buf[1], (const int *)buf[2],
                     (const int2 *)buf[3], (const int *)buf[4], (const int2 *)buf[5]};
    const int G = (int)std::min<long long>((nnb * 64 + WG - 1) / WG, 16384);
    for (int f = 0; f < n_fields; f++) kf_transfer<<<G, WG, 0, h->stream>>>(T, d_old[f], d_new[f]);
    if (hipGetLastError() != hipSuccess || hipStreamSynchronize(h->stream) != hipSuccess)
      rc = fail(h, CUP2D_PS_EHIP, "transfer kernel failed");
  }
  (void)hipStreamSynchronize(h->stream);
  for (void *p : buf)
    if (p) (void)hipFree(p);
  return rc;
}

int cup2d_ps_transfer_device(cup2d_ps *h, const cup2d_mesh_desc *old_mesh,
                             const cup2d_mesh_desc *new_mesh, const int *src_kind, const long long *src,
                             int n_fields, const double *const *d_old, double *const *d_new) {
  return guarded(h, [&] { return transfer_body(h, old_mesh, new_mesh, src_kind, src, n_fields, d_old, d_new); });
}

long long cup2d_op_csr(const cup2d_mesh_desc *m, int mean_constraint, long long *row_ptr, int *col,
                       double *val, long long cap, long long *mean_row) {
  try {  // errors are -1 here: a positive value is the entry count
    return op_csr_body(m, mean_constraint, row_ptr, col, val, cap, mean_row);
  } catch (const std::exception &e) {
    fail(nullptr, CUP2D_PS_EINVAL, e.what());
    return -1;
  }
}

}  // extern "C"
