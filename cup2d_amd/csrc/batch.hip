// batch.hip -- many independent AMR Poisson solves advanced together.
//
// Each instance is the solve of ExpAMRSolver::solve (ExpAMRSolver.cpp:354-416)
// on its own mesh: BiCGSTAB (BiCGSTAB.cu:412-620) with the 64x64 block
// preconditioner and h^2-weighted mean removal.  The instances' blocks are
// concatenated into one block-diagonal operator and every iteration advances
// all unfinished instances at once:
//
//  * work units: "items" (a run of up to `ib` blocks of one instance) for the
//    SpMV and vector kernels, a workgroup per item; "tiles" (16 blocks of one
//    instance) for the preconditioner kernels, a wavefront per tile.  The
//    kernels walk the unit lists grid-stride with fixed grids (so a chunk of
//    iterations is captured once as a hipGraph) and skip the units of
//    instances whose device-side state says done: when few instances remain,
//    their units are still spread over the whole GPU.
//  * per-instance scalars: after each reduction point one kernel with a
//    workgroup per instance re-reduces that instance's item partials in a
//    fixed order and advances its scalar recurrence (deterministic, no
//    atomics or fences).
//  * the item size `ib` is picked by a timing autotuner and cached per
//    batch shape.
//
// Precisions: FP64, or FP32-stored Krylov vectors with FP64 arithmetic and
// the BF16 MFMA preconditioner inside per-instance FP64 iterative refinement
// (the scheme of the single-solve mixed path).

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <tuple>
#include <vector>

#include "../../include/cup2d_poisson.h"
#include "assemble_device.h"
#include "assembly.h"
#include "kernels.h"

using namespace cup2d;

namespace {

constexpr int NQB = 7;        // update totals (NQ5) + sum h^2 x (initial mean row)
constexpr int BGRID = 2048;   // workgroups of the item / tile kernels (grid-stride beyond)
constexpr int MAX_INST = 4096;
constexpr int MAX_ITEM_BLOCKS = 256;  // item_blocks keeps one block header per lane: 64 * WAVES

struct BView {
  OpView A;                          // concatenated operator (A.mean_row unused)
  const int4 *item;                  // {instance, first block / explicit row, count, 0};
                                     // count < 0: explicit item of -count rows
  const int *item_begin;             // [K+1] first item of each instance
  const int4 *tile;                  // {instance, first block, blocks, 0}
  const long long *mean_row;         // [K] global row of the mean constraint or -1
  int K, n_items, n_tiles;
  int xg_log2;                       // operator kernels' XCD run length: 2^xg_log2 items (0: plain order)
};

// Workgroup reduction of NQ quantities; thread 0 stores them at
// part[q*NI + idx].
template <int NQ, unsigned MAXQ>
__device__ void store_item_partials(double (&v)[NQ], double *part, int NI, int idx) {
  __shared__ double red[NQ][WAVES];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
  for (int q = 0; q < NQ; q++) v[q] = ((MAXQ >> q) & 1) ? wave_max(v[q]) : wave_sum(v[q]);
  if (lane == 0)
#pragma unroll
    for (int q = 0; q < NQ; q++) red[q][w] = v[q];
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int q = 0; q < NQ; q++) {
      double a = red[q][0];
      for (int i = 1; i < WAVES; i++) a = ((MAXQ >> q) & 1) ? nmax(a, red[q][i]) : a + red[q][i];
      part[(long)q * NI + idx] = a;
    }
  }
  __syncthreads();  // red[] is reused by the workgroup's next item
}

// Totals of partials i0..i1-1 of each quantity (fixed order; every thread
// gets them).
template <int NQ, unsigned MAXQ>
__device__ void range_totals(const double *__restrict__ part, int NI, int i0, int i1,
                             double (&out)[NQ]) {
  __shared__ double fin[NQ][WAVES];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  double a[NQ];
#pragma unroll
  for (int q = 0; q < NQ; q++) a[q] = 0.0;
  for (int i = i0 + threadIdx.x; i < i1; i += WG) {
    double p[NQ];  // the NQ loads of a partial are in flight together
#pragma unroll
    for (int q = 0; q < NQ; q++) p[q] = part[(long)q * NI + i];
#pragma unroll
    for (int q = 0; q < NQ; q++) a[q] = ((MAXQ >> q) & 1) ? nmax(a[q], p[q]) : a[q] + p[q];
  }
#pragma unroll
  for (int q = 0; q < NQ; q++) {
    const double r = ((MAXQ >> q) & 1) ? wave_max(a[q]) : wave_sum(a[q]);
    if (lane == 0) fin[q][w] = r;
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < NQ; q++) {
    double r = fin[q][0];
    for (int i = 1; i < WAVES; i++) r = ((MAXQ >> q) & 1) ? nmax(r, fin[q][i]) : r + fin[q][i];
    out[q] = r;
  }
}

// ----------------------------------------------------------------------------
// item kernels.  A block item (count > 0) is taken by a workgroup, a
// wavefront per UB blocks at a time (lane = cell); an explicit item
// (count < 0: -count explicit rows from the first one) by a workgroup, a
// thread per row.  Every item of a running instance stores its partials,
// zeros where it has nothing to add.
// ----------------------------------------------------------------------------

// blocks per wavefront in flight in the operator kernels: FP32 vectors carry
// half the bytes per load, so twice the blocks
constexpr int UB64 = 4, UB32 = 4;
template <class TV>
constexpr int ub_of() { return sizeof(TV) == 4 ? UB32 : UB64; }

// Which instances a kernel works on.
enum Filter { F_ALL, F_RUNNING, F_LIVE, F_REPLACE };
__device__ __forceinline__ bool wanted(const Scal &s, int f) {
  switch (f) {
    case F_RUNNING: return !s.idle;                 // iterating
    case F_LIVE: return !s.done;                    // not finished
    case F_REPLACE: return s.replace && !s.done;    // replacement pending
    default: return true;
  }
}

// Block headers of a block item, one per lane: lane c of wave w holds block
// ((c / UB) * WAVES + w) * UB + c % UB of the item (items have at most
// 64 * WAVES blocks), the order in which item_blocks visits them; empty for
// explicit items and past the item's end.
struct Heads {
  int4 n;
  unsigned long long m;
};
template <int UB>
__device__ __forceinline__ Heads load_heads(const BView &V, int4 mt) {
  Heads h{make_int4(-1, -1, -1, -1), 0ull};
  const int c = lane_id(), w = wave_id();
  const int jl = ((c / UB) * WAVES + w) * UB + (c % UB);
  if (jl < mt.z) {
    h.n = V.A.nbr[mt.y + jl];
    h.m = V.A.mask[mt.y + jl];
  }
  return h;
}

__device__ __forceinline__ int4 readlane4(int4 v, int l) {
  return make_int4(__builtin_amdgcn_readlane(v.x, l), __builtin_amdgcn_readlane(v.y, l),
                   __builtin_amdgcn_readlane(v.z, l), __builtin_amdgcn_readlane(v.w, l));
}

// Calls body(item, meta, heads) for the workgroup's items (grid-stride)
// whose instance passes `filter` (every item when st is null).  The metas of
// up to 64 of the workgroup's items are loaded at once (one per lane), and
// with UB > 0 the next item's block headers (load_heads<UB>) are loaded
// while the current item is processed, so an item's data loads wait on no
// other load.  body may synchronize the workgroup (the skip is uniform).
template <int UB = 0, bool XCD = false, class F>
__device__ __forceinline__ void running_items(const BView &V, const Scal *st, F &&body,
                                              int filter = F_RUNNING) {
  // XCD: with a grid of a multiple of 8 workgroups (workgroup g runs on XCD
  // g % 8), the items go to the XCDs in runs of XG = 2^xg_log2 consecutive
  // items, run r to XCD r % 8, so that neighbouring items (whose blocks read
  // each other's faces) mostly share one XCD's L2 while every instance --
  // the last ones running included -- still spreads over all eight XCDs.
  // Item t of XCD x is (t / XG) 8 XG + x XG + t % XG; XG = 1 is the plain
  // grid-stride order.
  const int NG = (int)gridDim.x, c = lane_id();
  const bool xm = XCD && (NG & 7) == 0;
  const int sh = xm ? V.xg_log2 : 0, XG = 1 << sh;
  const int G = xm ? NG >> 3 : NG, x = (int)blockIdx.x & 7;
  int n_mine = V.n_items, first = blockIdx.x;
  if (xm) {
    const int rem = V.n_items & (8 * XG - 1);
    n_mine = (V.n_items >> (sh + 3) << sh) + min(XG, max(0, rem - x * XG));
    first = (int)blockIdx.x >> 3;
  }
  auto item_of = [&](int t) { return xm ? ((t >> sh) << (sh + 3)) + (x << sh) + (t & (XG - 1)) : t; };
  for (int base = first; base < n_mine; base += 64 * G) {
    const int cnt = min(64, (n_mine - base + G - 1) / G);
    const int mine = item_of(base + c * G);
    const int4 ml = V.item[c < cnt ? mine : item_of(base)];
    auto want = [&](int4 mt) { return !st || wanted(st[mt.x], filter); };
    int4 mt = readlane4(ml, 0);
    Heads hd{make_int4(-1, -1, -1, -1), 0ull};
    if constexpr (UB > 0)
      if (want(mt)) hd = load_heads<UB>(V, mt);
    for (int k = 0; k < cnt; k++) {
      const int4 nmt = readlane4(ml, k + 1 < cnt ? k + 1 : k);
      Heads nhd{make_int4(-1, -1, -1, -1), 0ull};
      if constexpr (UB > 0)
        if (k + 1 < cnt && want(nmt)) nhd = load_heads<UB>(V, nmt);
      if (want(mt)) body(item_of(base + k * G), mt, hd);
      mt = nmt;
      hd = nhd;
    }
  }
}

// Runs f(R, u, bl) for the item's blocks, UB = ub_of<TV>() per wave at a
// time, after issuing every load of those blocks (stencil, h^2 and aux[i]);
// hd holds the item's block headers (load_heads<UB>), read back with
// readlane.
template <class TV, class F>
__device__ __forceinline__ void item_blocks(const BView &V, const TV *__restrict__ x,
                                            const TV *__restrict__ aux, int b0, int nbk,
                                            const Heads &hd, F &&f) {
  constexpr int UB = ub_of<TV>();
  const int c = lane_id(), w = wave_id();
  const unsigned hm_lo = (unsigned)hd.m, hm_hi = (unsigned)(hd.m >> 32);
  for (int s = 0; (s * WAVES + w) * UB < nbk; s++) {
    const int j0 = (s * WAVES + w) * UB;
    int bl[UB];
    bool ok[UB];
    BlockRows<UB, TV> R;
#pragma unroll
    for (int u = 0; u < UB; u++) {
      ok[u] = j0 + u < nbk;
      bl[u] = b0 + (ok[u] ? j0 + u : j0);
      const int src = s * UB + u;
      R.nb[u] = readlane4(hd.n, src);
      R.m[u] = (unsigned long long)(unsigned)__builtin_amdgcn_readlane((int)hm_lo, src) |
               ((unsigned long long)(unsigned)__builtin_amdgcn_readlane((int)hm_hi, src) << 32);
    }
    R.load_data(V.A, x, aux, bl, c);
#pragma unroll
    for (int u = 0; u < UB; u++)
      if (ok[u]) f(R, u, bl[u]);
  }
}

// ----------------------------------------------------------------------------
// Row kernels of the operator: two threads per block row, four cells each.
// Thread t of the workgroup owns half h = t & 1 of row iy = (t >> 1) & 7 of
// block t >> 4 of a pass of RB = 16 blocks of the item: its four values come
// in with 16-byte loads (contiguous over the workgroup), the horizontal
// neighbours inside the half row stay in registers, the one across the
// middle of the row is swapped with the partner thread (DPP, no LDS), the
// one across the block's x face (the West face for h = 0, East for h = 1) is
// a single load, and the half rows above and below come from an LDS tile
// holding, per block, its eight rows plus the halo rows across its y faces.
// A wall face's halo is the cell itself (its term x_nb - x_c vanishes), so
// the stencil is the same five terms for every plain cell; explicit
// (coarse-fine) cells are left to the explicit items.
// ----------------------------------------------------------------------------
constexpr int RB = WG / 16;       // blocks per pass: 16 threads per block
constexpr int TROWS = RB * 10;    // LDS rows per chunk column: row below, 8 rows, row above

template <class T>
struct Chunk;
template <>
struct Chunk<double> {
  typedef double2 type;
  static constexpr int n = 4;     // 16-byte chunks per row of 8
};
template <>
struct Chunk<float> {
  typedef float4 type;
  static constexpr int n = 2;
};

__device__ __forceinline__ void split(double2 c, double *v) { v[0] = c.x; v[1] = c.y; }
__device__ __forceinline__ void split(float4 c, float *v) { v[0] = c.x; v[1] = c.y; v[2] = c.z; v[3] = c.w; }
__device__ __forceinline__ double2 join(const double *v) { return make_double2(v[0], v[1]); }
__device__ __forceinline__ float4 join(const float *v) { return make_float4(v[0], v[1], v[2], v[3]); }

// the four cells base .. base + 3 (16-byte aligned)
template <class T>
__device__ __forceinline__ void ld4c(const T *__restrict__ x, long base, T (&v)[4]) {
  typedef typename Chunk<T>::type C;
  constexpr int NC = Chunk<T>::n / 2, W = 4 / NC;
  const C *p = reinterpret_cast<const C *>(x + base);
  C c[NC];
#pragma unroll
  for (int k = 0; k < NC; k++) c[k] = p[k];
#pragma unroll
  for (int k = 0; k < NC; k++) split(c[k], v + W * k);
}
template <class T>
__device__ __forceinline__ void st4c(T *__restrict__ x, long base, const T (&v)[4]) {
  typedef typename Chunk<T>::type C;
  constexpr int NC = Chunk<T>::n / 2, W = 4 / NC;
  C *p = reinterpret_cast<C *>(x + base);
#pragma unroll
  for (int k = 0; k < NC; k++) p[k] = join(v + W * k);
}
// LDS tile: chunk k of row r at tile[k * TROWS + r]; half h of a row is
// chunks h NC / 2 .. (h + 1) NC / 2 - 1.  The threads of a row pair write
// chunk columns a multiple of the bank period apart (2-way conflicts); a
// padded column stride that removes them measured no faster.
template <class T>
__device__ __forceinline__ void st_tile(typename Chunk<T>::type *tile, int r, int h, const T (&v)[4]) {
  constexpr int NC = Chunk<T>::n / 2, W = 4 / NC;
#pragma unroll
  for (int k = 0; k < NC; k++) tile[(h * NC + k) * TROWS + r] = join(v + W * k);
}
template <class T>
__device__ __forceinline__ void ld_tile(const typename Chunk<T>::type *tile, int r, int h, T (&v)[4]) {
  constexpr int NC = Chunk<T>::n / 2, W = 4 / NC;
#pragma unroll
  for (int k = 0; k < NC; k++) split(tile[(h * NC + k) * TROWS + r], v + W * k);
}

// v of the partner thread (lane ^ 1), through DPP
__device__ __forceinline__ double swap_pair(double v) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_mov_dpp((int)b, 0xb1, 0xf, 0xf, false);  // quad_perm [1,0,3,2]
  const int hi = __builtin_amdgcn_mov_dpp((int)(b >> 32), 0xb1, 0xf, 0xf, false);
  return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}
__device__ __forceinline__ float swap_pair(float v) {
  return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0xb1, 0xf, 0xf, false));
}

// LDS of the row kernels' tile
template <class T>
constexpr int row_tile_bytes() { return Chunk<T>::n * TROWS * 16; }

// Calls f(base, y, x, a, h2, emask) for each half row of the item's blocks
// (b0 .. b0 + nbk - 1): base = first global index of the four cells, y =
// their plain stencil (FP64), x / a = the four values of x / aux, h2 = h^2
// of the block, emask = the four cells' explicit-cell bits.  Every thread
// of the workgroup must call it (it synchronizes the workgroup).
template <class T, bool AUX = true, class F>
__device__ __forceinline__ void row_blocks(const BView &V, const T *__restrict__ x,
                                           const T *__restrict__ aux, int b0, int nbk,
                                           unsigned char *tile_raw, F &&f) {
  typedef typename Chunk<T>::type C;
  C *tile = reinterpret_cast<C *>(tile_raw);
  const int t = threadIdx.x, jl = t >> 4, iy = (t >> 1) & 7, hf = t & 1;
  const bool edge = iy == 0 || iy == 7;
  for (int s0 = 0; s0 < nbk; s0 += RB) {
    const int j = s0 + jl;
    const bool ok = j < nbk;
    const int b = b0 + (ok ? j : s0);
    const int4 nb = V.A.nbr[b];
    const unsigned long long m = V.A.mask[b];
    const double h2 = V.A.h2[b];
    const long base = (long)b * 64 + iy * 8 + 4 * hf;
    T xr[4], ar[4], hr[4];
    ld4c(x, base, xr);
    if constexpr (AUX) ld4c(aux, base, ar);
    // across the x face: West (col 7 of nb.x) for the low half, East (col 0
    // of nb.y) for the high half; the cell itself at a wall
    const int nx = hf ? nb.y : nb.x;
    const T xf = x[nx >= 0 ? (long)nx * 64 + iy * 8 + (hf ? 0 : 7) : base + (hf ? 3 : 0)];
    if (edge) {
      const int nv = iy == 0 ? nb.z : nb.w;
      ld4c(x, nv >= 0 ? (long)nv * 64 + (iy == 0 ? 56 : 0) + 4 * hf : base, hr);
    }
    st_tile(tile, 10 * jl + iy + 1, hf, xr);
    if (edge) st_tile(tile, 10 * jl + (iy == 0 ? 0 : 9), hf, hr);
    // the partner's cell next to the middle of the row
    const T xm = swap_pair(hf ? xr[0] : xr[3]);
    __syncthreads();
    T dn[4], up[4];
    ld_tile(tile, 10 * jl + iy, hf, dn);
    ld_tile(tile, 10 * jl + iy + 2, hf, up);
    __syncthreads();  // the tile is rewritten by the next pass
    const double lo = hf ? (double)xm : (double)xf;   // left of cell 0
    const double hi = hf ? (double)xf : (double)xm;   // right of cell 3
    double y[4];
#pragma unroll
    for (int k = 0; k < 4; k++) {
      const double l = k > 0 ? (double)xr[k - 1] : lo;
      const double r = k < 3 ? (double)xr[k + 1] : hi;
      y[k] = (((double)dn[k] + l) + (r + (double)up[k])) - 4.0 * (double)xr[k];
    }
    if (ok) f(base, y, xr, ar, h2, (unsigned)(m >> (8 * iy + 4 * hf)) & 0xfu);
  }
}

// Stores the plain ones of four cells (all with one set of 16-byte stores
// when none is explicit).
template <class T>
__device__ __forceinline__ void st_plain(T *__restrict__ y, long base, const T (&v)[4], unsigned em) {
  if (em == 0) {
    st4c(y, base, v);
  } else {
#pragma unroll
    for (int k = 0; k < 4; k++)
      if (!((em >> k) & 1u)) y[base + k] = v[k];
  }
}

// r = b - A x (FP64), stored as T to r and rhat; p = nu = 0; x_opt = x
// (FP64) or the correction and its best = 0 (mixed: xcor, xopt FP32);
// partials of the update totals and of sum h^2 x, for the instances that
// pass `filter` (the start of a solve, or a residual replacement).  The
// mean row's residual is finished by kb_eval.
template <class T>
__global__ __launch_bounds__(WG) void kb_init(BView V, const double *__restrict__ b,
                                              const double *__restrict__ x,
                                              T *__restrict__ xopt, T *__restrict__ xcor, T *__restrict__ r,
                                              T *__restrict__ rhat, T *__restrict__ p,
                                              T *__restrict__ nu, const Scal *st, double *p5,
                                              int filter) {
  const int c = lane_id();
  running_items<UB64>(V, st, [&](int item, int4 mt, const Heads &hd) {
    const long long mr = V.mean_row[mt.x];
    double v[NQB] = {0, 0, 0, 0, 0, 0, 0};
    auto store_r = [&](long i, double y, double bi) {
      const double ri = (i == mr) ? 0.0 : bi - y;
      const T rt = (T)ri;
      r[i] = rt;
      rhat[i] = rt;
      const double rd = (double)rt;
      v[0] = nmax(v[0], fabs(rd));
      v[1] += rd * rd;
      v[4] += rd;
    };
    if (mt.z > 0) {
      item_blocks<double>(V, x, b, mt.y, mt.z, hd, [&](const auto &R, int u, int) {
        bool plain;
        const double y = R.y(u, c, plain);
        const long i = R.i(u, c);
        p[i] = T(0);
        nu[i] = T(0);
        if constexpr (sizeof(T) == 8) {
          xopt[i] = R.xc(u);
        } else {  // the pass's correction and its best start at 0 (x itself)
          xopt[i] = T(0);
          xcor[i] = T(0);
        }
        v[6] += R.hh[u] * R.xc(u);
        if (plain) store_r(i, y, R.a(u));
      });
    } else if ((int)threadIdx.x < -mt.z) {
      const int e = mt.y + threadIdx.x, row = V.A.erow[e];
      store_r(row, explicit_row<double>(V.A, x, e, mt.w), b[row]);
    }
    v[2] = v[1];
    v[3] = v[1];
    v[5] = v[4];
    store_item_partials<NQB, 1u>(v, p5, V.n_items, item);
  }, filter);
}

// nu = A z1; partials of (rhat, nu) and sum h^2 z1
template <class T>
__global__ __launch_bounds__(WG) void kb_spmv_nu(BView V, const T *__restrict__ z1,
                                                 T *__restrict__ nu, const T *__restrict__ rhat,
                                                 const Scal *st, double *p2) {
  __shared__ __align__(16) unsigned char tile[row_tile_bytes<T>()];
  running_items<0, true>(V, st, [&](int item, int4 mt, const Heads &) {
    const long long mr = V.mean_row[mt.x];
    double v[2] = {0.0, 0.0};
    if (mt.z > 0) {
      row_blocks<T>(V, z1, rhat, mt.y, mt.z, tile,
                    [&](long base, const double (&y)[4], const T (&xr)[4], const T (&ar)[4], double h2,
                        unsigned em) {
                      double hs = 0.0;
                      T yt[4];
#pragma unroll
                      for (int k = 0; k < 4; k++) {
                        hs += (double)xr[k];
                        yt[k] = (T)y[k];
                      }
                      if (base == mr) yt[0] = T(0);  // the mean row (cell 0 of its block): kb_alpha
#pragma unroll
                      for (int k = 0; k < 4; k++)
                        v[0] += ((em >> k) & 1u) ? 0.0 : (double)ar[k] * (double)yt[k];
                      v[1] += h2 * hs;
                      st_plain(nu, base, yt, em);
                    });
    } else if ((int)threadIdx.x < -mt.z) {
      const int e = mt.y + threadIdx.x, row = V.A.erow[e];
      double y = explicit_row<T>(V.A, z1, e, mt.w);
      if (row == mr) y = 0.0;  // completed by kb_alpha
      const T yt = (T)y;
      nu[row] = yt;
      v[0] += (double)rhat[row] * (double)yt;
    }
    store_item_partials<2, 0u>(v, p2, V.n_items, item);
  });
}

// t = A z2; partials of (t, s), (t, t) and sum h^2 z2
template <class T>
__global__ __launch_bounds__(WG) void kb_spmv_t(BView V, const T *__restrict__ z2,
                                                T *__restrict__ t, const T *__restrict__ sv,
                                                const Scal *st, double *p4) {
  __shared__ __align__(16) unsigned char tile[row_tile_bytes<T>()];
  running_items<0, true>(V, st, [&](int item, int4 mt, const Heads &) {
    const long long mr = V.mean_row[mt.x];
    double v[3] = {0.0, 0.0, 0.0};
    if (mt.z > 0) {
      row_blocks<T>(V, z2, sv, mt.y, mt.z, tile,
                    [&](long base, const double (&y)[4], const T (&xr)[4], const T (&ar)[4], double h2,
                        unsigned em) {
                      double hs = 0.0;
                      T yt[4];
#pragma unroll
                      for (int k = 0; k < 4; k++) {
                        hs += (double)xr[k];
                        yt[k] = (T)y[k];
                      }
                      if (base == mr) yt[0] = T(0);  // the mean row: kb_omega
#pragma unroll
                      for (int k = 0; k < 4; k++) {
                        const double yd = ((em >> k) & 1u) ? 0.0 : (double)yt[k];
                        v[0] += yd * (double)ar[k];
                        v[1] += yd * yd;
                      }
                      v[2] += h2 * hs;
                      st_plain(t, base, yt, em);
                    });
    } else if ((int)threadIdx.x < -mt.z) {
      const int e = mt.y + threadIdx.x, row = V.A.erow[e];
      double y = explicit_row<T>(V.A, z2, e, mt.w);
      if (row == mr) y = 0.0;  // completed by kb_omega
      const T yt = (T)y;
      t[row] = yt;
      const double yd = (double)yt;
      v[0] += yd * (double)sv[row];
      v[1] += yd * yd;
    }
    store_item_partials<3, 0u>(v, p4, V.n_items, item);
  });
}

// y = A x over every instance (the operator alone, FP64); partials of
// sum h^2 x for the mean rows
__global__ __launch_bounds__(WG) void kb_spmv(BView V, const double *__restrict__ x,
                                              double *__restrict__ y, double *pm) {
  __shared__ __align__(16) unsigned char tile[row_tile_bytes<double>()];
  running_items<0, true>(V, nullptr, [&](int item, int4 mt, const Heads &) {
    double v[1] = {0.0};
    if (mt.z > 0) {
      row_blocks<double, false>(V, x, x, mt.y, mt.z, tile,
                         [&](long base, const double (&yr)[4], const double (&xr)[4], const double (&)[4],
                             double h2, unsigned em) {
                           double hs = 0.0;
#pragma unroll
                           for (int k = 0; k < 4; k++) hs += xr[k];
                           v[0] += h2 * hs;
                           st_plain(y, base, yr, em);
                         });
    } else if ((int)threadIdx.x < -mt.z) {
      const int e = mt.y + threadIdx.x;
      y[V.A.erow[e]] = explicit_row<double>(V.A, x, e, mt.w);
    }
    store_item_partials<1, 0u>(v, pm, V.n_items, item);
  });
}

// y[mean row of k] = sum h^2 x over instance k
__global__ __launch_bounds__(WG) void kb_meanrow(BView V, const double *pm, double *__restrict__ y) {
  const int k = blockIdx.x;
  const long long mr = V.mean_row[k];
  if (mr < 0) return;
  double o[1];
  range_totals<1, 0u>(pm, V.n_items, V.item_begin[k], V.item_begin[k + 1], o);
  if (threadIdx.x == 0) y[mr] = o[0];
}

// Two consecutive cells of a vector (16-byte loads for FP64, 8-byte for
// FP32), as doubles.
__device__ __forceinline__ void ld2(const double *__restrict__ a, long i, double (&v)[2]) {
  const double2 c = *reinterpret_cast<const double2 *>(a + i);
  v[0] = c.x;
  v[1] = c.y;
}
__device__ __forceinline__ void ld2(const float *__restrict__ a, long i, double (&v)[2]) {
  const float2 c = *reinterpret_cast<const float2 *>(a + i);
  v[0] = (double)c.x;
  v[1] = (double)c.y;
}
__device__ __forceinline__ void st2(double *__restrict__ a, long i, const double (&v)[2]) {
  *reinterpret_cast<double2 *>(a + i) = make_double2(v[0], v[1]);
}
__device__ __forceinline__ void st2(float *__restrict__ a, long i, const float (&v)[2]) {
  *reinterpret_cast<float2 *>(a + i) = make_float2(v[0], v[1]);
}
__device__ __forceinline__ void st2(float *__restrict__ a, long i, const double (&v)[2]) {
  *reinterpret_cast<float2 *>(a + i) = make_float2((float)v[0], (float)v[1]);
}

// x_opt = x when the previous iterate was the best so far; x += alpha z1 +
// omega z2; r = s - omega t; update partials.  FP64: x is the iterate.
// Mixed: x is the pass's FP32 correction e (the iterate is the FP64 x of the
// last residual replacement plus e; kb_fold adds it in), x_opt the best
// correction -- 12 B per cell of the 24 an FP64 x and x_opt would move.  Four
// pairs of consecutive cells per thread per pass (vector loads), all loads
// first: a 32-block item (2048 cells) is one pass of a workgroup; a thread
// takes its cells in the same order for any number of pairs per pass, so the
// partial sums do not depend on it (four measured ~3 % faster than two in
// FP64, ~2 % in FP32).
template <class T>
__global__ __launch_bounds__(WG) void kb_xrupdate(BView V, T *__restrict__ x,
                                                  T *__restrict__ xopt,
                                                  const T *__restrict__ z1,
                                                  const T *__restrict__ z2, T *__restrict__ r,
                                                  const T *__restrict__ t,
                                                  const T *__restrict__ rhat, const Scal *st,
                                                  double *p5) {
  constexpr int NP = 4;  // pairs per thread per pass
  running_items(V, st, [&](int item, int4 mt, const Heads &) {
    const double alpha = st[mt.x].alpha, omega = st[mt.x].omega;
    const int improved = st[mt.x].improved;
    double v[NQB] = {0, 0, 0, 0, 0, 0, 0};
    const long i0 = (long)mt.y * 64, ncell = mt.z > 0 ? (long)mt.z * 64 : 0;
    for (long off = 2 * (long)threadIdx.x; off < ncell; off += 2 * NP * WG) {
      long ii[NP];
      bool ok[NP];
      double xo[NP][2], a1[NP][2], a2[NP][2], rr[NP][2], tt[NP][2], rh[NP][2];
#pragma unroll
      for (int q = 0; q < NP; q++) {
        ok[q] = off + 2 * q * WG < ncell;
        ii[q] = i0 + (ok[q] ? off + 2 * q * WG : off);
        ld2(x, ii[q], xo[q]);
        ld2(z1, ii[q], a1[q]);
        ld2(z2, ii[q], a2[q]);
        ld2(r, ii[q], rr[q]);
        ld2(t, ii[q], tt[q]);
        ld2(rhat, ii[q], rh[q]);
      }
#pragma unroll
      for (int q = 0; q < NP; q++) {
        if (!ok[q]) continue;
        if (improved) st2(xopt, ii[q], xo[q]);
        double xn[2];
        T rt[2];
#pragma unroll
        for (int k = 0; k < 2; k++) {
          xn[k] = xo[q][k] + alpha * a1[q][k] + omega * a2[q][k];
          rt[k] = (T)(rr[q][k] - omega * tt[q][k]);
          const double rd = (double)rt[k];
          v[0] = nmax(v[0], fabs(rd));
          v[1] += rd * rd;
          v[2] += rh[q][k] * rd;
          v[3] += rh[q][k] * rh[q][k];
          v[4] += rd;
          v[5] += rh[q][k];
        }
        st2(x, ii[q], xn);
        st2(r, ii[q], rt);
      }
    }
    store_item_partials<NQB, 1u>(v, p5, V.n_items, item);
  });
}

// cell range of a block item (empty for an explicit item)
__device__ __forceinline__ void item_cells(const int4 &mt, long &i0, long &i1) {
  i0 = (long)mt.y * 64;
  i1 = i0 + (mt.z > 0 ? (long)mt.z * 64 : 0);
}

// x = x_opt on the instances (passing `filter`) whose last iterate is not
// their best
__global__ __launch_bounds__(WG) void kb_select(BView V, double *__restrict__ x,
                                                const double *__restrict__ xopt, const Scal *st,
                                                int filter) {
  for (int item = blockIdx.x; item < V.n_items; item += gridDim.x) {
    const int4 mt = V.item[item];
    if (st[mt.x].improved || !wanted(st[mt.x], filter)) continue;
    long i0, i1;
    item_cells(mt, i0, i1);
    for (long i = i0 + threadIdx.x; i < i1; i += WG) x[i] = xopt[i];
  }
}

// Mixed: x += the pass's FP32 correction on the instances (passing
// `filter`): the last one when the last iterate is the best, else the best.
__global__ __launch_bounds__(WG) void kb_fold(BView V, double *__restrict__ x,
                                              const float *__restrict__ e,
                                              const float *__restrict__ eopt, const Scal *st,
                                              int filter) {
  for (int item = blockIdx.x; item < V.n_items; item += gridDim.x) {
    const int4 mt = V.item[item];
    if (!wanted(st[mt.x], filter)) continue;
    const float *__restrict__ c = st[mt.x].improved ? e : eopt;
    long i0, i1;
    item_cells(mt, i0, i1);
    for (long i = i0 + threadIdx.x; i < i1; i += WG) x[i] += (double)c[i];
  }
}

// partials of sum h^2 x and sum h^2 (mean removal, ExpAMRSolver.cpp:395-411)
__global__ __launch_bounds__(WG) void kb_fmean(BView V, const double *__restrict__ x, double *pf) {
  for (int item = blockIdx.x; item < V.n_items; item += gridDim.x) {
    long i0, i1;
    item_cells(V.item[item], i0, i1);
    double v[2] = {0.0, 0.0};
    for (long i = i0 + threadIdx.x; i < i1; i += WG) {
      const double hh = V.A.h2[i >> 6];
      v[0] += x[i] * hh;
      v[1] += hh;
    }
    store_item_partials<2, 0u>(v, pf, V.n_items, item);
  }
}

// y = x over n doubles: the streaming ceiling of the vector kernels' access
// pattern (8-byte loads, four per thread in flight), which calibrates the
// bandwidth and traffic figures of the solver kernels.
__global__ __launch_bounds__(WG) void kb_copy(const double *__restrict__ x, double *__restrict__ y,
                                              long n) {
  for (long base = (long)blockIdx.x * 4 * WG; base < n; base += (long)gridDim.x * 4 * WG) {
    double v[4];
#pragma unroll
    for (int q = 0; q < 4; q++) {
      const long i = base + q * WG + threadIdx.x;
      v[q] = i < n ? x[i] : 0.0;
    }
#pragma unroll
    for (int q = 0; q < 4; q++) {
      const long i = base + q * WG + threadIdx.x;
      if (i < n) y[i] = v[q];
    }
  }
}

// out = x - mean_k
__global__ __launch_bounds__(WG) void kb_fwrite(BView V, const double *__restrict__ x,
                                                double *__restrict__ out, const Scal *st) {
  for (int item = blockIdx.x; item < V.n_items; item += gridDim.x) {
    const int4 mt = V.item[item];
    const double mean = st[mt.x].mean;
    long i0, i1;
    item_cells(mt, i0, i1);
    for (long i = i0 + threadIdx.x; i < i1; i += WG) out[i] = x[i] - mean;
  }
}

// The pressure step around the solve, per instance (PressureSingle.cpp:842-879):
// b = rhs - L p_old (y = L p_old from kb_spmv without the mean-row
// replacement; the solve ignores b at a mean row), and after the solve
// p = dp + (p_old - <dp>_h2).
__global__ __launch_bounds__(WG) void kb_rhs_sub(long n, const double *__restrict__ rhs,
                                                 double *__restrict__ b) {
  for (long i = (long)blockIdx.x * WG + threadIdx.x; i < n; i += (long)gridDim.x * WG)
    b[i] = rhs[i] - b[i];
}
__global__ __launch_bounds__(WG) void kb_padd(BView V, const double *__restrict__ dp,
                                              double *__restrict__ p, const Scal *st) {
  for (int item = blockIdx.x; item < V.n_items; item += gridDim.x) {
    const int4 mt = V.item[item];
    const double avg = st[mt.x].mean;
    long i0, i1;
    item_cells(mt, i0, i1);
    for (long i = i0 + threadIdx.x; i < i1; i += WG) p[i] = dp[i] + (p[i] - avg);
  }
}

// ----------------------------------------------------------------------------
// tile kernels (preconditioner; a wavefront per 16-block tile)
// ----------------------------------------------------------------------------

// Calls body(meta) for the wave's tiles (grid-stride over waves) whose
// instance is iterating; the next tile's meta is loaded while the current
// one is processed.
template <class F>
__device__ __forceinline__ void running_tiles(const BView &V, const Scal *st, F &&body) {
  const int stride = gridDim.x * WAVES;
  int tile = blockIdx.x * WAVES + wave_id();
  if (tile >= V.n_tiles) return;
  int4 tm = V.tile[tile];
  for (;;) {
    const int next = tile + stride;
    const bool more = next < V.n_tiles;
    const int4 ntm = V.tile[more ? next : tile];
    if (!st[tm.x].idle) body(tm);
    if (!more) break;
    tile = next;
    tm = ntm;
  }
}

// The thread's four cells of tile tm (see tile_prec64): first index and
// whether its block exists (a lane past the tile's end addresses the
// tile's first block).
__device__ __forceinline__ long btile_cells(int4 tm, int w, int lane, bool &ok) {
  ok = (lane & 15) < tm.z;
  return (long)(tm.y + (ok ? (lane & 15) : 0)) * 64 + cell4(w, lane);
}

// FP64: p = r + beta (p - omega nu), or on restart p = rhat = r; z1 = P p
// on MFMA (a workgroup per 16-block tile, tiles grid-stride; the next tile's
// loads are in flight during the current tile's product)
__global__ __launch_bounds__(WG) void kb_pupdate64(BView V, const double *__restrict__ pf,
                                                   const double *__restrict__ r,
                                                   double *__restrict__ rhat,
                                                   double *__restrict__ p,
                                                   const double *__restrict__ nu,
                                                   double *__restrict__ z1, const Scal *st) {
  __shared__ double tile[TILE][TROW];
  const int lane = lane_id(), w = wave_id();
  int t = blockIdx.x;
  if (t >= V.n_tiles) return;
  PFrag64 F;
  load_pfrag64(F, pf, w, lane);
  int4 tm = V.tile[t];
  bool run = !st[tm.x].idle;  // uniform over the workgroup
  bool ok = false;
  long i0 = 0;
  double rv[4], pv[4], nv[4];
  auto load = [&](int4 m) {
    i0 = btile_cells(m, w, lane, ok);
    ld4(r, i0, rv);
    ld4(p, i0, pv);
    ld4(nu, i0, nv);
  };
  if (run) load(tm);
  for (;;) {
    const int nt = t + (int)gridDim.x;
    const bool more = nt < V.n_tiles;
    const int4 ntm = V.tile[more ? nt : t];
    const bool nrun = more && !st[ntm.x].idle;
    if (run) {
      const int k = tm.x;
      const double beta = st[k].beta, omega = st[k].omega;
      const int restart = st[k].restart;
      double pi[4];
#pragma unroll
      for (int q = 0; q < 4; q++) pi[q] = restart ? rv[q] : rv[q] + beta * (pv[q] - omega * nv[q]);
      if (ok) {
        if (restart) st4(rhat, i0, rv);
        st4(p, i0, pi);
      }
      tile_put(tile, w, lane, ok, pi);
      const bool oo = ok;
      const long io = i0;
      if (nrun) load(ntm);  // in flight during the product
      __syncthreads();
      const f64x4 z = tile_prec64(tile, F, lane);
      __syncthreads();  // the tile is rewritten by the next tile
      if (oo) st4(z1, io, z);
    } else if (nrun) {
      load(ntm);
    }
    if (!more) break;
    t = nt;
    tm = ntm;
    run = nrun;
  }
}

// FP64: s = r - alpha nu (stored in r); z2 = P s on MFMA (as kb_pupdate64)
__global__ __launch_bounds__(WG) void kb_supdate64(BView V, const double *__restrict__ pf,
                                                   double *__restrict__ r,
                                                   const double *__restrict__ nu,
                                                   double *__restrict__ z2, const Scal *st) {
  __shared__ double tile[TILE][TROW];
  const int lane = lane_id(), w = wave_id();
  int t = blockIdx.x;
  if (t >= V.n_tiles) return;
  PFrag64 F;
  load_pfrag64(F, pf, w, lane);
  int4 tm = V.tile[t];
  bool run = !st[tm.x].idle;
  bool ok = false;
  long i0 = 0;
  double rv[4], nv[4];
  auto load = [&](int4 m) {
    i0 = btile_cells(m, w, lane, ok);
    ld4(r, i0, rv);
    ld4(nu, i0, nv);
  };
  if (run) load(tm);
  for (;;) {
    const int nt = t + (int)gridDim.x;
    const bool more = nt < V.n_tiles;
    const int4 ntm = V.tile[more ? nt : t];
    const bool nrun = more && !st[ntm.x].idle;
    if (run) {
      const double alpha = st[tm.x].alpha;
      double si[4];
#pragma unroll
      for (int q = 0; q < 4; q++) si[q] = rv[q] - alpha * nv[q];
      if (ok) st4(r, i0, si);
      tile_put(tile, w, lane, ok, si);
      const bool oo = ok;
      const long io = i0;
      if (nrun) load(ntm);
      __syncthreads();
      const f64x4 z = tile_prec64(tile, F, lane);
      __syncthreads();
      if (oo) st4(z2, io, z);
    } else if (nrun) {
      load(ntm);
    }
    if (!more) break;
    t = nt;
    tm = ntm;
    run = nrun;
  }
}

// FP32 storage, BF16 MFMA: r -= mu (projection); p update or restart;
// z1 = P p
__global__ __launch_bounds__(WG) void kb_pupdate32(BView V, const bf16x8 *__restrict__ pf,
                                                   float *__restrict__ r, float *__restrict__ rhat,
                                                   float *__restrict__ p,
                                                   const float *__restrict__ nu,
                                                   float *__restrict__ z1, const Scal *st) {
  const int lane = lane_id(), q = lane >> 4, j = lane & 15;
  PFrag P;
  load_pfrag(P, pf, lane);
  running_tiles(V, st, [&](int4 tm) {
    const int k = tm.x, b0 = tm.y, tnb = tm.z;
    const double beta = st[k].beta, omega = st[k].omega, mu = st[k].mu;
    const int restart = st[k].restart;
    const bool ok = j < tnb;
    const long base = (long)(b0 + (ok ? j : 0)) * 64;
    float pin[16];
    MfIn in;
    mf_load(r, p, nu, base, q, in);
    mf_pupdate(in, r, rhat, p, base, q, ok, beta, omega, mu, restart, pin);
    float z[16];
    prec_mfma(P, pin, z);
    if (ok) st16_out(z1, base, q, z);
  });
}

// FP32 storage, BF16 MFMA: s = r - alpha nu (stored in r); z2 = P s
__global__ __launch_bounds__(WG) void kb_supdate32(BView V, const bf16x8 *__restrict__ pf,
                                                   float *__restrict__ r,
                                                   const float *__restrict__ nu,
                                                   float *__restrict__ z2, const Scal *st) {
  const int lane = lane_id(), q = lane >> 4, j = lane & 15;
  PFrag P;
  load_pfrag(P, pf, lane);
  running_tiles(V, st, [&](int4 tm) {
    const int k = tm.x, b0 = tm.y, tnb = tm.z;
    const double alpha = st[k].alpha;
    const bool ok = j < tnb;
    const long base = (long)(b0 + (ok ? j : 0)) * 64;
    float rv[16], nv[16], sv[16];
    ld16(r, base, q, rv);
    ld16(nu, base, q, nv);
#pragma unroll
    for (int kk = 0; kk < 16; kk++) sv[kk] = ok ? (float)((double)rv[kk] - alpha * (double)nv[kk]) : 0.f;
    if (ok) st16_in(r, base, q, sv);
    float z[16];
    prec_mfma(P, sv, z);
    if (ok) st16_out(z2, base, q, z);
  });
}

// ----------------------------------------------------------------------------
// per-instance scalar kernels (a workgroup per instance)
// ----------------------------------------------------------------------------

// alpha_k = rho_k / ((rhat, nu)_k + eps) (BiCGSTAB.cu:523-531); completes
// the mean row of nu
template <class T>
__global__ __launch_bounds__(WG) void kb_alpha(BView V, Scal *st, const double *p2,
                                               T *__restrict__ nu, const T *__restrict__ rhat) {
  const int k = blockIdx.x;
  // the state and the mean row's values are read before the reduction
  const int idle = st[k].idle;
  const double rho = st[k].rho, eps = st[k].eps;
  const long long mr = V.mean_row[k];
  const double rh_mr = mr >= 0 ? (double)rhat[mr] : 0.0;
  double o[2];
  range_totals<2, 0u>(p2, V.n_items, V.item_begin[k], V.item_begin[k + 1], o);
  if (idle) return;
  if (threadIdx.x == 0) {
    double rn = o[0];
    if (mr >= 0) {
      const T m = (T)o[1];
      nu[mr] = m;
      rn += rh_mr * (double)m;
    }
    st[k].alpha = rho / (rn + eps);
    st[k].red[2] = rn;
  }
}

// omega_k = (t, s)_k / ((t, t)_k + eps) (BiCGSTAB.cu:548-556); completes the
// mean row of t
template <class T>
__global__ __launch_bounds__(WG) void kb_omega(BView V, Scal *st, const double *p4,
                                               T *__restrict__ t, const T *__restrict__ sv) {
  const int k = blockIdx.x;
  // the state and the mean row's values are read before the reduction
  const int idle = st[k].idle;
  const double eps = st[k].eps;
  const long long mr = V.mean_row[k];
  const double s_mr = mr >= 0 ? (double)sv[mr] : 0.0;
  double o[3];
  range_totals<3, 0u>(p4, V.n_items, V.item_begin[k], V.item_begin[k + 1], o);
  if (idle) return;
  if (threadIdx.x == 0) {
    double ts = o[0], tt = o[1];
    if (mr >= 0) {
      const T m = (T)o[2];
      t[mr] = m;
      const double md = (double)m;
      ts += md * s_mr;
      tt += md * md;
    }
    st[k].omega = ts / (tt + eps);
    st[k].red[3] = ts;
    st[k].red[4] = tt;
  }
}

// Instance states from the update partials.  mode 0: end of an iteration
// (iterating instances); mode 1: after kb_init at the start of a solve
// (unfinished instances: their initial state); mode 2: after kb_init of a
// residual replacement (instances with one pending: restart from it).  In
// modes 1 and 2 the mean row of r is completed here: b = 0 on that row
// (ExpAMRSolver.cpp:340-345) and (A x)_row = sum h^2 x.  `init_check`
// (mode 1): an initial residual already within tolerance ends the instance.
// A non-finite initial residual ends it unconverged.
template <class T>
__global__ __launch_bounds__(WG) void kb_eval(BView V, Scal *st, const double *p5,
                                              T *__restrict__ r, T *__restrict__ rhat, int mode,
                                              int init_check) {
  const int k = blockIdx.x;
  // the state is staged in LDS (a word per thread) while the partials are
  // reduced; the reduction's barrier publishes it
  __shared__ Scal prev;
  static_assert(sizeof(Scal) % 8 == 0 && sizeof(Scal) / 8 <= WG, "Scal staging");
  if (threadIdx.x < sizeof(Scal) / 8)
    reinterpret_cast<unsigned long long *>(&prev)[threadIdx.x] =
        reinterpret_cast<const unsigned long long *>(&st[k])[threadIdx.x];
  const long long mr = V.mean_row[k];
  double o[NQB];
  range_totals<NQB, 1u>(p5, V.n_items, V.item_begin[k], V.item_begin[k + 1], o);
  if (!wanted(prev, mode == 0 ? F_RUNNING : (mode == 1 ? F_LIVE : F_REPLACE))) return;
  if (threadIdx.x == 0) {
    double oo[NQ5];
    for (int q = 0; q < NQ5; q++) oo[q] = o[q];
    if (mode != 0 && mr >= 0) {
      const T rm = (T)(0.0 - o[6]);
      r[mr] = rm;
      rhat[mr] = rm;
      const double rd = (double)rm;
      oo[0] = nmax(oo[0], fabs(rd));
      oo[1] += rd * rd;
      oo[2] += rd * rd;
      oo[3] += rd * rd;
      oo[4] += rd;
      oo[5] += rd;
    }
    Scal s = evaluate(prev, oo, mode == 2);
    if (mode != 0 && (s.err != s.err || s.err > 1.79e308)) {
      s.done = s.idle = 1;
      s.converged = 0;
    } else if (mode == 1 && (s.err == 0.0 || (init_check && (s.err <= s.tol_abs || s.tol_rel >= 1.0)))) {
      s.converged = s.done = s.idle = 1;
    }
    st[k] = s;
  }
}

// all_done = no instance is unfinished
__global__ __launch_bounds__(WG) void kb_alldone(int K, const Scal *st, int *all_done) {
  int busy = 0;
  for (int k = threadIdx.x; k < K; k += WG) busy |= st[k].done ? 0 : 1;
  busy = __syncthreads_or(busy);
  if (threadIdx.x == 0) *all_done = busy ? 0 : 1;
}

// mean_k = sum h^2 x / sum h^2
__global__ __launch_bounds__(WG) void kb_mean(BView V, Scal *st, const double *pf) {
  const int k = blockIdx.x;
  double o[2];
  range_totals<2, 0u>(pf, V.n_items, V.item_begin[k], V.item_begin[k + 1], o);
  if (threadIdx.x == 0) st[k].mean = o[0] / o[1];
}

// partials of |b - y|^2, |b - y|_inf, |b|^2, |b|_inf per item (b's mean
// row taken as 0); y holds A x.  Explicit items hold no cells of their own.
__global__ __launch_bounds__(WG) void kb_rpart(BView V, const double *__restrict__ b,
                                               const double *__restrict__ y, double *part) {
  for (int item = blockIdx.x; item < V.n_items; item += gridDim.x) {
    const int4 mt = V.item[item];
    const long long mr = V.mean_row[mt.x];
    long i0, i1;
    item_cells(mt, i0, i1);
    double v[4] = {0.0, 0.0, 0.0, 0.0};
    for (long i = i0 + threadIdx.x; i < i1; i += WG) {
      const double bi = (i == mr) ? 0.0 : b[i];
      const double ri = bi - y[i];
      v[0] += ri * ri;
      v[1] = nmax(v[1], fabs(ri));
      v[2] += bi * bi;
      v[3] = nmax(v[3], fabs(bi));
    }
    store_item_partials<4, 0xAu>(v, part, V.n_items, item);
  }
}
// out[4k .. 4k+3] = the totals of instance k
__global__ __launch_bounds__(WG) void kb_rtot(BView V, const double *part, double *out) {
  const int k = blockIdx.x;
  double o[4];
  range_totals<4, 0xAu>(part, V.n_items, V.item_begin[k], V.item_begin[k + 1], o);
  if (threadIdx.x < 4) out[4 * k + threadIdx.x] = o[threadIdx.x];
}

thread_local std::string g_berr;  // failures without a handle, per calling thread

}  // namespace

// ----------------------------------------------------------------------------
// host side
// ----------------------------------------------------------------------------

struct BBuf {
  void *p;
  size_t bytes;
};

struct cup2d_psb {
  cup2d_ps_config cfg{};
  std::string err;
  bool mixed = false;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  bool ready = false;
  int K = 0;
  long long nb = 0, n = 0, nnz = 0;
  int ib = 0;                       // blocks per item
  int xg_log2 = 3;                  // XCD run length of the operator kernels' items: 2^xg_log2 (0: plain)
  std::vector<long long> boff;      // [K+1] block offset of each instance
  std::vector<long long> eoff;      // [K+1] explicit-row offset of each instance
  std::vector<long long> h_mean_row;
  std::vector<int> item_inst, item_b0, item_nb, item_w, item_begin;
  std::vector<int> tile_inst, tile_b0, tile_nb;
  // device: operator
  int4 *d_nbr = nullptr;
  int2 *d_fid = nullptr;            // [4 nb] coarse-fine neighbours of each face
  unsigned long long *d_mask = nullptr;
  int *d_erow = nullptr, *d_ecol = nullptr;
  uint16_t *d_ecode = nullptr;
  double *d_etab = nullptr, *d_h2 = nullptr, *d_P = nullptr;
  std::vector<uint8_t> elen;        // [n_exc] entries of each explicit row
  int ell_w = 0;
  bf16x8 *d_pfrag = nullptr;
  long long *d_mean_row = nullptr;
  double *d_h = nullptr;            // [nb] cell size per block
  // device: work units
  int4 *d_item = nullptr;
  int *d_item_begin = nullptr;
  int4 *d_tile = nullptr;
  double *d_p2 = nullptr, *d_p4 = nullptr, *d_p5 = nullptr, *d_pm = nullptr;
  // device: vectors and state
  double *d_x = nullptr, *d_xopt = nullptr, *d_b = nullptr, *d_xio = nullptr;
  void *v_r = nullptr, *v_rhat = nullptr, *v_p = nullptr, *v_nu = nullptr, *v_z1 = nullptr,
       *v_z2 = nullptr, *v_t = nullptr;  // Krylov vectors, FP64 or FP32
  Scal *d_st = nullptr;
  int *d_all_done = nullptr;
  int *h_all_done = nullptr;        // pinned
  std::vector<Scal> h_st;
  hipGraphExec_t gexec = nullptr;
  int gchunk = 0;
  double *d_rt = nullptr;           // [4K] residual norms of the returned solutions
  std::vector<double> h_rt;
  // per-phase timers (cfg.profile): an event after every launch, tagged
  // with the launch's phase
  std::vector<hipEvent_t> pev;
  std::vector<int> pph;
  size_t pused = 0;
  double pms[5] = {0, 0, 0, 0, 0};
  std::vector<BBuf> allocs;       // mesh-dependent device memory
  std::vector<BBuf> unit_allocs;  // work-unit lists and partials (depend on ib)
  std::vector<BBuf> spare;        // released buffers, reused by size (mesh changes, retuning)
};

namespace {

int bfail(cup2d_psb *h, int code, const std::string &msg) {
  if (h) h->err = msg;
  else g_berr = msg;
  return code;
}

// Runs an entry point's body so that no C++ exception crosses the C ABI: a
// failed host allocation becomes CUP2D_PS_ENOMEM, any other exception EINVAL,
// with its text as the last error.
template <class F>
int bguarded(cup2d_psb *h, F &&body) {
  try {
    return body();
  } catch (const std::bad_alloc &) {
    return bfail(h, CUP2D_PS_ENOMEM, "host allocation failed");
  } catch (const std::exception &e) {
    return bfail(h, CUP2D_PS_EINVAL, e.what());
  }
}

#define BCHK(h, expr)                                                                 \
  do {                                                                                \
    hipError_t _e = (expr);                                                           \
    if (_e != hipSuccess)                                                             \
      return bfail((h), _e == hipErrorOutOfMemory ? CUP2D_PS_ENOMEM : CUP2D_PS_EHIP,  \
                   std::string(#expr) + ": " + hipGetErrorString(_e));                \
  } while (0)

template <class T>
int balloc(cup2d_psb *h, std::vector<BBuf> &pool, T **p, size_t count) {
  *p = nullptr;
  const size_t bytes = std::max<size_t>(count, 1) * sizeof(T);
  // the smallest spare buffer that holds it and is not more than twice as large
  size_t best = h->spare.size();
  for (size_t i = 0; i < h->spare.size(); i++) {
    const size_t b = h->spare[i].bytes;
    if (b >= bytes && b <= 2 * bytes + 4096 && (best == h->spare.size() || b < h->spare[best].bytes))
      best = i;
  }
  BBuf buf{nullptr, bytes};
  if (best < h->spare.size()) {
    buf = h->spare[best];
    h->spare.erase(h->spare.begin() + best);
  } else {
    BCHK(h, hipMalloc(&buf.p, bytes));
  }
  pool.push_back(buf);
  *p = static_cast<T *>(buf.p);
  return CUP2D_PS_OK;
}

template <class T>
int bupload(cup2d_psb *h, std::vector<BBuf> &pool, T **p, const std::vector<T> &v) {
  int rc = balloc(h, pool, p, v.size());
  if (rc) return rc;
  if (!v.empty()) BCHK(h, hipMemcpy(*p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
  return CUP2D_PS_OK;
}

void bfree_pool(std::vector<BBuf> &pool) {
  for (const BBuf &b : pool) (void)hipFree(b.p);
  pool.clear();
}

// Buffers of a pool back to the spare list (callers have synchronised the
// stream, so they are handed out again in stream order).
void brecycle(cup2d_psb *h, std::vector<BBuf> &pool) {
  h->spare.insert(h->spare.end(), pool.begin(), pool.end());
  pool.clear();
}

// Frees spare buffers, largest first, until they hold no more than the live
// ones.
void btrim(cup2d_psb *h) {
  size_t live = 0, kept = 0;
  for (const BBuf &b : h->allocs) live += b.bytes;
  for (const BBuf &b : h->unit_allocs) live += b.bytes;
  for (const BBuf &b : h->spare) kept += b.bytes;
  std::sort(h->spare.begin(), h->spare.end(),
            [](const BBuf &a, const BBuf &b) { return a.bytes > b.bytes; });
  size_t i = 0;
  for (; i < h->spare.size() && kept > live; i++) {
    (void)hipFree(h->spare[i].p);
    kept -= h->spare[i].bytes;
  }
  h->spare.erase(h->spare.begin(), h->spare.begin() + i);
}

void drop_graph(cup2d_psb *h) {
  if (h->gexec) {
    (void)hipGraphExecDestroy(h->gexec);
    h->gexec = nullptr;
  }
}

// Profiling (cfg.profile), as in the single-solve handle: an event after
// the launch just made, tagged with its phase (-1: a start mark).
void bptick(cup2d_psb *h, int ph) {
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
void bpcollect(cup2d_psb *h) {
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
bool bgraphs_on(const cup2d_psb *h) { return h->cfg.use_graph && !h->cfg.profile; }

BView bview(const cup2d_psb *h) {
  BView V;
  V.A.nb = (int)h->nb;
  V.A.mean_row = -1;
  V.A.nbr = h->d_nbr;
  V.A.mask = h->d_mask;
  V.A.n_exc = (int)h->eoff[h->K];
  V.A.ell_w = h->ell_w;
  V.A.erow = h->d_erow;
  V.A.ecol = h->d_ecol;
  V.A.ecode = h->d_ecode;
  V.A.etab = h->d_etab;
  V.A.ewid = nullptr;                 // explicit items carry their width
  V.A.h2 = h->d_h2;
  V.item = h->d_item;
  V.item_begin = h->d_item_begin;
  V.tile = h->d_tile;
  V.mean_row = h->d_mean_row;
  V.K = h->K;
  V.n_items = (int)h->item_inst.size();
  V.n_tiles = (int)h->tile_inst.size();
  V.xg_log2 = h->xg_log2;
  return V;
}

// item kernels: a multiple of 8 workgroups from 8 items on (the operator
// kernels' XCD order needs it; the grid-stride loops cover every item)
int grid_items(const cup2d_psb *h) {
  const size_t g = std::min<size_t>(BGRID, std::max<size_t>(1, h->item_inst.size()));
  return (int)(g >= 8 ? g & ~(size_t)7 : g);
}
// FP64 preconditioner kernels: a workgroup per tile
int grid_tiles64(const cup2d_psb *h) {
  return (int)std::min<size_t>(BGRID, std::max<size_t>(1, h->tile_inst.size()));
}
int grid_tiles(const cup2d_psb *h) {
  return (int)std::min<size_t>(BGRID, std::max<size_t>(1, (h->tile_inst.size() + WAVES - 1) / WAVES));
}

// The iterate's buffers of the iteration kernels: FP64 x and x_opt, or, in
// mixed precision, the pass's FP32 correction and its best, carved out of
// the FP64 x_opt buffer (2n floats; x_opt is not needed there: the best
// iterate is the last replacement's x plus the best correction).
template <class T>
T *iter_x(const cup2d_psb *h) {
  if constexpr (sizeof(T) == 8) return h->d_x;
  else return reinterpret_cast<float *>(h->d_xopt);
}
template <class T>
T *iter_xopt(const cup2d_psb *h) {
  if constexpr (sizeof(T) == 8) return h->d_xopt;
  else return reinterpret_cast<float *>(h->d_xopt) + h->n;
}

// Item and tile lists for `ib` blocks per item (tiles are TILE blocks; an
// instance's last item and tile may be shorter); each instance's explicit
// rows follow its block items as items of up to WG rows.
int build_units(cup2d_psb *h, int ib) {
  drop_graph(h);
  (void)hipStreamSynchronize(h->stream);
  brecycle(h, h->unit_allocs);
  h->ib = ib;
  h->item_inst.clear();
  h->item_b0.clear();
  h->item_nb.clear();
  h->item_w.clear();
  h->item_begin.assign(1, 0);
  h->tile_inst.clear();
  h->tile_b0.clear();
  h->tile_nb.clear();
  for (int k = 0; k < h->K; k++) {
    const long long b0 = h->boff[k], b1 = h->boff[k + 1];
    for (long long b = b0; b < b1; b += ib) {
      h->item_inst.push_back(k);
      h->item_b0.push_back((int)b);
      h->item_nb.push_back((int)std::min<long long>(ib, b1 - b));
      h->item_w.push_back(0);
    }
    for (long long e = h->eoff[k]; e < h->eoff[k + 1]; e += WG) {
      const long long e1 = std::min<long long>(e + WG, h->eoff[k + 1]);
      h->item_inst.push_back(k);
      h->item_b0.push_back((int)e);
      h->item_nb.push_back(-(int)(e1 - e));
      int w = 0;
      for (long long j = e; j < e1; j++) w = std::max(w, (int)h->elen[j]);
      h->item_w.push_back(w);
    }
    h->item_begin.push_back((int)h->item_inst.size());
    for (long long b = b0; b < b1; b += TILE) {
      h->tile_inst.push_back(k);
      h->tile_b0.push_back((int)b);
      h->tile_nb.push_back((int)std::min<long long>(TILE, b1 - b));
    }
  }
  const size_t NI = h->item_inst.size();
  auto &pool = h->unit_allocs;
  int rc;
  std::vector<int4> tile4(h->tile_inst.size());
  for (size_t i = 0; i < tile4.size(); i++)
    tile4[i] = make_int4(h->tile_inst[i], h->tile_b0[i], h->tile_nb[i], 0);
  std::vector<int4> item4(NI);
  for (size_t i = 0; i < NI; i++)
    item4[i] = make_int4(h->item_inst[i], h->item_b0[i], h->item_nb[i], h->item_w[i]);
  if ((rc = bupload(h, pool, &h->d_item, item4)) ||
      (rc = bupload(h, pool, &h->d_item_begin, h->item_begin)) ||
      (rc = bupload(h, pool, &h->d_tile, tile4)) ||
      (rc = balloc(h, pool, &h->d_p2, 2 * NI)) || (rc = balloc(h, pool, &h->d_p4, 3 * NI)) ||
      (rc = balloc(h, pool, &h->d_p5, (size_t)NQB * NI)) || (rc = balloc(h, pool, &h->d_pm, 2 * NI)))
    return rc;
  return CUP2D_PS_OK;
}

// One BiCGSTAB iteration of every iterating instance: 8 kernels.
template <class T>
void launch_iteration_b(cup2d_psb *h) {
  const BView V = bview(h);
  const int Gi = grid_items(h), Gt = grid_tiles(h), K = h->K;
  hipStream_t s = h->stream;
  T *r = (T *)h->v_r, *rhat = (T *)h->v_rhat, *p = (T *)h->v_p, *nu = (T *)h->v_nu,
    *z1 = (T *)h->v_z1, *z2 = (T *)h->v_z2, *t = (T *)h->v_t;
  Scal *st = h->d_st;
  if constexpr (sizeof(T) == 8)
    kb_pupdate64<<<grid_tiles64(h), WG, 0, s>>>(V, h->d_P, r, rhat, p, nu, z1, st);
  else
    kb_pupdate32<<<Gt, WG, 0, s>>>(V, h->d_pfrag, r, rhat, p, nu, z1, st);
  bptick(h, CUP2D_PS_PHASE_PREC);
  kb_spmv_nu<T><<<Gi, WG, 0, s>>>(V, z1, nu, rhat, st, h->d_p2);
  bptick(h, CUP2D_PS_PHASE_SPMV);
  kb_alpha<T><<<K, WG, 0, s>>>(V, st, h->d_p2, nu, rhat);
  bptick(h, CUP2D_PS_PHASE_SCALAR);
  if constexpr (sizeof(T) == 8)
    kb_supdate64<<<grid_tiles64(h), WG, 0, s>>>(V, h->d_P, r, nu, z2, st);
  else
    kb_supdate32<<<Gt, WG, 0, s>>>(V, h->d_pfrag, r, nu, z2, st);
  bptick(h, CUP2D_PS_PHASE_PREC);
  kb_spmv_t<T><<<Gi, WG, 0, s>>>(V, z2, t, r, st, h->d_p4);
  bptick(h, CUP2D_PS_PHASE_SPMV);
  kb_omega<T><<<K, WG, 0, s>>>(V, st, h->d_p4, t, r);
  bptick(h, CUP2D_PS_PHASE_SCALAR);
  kb_xrupdate<T><<<Gi, WG, 0, s>>>(V, iter_x<T>(h), iter_xopt<T>(h), z1, z2, r, t, rhat, st, h->d_p5);
  bptick(h, CUP2D_PS_PHASE_UPDATE);
  kb_eval<T><<<K, WG, 0, s>>>(V, st, h->d_p5, r, rhat, 0, 0);
  bptick(h, CUP2D_PS_PHASE_SCALAR);
}

int chunk_b(const cup2d_psb *h) { return h->cfg.chunk > 0 ? h->cfg.chunk : 8; }

// Residual replacement step (mixed precision): the instances with one
// pending add their best correction to x, recompute r = b - A x in FP64 from
// it and restart with a zero correction.
void launch_replacement_b(cup2d_psb *h) {
  const BView V = bview(h);
  const int Gi = grid_items(h);
  hipStream_t s = h->stream;
  float *r = (float *)h->v_r, *rhat = (float *)h->v_rhat, *p = (float *)h->v_p,
        *nu = (float *)h->v_nu;
  float *e = iter_x<float>(h), *eopt = iter_xopt<float>(h);
  kb_fold<<<Gi, WG, 0, s>>>(V, h->d_x, e, eopt, h->d_st, F_REPLACE);
  kb_init<float><<<Gi, WG, 0, s>>>(V, h->d_b, h->d_x, eopt, e, r, rhat, p, nu, h->d_st, h->d_p5,
                                   F_REPLACE);
  kb_eval<float><<<h->K, WG, 0, s>>>(V, h->d_st, h->d_p5, r, rhat, 2, 0);
  bptick(h, CUP2D_PS_PHASE_SETUP);
}

void launch_chunk_b(cup2d_psb *h, int C) {
  if (h->mixed) launch_replacement_b(h);
  for (int i = 0; i < C; i++) {
    if (h->mixed) launch_iteration_b<float>(h);
    else launch_iteration_b<double>(h);
  }
  kb_alldone<<<1, WG, 0, h->stream>>>(h->K, h->d_st, h->d_all_done);
  bptick(h, CUP2D_PS_PHASE_SCALAR);
}

int ensure_graph_b(cup2d_psb *h) {
  const int C = chunk_b(h);
  if (h->gexec && h->gchunk == C) return CUP2D_PS_OK;
  drop_graph(h);
  hipGraph_t g = nullptr;
  BCHK(h, hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
  launch_chunk_b(h, C);
  BCHK(h, hipStreamEndCapture(h->stream, &g));
  hipError_t e = hipGraphInstantiate(&h->gexec, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  if (e != hipSuccess)
    return bfail(h, CUP2D_PS_EHIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
  h->gchunk = C;
  return CUP2D_PS_OK;
}

int read_all_done(cup2d_psb *h, int &done) {
  BCHK(h, hipMemcpyAsync(h->h_all_done, h->d_all_done, sizeof(int), hipMemcpyDeviceToHost, h->stream));
  BCHK(h, hipStreamSynchronize(h->stream));
  done = *h->h_all_done;
  return CUP2D_PS_OK;
}

// Chunks of iterations until every instance is done; `max_iter` bounds the
// launches (each instance's own cap is in its state), `max_chunks` > 0 stops
// earlier (timing runs).
int run_chunks_b(cup2d_psb *h, int max_iter, int max_chunks) {
  const int C = chunk_b(h);
  if (bgraphs_on(h)) {
    int rc = ensure_graph_b(h);
    if (rc) return rc;
  }
  int launched = 0, chunks = 0;
  for (;;) {
    if (bgraphs_on(h)) {
      BCHK(h, hipGraphLaunch(h->gexec, h->stream));
    } else {
      launch_chunk_b(h, C);
      BCHK(h, hipGetLastError());
    }
    launched += C;
    chunks++;
    int done = 0;
    int rc = read_all_done(h, done);
    if (rc) return rc;
    bpcollect(h);
    // every instance stops at its own iteration cap; pauses for residual
    // replacement skip iterations, hence the margin
    if (done || launched > 2 * max_iter + 2 * C) break;
    if (max_chunks > 0 && chunks >= max_chunks) break;
  }
  return CUP2D_PS_OK;
}

Scal marker_b(const cup2d_psb *h, int k, double tol_abs, double tol_rel, int max_restarts,
              int max_iter) {
  Scal s{};
  s.n = 64.0 * double(h->boff[k + 1] - h->boff[k]);
  s.project = (h->mixed && h->h_mean_row[k] < 0) ? 1 : 0;
  if (h->mixed) {
    // FP32-stored residual: FP64 residual replacement and restart from the
    // best iterate after each reduction by inner_tol (default 1e-2) of the
    // recurrence residual, or a stall of STALL_BASE iterations (doubled per
    // replacement that gained less than 2x)
    s.rep_factor = h->cfg.inner_tol > 0 ? h->cfg.inner_tol : 1e-2;
    s.stall = STALL_BASE;
  }
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

// One pass over the states in h_st (marker states start, done states are
// left alone): r0 = b - A x in FP64 for the starting instances, iterations
// until every instance is done, x = x_opt where the last iterate is not the
// best.  Returns with h_st holding the final states.
template <class T>
int run_pass(cup2d_psb *h, const double *d_b, int max_iter, int max_chunks) {
  const BView V = bview(h);
  const int Gi = grid_items(h), K = h->K;
  T *r = (T *)h->v_r, *rhat = (T *)h->v_rhat, *p = (T *)h->v_p, *nu = (T *)h->v_nu;
  BCHK(h, hipMemcpyAsync(h->d_st, h->h_st.data(), K * sizeof(Scal), hipMemcpyHostToDevice, h->stream));
  kb_init<T><<<Gi, WG, 0, h->stream>>>(V, d_b, h->d_x, iter_xopt<T>(h),
                                       sizeof(T) == 8 ? nullptr : iter_x<T>(h), r, rhat, p, nu,
                                       h->d_st, h->d_p5, F_LIVE);
  kb_eval<T><<<K, WG, 0, h->stream>>>(V, h->d_st, h->d_p5, r, rhat, 1, sizeof(T) == 4 ? 1 : 0);
  kb_alldone<<<1, WG, 0, h->stream>>>(K, h->d_st, h->d_all_done);
  bptick(h, CUP2D_PS_PHASE_SETUP);
  BCHK(h, hipGetLastError());
  int done = 0;
  int rc = read_all_done(h, done);
  if (rc) return rc;
  if (!done && max_iter > 0) {
    if ((rc = run_chunks_b(h, max_iter, max_chunks))) return rc;
  }
  if constexpr (sizeof(T) == 8)
    kb_select<<<Gi, WG, 0, h->stream>>>(V, h->d_x, h->d_xopt, h->d_st, F_ALL);
  else
    kb_fold<<<Gi, WG, 0, h->stream>>>(V, h->d_x, iter_x<float>(h), iter_xopt<float>(h), h->d_st, F_ALL);
  bptick(h, CUP2D_PS_PHASE_SETUP);
  BCHK(h, hipGetLastError());
  BCHK(h, hipMemcpyAsync(h->h_st.data(), h->d_st, K * sizeof(Scal), hipMemcpyDeviceToHost, h->stream));
  BCHK(h, hipStreamSynchronize(h->stream));
  return CUP2D_PS_OK;
}

// One BiCGSTAB run per instance from r0 = b - A x0 (FP64).  Mixed precision:
// FP32-stored Krylov vectors with FP64 residual replacement and restart
// (launch_replacement_b at the start of every chunk), so the stopping test is
// met on an FP64 residual.
int solve_b_body(cup2d_psb *h, const double *d_b, const double *d_x0, double *d_out,
                 double tol_abs, double tol_rel, int max_restarts, cup2d_ps_stats *stats) {
  if (!h->ready) return bfail(h, CUP2D_PS_ENOMESH, "solve before cup2d_psb_set_meshes");
  BCHK(h, hipSetDevice(h->cfg.device));
  const int K = h->K, max_iter = h->cfg.max_iter;
  for (double &m : h->pms) m = 0.0;
  h->pused = 0;
  bptick(h, -1);
  BCHK(h, hipEventRecord(h->ev0, h->stream));
  BCHK(h, hipMemcpyAsync(h->d_x, d_x0, h->n * sizeof(double), hipMemcpyDeviceToDevice, h->stream));
  if (d_b != h->d_b)  // the replacement step of the captured graph reads the handle's copy
    BCHK(h, hipMemcpyAsync(h->d_b, d_b, h->n * sizeof(double), hipMemcpyDeviceToDevice, h->stream));
  std::vector<int> iters(K, 0), restarts(K, 0), reps(K, 0), conv(K, 0);
  std::vector<double> err0(K, -1.0), errf(K, 0.0);
  for (int k = 0; k < K; k++) h->h_st[k] = marker_b(h, k, tol_abs, tol_rel, max_restarts, max_iter);
  int rc = h->mixed ? run_pass<float>(h, h->d_b, max_iter, -1) : run_pass<double>(h, h->d_b, max_iter, -1);
  if (rc) return rc;
  for (int k = 0; k < K; k++) {
    const Scal &s = h->h_st[k];
    iters[k] = std::max(s.iter, 0);
    restarts[k] = s.restarts;
    reps[k] = s.n_rep;
    conv[k] = s.converged;
    err0[k] = s.err0;
    // mixed: converged means the last FP64 residual met the target; it is
    // the returned iterate's
    errf[k] = (h->mixed && s.converged) ? s.err_rep : (s.iter > 0 ? s.err_opt : s.err0);
  }
  const BView V = bview(h);
  const int Gi = grid_items(h);
  kb_fmean<<<Gi, WG, 0, h->stream>>>(V, h->d_x, h->d_pm);
  kb_mean<<<K, WG, 0, h->stream>>>(V, h->d_st, h->d_pm);
  kb_fwrite<<<Gi, WG, 0, h->stream>>>(V, h->d_x, d_out, h->d_st);
  BCHK(h, hipGetLastError());
  BCHK(h, hipEventRecord(h->ev1, h->stream));
  // the returned solutions' true residuals, recomputed in FP64 (x_opt is
  // free after the solve and holds A x)
  kb_spmv<<<Gi, WG, 0, h->stream>>>(V, d_out, h->d_xopt, h->d_pm);
  kb_meanrow<<<K, WG, 0, h->stream>>>(V, h->d_pm, h->d_xopt);
  kb_rpart<<<Gi, WG, 0, h->stream>>>(V, d_b, h->d_xopt, h->d_p5);
  kb_rtot<<<K, WG, 0, h->stream>>>(V, h->d_p5, h->d_rt);
  bptick(h, CUP2D_PS_PHASE_SETUP);
  BCHK(h, hipGetLastError());
  h->h_rt.assign(4 * (size_t)K, 0.0);
  BCHK(h, hipMemcpyAsync(h->h_rt.data(), h->d_rt, 4 * K * sizeof(double), hipMemcpyDeviceToHost, h->stream));
  BCHK(h, hipStreamSynchronize(h->stream));
  bpcollect(h);
  float ms = 0;
  BCHK(h, hipEventElapsedTime(&ms, h->ev0, h->ev1));
  if (stats)
    for (int k = 0; k < K; k++) {
      std::memset(&stats[k], 0, sizeof(cup2d_ps_stats));
      stats[k].iterations = iters[k];
      stats[k].restarts = restarts[k];
      stats[k].converged = conv[k];
      stats[k].refinements = reps[k];
      stats[k].err_init = err0[k];
      stats[k].err_final = errf[k];
      stats[k].ms = ms;
      const double *rt = &h->h_rt[4 * (size_t)k];
      stats[k].true_rel2 = rt[2] > 0 ? std::sqrt(rt[0]) / std::sqrt(rt[2]) : 0.0;
      stats[k].true_relinf = rt[3] > 0 ? rt[1] / rt[3] : 0.0;
      for (int q = 0; q < 5; q++) stats[k].ms_phase[q] = h->cfg.profile ? h->pms[q] : 0.0;
    }
  return CUP2D_PS_OK;
}

// ms per batched iteration with the current units: a warm-up run, then a
// timed run of 4 chunks with zero tolerances (x starts at 0)
int time_units(cup2d_psb *h, const double *d_b, double &ms_per_iter) {
  const int K = h->K, C = chunk_b(h);
  float ms = 0;
  for (int rep = 0; rep < 2; rep++) {
    for (int k = 0; k < K; k++) h->h_st[k] = marker_b(h, k, 0.0, 0.0, 1 << 20, 1 << 20);
    BCHK(h, hipMemsetAsync(h->d_x, 0, h->n * sizeof(double), h->stream));
    BCHK(h, hipEventRecord(h->ev0, h->stream));
    int rc = h->mixed ? run_pass<float>(h, d_b, 4 * C, 4) : run_pass<double>(h, d_b, 4 * C, 4);
    if (rc) return rc;
    BCHK(h, hipEventRecord(h->ev1, h->stream));
    BCHK(h, hipEventSynchronize(h->ev1));
    BCHK(h, hipEventElapsedTime(&ms, h->ev0, h->ev1));
  }
  ms_per_iter = ms / (4.0 * C);
  return CUP2D_PS_OK;
}

int solve_b(cup2d_psb *h, const double *d_b, const double *d_x0, double *d_out, double tol_abs,
            double tol_rel, int max_restarts, cup2d_ps_stats *stats) {
  return bguarded(h, [&] {
    return solve_b_body(h, d_b, d_x0, d_out, tol_abs, tol_rel, max_restarts, stats);
  });
}

// Tuned item sizes of this process, by (precision, instances, blocks).
// Handles on different threads may tune and read it at once: g_tuned_mu.
std::map<std::tuple<int, int, long long>, std::pair<int, int>> g_tuned;  // (ib, xg_log2)
std::mutex g_tuned_mu;

bool config_ok(const cup2d_ps_config &c) {
  return c.max_iter > 0 && c.chunk >= 0 && c.eps >= 0 && c.breakdown_tol >= 0 &&
         c.inner_tol >= 0 && c.inner_tol < 1 &&
         (c.precision == CUP2D_PS_FP64 || c.precision == CUP2D_PS_MIXED) &&
         (c.norm == CUP2D_PS_NORM_INF || c.norm == CUP2D_PS_NORM_L2) && c.item_blocks >= 0 &&
         c.item_blocks <= MAX_ITEM_BLOCKS && c.schedule == 0 && (c.profile == 0 || c.profile == 1) &&
         (c.xcd_run == 0 || (c.xcd_run > 0 && c.xcd_run <= 32 && (c.xcd_run & (c.xcd_run - 1)) == 0));
}

}  // namespace

extern "C" {

static int create_body(const cup2d_ps_config *cfg, cup2d_psb **out) {
  if (!out) return bfail(nullptr, CUP2D_PS_EINVAL, "out is NULL");
  *out = nullptr;
  cup2d_ps_config c;
  if (cfg) c = *cfg;
  else cup2d_ps_config_default(&c);
  if (!config_ok(c)) return bfail(nullptr, CUP2D_PS_EINVAL, "invalid configuration");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0)
    return bfail(nullptr, CUP2D_PS_ENODEV, "no HIP device");
  if (c.device < 0 || c.device >= ndev) return bfail(nullptr, CUP2D_PS_EINVAL, "bad device ordinal");
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, c.device) != hipSuccess)
    return bfail(nullptr, CUP2D_PS_ENODEV, "hipGetDeviceProperties failed");
  if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
    return bfail(nullptr, CUP2D_PS_ENODEV, std::string("built for gfx950, device is ") + prop.gcnArchName);
  auto h = std::make_unique<cup2d_psb>();
  h->cfg = c;
  h->mixed = c.precision == CUP2D_PS_MIXED;
  if (hipSetDevice(c.device) != hipSuccess ||
      hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreate(&h->ev0) != hipSuccess || hipEventCreate(&h->ev1) != hipSuccess ||
      hipHostMalloc((void **)&h->h_all_done, sizeof(int), hipHostMallocDefault) != hipSuccess) {
    cup2d_psb_destroy(h.release());
    return bfail(nullptr, CUP2D_PS_EHIP, "stream/event/pinned allocation failed");
  }
  *out = h.release();
  return CUP2D_PS_OK;
}

int cup2d_psb_destroy(cup2d_psb *h) {
  if (!h) return CUP2D_PS_OK;
  (void)hipSetDevice(h->cfg.device);
  if (h->stream) (void)hipStreamSynchronize(h->stream);
  drop_graph(h);
  bfree_pool(h->unit_allocs);
  bfree_pool(h->allocs);
  bfree_pool(h->spare);
  if (h->h_all_done) (void)hipHostFree(h->h_all_done);
  for (hipEvent_t e : h->pev) (void)hipEventDestroy(e);
  if (h->ev0) (void)hipEventDestroy(h->ev0);
  if (h->ev1) (void)hipEventDestroy(h->ev1);
  if (h->stream) (void)hipStreamDestroy(h->stream);
  delete h;
  return CUP2D_PS_OK;
}

static int build_host(cup2d_psb *h, int n_inst, const cup2d_mesh_desc *meshes);
static int build_device(cup2d_psb *h, int n_inst, const cup2d_mesh_desc *meshes);
static int finish_meshes(cup2d_psb *h, int n_inst);

static int set_meshes_body(cup2d_psb *h, int n_inst, const cup2d_mesh_desc *meshes, bool on_device) {
  if (!h) return bfail(nullptr, CUP2D_PS_EINVAL, "handle is NULL");
  if (!meshes || n_inst <= 0 || n_inst > MAX_INST)
    return bfail(h, CUP2D_PS_EINVAL, "need 1.." + std::to_string(MAX_INST) + " meshes");
  BCHK(h, hipSetDevice(h->cfg.device));
  BCHK(h, hipStreamSynchronize(h->stream));
  drop_graph(h);
  brecycle(h, h->unit_allocs);
  brecycle(h, h->allocs);
  h->ready = false;
  int rc = on_device ? build_device(h, n_inst, meshes) : build_host(h, n_inst, meshes);
  if (rc) return rc;
  // cell size per block, h0 / 2^level (BlockInfo::h), for the projection operators
  std::vector<double> hb;
  hb.reserve((size_t)h->nb);
  for (int k = 0; k < n_inst; k++) {
    const cup2d_mesh_desc &m = meshes[k];
    const double h0 = m.extent / (double(std::max(m.bpdx, m.bpdy)) * BS);
    for (long long b = 0; b < m.n_blocks; b++) hb.push_back(h0 / double(1LL << m.level[b]));
  }
  if ((rc = bupload(h, h->allocs, &h->d_h, hb))) return rc;
  return finish_meshes(h, n_inst);
}

// The concatenated operator assembled on host threads.
static int build_host(cup2d_psb *h, int n_inst, const cup2d_mesh_desc *meshes) {
  // the instances' operators are independent: built on host threads, without
  // the plain CSR copy (the batch never reads it)
  std::vector<cup2d::Operator> ops(n_inst);
  std::vector<std::string> errs(n_inst);
  std::vector<char> ok(n_inst, 0);
  {
    const int nt = (int)std::max(1u, std::min<unsigned>(16u, std::thread::hardware_concurrency()));
    auto work = [&](int t) {
      for (int k = t; k < n_inst; k += nt) {
        const cup2d_mesh_desc &m = meshes[k];
        MeshDesc md{m.bpdx, m.bpdy, m.levels, m.periodic_x, m.periodic_y, m.extent, m.n_blocks,
                    m.level, m.bx, m.by};
        try {
          ok[k] = build_operator(md, h->cfg.mean_constraint, ops[k], errs[k], false) ? 1 : 0;
        } catch (const std::exception &e) {  // e.g. out of host memory
          errs[k] = e.what();
          ok[k] = 0;
        }
      }
    };
    std::vector<std::thread> pool;
    std::vector<int> mine(1, 0);  // strides this thread runs: its own, and any not started
    for (int t = 1; t < std::min(nt, n_inst); t++) {
      try {
        pool.emplace_back(work, t);
      } catch (...) {  // no exception may cross the C ABI: run that stride here
        mine.push_back(t);
      }
    }
    for (int t : mine) work(t);
    for (std::thread &th : pool) th.join();
  }
  std::vector<long long> boff(n_inst + 1, 0);
  long long nnz = 0;
  for (int k = 0; k < n_inst; k++) {
    if (!ok[k]) return bfail(h, CUP2D_PS_EINVAL, "mesh " + std::to_string(k) + ": " + errs[k]);
    boff[k + 1] = boff[k] + ops[k].nb;
    nnz += ops[k].nnz();
  }
  if (boff[n_inst] * 64 >= (1LL << 31) - 64)
    return bfail(h, CUP2D_PS_EINVAL, "batch too large for 32-bit cell indices");
  h->K = n_inst;
  h->boff = boff;
  h->nnz = nnz;
  h->nb = boff[n_inst];
  h->n = h->nb * 64;
  // concatenated operator: block ids, explicit rows and their columns shifted
  // by each instance's offsets
  int W = 0;
  for (int k = 0; k < n_inst; k++) W = std::max(W, ell_width(ops[k]));
  if (W > ELL_MAX)
    return bfail(h, CUP2D_PS_EINVAL, "explicit row with more than " + std::to_string(ELL_MAX) + " entries");
  std::vector<int4> nbr;
  std::vector<int2> fid;
  std::vector<unsigned long long> mask;
  std::vector<double> h2;
  EllRows ell;
  ell.W = W;
  std::string eerr;
  nbr.reserve(h->nb);
  fid.reserve(4 * h->nb);
  mask.reserve(h->nb);
  h2.reserve(h->nb);
  h->h_mean_row.assign(n_inst, -1);
  h->eoff.assign(n_inst + 1, 0);
  for (int k = 0; k < n_inst; k++) {
    const cup2d::Operator &op = ops[k];
    const int bo = (int)boff[k];
    const long long ro = boff[k] * 64;
    auto shift = [bo](int v) { return v >= 0 ? v + bo : v; };
    for (long long b = 0; b < op.nb; b++) {
      int4 v;
      v.x = shift(op.nbr[4 * b]);
      v.y = shift(op.nbr[4 * b + 1]);
      v.z = shift(op.nbr[4 * b + 2]);
      v.w = shift(op.nbr[4 * b + 3]);
      nbr.push_back(v);
      for (int f = 0; f < 4; f++)  // block ids shifted; a coarse face's half code (< 0) kept
        fid.push_back(make_int2(shift(op.fid[8 * b + 2 * f]), shift(op.fid[8 * b + 2 * f + 1])));
      mask.push_back(op.exc_mask[b]);
      h2.push_back(op.h2[b]);
    }
    if (!ell.append(op, ro, eerr))
      return bfail(h, CUP2D_PS_EINVAL, "mesh " + std::to_string(k) + ": " + eerr);
    h->eoff[k + 1] = (long long)ell.erow.size();
    if (op.mean_row >= 0) h->h_mean_row[k] = op.mean_row + ro;
  }
  h->ell_w = W;
  h->elen = std::move(ell.len);
  ops.clear();
  auto &pool = h->allocs;
  int rc;
  if ((rc = bupload(h, pool, &h->d_nbr, nbr)) || (rc = bupload(h, pool, &h->d_mask, mask)) ||
      (rc = bupload(h, pool, &h->d_fid, fid)) || (rc = bupload(h, pool, &h->d_erow, ell.erow)) ||
      (rc = bupload(h, pool, &h->d_ecol, ell_transpose(ell.ecol_rm, W))) ||
      (rc = bupload(h, pool, &h->d_ecode, ell_transpose(ell.ecode_rm, W))) ||
      (rc = bupload(h, pool, &h->d_etab, ell.values)) || (rc = bupload(h, pool, &h->d_h2, h2)))
    return rc;
  return CUP2D_PS_OK;
}

// The same operator assembled by the device-side assembly, all meshes in one
// pipeline (assemble_device.hip); only the meshes' block lists cross PCIe.
static int build_device(cup2d_psb *h, int n_inst, const cup2d_mesh_desc *meshes) {
  std::vector<MeshDesc> md(n_inst);
  for (int k = 0; k < n_inst; k++) {
    const cup2d_mesh_desc &m = meshes[k];
    md[k] = MeshDesc{m.bpdx, m.bpdy, m.levels, m.periodic_x, m.periodic_y, m.extent, m.n_blocks,
                     m.level, m.bx, m.by};
  }
  auto &pool = h->allocs;
  const DeviceAlloc alloc = [h, &pool](size_t bytes) -> void * {
    uint8_t *p = nullptr;
    return balloc(h, pool, &p, bytes) ? nullptr : p;
  };
  std::vector<BBuf> temps;  // the assembly's work arrays, spare again afterwards
  const DeviceAlloc temp = [h, &temps](size_t bytes) -> void * {
    uint8_t *p = nullptr;
    return balloc(h, temps, &p, bytes) ? nullptr : p;
  };
  DeviceOperator D;
  std::string err;
  int bad = -1;
  const int rc = assemble_on_device(md.data(), n_inst, h->cfg.mean_constraint, h->stream, alloc, temp,
                                    D, bad, err);
  (void)hipStreamSynchronize(h->stream);
  brecycle(h, temps);
  if (rc) return bfail(h, rc, bad >= 0 ? "mesh " + std::to_string(bad) + ": " + err : err);
  h->K = n_inst;
  h->boff = D.boff;
  h->nnz = D.nnz;
  h->nb = D.nb;
  h->n = D.n;
  h->ell_w = D.ell_w;
  h->eoff = D.eoff;
  h->elen = std::move(D.elen);
  h->h_mean_row = D.mean_rows;
  h->d_nbr = D.nbr;
  h->d_fid = D.fid;
  h->d_mask = D.mask;
  h->d_h2 = D.h2;
  h->d_erow = D.erow;
  h->d_ecol = D.ecol;
  h->d_ecode = D.ecode;
  h->d_etab = D.etab;
  return CUP2D_PS_OK;
}

// Preconditioner, work vectors and work units: the part of a mesh change
// that does not depend on where the operator was assembled.
static int finish_meshes(cup2d_psb *h, int n_inst) {
  auto &pool = h->allocs;
  int rc;
  static const std::vector<double> P = make_pinv();  // the same for every mesh
  static const std::vector<uint16_t> pf = make_pfrag(P);
  static const std::vector<double> pf64 = make_pfrag64(P);
  if ((rc = bupload(h, pool, &h->d_P, pf64)) || (rc = bupload(h, pool, &h->d_mean_row, h->h_mean_row)))
    return rc;
  if ((rc = balloc(h, pool, &h->d_pfrag, pf.size() / 8))) return rc;
  BCHK(h, hipMemcpy(h->d_pfrag, pf.data(), pf.size() * sizeof(uint16_t), hipMemcpyHostToDevice));
  const size_t n = (size_t)h->n, es = h->mixed ? sizeof(float) : sizeof(double);
  double **dv[] = {&h->d_x, &h->d_xopt, &h->d_b, &h->d_xio};
  for (double **v : dv)
    if ((rc = balloc(h, pool, v, n))) return rc;
  void **kv[] = {&h->v_r, &h->v_rhat, &h->v_p, &h->v_nu, &h->v_z1, &h->v_z2, &h->v_t};
  for (void **v : kv) {
    char *p = nullptr;
    if ((rc = balloc(h, pool, &p, n * es))) return rc;
    *v = p;
  }
  if ((rc = balloc(h, pool, &h->d_st, (size_t)n_inst)) || (rc = balloc(h, pool, &h->d_all_done, 1)) ||
      (rc = balloc(h, pool, &h->d_rt, 4 * (size_t)n_inst)))
    return rc;
  h->h_st.assign(n_inst, Scal{});
  int ib = h->cfg.item_blocks;
  {
    std::lock_guard<std::mutex> lock(g_tuned_mu);
    auto it = g_tuned.find({(int)h->mixed, n_inst, h->nb});
    if (ib <= 0) ib = it != g_tuned.end() ? it->second.first : 8;
    const int xr = h->cfg.xcd_run;
    h->xg_log2 = xr > 0 ? __builtin_ctz((unsigned)xr) : (it != g_tuned.end() ? it->second.second : 3);
  }
  if ((rc = build_units(h, ib))) return rc;
  btrim(h);
  h->ready = true;
  return CUP2D_PS_OK;
}

int cup2d_psb_offsets(const cup2d_psb *h, long long *offsets) {
  if (!h || !offsets) return CUP2D_PS_EINVAL;
  if (!h->ready) return CUP2D_PS_ENOMESH;
  for (int k = 0; k <= h->K; k++) offsets[k] = h->boff[k] * 64;
  return CUP2D_PS_OK;
}

int cup2d_psb_solve_device(cup2d_psb *h, const double *d_b, double *d_x, double tol_abs,
                           double tol_rel, int max_restarts, cup2d_ps_stats *stats) {
  if (!h) return bfail(nullptr, CUP2D_PS_EINVAL, "handle is NULL");
  if (!d_b || !d_x) return bfail(h, CUP2D_PS_EINVAL, "NULL vector");
  if (!(tol_abs >= 0) || !(tol_rel >= 0) || max_restarts < 0)
    return bfail(h, CUP2D_PS_EINVAL, "negative tolerance or restart count");
  return solve_b(h, d_b, d_x, d_x, tol_abs, tol_rel, max_restarts, stats);
}

int cup2d_psb_solve(cup2d_psb *h, const double *b, double *x, double tol_abs, double tol_rel,
                    int max_restarts, cup2d_ps_stats *stats) {
  if (!h) return bfail(nullptr, CUP2D_PS_EINVAL, "handle is NULL");
  if (!b || !x) return bfail(h, CUP2D_PS_EINVAL, "NULL vector");
  if (!h->ready) return bfail(h, CUP2D_PS_ENOMESH, "solve before cup2d_psb_set_meshes");
  if (!(tol_abs >= 0) || !(tol_rel >= 0) || max_restarts < 0)
    return bfail(h, CUP2D_PS_EINVAL, "negative tolerance or restart count");
  BCHK(h, hipSetDevice(h->cfg.device));
  const size_t bytes = (size_t)h->n * sizeof(double);
  BCHK(h, hipMemcpyAsync(h->d_b, b, bytes, hipMemcpyHostToDevice, h->stream));
  BCHK(h, hipMemcpyAsync(h->d_xio, x, bytes, hipMemcpyHostToDevice, h->stream));
  int rc = solve_b(h, h->d_b, h->d_xio, h->d_xio, tol_abs, tol_rel, max_restarts, stats);
  if (rc) return rc;
  BCHK(h, hipMemcpyAsync(x, h->d_xio, bytes, hipMemcpyDeviceToHost, h->stream));
  BCHK(h, hipStreamSynchronize(h->stream));
  return CUP2D_PS_OK;
}

static int autotune_body(cup2d_psb *h, const double *d_b, int *item_blocks, double *ms_per_iter) {
  if (!h) return bfail(nullptr, CUP2D_PS_EINVAL, "handle is NULL");
  if (!d_b) return bfail(h, CUP2D_PS_EINVAL, "NULL vector");
  if (!h->ready) return bfail(h, CUP2D_PS_ENOMESH, "autotune before cup2d_psb_set_meshes");
  BCHK(h, hipSetDevice(h->cfg.device));
  if (d_b != h->d_b)  // the captured graph reads the handle's copy of b
    BCHK(h, hipMemcpyAsync(h->d_b, d_b, h->n * sizeof(double), hipMemcpyDeviceToDevice, h->stream));
  // the schedule: blocks per item, and the XCD run length of the operator
  // kernels' items (1: plain grid-stride order, or runs of 4 .. 32 items)
  static const int cand[] = {4, 8, 16, 32, 64};
  static const int xgs[] = {0, 2, 3, 4, 5};
  int best = h->ib, best_xg = h->xg_log2;
  double best_ms = 1e300;
  for (int ib : cand) {
    int rc = build_units(h, ib);
    if (rc) return rc;
    for (int xg : xgs) {
      h->xg_log2 = xg;
      double ms = 0;
      if ((rc = time_units(h, h->d_b, ms))) return rc;
      if (ms < best_ms) {
        best_ms = ms;
        best = ib;
        best_xg = xg;
      }
    }
  }
  h->xg_log2 = best_xg;
  int rc = build_units(h, best);
  if (rc) return rc;
  btrim(h);
  {
    std::lock_guard<std::mutex> lock(g_tuned_mu);
    g_tuned[{(int)h->mixed, h->K, h->nb}] = {best, best_xg};
  }
  if (item_blocks) *item_blocks = best;
  if (ms_per_iter) *ms_per_iter = best_ms;
  return CUP2D_PS_OK;
}

// One batched iteration kernel alone, every instance running: inputs x = 0
// and a fixed b, one chunk of iterations to fill every partial, then `reps`
// launches of kernel `which` between HIP events on the handle's stream.
extern "C++" {
template <class T>
static int kernel_time_b(cup2d_psb *h, int which, int reps, double *ms_per_launch) {
  const size_t n = (size_t)h->n;
  std::vector<double> bh(n);
  for (size_t i = 0; i < n; i++) bh[i] = double(int(i % 7) - 3);
  BCHK(h, hipMemcpyAsync(h->d_b, bh.data(), n * sizeof(double), hipMemcpyHostToDevice, h->stream));
  BCHK(h, hipMemsetAsync(h->d_x, 0, n * sizeof(double), h->stream));
  for (int k = 0; k < h->K; k++) {
    h->h_st[k] = marker_b(h, k, 0.0, 0.0, 1 << 20, 1 << 20);
    h->h_st[k].rep_factor = 0.0;  // no pauses: every instance iterates in every launch
    h->h_st[k].stall = 0;
  }
  int rc = run_pass<T>(h, h->d_b, chunk_b(h), 1);
  if (rc) return rc;
  BCHK(h, hipStreamSynchronize(h->stream));  // bh is read before it goes out of scope
  const BView V = bview(h);
  const int Gi = grid_items(h), Gt = grid_tiles(h), K = h->K;
  hipStream_t s = h->stream;
  T *r = (T *)h->v_r, *rhat = (T *)h->v_rhat, *p = (T *)h->v_p, *nu = (T *)h->v_nu,
    *z1 = (T *)h->v_z1, *z2 = (T *)h->v_z2, *t = (T *)h->v_t;
  Scal *st = h->d_st;
  auto launch = [&]() {
    switch (which) {
      case 0:
        if constexpr (sizeof(T) == 8) kb_pupdate64<<<grid_tiles64(h), WG, 0, s>>>(V, h->d_P, r, rhat, p, nu, z1, st);
        else kb_pupdate32<<<Gt, WG, 0, s>>>(V, h->d_pfrag, r, rhat, p, nu, z1, st);
        break;
      case 1: kb_spmv_nu<T><<<Gi, WG, 0, s>>>(V, z1, nu, rhat, st, h->d_p2); break;
      case 2:
        if constexpr (sizeof(T) == 8) kb_supdate64<<<grid_tiles64(h), WG, 0, s>>>(V, h->d_P, r, nu, z2, st);
        else kb_supdate32<<<Gt, WG, 0, s>>>(V, h->d_pfrag, r, nu, z2, st);
        break;
      case 3: kb_spmv_t<T><<<Gi, WG, 0, s>>>(V, z2, t, r, st, h->d_p4); break;
      case 4: kb_xrupdate<T><<<Gi, WG, 0, s>>>(V, iter_x<T>(h), iter_xopt<T>(h), z1, z2, r, t, rhat, st, h->d_p5); break;
      case 5: kb_alpha<T><<<K, WG, 0, s>>>(V, st, h->d_p2, nu, rhat); break;
      case 6: kb_omega<T><<<K, WG, 0, s>>>(V, st, h->d_p4, t, r); break;
      default: kb_eval<T><<<K, WG, 0, s>>>(V, st, h->d_p5, r, rhat, 0, 0);
    }
  };
  launch();  // warm-up, untimed
  BCHK(h, hipEventRecord(h->ev0, s));
  for (int k = 0; k < reps; k++) launch();
  BCHK(h, hipEventRecord(h->ev1, s));
  BCHK(h, hipEventSynchronize(h->ev1));
  BCHK(h, hipGetLastError());
  float ms = 0;
  BCHK(h, hipEventElapsedTime(&ms, h->ev0, h->ev1));
  *ms_per_launch = ms / reps;
  return CUP2D_PS_OK;
}
}  // extern "C++"

static int pressure_step_b(cup2d_psb *h, const double *d_rhs, double *d_p, double tol_abs,
                           double tol_rel, int max_restarts, cup2d_ps_stats *stats) {
  if (!h) return bfail(nullptr, CUP2D_PS_EINVAL, "handle is NULL");
  if (!d_rhs || !d_p) return bfail(h, CUP2D_PS_EINVAL, "NULL vector");
  if (d_rhs == d_p) return bfail(h, CUP2D_PS_EINVAL, "rhs and p must not alias");
  if (!h->ready) return bfail(h, CUP2D_PS_ENOMESH, "pressure step before cup2d_psb_set_meshes");
  if (!(tol_abs >= 0) || !(tol_rel >= 0) || max_restarts < 0)
    return bfail(h, CUP2D_PS_EINVAL, "negative tolerance or restart count");
  BCHK(h, hipSetDevice(h->cfg.device));
  const BView V = bview(h);
  const int Gi = grid_items(h);
  kb_spmv<<<Gi, WG, 0, h->stream>>>(V, d_p, h->d_b, h->d_pm);  // L p_old (no mean-row fix)
  kb_rhs_sub<<<Gi, WG, 0, h->stream>>>((long)h->n, d_rhs, h->d_b);
  BCHK(h, hipGetLastError());
  BCHK(h, hipMemsetAsync(h->d_xio, 0, (size_t)h->n * sizeof(double), h->stream));  // dp0 = 0
  const int rc = solve_b(h, h->d_b, h->d_xio, h->d_xio, tol_abs, tol_rel, max_restarts, stats);
  if (rc) return rc;
  kb_fmean<<<Gi, WG, 0, h->stream>>>(V, h->d_xio, h->d_pm);
  kb_mean<<<h->K, WG, 0, h->stream>>>(V, h->d_st, h->d_pm);
  kb_padd<<<Gi, WG, 0, h->stream>>>(V, h->d_xio, d_p, h->d_st);
  BCHK(h, hipGetLastError());
  BCHK(h, hipStreamSynchronize(h->stream));
  return CUP2D_PS_OK;
}

int cup2d_psb_pressure_step_device(cup2d_psb *h, const double *d_rhs, double *d_p, double tol_abs,
                                   double tol_rel, int max_restarts, cup2d_ps_stats *stats) {
  return bguarded(h, [&] { return pressure_step_b(h, d_rhs, d_p, tol_abs, tol_rel, max_restarts, stats); });
}

static int projection_checks_b(cup2d_psb *h, double dt) {
  if (!h->ready) return bfail(h, CUP2D_PS_ENOMESH, "no meshes");
  if (!(dt > 0)) return bfail(h, CUP2D_PS_EINVAL, "dt must be positive");
  return CUP2D_PS_OK;
}

static int cell_grid_b(const cup2d_psb *h) {
  return (int)std::min<long long>((h->n + WG - 1) / WG, 16384);
}

int cup2d_psb_pressure_rhs_device(cup2d_psb *h, const double *d_u, const double *d_v,
                                  const double *d_udef_u, const double *d_udef_v,
                                  const double *d_chi, double dt, double *d_rhs) {
  return bguarded(h, [&]() -> int {
    if (!h) return bfail(nullptr, CUP2D_PS_EINVAL, "handle is NULL");
    if (!d_u || !d_v || !d_udef_u || !d_udef_v || !d_chi || !d_rhs)
      return bfail(h, CUP2D_PS_EINVAL, "NULL vector");
    // the kernel reads neighbours of every input while writing rhs
    if (d_rhs == d_u || d_rhs == d_v || d_rhs == d_udef_u || d_rhs == d_udef_v || d_rhs == d_chi)
      return bfail(h, CUP2D_PS_EINVAL, "rhs must not alias an input");
    if (const int rc = projection_checks_b(h, dt)) return rc;
    BCHK(h, hipSetDevice(h->cfg.device));
    k_pressure_rhs<<<cell_grid_b(h), WG, 0, h->stream>>>((int)h->nb, h->d_nbr, h->d_mask, h->d_h, d_u,
                                                         d_v, d_udef_u, d_udef_v, d_chi, dt, d_rhs);
    const int n_exc = (int)h->eoff[h->K];
    if (n_exc > 0)
      k_pressure_rhs_cf<<<exc_grid(n_exc), WG, 0, h->stream>>>(n_exc, h->d_erow, h->d_nbr, h->d_fid,
                                                               h->d_h, d_u, d_v, d_udef_u, d_udef_v,
                                                               d_chi, dt, d_rhs);
    BCHK(h, hipGetLastError());
    BCHK(h, hipStreamSynchronize(h->stream));
    return CUP2D_PS_OK;
  });
}

int cup2d_psb_pressure_correct_device(cup2d_psb *h, const double *d_p, double dt, double *d_u,
                                      double *d_v) {
  return bguarded(h, [&]() -> int {
    if (!h) return bfail(nullptr, CUP2D_PS_EINVAL, "handle is NULL");
    if (!d_p || !d_u || !d_v) return bfail(h, CUP2D_PS_EINVAL, "NULL vector");
    if (d_u == d_v || d_p == d_u || d_p == d_v) return bfail(h, CUP2D_PS_EINVAL, "p, u and v must not alias");
    if (const int rc = projection_checks_b(h, dt)) return rc;
    BCHK(h, hipSetDevice(h->cfg.device));
    k_pressure_correct<<<cell_grid_b(h), WG, 0, h->stream>>>((int)h->nb, h->d_nbr, h->d_mask, h->d_h, d_p,
                                                             dt, d_u, d_v);
    const int n_exc = (int)h->eoff[h->K];
    if (n_exc > 0)
      k_pressure_correct_cf<<<exc_grid(n_exc), WG, 0, h->stream>>>(n_exc, h->d_erow, h->d_nbr, h->d_fid,
                                                                   h->d_h, d_p, dt, d_u, d_v);
    BCHK(h, hipGetLastError());
    BCHK(h, hipStreamSynchronize(h->stream));
    return CUP2D_PS_OK;
  });
}

int cup2d_psb_spmv_device(cup2d_psb *h, const double *d_x, double *d_y) {
  if (!h) return bfail(nullptr, CUP2D_PS_EINVAL, "handle is NULL");
  if (!d_x || !d_y) return bfail(h, CUP2D_PS_EINVAL, "NULL vector");
  if (!h->ready) return bfail(h, CUP2D_PS_ENOMESH, "spmv before cup2d_psb_set_meshes");
  BCHK(h, hipSetDevice(h->cfg.device));
  const BView V = bview(h);
  kb_spmv<<<grid_items(h), WG, 0, h->stream>>>(V, d_x, d_y, h->d_pm);
  kb_meanrow<<<h->K, WG, 0, h->stream>>>(V, h->d_pm, d_y);
  BCHK(h, hipGetLastError());
  BCHK(h, hipStreamSynchronize(h->stream));
  return CUP2D_PS_OK;
}

int cup2d_psb_spmv_time(cup2d_psb *h, const double *d_x, double *d_y, int reps,
                        double *ms_per_launch) {
  if (!h) return bfail(nullptr, CUP2D_PS_EINVAL, "handle is NULL");
  if (!d_x || !d_y || !ms_per_launch || reps <= 0) return bfail(h, CUP2D_PS_EINVAL, "bad argument");
  if (!h->ready) return bfail(h, CUP2D_PS_ENOMESH, "spmv before cup2d_psb_set_meshes");
  BCHK(h, hipSetDevice(h->cfg.device));
  const BView V = bview(h);
  const int G = grid_items(h);
  kb_spmv<<<G, WG, 0, h->stream>>>(V, d_x, d_y, h->d_pm);  // warm-up, untimed
  BCHK(h, hipEventRecord(h->ev0, h->stream));
  for (int k = 0; k < reps; k++) kb_spmv<<<G, WG, 0, h->stream>>>(V, d_x, d_y, h->d_pm);
  BCHK(h, hipEventRecord(h->ev1, h->stream));
  BCHK(h, hipEventSynchronize(h->ev1));
  BCHK(h, hipGetLastError());
  float ms = 0;
  BCHK(h, hipEventElapsedTime(&ms, h->ev0, h->ev1));
  *ms_per_launch = ms / reps;
  return CUP2D_PS_OK;
}

int cup2d_psb_kernel_time(cup2d_psb *h, int which, int reps, double *ms_per_launch) {
  if (!h) return bfail(nullptr, CUP2D_PS_EINVAL, "handle is NULL");
  if (!ms_per_launch || reps <= 0 || which < 0 || which > 7) return bfail(h, CUP2D_PS_EINVAL, "bad argument");
  if (!h->ready) return bfail(h, CUP2D_PS_ENOMESH, "kernel timing before cup2d_psb_set_meshes");
  return bguarded(h, [&]() -> int {
    BCHK(h, hipSetDevice(h->cfg.device));
    return h->mixed ? kernel_time_b<float>(h, which, reps, ms_per_launch)
                    : kernel_time_b<double>(h, which, reps, ms_per_launch);
  });
}

int cup2d_psb_copy_time(cup2d_psb *h, const double *d_x, double *d_y, int reps, double *ms_per_launch) {
  if (!h) return bfail(nullptr, CUP2D_PS_EINVAL, "handle is NULL");
  if (!d_x || !d_y || !ms_per_launch || reps <= 0) return bfail(h, CUP2D_PS_EINVAL, "bad argument");
  if (!h->ready) return bfail(h, CUP2D_PS_ENOMESH, "copy before cup2d_psb_set_meshes");
  BCHK(h, hipSetDevice(h->cfg.device));
  const long n = (long)h->n;
  const int G = (int)std::min<long>(16384, (n + 4 * WG - 1) / (4 * WG));
  kb_copy<<<G, WG, 0, h->stream>>>(d_x, d_y, n);  // warm-up, untimed
  BCHK(h, hipEventRecord(h->ev0, h->stream));
  for (int k = 0; k < reps; k++) kb_copy<<<G, WG, 0, h->stream>>>(d_x, d_y, n);
  BCHK(h, hipEventRecord(h->ev1, h->stream));
  BCHK(h, hipEventSynchronize(h->ev1));
  BCHK(h, hipGetLastError());
  float ms = 0;
  BCHK(h, hipEventElapsedTime(&ms, h->ev0, h->ev1));
  *ms_per_launch = ms / reps;
  return CUP2D_PS_OK;
}

int cup2d_psb_sizes(const cup2d_psb *h, long long *n, long long *nnz, long long *n_explicit) {
  if (!h) return CUP2D_PS_EINVAL;
  if (!h->ready) return CUP2D_PS_ENOMESH;
  if (n) *n = h->n;
  if (nnz) *nnz = h->nnz;
  if (n_explicit) *n_explicit = h->eoff[h->K];
  return CUP2D_PS_OK;
}

int cup2d_psb_item_blocks(const cup2d_psb *h) { return h ? h->ib : 0; }
int cup2d_psb_xcd_run(const cup2d_psb *h) { return h ? 1 << h->xg_log2 : 0; }

const char *cup2d_psb_last_error(const cup2d_psb *h) {
  return h ? h->err.c_str() : g_berr.c_str();
}

int cup2d_psb_create(const cup2d_ps_config *cfg, cup2d_psb **out) {
  return bguarded(nullptr, [&] { return create_body(cfg, out); });
}

int cup2d_psb_set_meshes(cup2d_psb *h, int n_inst, const cup2d_mesh_desc *meshes) {
  return bguarded(h, [&] { return set_meshes_body(h, n_inst, meshes, false); });
}

int cup2d_psb_set_meshes_on_device(cup2d_psb *h, int n_inst, const cup2d_mesh_desc *meshes) {
  return bguarded(h, [&] { return set_meshes_body(h, n_inst, meshes, true); });
}

int cup2d_psb_autotune(cup2d_psb *h, const double *d_b, int *item_blocks, double *ms_per_iter) {
  return bguarded(h, [&] { return autotune_body(h, d_b, item_blocks, ms_per_iter); });
}

}  // extern "C"
