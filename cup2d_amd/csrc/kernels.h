// kernels.h -- device-side building blocks shared by the solver kernels.
#pragma once

#include <hip/hip_runtime.h>

namespace cup2d {

constexpr int WAVES = 4;            // wavefronts per workgroup
constexpr int WG = 64 * WAVES;      // threads per workgroup
constexpr int TILE = 16;            // blocks per wavefront in the MFMA preconditioner

// Device view of the operator.  Rows without a coarse-fine face are the
// plain 5-point stencil sum_f (x_nb(f) - x_self) over the non-wall faces,
// given by one same-level neighbour id per block face (nbr: -1 wall, -2
// coarse-fine).  Rows with a coarse-fine face ("explicit rows", bit c of
// mask[b]) are stored as ELL: row e has ell_w entries ecol/eval[w * n_exc +
// e] (padded with zero values).  The coefficients are FP64 in both
// precisions: FP32 coefficients would break the operator's exact zero row
// sums (the Neumann null space) and make the singular system inconsistent.
struct OpView {
  int nb;
  int mean_row;                          // -1 when the mean row is not replaced
  const int4 *nbr;                       // [nb] same-level neighbour per face (Xm, Xp, Ym, Yp)
  const unsigned long long *mask;        // [nb] explicit rows of each block
  int n_exc, ell_w;
  const int *erow;                       // [n_exc] global row of each explicit row
  const int *ecol;                       // [ell_w * n_exc] column (padding: own row)
  const unsigned short *ecode;           // [ell_w * n_exc] coefficient, as an index into etab
  const double *etab;                    // distinct coefficients (etab[0] = 0.0, the padding)
  const unsigned char *ewid;             // [ceil(n_exc / 64)] widest row of each run of 64
  const double *h2;                      // [nb]
};

constexpr int ELL_MAX = 16;              // widest explicit row the kernels take

// Solver scalars, resident on the device for the whole solve.  The solver
// keeps two copies, indexed by iteration parity: iteration i reads and
// updates slot i%2 only through fields no other workgroup of the same launch
// reads, and derives slot i%2 from slot (i-1)%2 plus the partial sums of the
// previous kernel.  Kernel boundaries publish the writes; no kernel of the
// iteration needs a device-wide fence.
struct Scal {
  double alpha, beta, omega, rho, rho_prev;
  double err, err0, err_opt;
  double mean;                           // h^2-weighted mean of the solution
  double mu;                             // mean of r removed at the top of the iteration
  double err_ref;                        // last error that improved 2x on its predecessor
  double n;                              // number of unknowns
  double tol_abs, tol_rel, eps, bd_tol;
  double red[8];                         // last reduction results (diagnostics)
  int iter, max_iter, restarts, max_restarts;
  int done, converged, improved, restart;
  int norm;
  int project;                           // remove the 1-component of r every iteration
  int iter_ref;                          // iteration of err_ref
  int stall;                             // end the solve after this many iterations without a 2x gain (0: never)
  // residual replacement (FP32-stored Krylov vectors): once the recurrence
  // residual has fallen by rep_factor, stalls, or claims convergence, the
  // solve pauses; the next replacement step takes its best iterate,
  // recomputes r = b - A x in FP64 and restarts BiCGSTAB from it.
  // Convergence is accepted only on such a residual.  Off when rep_factor
  // is 0.
  double rep_factor;                     // replace when err < rep_factor * err_rep
  double err_rep;                        // FP64 residual norm at the last replacement
  int iter_rep;                          // iteration of the last replacement
  int replace;                           // a replacement is pending
  int n_rep;                             // replacements so far
  int idle;                              // done, or waiting for its replacement
};

__device__ __forceinline__ int wave_id() { return __builtin_amdgcn_readfirstlane(threadIdx.x >> 6); }
__device__ __forceinline__ int lane_id() { return threadIdx.x & 63; }

// Orders the wave's own LDS writes before its later LDS reads (lanes of one
// wavefront exchange values through LDS without a workgroup barrier).
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}
// max that propagates NaN (fmax would drop it and hide a failed iteration)
__device__ __forceinline__ double nmax(double a, double b) {
  return (a != a || b != b) ? __builtin_nan("") : fmax(a, b);
}
__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) v = nmax(v, __shfl_xor(v, o, 64));
  return v;
}

// Workgroup reduction of NQ per-thread quantities (bit q of MAXQ: max, else
// sum); thread 0 stores them to part[q*gridDim.x + blockIdx.x].  No fence:
// the consumer is a later kernel.
template <int NQ, unsigned MAXQ>
__device__ void store_partials(double (&v)[NQ], double *part) {
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
      part[q * gridDim.x + blockIdx.x] = a;
    }
  }
}

// Every workgroup reduces the G partials of each quantity written by the
// previous kernel, in a fixed order (identical totals in all workgroups);
// every thread receives the totals.
template <int NQ, unsigned MAXQ>
__device__ void load_totals(const double *__restrict__ part, int G, double (&out)[NQ]) {
  __shared__ double fin[NQ][WAVES];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  double a[NQ];
#pragma unroll
  for (int q = 0; q < NQ; q++) a[q] = 0.0;
  // up to LR partials per thread per pass, all their loads issued before the
  // first is added (one round trip for G <= LR * WG); added in the order of
  // a one-at-a-time loop, so the totals do not depend on LR
  constexpr int LR = 3;
  for (int i0 = threadIdx.x; i0 < G; i0 += LR * WG) {
    double p[LR][NQ];
#pragma unroll
    for (int r = 0; r < LR; r++)
      if (i0 + r * WG < G)
#pragma unroll
        for (int q = 0; q < NQ; q++) p[r][q] = part[q * G + i0 + r * WG];
#pragma unroll
    for (int r = 0; r < LR; r++)
      if (i0 + r * WG < G)
#pragma unroll
        for (int q = 0; q < NQ; q++) a[q] = ((MAXQ >> q) & 1) ? nmax(a[q], p[r][q]) : a[q] + p[r][q];
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

// Loads of the stencil row of cell c of block b (lane = cell), kept in the
// storage type: its own value and the values across the block faces it
// touches (ix = 0 or 7: the horizontal neighbour block, iy = 0 or 7: the
// vertical one).  All loads are independent and unconditional -- a lane
// with no neighbour block across a face reads a cell of its own block, which
// the wave loads anyway, and stencil_eval ignores it -- so a wave has them in
// flight together and none waits behind a branch.
template <class TV>
__device__ __forceinline__ void stencil_load(const TV *__restrict__ x, int b, int4 nb, int c,
                                             TV &tc, TV &th, TV &tv) {
  const int ix = c & 7, iy = c >> 3;
  const int hb = (ix == 0) ? nb.x : ((ix == 7) ? nb.y : -1);
  const int vb = (iy == 0) ? nb.z : ((iy == 7) ? nb.w : -1);
  const int hc = (ix == 0) ? iy * 8 + 7 : iy * 8;
  const int vc = (iy == 0) ? 56 + ix : ix;
  tc = x[(long)b * 64 + c];
  th = x[(long)(hb >= 0 ? hb : b) * 64 + hc];
  tv = x[(long)(vb >= 0 ? vb : b) * 64 + vc];
}

// 5-point stencil row of cell c (FP64 arithmetic): sum over the non-wall
// faces of (x_nb - x_c), in the face order Ym, Xm, Xp, Yp.  Neighbours
// inside the block come from the other lanes of the wave (shuffled in the
// storage type), those across a face from th / tv (walls and coarse-fine
// faces are skipped).  Meaningless for explicit rows, which the caller
// overwrites.
template <class TV>
__device__ __forceinline__ double stencil_eval(int4 nb, int c, TV tc, TV th, TV tv) {
  const int ix = c & 7, iy = c >> 3;
  const double xl = (double)__shfl(tc, (c - 1) & 63), xr = (double)__shfl(tc, (c + 1) & 63);
  const double xd = (double)__shfl(tc, (c - 8) & 63), xu = (double)__shfl(tc, (c + 8) & 63);
  const double xc = (double)tc, xh = (double)th, xv = (double)tv;
  // branch-free: a face contributes its neighbour when it lies inside the
  // block or across a same-level face (adding +0.0 otherwise leaves the sum
  // unchanged, so the order and the result are those of the plain loop)
  const bool od = iy > 0 || nb.z >= 0, ol = ix > 0 || nb.x >= 0;
  const bool orr = ix < 7 || nb.y >= 0, ou = iy < 7 || nb.w >= 0;
  double y = 0.0;
  y += od ? (iy > 0 ? xd : xv) : 0.0;
  y += ol ? (ix > 0 ? xl : xh) : 0.0;
  y += orr ? (ix < 7 ? xr : xh) : 0.0;
  y += ou ? (iy < 7 ? xu : xv) : 0.0;
  const int deg = int(od) + int(ol) + int(orr) + int(ou);
  return y - double(deg) * xc;
}

// Explicit row e (one thread): sum_w val * x[col] in FP64 over its first
// `width` entries (wave-uniform, at least the row's length; entries past the
// row's length are (own row, 0.0)), in ascending column order; loads issued
// eight entries at a time.
// ELL entries w0 .. w0 + 7 of explicit row e (entries past `width` repeat
// entry w0: the loads stay unconditional)
struct Ent8 {
  int col[8];
  unsigned code[8];
};
__device__ __forceinline__ void load_ent8(const OpView &A, int e, int w0, int width, Ent8 &E) {
#pragma unroll
  for (int w = 0; w < 8; w++) {
    const long k = (long)(w0 + w < width ? w0 + w : w0) * A.n_exc + e;
    E.col[w] = A.ecol[k];
    E.code[w] = A.ecode[k];
  }
}
// acc + the entries' products, in entry order
template <class TV>
__device__ __forceinline__ double sum_ent8(const OpView &A, const TV *__restrict__ x, const Ent8 &E,
                                           int w0, int width, double acc) {
  double xv[8], val[8];
#pragma unroll
  for (int w = 0; w < 8; w++) {
    xv[w] = (double)x[E.col[w]];
    val[w] = A.etab[E.code[w]];
  }
#pragma unroll
  for (int w = 0; w < 8; w++)
    if (w0 + w < width) acc += val[w] * xv[w];
  return acc;
}

// Explicit row e (width <= ELL_MAX = 16 entries): the index loads of both
// halves of a wide row are issued before the first half's gathers, so a row
// costs two dependent round trips whatever its width.
template <class TV>
__device__ __forceinline__ double explicit_row(const OpView &A, const TV *__restrict__ x, int e,
                                               int width) {
  Ent8 lo;
  load_ent8(A, e, 0, width, lo);
  if (width <= 8) return sum_ent8(A, x, lo, 0, width, 0.0);
  Ent8 hi;
  load_ent8(A, e, 8, width, hi);
  return sum_ent8(A, x, hi, 8, width, sum_ent8(A, x, lo, 0, width, 0.0));
}

// Width of explicit row e's run of 64 rows (e..e+63 aligned at 64 are one
// wave's rows in the single-mesh kernels, so the value is wave-uniform).
__device__ __forceinline__ int ell_run_width(const OpView &A, int e) {
  return __builtin_amdgcn_readfirstlane((int)A.ewid[e >> 6]);
}

// The wave's U blocks b[u] (lane = cell): headers, then every load of the
// U blocks at once -- the stencil loads, h^2 of the block and aux[i] (a
// second vector of the same type the caller combines with the result;
// callers without one pass x).  Values stay in the storage type until used.
template <int U, class TV>
struct BlockRows {
  int bk[U];
  int4 nb[U];
  unsigned long long m[U];
  TV tc[U], th[U], tv[U], ta[U];
  double hh[U];

  __device__ __forceinline__ void load(const OpView &A, const TV *__restrict__ x,
                                       const TV *__restrict__ aux, const int (&b)[U], int c) {
#pragma unroll
    for (int u = 0; u < U; u++) {
      nb[u] = A.nbr[b[u]];
      m[u] = A.mask[b[u]];
    }
    load_data(A, x, aux, b, c);
  }
  // the same with nb[] and m[] already set
  __device__ __forceinline__ void load_data(const OpView &A, const TV *__restrict__ x,
                                            const TV *__restrict__ aux, const int (&b)[U], int c) {
#pragma unroll
    for (int u = 0; u < U; u++) {
      bk[u] = b[u];
      stencil_load<TV>(x, b[u], nb[u], c, tc[u], th[u], tv[u]);
      ta[u] = aux[(long)b[u] * 64 + c];
      hh[u] = A.h2[b[u]];
    }
  }
  __device__ __forceinline__ long i(int u, int c) const { return (long)bk[u] * 64 + c; }
  __device__ __forceinline__ double xc(int u) const { return (double)tc[u]; }
  __device__ __forceinline__ double a(int u) const { return (double)ta[u]; }
  // stencil value of block u; `plain` is false on explicit rows
  __device__ __forceinline__ double y(int u, int c, bool &plain) const {
    plain = !((m[u] >> c) & 1ull);
    return stencil_eval<TV>(nb[u], c, tc[u], th[u], tv[u]);
  }
};

// ---------------------------------------------------------------------------
// FP64 MFMA block preconditioner: Z(64 x 16 blocks) = P(64x64) R(64 x 16)
// with v_mfma_f64_16x16x4f64 (A 16x4, B 4x16, D 16x16: lane l holds
// A[l & 15][l >> 4], B[l >> 4][l & 15] and, in register v,
// D[(l >> 4) + 4 v][l & 15] -- checked on the device).
// A workgroup takes a tile of TILE = 16 blocks; wave w computes output rows
// 16 w .. 16 w + 15 of every block of the tile in 16 K steps.  K step ks
// covers input cells 16 q + ks (q = 0..3 is the lane's k slot), so lane
// (j = l & 15, q = l >> 4) supplies cells 16 q .. 16 q + 15 of block j.  The
// A rows are ordered o(m) = 4 (m & 3) + (m >> 2), so that D row q + 4 v is
// output cell 16 w + 4 q + v:
//   A operand of lane l, wave w, step ks: P[16 w + o(l & 15)][16 (l >> 4) + ks]
//   D register v of lane l, wave w:       Z[16 w + 4 (l >> 4) + v][block l & 15]
// The cells a thread updates, loads and stores are its output cells
// 16 w + 4 q + v (v = 0..3); the inputs of the product go through an LDS
// tile (tile_put / tile_get) between the update and the product.
// ---------------------------------------------------------------------------
typedef double f64x4 __attribute__((ext_vector_type(4)));
constexpr int TROW = 65;                 // LDS tile row stride (doubles): conflict-free reads

struct PFrag64 {
  double a[16];
};

// the wave's 16 A-operand fragments (P rows 16 w .. 16 w + 15)
__device__ __forceinline__ void load_pfrag64(PFrag64 &F, const double *__restrict__ pf, int w,
                                             int lane) {
#pragma unroll
  for (int ks = 0; ks < 16; ks++) F.a[ks] = pf[(w * 16 + ks) * 64 + lane];
}

// first of the thread's four output cells (relative to its block)
__device__ __forceinline__ int cell4(int w, int lane) { return 16 * w + 4 * (lane >> 4); }

__device__ __forceinline__ void ld4(const double *__restrict__ a, long i, double (&v)[4]) {
  const double2 *p2 = reinterpret_cast<const double2 *>(a + i);
  const double2 u0 = p2[0], u1 = p2[1];
  v[0] = u0.x; v[1] = u0.y; v[2] = u1.x; v[3] = u1.y;
}
__device__ __forceinline__ void st4(double *__restrict__ a, long i, const double (&v)[4]) {
  double2 *p2 = reinterpret_cast<double2 *>(a + i);
  p2[0] = make_double2(v[0], v[1]);
  p2[1] = make_double2(v[2], v[3]);
}

// the thread's four values (cells cell4 + v of block j of the tile) into the
// LDS tile; zeros for a lane past the tile's last block
__device__ __forceinline__ void tile_put(double (*tile)[TROW], int w, int lane, bool ok,
                                         const double (&v)[4]) {
  const int j = lane & 15, c0 = cell4(w, lane);
#pragma unroll
  for (int k = 0; k < 4; k++) tile[j][c0 + k] = ok ? v[k] : 0.0;
}

// z (the thread's four output cells) = rows 16 w + 4 q + v of P times block
// j of the tile, after a workgroup barrier that published the tile
__device__ __forceinline__ f64x4 tile_prec64(const double (*tile)[TROW], const PFrag64 &F,
                                             int lane) {
  const int j = lane & 15, q = lane >> 4;
  double in[16];
#pragma unroll
  for (int ks = 0; ks < 16; ks++) in[ks] = tile[j][16 * q + ks];
  f64x4 acc = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
  for (int ks = 0; ks < 16; ks++) acc = __builtin_amdgcn_mfma_f64_16x16x4f64(F.a[ks], in[ks], acc, 0, 0, 0);
  return acc;
}
__device__ __forceinline__ void st4(double *__restrict__ a, long i, const f64x4 &v) {
  double2 *p2 = reinterpret_cast<double2 *>(a + i);
  p2[0] = make_double2(v[0], v[1]);
  p2[1] = make_double2(v[2], v[3]);
}

// ---------------------------------------------------------------------------
// BF16 MFMA block preconditioner: Z(64 x 16 blocks) = P(64x64) R(64 x 16)
// with v_mfma_f32_16x16x32_bf16.  A tile of TILE=16 blocks per wavefront:
// lane l works on block tb + (l & 15), quarter q = l >> 4.
//   input cells of the lane  : 32 ks + 8 q + j      (ks = 0,1; j = 0..7)
//   output cells of the lane : 16 mt + 4 q + v      (mt = 0..3; v = 0..3)
// A operand (P) of lane l for (mt, ks): P[16 mt + (l & 15)][32 ks + 8 q + j];
// B operand for ks: R[32 ks + 8 q + j][block l & 15];
// D of lane l, register v: Z[16 mt + 4 q + v][block l & 15].
// The FP32 input is split into three bf16 terms (hi + mid + lo, three MFMAs
// per step), which represent it to ~2^-24: a preconditioner whose input is
// rounded coarser than that is no longer close enough to linear for
// BiCGSTAB on the singular Neumann system.  P itself is the bf16 rounding of
// P_inv, a fixed linear preconditioner.
// ---------------------------------------------------------------------------
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

struct PFrag {
  bf16x8 a[4][2];
};

__device__ __forceinline__ void load_pfrag(PFrag &P, const bf16x8 *__restrict__ pf, int lane) {
#pragma unroll
  for (int mt = 0; mt < 4; mt++)
#pragma unroll
    for (int ks = 0; ks < 2; ks++) P.a[mt][ks] = pf[(mt * 2 + ks) * 64 + lane];
}

// in[ks*8 + j] = value at input cell 32 ks + 8 q + j; out[mt*4 + v] = value at
// output cell 16 mt + 4 q + v.
__device__ __forceinline__ void prec_mfma(const PFrag &P, const float (&in)[16], float (&out)[16]) {
  bf16x8 hi[2], md[2], lo[2];
#pragma unroll
  for (int ks = 0; ks < 2; ks++)
#pragma unroll
    for (int j = 0; j < 8; j++) {
      const float v = in[ks * 8 + j];
      const __bf16 h = (__bf16)v;
      const float r1 = v - (float)h;
      const __bf16 m = (__bf16)r1;
      hi[ks][j] = h;
      md[ks][j] = m;
      lo[ks][j] = (__bf16)(r1 - (float)m);
    }
#pragma unroll
  for (int mt = 0; mt < 4; mt++) {
    f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int ks = 0; ks < 2; ks++) {
      acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(P.a[mt][ks], lo[ks], acc, 0, 0, 0);
      acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(P.a[mt][ks], md[ks], acc, 0, 0, 0);
      acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(P.a[mt][ks], hi[ks], acc, 0, 0, 0);
    }
#pragma unroll
    for (int v = 0; v < 4; v++) out[mt * 4 + v] = acc[v];
  }
}

__device__ __forceinline__ int in_cell(int q, int k) { return 32 * (k >> 3) + 8 * q + (k & 7); }
__device__ __forceinline__ int out_cell(int q, int k) { return 16 * (k >> 2) + 4 * q + (k & 3); }

constexpr int NQ5 = 6;  // totals of the x/r update: |r|_inf, |r|^2, (rhat,r), |rhat|^2, sum r, sum rhat

// State at the top of iteration i from the state of iteration i-1 and the
// totals o of iteration i-1's update (BiCGSTAB.cu:471-509 and :572-600).
// prev.iter < 0 marks the state after r0 = b - A x0 (BiCGSTAB.cu:427-465);
// `replaced` marks the state after a residual replacement (r = b - A x of
// the best iterate, recomputed in FP64, rhat = r, p = nu = 0): BiCGSTAB
// restarts from it, with the iteration and restart counts kept.
// With `project` (FP32-stored residuals) the component of r along the left
// null vector 1 of the Neumann operator, mu = sum r / n, is removed: it is
// rounding noise no iterate can cancel, and left in place it stalls the
// iteration at ~1e-7 and then makes it diverge.  The totals are corrected
// for it here and the kernels subtract mu from r.  Deterministic:
// recomputing it from the same inputs gives the same state.
constexpr int STALL_BASE = 150;          // mixed precision: stall window, iterations without a 2x gain
constexpr int STALL_MAX = 150 << 5;      // after repeated unproductive replacements

__device__ inline Scal evaluate(const Scal &prev, const double (&o)[NQ5], bool replaced = false) {
  Scal s = prev;
  double rr = o[1], rhr = o[2], hh = o[3], amax = o[0];
  s.mu = 0.0;
  if (prev.project) {
    const double mu = o[4] / prev.n;
    s.mu = mu;
    rr = fmax(rr - prev.n * mu * mu, 0.0);
    rhr -= mu * o[5];
    amax += fabs(mu);  // bound of max |r_i - mu|
  }
  const double err = prev.norm ? sqrt(rr) : amax;
  const bool rep_mode = prev.rep_factor > 0.0;
  if (prev.iter < 0 || replaced) {
    if (prev.iter < 0) {
      s.err0 = err;
      s.iter = 0;
      s.restarts = s.n_rep = 0;
      s.converged = 0;
    } else {
      s.n_rep = prev.n_rep + 1;
      // the pass since the last replacement gained less than 2x in FP64:
      // it was cut short on a plateau BiCGSTAB would have left (hard
      // meshes need long runs between breakdowns), so wait twice as long
      // before the next stall replacement
      if (prev.stall > 0 && err > 0.5 * prev.err_rep) s.stall = min(2 * prev.stall, STALL_MAX);
    }
    s.err = s.err_opt = s.err_ref = s.err_rep = err;
    s.iter_ref = s.iter_rep = s.iter;
    s.replace = 0;
    s.alpha = s.omega = s.rho_prev = 1.0;
    s.rho = rhr;
    s.beta = (s.rho / (s.rho_prev + s.eps)) * (s.alpha / (s.omega + s.eps));
    s.improved = replaced ? 1 : 0;  // x is the selected best iterate
    s.restart = 0;
    s.done = (err == 0.0 || s.iter >= s.max_iter) ? 1 : 0;
    if (replaced && (err <= s.tol_abs || err / s.err0 <= s.tol_rel)) s.converged = s.done = 1;
    s.idle = s.done;
    return s;
  }
  if (prev.done) {
    s.restart = 0;
    s.mu = 0.0;
    return s;
  }
  s.err = err;
  s.iter = prev.iter + 1;
  s.improved = 0;
  if (err < prev.err_opt) {
    s.improved = 1;
    s.err_opt = err;
    if (err <= s.tol_abs || err / s.err0 <= s.tol_rel) {
      if (rep_mode) {
        s.replace = 1;  // the recurrence says converged: confirm on the FP64 residual
      } else {
        s.converged = 1;
        s.done = 1;
      }
    }
  }
  if (!s.done && s.iter >= s.max_iter) s.done = 1;
  if (rep_mode && err < prev.rep_factor * s.err_rep) s.replace = 1;
  // stagnation: mixed precision replaces the residual (the FP32 recurrence
  // reached its floor), otherwise the solve ends
  if (err < 0.5 * s.err_ref) {
    s.err_ref = err;
    s.iter_ref = s.iter;
  }
  if (!s.done && s.stall > 0 && s.iter - s.iter_ref > s.stall) {
    if (rep_mode) s.replace = 1;
    else s.done = 1;
  }
  s.rho_prev = prev.rho;
  s.rho = rhr;
  s.beta = (s.rho / (s.rho_prev + s.eps)) * (s.alpha / (s.omega + s.eps));
  // breakdown test of BiCGSTAB.cu:483; a beta that would overflow FP32
  // storage (omega ~ 0) or is not finite is handled the same way
  const bool breakdown = rhr * rhr < s.bd_tol * rr * hh || !(fabs(s.beta) < 1e30);
  s.restart = 0;
  if (!s.done && !s.replace && breakdown && s.max_restarts > 0) {
    s.restarts += 1;
    if (s.restarts >= s.max_restarts) {
      s.done = 1;
    } else {
      s.restart = 1;
      s.rho = rr;  // rhat = r
      s.rho_prev = s.alpha = s.omega = 1.0;
      s.beta = (s.rho / (s.rho_prev + s.eps)) * (s.alpha / (s.omega + s.eps));
    }
  }
  if (s.done) s.replace = 0;
  s.idle = s.done || s.replace;  // a pending replacement pauses the iteration
  s.red[5] = rr;
  s.red[6] = rhr;
  s.red[7] = hh;
  return s;
}




// 16 values of the lane's input cells (32 ks + 8 q + j) of one block
__device__ __forceinline__ void ld16(const float *__restrict__ a, long base, int q, float (&v)[16]) {
#pragma unroll
  for (int ks = 0; ks < 2; ks++) {
    const float4 *p4 = reinterpret_cast<const float4 *>(a + base + 32 * ks + 8 * q);
    const float4 u0 = p4[0], u1 = p4[1];
    v[ks * 8 + 0] = u0.x; v[ks * 8 + 1] = u0.y; v[ks * 8 + 2] = u0.z; v[ks * 8 + 3] = u0.w;
    v[ks * 8 + 4] = u1.x; v[ks * 8 + 5] = u1.y; v[ks * 8 + 6] = u1.z; v[ks * 8 + 7] = u1.w;
  }
}
__device__ __forceinline__ void st16_in(float *__restrict__ a, long base, int q, const float (&v)[16]) {
#pragma unroll
  for (int ks = 0; ks < 2; ks++) {
    float4 *p4 = reinterpret_cast<float4 *>(a + base + 32 * ks + 8 * q);
    p4[0] = make_float4(v[ks * 8 + 0], v[ks * 8 + 1], v[ks * 8 + 2], v[ks * 8 + 3]);
    p4[1] = make_float4(v[ks * 8 + 4], v[ks * 8 + 5], v[ks * 8 + 6], v[ks * 8 + 7]);
  }
}
// 16 values of the lane's output cells (16 mt + 4 q + v)
__device__ __forceinline__ void st16_out(float *__restrict__ a, long base, int q, const float (&v)[16]) {
#pragma unroll
  for (int mt = 0; mt < 4; mt++)
    *reinterpret_cast<float4 *>(a + base + 16 * mt + 4 * q) =
        make_float4(v[mt * 4 + 0], v[mt * 4 + 1], v[mt * 4 + 2], v[mt * 4 + 3]);
}

// Top-of-iteration update of the lane's 16 input cells of one block, FP32
// storage with FP64 arithmetic, in two steps so that the loads can be
// issued early: mf_load reads r, p and nu (p and nu unused after a
// restart); mf_pupdate then applies r -= mu (written back when mu != 0) and
// p = r + beta (p - omega nu), or rhat = p = r after a restart.  pin receives
// p; a lane with ok false (past the last block; base is any valid block)
// stores nothing and gets pin = 0.
struct MfIn {
  float r[16], p[16], nu[16];
};
__device__ __forceinline__ void mf_load(const float *__restrict__ r, const float *__restrict__ p,
                                        const float *__restrict__ nu, long base, int q, MfIn &in) {
  ld16(r, base, q, in.r);
  ld16(p, base, q, in.p);
  ld16(nu, base, q, in.nu);
}
__device__ __forceinline__ void mf_pupdate(MfIn &in, float *__restrict__ r,
                                           float *__restrict__ rhat, float *__restrict__ p,
                                           long base, int q, bool ok, double beta, double omega,
                                           double mu, int restart, float (&pin)[16]) {
  if (mu != 0.0)
#pragma unroll
    for (int k = 0; k < 16; k++) in.r[k] = (float)((double)in.r[k] - mu);
  if (restart) {
#pragma unroll
    for (int k = 0; k < 16; k++) pin[k] = in.r[k];
  } else {
#pragma unroll
    for (int k = 0; k < 16; k++)
      pin[k] = (float)((double)in.r[k] + beta * ((double)in.p[k] - omega * (double)in.nu[k]));
  }
  if (ok) {
    if (mu != 0.0) st16_in(r, base, q, in.r);
    if (restart) st16_in(rhat, base, q, in.r);
    st16_in(p, base, q, pin);
  } else {
#pragma unroll
    for (int k = 0; k < 16; k++) pin[k] = 0.f;
  }
}

// ----------------------------------------------------------------------------
// operators either side of the pressure solve (one thread per cell), shared
// by the single and the batched solver; defined per translation unit
// ----------------------------------------------------------------------------
namespace {

// Values of a cell's four neighbours (E, W, N, S) across plain faces:
// inside the block, in the same-level neighbour block, or the ghost of a
// domain face -- `wall_sign` times the cell's own value (the reference's
// labs mirror the boundary cell: -1 for the face-normal velocity component,
// Definitions.h:84-117; +1 for the zero-gradient pressure).  Cells on a
// coarse-fine face are computed by the *_cf kernels below instead.
struct Nb4 {
  double e, w, n, s;
};
__device__ __forceinline__ Nb4 neighbours(const double *__restrict__ f, long i, int4 nb, double xs,
                                          double ys) {
  const int c = (int)(i & 63), ix = c & 7, iy = c >> 3;
  const double self = f[i];
  Nb4 r;
  r.e = ix < 7 ? f[i + 1] : nb.y >= 0 ? f[(long)nb.y * 64 + iy * 8] : xs * self;
  r.w = ix > 0 ? f[i - 1] : nb.x >= 0 ? f[(long)nb.x * 64 + iy * 8 + 7] : xs * self;
  r.n = iy < 7 ? f[i + 8] : nb.w >= 0 ? f[(long)nb.w * 64 + ix] : ys * self;
  r.s = iy > 0 ? f[i - 8] : nb.z >= 0 ? f[(long)nb.z * 64 + 56 + ix] : ys * self;
  return r;
}

// rhs = h/(2 dt) [(u_E - u_W) + (v_N - v_S)] - h/(2 dt) chi [(ud_E - ud_W) + (vd_N - vd_S)]
// (updatePressureRHS, PressureSingle.cpp:336-349), in the reference's order of
// operations (no contraction).  Cells on a coarse-fine face (set in mask)
// are left to k_pressure_rhs_cf.
__global__ __launch_bounds__(WG) void k_pressure_rhs(int nb, const int4 *__restrict__ nbr,
                                                     const unsigned long long *__restrict__ mask,
                                                     const double *__restrict__ hb,
                                                     const double *__restrict__ u,
                                                     const double *__restrict__ v,
                                                     const double *__restrict__ ud,
                                                     const double *__restrict__ vd,
                                                     const double *__restrict__ chi, double dt,
                                                     double *__restrict__ rhs) {
#pragma clang fp contract(off)
  const long n = (long)nb * 64;
  for (long i = (long)blockIdx.x * WG + threadIdx.x; i < n; i += (long)gridDim.x * WG) {
    if ((mask[i >> 6] >> (i & 63)) & 1ull) continue;
    const int4 f = nbr[i >> 6];
    const double facDiv = 0.5 * hb[i >> 6] / dt;
    const Nb4 U = neighbours(u, i, f, -1.0, 1.0), V = neighbours(v, i, f, 1.0, -1.0);
    const Nb4 UD = neighbours(ud, i, f, -1.0, 1.0), VD = neighbours(vd, i, f, 1.0, -1.0);
    double t = facDiv * ((U.e - U.w) + (V.n - V.s));
    t += -facDiv * chi[i] * ((UD.e - UD.w) + (VD.n - VD.s));
    rhs[i] = t;
  }
}

// u += (-dt/2 h (p_E - p_W)) / h^2, v likewise (pressureCorrectionKernel and
// PressureSingle::pressureCorrection, PressureSingle.cpp:122-210), zero-gradient
// pressure ghosts, in the reference's order of operations (no contraction).
// Cells on a coarse-fine face are left to k_pressure_correct_cf.
__global__ __launch_bounds__(WG) void k_pressure_correct(int nb, const int4 *__restrict__ nbr,
                                                         const unsigned long long *__restrict__ mask,
                                                         const double *__restrict__ hb,
                                                         const double *__restrict__ p, double dt,
                                                         double *__restrict__ u,
                                                         double *__restrict__ v) {
#pragma clang fp contract(off)
  const long n = (long)nb * 64;
  for (long i = (long)blockIdx.x * WG + threadIdx.x; i < n; i += (long)gridDim.x * WG) {
    if ((mask[i >> 6] >> (i & 63)) & 1ull) continue;
    const double h = hb[i >> 6];
    const double pFac = -0.5 * dt * h, ih2 = 1.0 / h / h;
    const Nb4 P = neighbours(p, i, nbr[i >> 6], 1.0, 1.0);
    const double tu = pFac * (P.e - P.w), tv = pFac * (P.n - P.s);
    u[i] += tu * ih2;
    v[i] += tv * ih2;
  }
}

// ----------------------------------------------------------------------------
// Cells on a coarse-fine face, in face form: the operator is the sum over the
// faces Xm, Xp, Ym, Yp of s_f flux_f (s = -1 low, +1 high), flux_f being the
// face value the reference's kernels store for the flux correction
// (PressureSingle.cpp:350-397, :139-185).  Across a coarser face the
// neighbour value is the explicit operator's coarse-fine ghost
// (ExpAMRSolver.cpp:100-128, ExpAMRSolver.h:268-300); across a finer face
// flux_f is the sum of the two fine cells' fluxes (the flux correction:
// coarse and fine fluxes equal and opposite).  fid holds the coarse-fine
// neighbours of each face: (coarse block, -1 - half) or the two fine blocks.
// ----------------------------------------------------------------------------

__device__ __forceinline__ long gcell(long b, int ix, int iy) { return b * 64 + iy * 8 + ix; }

// grid of the *_cf kernels: a thread per coarse-fine cell, up to 1024 workgroups
inline int exc_grid(int n_exc) { return n_exc > 0 ? (n_exc + WG - 1) / WG < 1024 ? (n_exc + WG - 1) / WG : 1024 : 1; }

// 2/3 x_close - 1/5 x_far + 8/15 (x_c + sT D1 + D2), D1/D2 along the face
// around coarse cell c (tangential index t in its block; stride `st`
// between coarse cells along the face), one-sided at t = 3, 7 (backward) and
// 0, 4 (forward)
__device__ __forceinline__ double cf_ghost(const double *__restrict__ x, long close, long far_,
                                           long c, int t, int st, double sT) {
#pragma clang fp contract(off)
  const double xc = x[c];
  double d1, d2;
  if (t == 3 || t == 7) {
    const double a = x[c - 2 * st], bb = x[c - st];
    d1 = ((1.0 / 8.0) * a + (-1.0 / 2.0) * bb) + (3.0 / 8.0) * xc;
    d2 = ((1.0 / 32.0) * a + (-1.0 / 16.0) * bb) + (1.0 / 32.0) * xc;
  } else if (t == 0 || t == 4) {
    const double a = x[c + 2 * st], bb = x[c + st];
    d1 = ((-1.0 / 8.0) * a + (1.0 / 2.0) * bb) + (-3.0 / 8.0) * xc;
    d2 = ((1.0 / 32.0) * a + (-1.0 / 16.0) * bb) + (1.0 / 32.0) * xc;
  } else {
    const double a = x[c - st], bb = x[c + st];
    d1 = (-1.0 / 8.0) * a + (1.0 / 8.0) * bb;
    d2 = ((1.0 / 32.0) * a + (1.0 / 32.0) * bb) + (-1.0 / 16.0) * xc;
  }
  return ((2.0 / 3.0) * x[close] + (-1.0 / 5.0) * x[far_]) + (8.0 / 15.0) * ((xc + sT * d1) + d2);
}

// Where the ghost of fine cell (ix, iy) of block b across face f (coarser
// neighbour cb, `half` of its face) comes from.
struct CfSrc {
  long close, far_, c;
  int t, st;
  double sT;
};
__device__ __forceinline__ CfSrc fine_side(long b, int ix, int iy, int f, long cb, int half) {
  CfSrc r;
  const bool xf = f < 2;
  r.close = gcell(b, ix, iy);
  r.sT = ((xf ? iy : ix) % 2 == 0) ? -1.0 : 1.0;
  if (xf) {
    const int ixc = f == 0 ? 7 : 0, iyc = iy / 2 + 4 * half;
    r.far_ = gcell(b, f == 0 ? ix + 1 : ix - 1, iy);
    r.c = gcell(cb, ixc, iyc);
    r.t = iyc;
    r.st = 8;
  } else {
    const int iyc = f == 2 ? 7 : 0, ixc = ix / 2 + 4 * half;
    r.far_ = gcell(b, ix, f == 2 ? iy + 1 : iy - 1);
    r.c = gcell(cb, ixc, iyc);
    r.t = ixc;
    r.st = 1;
  }
  return r;
}
// ... and of fine cell hh (0: low, 1: high) behind face f of coarse cell
// (ix, iy) of block b, whose finer neighbours are f0, f1.
__device__ __forceinline__ CfSrc coarse_side(long b, int ix, int iy, int f, long f0, long f1, int hh) {
  CfSrc r;
  const bool xf = f < 2;
  const int tc = xf ? iy : ix;
  const long fb = tc >= 4 ? f1 : f0;
  const int t = (tc % 4) * 2 + hh;
  switch (f) {
    case 0: r.close = gcell(fb, 7, t); r.far_ = gcell(fb, 6, t); break;
    case 1: r.close = gcell(fb, 0, t); r.far_ = gcell(fb, 1, t); break;
    case 2: r.close = gcell(fb, t, 7); r.far_ = gcell(fb, t, 6); break;
    default: r.close = gcell(fb, t, 0); r.far_ = gcell(fb, t, 1); break;
  }
  r.c = gcell(b, ix, iy);
  r.t = tc;
  r.st = xf ? 8 : 1;
  r.sT = hh == 0 ? -1.0 : 1.0;
  return r;
}

// Face values of one cell: coef (nb + self) - (coef chi_self) (dnb + dself)
// (with_d) or coef (nb + self); coef(h) of the cell that owns the face.
struct CfField {
  const double *c, *d;  // the component the face reads, its deformation part (or null)
  double wall;          // the wall ghost is wall * self
};
template <class Coef>
__device__ __forceinline__ double cf_flux(const CfField &F, const double *__restrict__ chi,
                                          const double *__restrict__ hb, long i, int ix, int iy,
                                          int f, int4 nb4, int2 fd, Coef coef) {
#pragma clang fp contract(off)
  const bool with_d = F.d != nullptr;
  auto one = [&](double self, double nbv, double dself, double dnb, double cf, long k) {
    double t = cf * (nbv + self);
    if (with_d) t = t - (cf * chi[k]) * (dnb + dself);
    return t;
  };
  const long b = i >> 6;
  const double cself = coef(hb[b]);
  const bool edge = f == 0 ? ix == 0 : f == 1 ? ix == 7 : f == 2 ? iy == 0 : iy == 7;
  const double ds = with_d ? F.d[i] : 0.0;
  if (!edge) {
    const long j = f == 0 ? i - 1 : f == 1 ? i + 1 : f == 2 ? i - 8 : i + 8;
    return one(F.c[i], F.c[j], ds, with_d ? F.d[j] : 0.0, cself, i);
  }
  const int nv = f == 0 ? nb4.x : f == 1 ? nb4.y : f == 2 ? nb4.z : nb4.w;
  if (nv == -1) return one(F.c[i], F.wall * F.c[i], ds, F.wall * ds, cself, i);
  if (nv >= 0) {
    const long j = f == 0 ? gcell(nv, 7, iy) : f == 1 ? gcell(nv, 0, iy) : f == 2 ? gcell(nv, ix, 7)
                                                                        : gcell(nv, ix, 0);
    return one(F.c[i], F.c[j], ds, with_d ? F.d[j] : 0.0, cself, i);
  }
  if (fd.y < 0) {  // coarser neighbour
    const CfSrc g = fine_side(b, ix, iy, f, fd.x, -1 - fd.y);
    const double gv = cf_ghost(F.c, g.close, g.far_, g.c, g.t, g.st, g.sT);
    const double gd = with_d ? cf_ghost(F.d, g.close, g.far_, g.c, g.t, g.st, g.sT) : 0.0;
    return one(F.c[i], gv, ds, gd, cself, i);
  }
  double tot = 0.0;  // finer neighbours: the two fine cells' fluxes
#pragma unroll
  for (int hh = 0; hh < 2; hh++) {
    const CfSrc g = coarse_side(b, ix, iy, f, fd.x, fd.y, hh);
    const double gv = cf_ghost(F.c, g.close, g.far_, g.c, g.t, g.st, g.sT);
    const double gd = with_d ? cf_ghost(F.d, g.close, g.far_, g.c, g.t, g.st, g.sT) : 0.0;
    const double fl = one(F.c[g.close], gv, with_d ? F.d[g.close] : 0.0, gd, coef(hb[g.close >> 6]),
                          g.close);
    tot = hh == 0 ? fl : tot + fl;
  }
  return tot;
}

// The RHS at the cells on a coarse-fine face (explicit rows erow[0..n_exc)),
// one thread per cell.
__global__ __launch_bounds__(WG) void k_pressure_rhs_cf(int n_exc, const int *__restrict__ erow,
                                                        const int4 *__restrict__ nbr,
                                                        const int2 *__restrict__ fid,
                                                        const double *__restrict__ hb,
                                                        const double *__restrict__ u,
                                                        const double *__restrict__ v,
                                                        const double *__restrict__ ud,
                                                        const double *__restrict__ vd,
                                                        const double *__restrict__ chi, double dt,
                                                        double *__restrict__ rhs) {
#pragma clang fp contract(off)
  for (int e = blockIdx.x * WG + threadIdx.x; e < n_exc; e += gridDim.x * WG) {
    const long i = erow[e], b = i >> 6;
    const int ix = (int)(i & 7), iy = (int)((i >> 3) & 7);
    const int4 nb4 = nbr[b];
    auto facDiv = [dt](double h) { return 0.5 * h / dt; };
    const CfField FU{u, ud, -1.0}, FV{v, vd, -1.0};
    double acc = 0.0;
#pragma unroll 1
    for (int f = 0; f < 4; f++) {
      const double fl = cf_flux(f < 2 ? FU : FV, chi, hb, i, ix, iy, f, nb4, fid[4 * b + f], facDiv);
      const double term = (f == 0 || f == 2) ? -fl : fl;
      acc = f == 0 ? term : acc + term;
    }
    rhs[i] = acc;
  }
}

// The correction at the cells on a coarse-fine face, one thread per cell.
__global__ __launch_bounds__(WG) void k_pressure_correct_cf(int n_exc, const int *__restrict__ erow,
                                                            const int4 *__restrict__ nbr,
                                                            const int2 *__restrict__ fid,
                                                            const double *__restrict__ hb,
                                                            const double *__restrict__ p, double dt,
                                                            double *__restrict__ u,
                                                            double *__restrict__ v) {
#pragma clang fp contract(off)
  for (int e = blockIdx.x * WG + threadIdx.x; e < n_exc; e += gridDim.x * WG) {
    const long i = erow[e], b = i >> 6;
    const int ix = (int)(i & 7), iy = (int)((i >> 3) & 7);
    const int4 nb4 = nbr[b];
    auto pFac = [dt](double h) { return -0.5 * dt * h; };
    const CfField FP{p, nullptr, 1.0};
    double t[4];
#pragma unroll 1
    for (int f = 0; f < 4; f++) {
      const double fl = cf_flux(FP, nullptr, hb, i, ix, iy, f, nb4, fid[4 * b + f], pFac);
      t[f] = (f == 0 || f == 2) ? -fl : fl;
    }
    const double h = hb[b], ih2 = 1.0 / h / h;
    u[i] = u[i] + (t[0] + t[1]) * ih2;
    v[i] = v[i] + (t[2] + t[3]) * ih2;
  }
}

}  // namespace

}  // namespace cup2d
