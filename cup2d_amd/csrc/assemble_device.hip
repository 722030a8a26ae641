// assemble_device.hip -- the AMR Poisson operator assembled on the GPU, for
// one mesh or for a batch of independent meshes at once.
//
// Restates the host assembly of assembly.cpp (reference
// source/Poisson/ExpAMRSolver.cpp:100-313) as a short pipeline of kernels
// over the concatenated blocks of all meshes (global block ids, so columns
// come out shifted by each mesh's offset):
//   1. ka_insert / ka_dups: the blocks go into open-addressing hash tables
//      keyed by (level, i, j), one region per mesh -- the lookup that
//      getBlockInfoAll(level, Z) gives the reference -- with the host's
//      range and duplicate checks;
//   2. ka_faces: one thread per block classifies its four faces (same level,
//      wall, one level coarser, one level finer; ExpAMRSolver.cpp:221-249,
//      :141-177) and writes the neighbour table, the explicit-cell mask, h^2
//      and its count of plain-row nonzeros;
//   3. an exclusive scan of the explicit rows per block gives each row its
//      place in mesh order (block, then cell);
//   4. ka_rows: one thread per explicit row (its block by binary search in
//      the scan, its cell as the k-th set bit of the block's mask) builds the
//      row in registers (ghost interpolation ExpAMRSolver.cpp:100-128, D1/D2
//      ExpAMRSolver.h:268-300) with the host's sequence of additions and
//      stores it sorted by column;
//   5. a stable radix sort orders the rows by (mesh, length); a sort + unique
//      of the coefficients' bit patterns gives the coefficient table;
//   6. ka_ell writes the column-major ELL arrays, each coefficient's code (a
//      binary search in the table) and the width of every run of 64 rows.
// Floating-point contraction is off in this file: every coefficient is the
// host's sequence of IEEE products, quotients and sums, so the values are
// bit-identical.

#include "assemble_device.h"

#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <climits>
#include <type_traits>
#include <vector>

#include "../../include/cup2d_poisson.h"
#include "kernels.h"

#pragma clang fp contract(off)

namespace cup2d {
namespace {

constexpr int AWG = 256;                    // threads per workgroup
constexpr unsigned long long EMPTY = ~0ull;  // free hash slot
constexpr int RCAP = ELL_MAX;               // entries a row can hold; one more is an error

// One mesh of the batch.
struct Inst {
  long long boff;           // its first block
  unsigned long long toff;  // its first hash slot
  unsigned long long mask;  // its hash slots - 1 (a power of two minus one)
  double h0;                // cell size at level 0
  int levels, bpdx, bpdy, px, py;
};

struct Blocks {
  const int *lvl, *bx, *by, *inst;  // [nb]: level, index, mesh
  const Inst *I;                    // [K]
  long long nb;
};

struct Hash {
  unsigned long long *keys;  // pos_key, EMPTY when free
  int *vals;                 // smallest block with that key
};

// Failure reports and counters of one mesh, reduced with integer atomics
// (the result does not depend on the order of the updates).  Block numbers
// are local to the mesh.
struct Errs {
  int blk[3];  // smallest block with a level / index out of range, or a duplicate
  int face;    // smallest 4 b + f whose neighbour is neither same, coarser nor finer
  int corner;  // last block with index (0, 0), -1 when none
  int pad;
  unsigned long long nnz;
};

// Over all meshes.
struct Glob {
  int wide;    // smallest explicit row wider than RCAP, or INT_MAX
  int maxlen;  // widest explicit row
};

__device__ __forceinline__ unsigned long long pos_key(int l, long long i, long long j) {
  return ((unsigned long long)l << 58) | ((unsigned long long)i << 29) | (unsigned long long)j;
}

// murmur3's 64-bit finaliser
__device__ __forceinline__ unsigned long long mix64(unsigned long long z) {
  z ^= z >> 33;
  z *= 0xff51afd7ed558ccdull;
  z ^= z >> 33;
  z *= 0xc4ceb9fe1a85ec53ull;
  z ^= z >> 33;
  return z;
}

__device__ long long lookup(const Hash &H, const Inst &in, int l, long long i, long long j) {
  if (l < 0 || l >= in.levels) return -1;
  const unsigned long long key = pos_key(l, i, j);
  unsigned long long s = mix64(key) & in.mask;
  for (unsigned long long p = 0; p <= in.mask; p++, s = (s + 1) & in.mask) {
    const unsigned long long k = H.keys[in.toff + s];
    if (k == key) return H.vals[in.toff + s];
    if (k == EMPTY) return -1;
  }
  return -1;
}

__device__ __forceinline__ long long gidx(long long b, int ix, int iy) { return b * BLEN + iy * BS + ix; }

__global__ __launch_bounds__(AWG) void ka_init(Hash H, unsigned long long slots, Errs *E, int K, Glob *g) {
  const unsigned long long stride = (unsigned long long)gridDim.x * AWG;
  const unsigned long long t = blockIdx.x * (unsigned long long)AWG + threadIdx.x;
  for (unsigned long long s = t; s < slots; s += stride) {
    H.keys[s] = EMPTY;
    H.vals[s] = INT_MAX;
  }
  for (unsigned long long k = t; k < (unsigned long long)K; k += stride) {
    E[k].blk[0] = E[k].blk[1] = E[k].blk[2] = E[k].face = INT_MAX;
    E[k].corner = -1;
    E[k].pad = 0;
    E[k].nnz = 0;
  }
  if (t == 0) {
    g->wide = INT_MAX;
    g->maxlen = 0;
  }
}

// Block status: 0 valid, 1 level out of range, 2 index out of range (the
// order of build_operator's checks), 3 repeats an earlier block.
enum : int { ST_OK = 0, ST_LEVEL = 1, ST_INDEX = 2, ST_DUP = 3 };

// Failures are recorded per block with plain stores and reduced by
// ka_status through one atomic whose address is computed, not chosen by
// branches: error atomics on several divergent paths of one kernel (the
// hash insertion's atomicMin among them) were merged by the compiler into
// one instruction whose address register was overwritten on one of the
// paths.  For the same reason each kernel below has at most one atomic site
// per operation.
__global__ __launch_bounds__(AWG) void ka_insert(Blocks B, Hash H, int *status) {
  const long long b = blockIdx.x * (long long)AWG + threadIdx.x;
  if (b >= B.nb) return;
  const Inst in = B.I[B.inst[b]];
  const int l = B.lvl[b], i = B.bx[b], j = B.by[b];
  const bool lvl_ok = l >= 0 && l < in.levels;
  const int ls = lvl_ok ? l : 0;
  const long long nx = (long long)in.bpdx << ls, ny = (long long)in.bpdy << ls;
  const bool idx_ok = i >= 0 && i < nx && j >= 0 && j < ny;
  const int st = !lvl_ok ? ST_LEVEL : !idx_ok ? ST_INDEX : ST_OK;
  status[b] = st;
  if (st != ST_OK) return;
  const unsigned long long key = pos_key(l, i, j);
  unsigned long long s = mix64(key) & in.mask;
  for (unsigned long long p = 0; p <= in.mask; p++, s = (s + 1) & in.mask) {
    const unsigned long long prev = atomicCAS(&H.keys[in.toff + s], EMPTY, key);
    if (prev == EMPTY || prev == key) {
      atomicMin(&H.vals[in.toff + s], (int)b);  // the first occurrence owns the position
      return;
    }
  }
}

// A block that does not own its position repeats an earlier block; the
// smallest such block is the one the host's in-order insertion rejects.
__global__ __launch_bounds__(AWG) void ka_dups(Blocks B, Hash H, int *status) {
  const long long b = blockIdx.x * (long long)AWG + threadIdx.x;
  if (b >= B.nb || status[b] != ST_OK) return;
  const Inst in = B.I[B.inst[b]];
  if (lookup(H, in, B.lvl[b], B.bx[b], B.by[b]) != b) status[b] = ST_DUP;
}

// Smallest block of each failure kind, per mesh.
__global__ __launch_bounds__(AWG) void ka_status(Blocks B, const int *__restrict__ status, Errs *E) {
  const long long b = blockIdx.x * (long long)AWG + threadIdx.x;
  if (b >= B.nb) return;
  const int st = status[b];
  if (st != ST_OK) {
    const int k = B.inst[b];
    atomicMin(&E[k].blk[st - 1], (int)(b - B.I[k].boff));
  }
}

// Faces of block b: nbr (same-level id, -1 wall, -2 coarse-fine), fid (the
// coarser block, or the two finer blocks low/high along the face), the
// explicit cells, h^2, and the nonzeros of the plain rows.  Blocks of meshes
// with a block failure are skipped (no rows).
__global__ __launch_bounds__(AWG) void ka_faces(Blocks B, Hash H, int4 *nbr, int2 *fid,
                                                unsigned long long *mask, int *cnt, double *h2, Errs *E) {
  const long long b = blockIdx.x * (long long)AWG + threadIdx.x;
  if (b == 0) cnt[B.nb] = 0;
  int k = -1;  // the block's mesh, -1 when there is nothing to count
  long long plain = 0;
  if (b < B.nb) {
    const int kb = B.inst[b];
    const Errs ek = E[kb];
    if (ek.blk[0] != INT_MAX || ek.blk[1] != INT_MAX || ek.blk[2] != INT_MAX) {
      cnt[b] = 0;
    } else {
      k = kb;
      const Inst in = B.I[k];
      const int l = B.lvl[b];
      const long long i = B.bx[b], j = B.by[b];
      const long long nx = (long long)in.bpdx << l, ny = (long long)in.bpdy << l;
      const int di[4] = {-1, 1, 0, 0}, dj[4] = {0, 0, -1, 1};
      int v[4];
      int bad_face = INT_MAX;
#pragma unroll
      for (int f = 0; f < 4; f++) {
        long long I = i + di[f], J = j + dj[f];
        const bool outx = I < 0 || I >= nx, outy = J < 0 || J >= ny;
        int2 fd = make_int2(-1, -1);
        v[f] = -1;
        if (!((outx && !in.px) || (outy && !in.py))) {
          I = (I + nx) % nx;
          J = (J + ny) % ny;
          long long s;
          if ((s = lookup(H, in, l, I, J)) >= 0) {
            v[f] = (int)s;
            fd.x = (int)s;
          } else if ((s = lookup(H, in, l - 1, I >> 1, J >> 1)) >= 0) {
            v[f] = -2;
            fd.x = (int)s;
            fd.y = -1 - (int)((f < 2 ? j : i) & 1);  // the half of the coarse face
          } else {
            long long c[2];
            for (int hh = 0; hh < 2; hh++) {
              const long long ci = f < 2 ? 2 * I + (f == 0 ? 1 : 0) : 2 * I + hh;
              const long long cj = f < 2 ? 2 * J + hh : 2 * J + (f == 2 ? 1 : 0);
              c[hh] = lookup(H, in, l + 1, ci, cj);
            }
            if (c[0] < 0 || c[1] < 0) {
              bad_face = min(bad_face, (int)(4 * (b - in.boff) + f));
            } else {
              v[f] = -2;
              fd = make_int2((int)c[0], (int)c[1]);
            }
          }
        }
        fid[4 * b + f] = fd;
      }
      nbr[b] = make_int4(v[0], v[1], v[2], v[3]);
      const unsigned long long C0 = 0x0101010101010101ull, R0 = 0xffull;
      unsigned long long m = 0;
      if (v[0] == -2) m |= C0;
      if (v[1] == -2) m |= C0 << 7;
      if (v[2] == -2) m |= R0;
      if (v[3] == -2) m |= R0 << 56;
      mask[b] = m;
      cnt[b] = __popcll(m);
      // a plain row: the cell and each face that is inside the block or same-level
      for (int c = 0; c < BLEN; c++) {
        if ((m >> c) & 1) continue;
        const int ix = c & 7, iy = c >> 3;
        plain += 1 + (ix > 0 || v[0] >= 0) + (ix < BS - 1 || v[1] >= 0) + (iy > 0 || v[2] >= 0) +
                 (iy < BS - 1 || v[3] >= 0);
      }
      const double h = in.h0 / double(1LL << l);
      h2[b] = h * h;
      if (bad_face != INT_MAX) atomicMin(&E[k].face, bad_face);
      if (i == 0 && j == 0) atomicMax(&E[k].corner, (int)(b - in.boff));
    }
  }
  // the plain-row count: summed over the wavefront when all its blocks are
  // of one mesh (the usual case: meshes are contiguous), else per lane --
  // one atomic site either way
  const int k0 = __shfl(k, 0);
  const bool uni = __all(k == k0);
  if (uni)
    for (int o = 32; o > 0; o >>= 1) plain += __shfl_xor(plain, o);
  const bool issue = uni ? ((threadIdx.x & 63) == 0 && k0 >= 0) : k >= 0;
  const int kk = uni ? k0 : k;
  if (issue) atomicAdd(&E[kk].nnz, (unsigned long long)plain);
}

// Coefficients of one explicit row, per column in order of first appearance.
// Every slot is addressed with a compile-time index (unrolled loops), so
// the row stays in registers.
struct RowB {
  int c[RCAP];
  double v[RCAP];
  int n = 0;
  bool over = false;
  __device__ __forceinline__ RowB() {
#pragma unroll
    for (int k = 0; k < RCAP; k++) {
      c[k] = -1;
      v[k] = 0.0;
    }
  }
  __device__ __forceinline__ void add(long long col, double val) {
    bool found = false;
#pragma unroll
    for (int k = 0; k < RCAP; k++)
      if (k < n && c[k] == (int)col) {
        v[k] += val;
        found = true;
      }
    if (found) return;
    if (n == RCAP) {
      over = true;
      return;
    }
#pragma unroll
    for (int k = 0; k < RCAP; k++)
      if (k == n) {
        c[k] = (int)col;
        v[k] = val;
      }
    n++;
  }
};

// One coarse-fine ghost: sI (2/3 close - 1/5 far + 8/15 (c + sT D1(c) + D2(c))),
// with one-sided differences at the tangential indices 3, 7 (backward) and
// 0, 4 (forward).  The additions are those of assembly.cpp's ghost(), in the
// same order.
template <class Step>
__device__ __forceinline__ void ghost(RowB &row, long long close, long long far_, long long c, int t, Step step,
                      double sI, double sT) {
  row.add(close, sI * 2. / 3.);
  row.add(far_, -sI * 1. / 5.);
  const double w = sI * 8. / 15.;
  row.add(c, w);
  // The three cases add to the same pattern of cells (two tangential
  // neighbours d0, d1 and c, twice), so one sequence of additions with
  // selected offsets and weights serves all of them; each coefficient is
  // still (sT w) * k or w * k with the host's constant k.
  const bool bwd = (t == BS - 1) || (t == BS / 2 - 1);
  const bool fwd = (t == 0) || (t == BS / 2);
  const int d0 = bwd ? -2 : fwd ? 2 : -1;
  const int d1 = bwd ? -1 : 1;
  const double k0 = bwd ? 1. / 8. : -1. / 8.;
  const double k1 = bwd ? -1. / 2. : fwd ? 1. / 2. : 1. / 8.;
  const double k2 = bwd ? 3. / 8. : fwd ? -3. / 8. : 0.;
  const double k4 = (bwd || fwd) ? -1. / 16. : 1. / 32.;
  const double k5 = (bwd || fwd) ? 1. / 32. : -1. / 16.;
  const double sw = sT * w;
  row.add(step(d0), sw * k0);  // first tangential difference (D1)
  row.add(step(d1), sw * k1);
  row.add(c, sw * k2);
  row.add(step(d0), w * (1. / 32.));  // second tangential difference (D2)
  row.add(step(d1), w * k4);
  row.add(c, w * k5);
}

// Row of cell `lane` of block b, stored at explicit row er.
__device__ __forceinline__ void build_row(const Blocks &B, const int4 *__restrict__ nbr,
                                       const int2 *__restrict__ fid, long long b, int lane, int er,
                                       unsigned *rlen, int *rrow, int *rcol, double *rval, int &wide,
                                       unsigned long long &entries) {
  const int ix = lane & 7, iy = lane >> 3;
  const long long self = gidx(b, ix, iy);
  const int4 n4 = nbr[b];
  RowB row;
  for (int f = 0; f < 4; f++) {
    // selects, not arrays indexed by f: nothing of the row leaves registers
    const int nvf = f == 0 ? n4.x : f == 1 ? n4.y : f == 2 ? n4.z : n4.w;
    const bool edgef = f == 0 ? ix == 0 : f == 1 ? ix == BS - 1 : f == 2 ? iy == 0 : iy == BS - 1;
    if (!edgef) {
      const long long inner = f == 0 ? gidx(b, ix - 1, iy) : f == 1 ? gidx(b, ix + 1, iy)
                            : f == 2 ? gidx(b, ix, iy - 1) : gidx(b, ix, iy + 1);
      row.add(inner, 1);
      row.add(self, -1);
      continue;
    }
    if (nvf == -1) continue;  // wall
    const bool xf = f < 2;
    if (nvf >= 0) {
      const long long s = nvf;
      const long long c = f == 0 ? gidx(s, BS - 1, iy) : f == 1 ? gidx(s, 0, iy)
                        : f == 2 ? gidx(s, ix, BS - 1) : gidx(s, ix, 0);
      row.add(c, 1);
      row.add(self, -1);
      continue;
    }
    const int2 fd = fid[4 * b + f];
    if (fd.y < 0) {  // coarser neighbour
      int ixc, iyc;
      long long inward;
      if (xf) {
        ixc = f == 0 ? BS - 1 : 0;
        iyc = iy / 2 + ((B.by[b] & 1) ? BS / 2 : 0);
        inward = gidx(b, f == 0 ? ix + 1 : ix - 1, iy);
      } else {
        iyc = f == 2 ? BS - 1 : 0;
        ixc = ix / 2 + ((B.bx[b] & 1) ? BS / 2 : 0);
        inward = gidx(b, ix, f == 2 ? iy + 1 : iy - 1);
      }
      const double sT = ((xf ? iy : ix) % 2 == 0) ? -1. : 1.;
      const long long cb = fd.x;
      auto step = [&](int d) { return xf ? gidx(cb, ixc, iyc + d) : gidx(cb, ixc + d, iyc); };
      ghost(row, self, inward, gidx(cb, ixc, iyc), xf ? iyc : ixc, step, 1., sT);
      row.add(self, -1);
    } else {  // two finer neighbours
      const long long fbk = ((xf ? iy : ix) >= BS / 2) ? fd.y : fd.x;
      const int t0 = ((xf ? iy : ix) % (BS / 2)) * 2;
      auto step = [&](int d) { return xf ? gidx(b, ix, iy + d) : gidx(b, ix + d, iy); };
#pragma unroll 1
      for (int hh = 0; hh < 2; hh++) {
        const int t = t0 + hh;
        long long close, far_;
        switch (f) {
          case 0: close = gidx(fbk, BS - 1, t); far_ = gidx(fbk, BS - 2, t); break;
          case 1: close = gidx(fbk, 0, t); far_ = gidx(fbk, 1, t); break;
          case 2: close = gidx(fbk, t, BS - 1); far_ = gidx(fbk, t, BS - 2); break;
          default: close = gidx(fbk, t, 0); far_ = gidx(fbk, t, 1); break;
        }
        row.add(close, 1);
        ghost(row, close, far_, self, xf ? iy : ix, step, -1., hh == 0 ? -1. : 1.);
      }
    }
  }
  rrow[er] = (int)self;
  if (row.over) {
    wide = er;
    rlen[er] = RCAP + 1;
    return;
  }
  // ascending columns: an entry's place is the number of smaller columns
  // (the columns of a row are distinct)
  rlen[er] = (unsigned)row.n;
#pragma unroll
  for (int k = 0; k < RCAP; k++) {
    if (k >= row.n) continue;
    int rank = 0;
#pragma unroll
    for (int q = 0; q < RCAP; q++) rank += (q < row.n && row.c[q] < row.c[k]) ? 1 : 0;
    rcol[(long long)er * RCAP + rank] = row.c[k];
    rval[(long long)er * RCAP + rank] = row.v[k];
  }
  entries = (unsigned long long)row.n;
}

__global__ __launch_bounds__(AWG) void ka_rows(Blocks B, const int4 *__restrict__ nbr,
                                               const int2 *__restrict__ fid,
                                               const unsigned long long *__restrict__ mask,
                                               const int *__restrict__ bptr, int n_exc, unsigned *rlen,
                                               unsigned *rkey, int *rrow, int *rcol, double *rval,
                                               Glob *g) {
  const int er = blockIdx.x * AWG + threadIdx.x;
  int wide = INT_MAX, len = 0;
  if (er < n_exc) {
    // block: the last b with bptr[b] <= er; cell: the k-th explicit cell of b
    long long lo = 0, hi = B.nb - 1;
    while (lo < hi) {
      const long long mid = (lo + hi + 1) >> 1;
      if (bptr[mid] <= er) lo = mid;
      else hi = mid - 1;
    }
    unsigned long long m = mask[lo];
    for (int k = er - bptr[lo]; k > 0; k--) m &= m - 1;  // drop the k lowest set bits
    unsigned long long entries = 0;
    build_row(B, nbr, fid, lo, __ffsll((long long)m) - 1, er, rlen, rrow, rcol, rval, wide, entries);
    len = (int)rlen[er];
    rkey[er] = ((unsigned)B.inst[lo] << 5) | (unsigned)len;  // sort key: mesh, then length
  }
  // one lane per wavefront reports, in straight-line code
  for (int o = 32; o > 0; o >>= 1) {
    len = max(len, __shfl_xor(len, o));
    wide = min(wide, __shfl_xor(wide, o));
  }
  if ((threadIdx.x & 63) == 0) {
    atomicMax(&g->maxlen, len);
    atomicMin(&g->wide, wide);
  }
}

__global__ __launch_bounds__(AWG) void ka_iota(int n, int *a) {
  const int i = blockIdx.x * AWG + threadIdx.x;
  if (i < n) a[i] = i;
}

// Bit patterns of every row slot (0.0 for the unused ones) plus one 0.0, so
// that code 0 is 0.0.
__global__ __launch_bounds__(AWG) void ka_keys(long long n_exc, const unsigned *__restrict__ rlen,
                                               const double *__restrict__ rval,
                                               unsigned long long *keys) {
  const long long q = blockIdx.x * (long long)AWG + threadIdx.x;
  if (q > n_exc * RCAP) return;
  const long long er = q / RCAP;
  const int w = (int)(q % RCAP);
  keys[q] = (er < n_exc && w < (int)rlen[er]) ? (unsigned long long)__double_as_longlong(rval[q]) : 0ull;
}

// ELL slot s (rows in (mesh, length) order): row, columns, codes; the width
// of each run of 64 rows (a wavefront's rows) and each row's length.
__global__ __launch_bounds__(AWG) void ka_ell(int n_exc, int W, const int *__restrict__ order,
                                              const unsigned *__restrict__ rlen,
                                              const int *__restrict__ rrow, const int *__restrict__ rcol,
                                              const double *__restrict__ rval,
                                              const unsigned long long *__restrict__ table, int n_codes,
                                              int *erow, int *ecol, uint16_t *ecode, uint8_t *ewid,
                                              uint8_t *elen) {
  const int s = blockIdx.x * AWG + threadIdx.x;
  int len = 0;
  if (s < n_exc) {
    const int r = order[s];
    len = (int)rlen[r];
    const int row = rrow[r];
    erow[s] = row;
    elen[s] = (uint8_t)len;
    for (int w = 0; w < W; w++) {
      int col = row, code = 0;
      if (w < len) {
        col = rcol[(long long)r * RCAP + w];
        const unsigned long long key =
            (unsigned long long)__double_as_longlong(rval[(long long)r * RCAP + w]);
        int lo = 0, hi = n_codes - 1;  // the key is in the table
        while (lo < hi) {
          const int mid = (lo + hi) >> 1;
          if (table[mid] < key) lo = mid + 1;
          else hi = mid;
        }
        code = lo;
      }
      ecol[(long long)w * n_exc + s] = col;
      ecode[(long long)w * n_exc + s] = (uint16_t)code;
    }
  }
  for (int o = 32; o > 0; o >>= 1) len = max(len, __shfl_xor(len, o));
  if ((threadIdx.x & 63) == 0 && s < n_exc) ewid[s >> 6] = (uint8_t)len;
}

// First explicit row of each mesh: eoff[k] = bptr[boff[k]], k = 0..K.
__global__ __launch_bounds__(AWG) void ka_eoff(int K, const long long *__restrict__ boff,
                                               const int *__restrict__ bptr, long long *eoff) {
  const int k = blockIdx.x * AWG + threadIdx.x;
  if (k <= K) eoff[k] = bptr[boff[k]];
}

// Work arrays of one assembly, from the caller's allocator.
struct Temps {
  const DeviceAlloc &a;
  template <class T>
  hipError_t get(T **out, size_t count) {
    *out = static_cast<T *>(a(std::max<size_t>(count, 1) * sizeof(T)));
    return *out ? hipSuccess : hipErrorOutOfMemory;
  }
};

// The host's message for a mesh's first failure (build_operator's order:
// block checks, then faces).
std::string mesh_error(const Errs &e) {
  const int eb = std::min({e.blk[0], e.blk[1], e.blk[2]});
  if (eb != INT_MAX)
    return "block " + std::to_string(eb) +
           (eb == e.blk[0] ? ": level out of range" : eb == e.blk[1] ? ": index out of range" : ": duplicate");
  return "block " + std::to_string(e.face / 4) + " face " + std::to_string(e.face % 4) +
         ": neighbour is neither same-level, coarser nor finer (mesh not 2:1 balanced)";
}

}  // namespace

#define ACHK(expr)                                                                      \
  do {                                                                                  \
    hipError_t _e = (expr);                                                             \
    if (_e != hipSuccess) {                                                             \
      err = std::string(#expr) + ": " + hipGetErrorString(_e);                          \
      return _e == hipErrorOutOfMemory ? CUP2D_PS_ENOMEM : CUP2D_PS_EHIP;               \
    }                                                                                   \
  } while (0)

int assemble_on_device(const MeshDesc *meshes, int K, int mean_constraint, hipStream_t st,
                       const DeviceAlloc &alloc, const DeviceAlloc &temp, DeviceOperator &out,
                       int &bad_mesh, std::string &err) {
  bad_mesh = -1;
  if (!meshes || K <= 0 || K > MAX_MESHES) {
    err = "need 1.." + std::to_string(MAX_MESHES) + " meshes";
    return CUP2D_PS_EINVAL;
  }
  out = DeviceOperator();
  std::vector<Inst> inst(K);
  long long nb = 0;
  unsigned long long slots = 0;
  for (int k = 0; k < K; k++) {
    const MeshDesc &m = meshes[k];
    if (m.n_blocks <= 0 || m.bpdx <= 0 || m.bpdy <= 0 || m.levels <= 0 || m.levels > 24 ||
        m.extent <= 0 || !m.level || !m.bx || !m.by) {
      bad_mesh = k;
      err = "invalid mesh description";
      return CUP2D_PS_EINVAL;
    }
    if (m.n_blocks * BLEN >= (1LL << 31) / 8) {
      bad_mesh = k;
      err = "mesh too large for 32-bit column indices";
      return CUP2D_PS_EINVAL;
    }
    unsigned long long sk = 64;
    while (sk < 2ull * (unsigned long long)m.n_blocks) sk <<= 1;
    inst[k] = Inst{nb, slots, sk - 1, m.extent / (double(std::max(m.bpdx, m.bpdy)) * BS), m.levels,
                   m.bpdx, m.bpdy, m.periodic_x, m.periodic_y};
    nb += m.n_blocks;
    slots += sk;
  }
  if (nb * BLEN >= (1LL << 31) - 64) {
    err = "batch too large for 32-bit cell indices";
    return CUP2D_PS_EINVAL;
  }
  out.nb = nb;
  out.n = nb * BLEN;
  auto give = [&](auto **p, size_t count) {
    *p = static_cast<std::remove_pointer_t<decltype(p)>>(
        alloc(std::max<size_t>(count, 1) * sizeof(**p)));
    return *p != nullptr;
  };
  Temps T{temp};

  // ---- blocks into the hash tables -------------------------------------------------
  std::vector<int> owner(nb);
  std::vector<long long> boff(K + 1);
  for (int k = 0; k < K; k++) {
    boff[k] = inst[k].boff;
    std::fill(owner.begin() + inst[k].boff, owner.begin() + inst[k].boff + meshes[k].n_blocks, k);
  }
  boff[K] = nb;
  int *d_lvl, *d_bx, *d_by, *d_inst, *d_status;
  Inst *d_I;
  long long *d_boff;
  Hash H;
  Errs *d_E;
  Glob *d_g;
  ACHK(T.get(&d_lvl, nb));
  ACHK(T.get(&d_bx, nb));
  ACHK(T.get(&d_by, nb));
  ACHK(T.get(&d_inst, nb));
  ACHK(T.get(&d_status, nb));
  ACHK(T.get(&d_I, K));
  ACHK(T.get(&d_boff, K + 1));
  ACHK(T.get(&H.keys, slots));
  ACHK(T.get(&H.vals, slots));
  ACHK(T.get(&d_E, K));
  ACHK(T.get(&d_g, 1));
  for (int k = 0; k < K; k++) {
    const size_t o = (size_t)inst[k].boff, c = (size_t)meshes[k].n_blocks * sizeof(int);
    ACHK(hipMemcpyAsync(d_lvl + o, meshes[k].level, c, hipMemcpyHostToDevice, st));
    ACHK(hipMemcpyAsync(d_bx + o, meshes[k].bx, c, hipMemcpyHostToDevice, st));
    ACHK(hipMemcpyAsync(d_by + o, meshes[k].by, c, hipMemcpyHostToDevice, st));
  }
  ACHK(hipMemcpyAsync(d_inst, owner.data(), nb * sizeof(int), hipMemcpyHostToDevice, st));
  ACHK(hipMemcpyAsync(d_I, inst.data(), K * sizeof(Inst), hipMemcpyHostToDevice, st));
  ACHK(hipMemcpyAsync(d_boff, boff.data(), (K + 1) * sizeof(long long), hipMemcpyHostToDevice, st));
  const Blocks B{d_lvl, d_bx, d_by, d_inst, d_I, nb};
  const int gb = (int)((nb + AWG - 1) / AWG);
  ka_init<<<(int)std::min<unsigned long long>(1024, slots / AWG + 1), AWG, 0, st>>>(H, slots, d_E, K, d_g);
  ka_insert<<<gb, AWG, 0, st>>>(B, H, d_status);
  ka_dups<<<gb, AWG, 0, st>>>(B, H, d_status);
  ka_status<<<gb, AWG, 0, st>>>(B, d_status, d_E);
  ACHK(hipGetLastError());

  // ---- faces, explicit cells, h^2 -----------------------------------------------
  if (!give(&out.nbr, nb) || !give(&out.mask, nb) || !give(&out.h2, nb) || !give(&out.fid, 4 * nb)) {
    err = "device allocation failed";
    return CUP2D_PS_ENOMEM;
  }
  int2 *const d_fid = out.fid;
  int *d_cnt, *d_bptr;
  ACHK(T.get(&d_cnt, nb + 1));
  ACHK(T.get(&d_bptr, nb + 1));
  ka_faces<<<gb, AWG, 0, st>>>(B, H, out.nbr, d_fid, out.mask, d_cnt, out.h2, d_E);
  ACHK(hipGetLastError());
  size_t tb = 0;
  char *d_tmp = nullptr;
  ACHK(hipcub::DeviceScan::ExclusiveSum(nullptr, tb, d_cnt, d_bptr, (int)(nb + 1), st));
  ACHK(T.get(&d_tmp, tb));
  ACHK(hipcub::DeviceScan::ExclusiveSum(d_tmp, tb, d_cnt, d_bptr, (int)(nb + 1), st));
  std::vector<Errs> E(K);
  int n_exc = 0;
  ACHK(hipMemcpyAsync(E.data(), d_E, K * sizeof(Errs), hipMemcpyDeviceToHost, st));
  ACHK(hipMemcpyAsync(&n_exc, d_bptr + nb, sizeof(int), hipMemcpyDeviceToHost, st));
  ACHK(hipStreamSynchronize(st));
  for (int k = 0; k < K; k++)
    if (std::min({E[k].blk[0], E[k].blk[1], E[k].blk[2], E[k].face}) != INT_MAX) {
      bad_mesh = k;
      err = mesh_error(E[k]);
      return CUP2D_PS_EINVAL;
    }
  out.boff = boff;
  out.mean_rows.assign(K, -1);
  for (int k = 0; k < K; k++)
    if (mean_constraint && E[k].corner >= 0) out.mean_rows[k] = (boff[k] + E[k].corner) * BLEN;
  out.corner_block = E[0].corner;
  out.mean_row = out.mean_rows[0];
  out.n_exc = n_exc;
  out.eoff.assign(K + 1, 0);
  if (n_exc == 0) {
    const double zero = 0.0;
    out.n_codes = 1;
    for (const Errs &e : E) out.nnz += (long long)e.nnz;
    if (!give(&out.erow, 1) || !give(&out.ecol, 1) || !give(&out.ecode, 1) || !give(&out.etab, 1) ||
        !give(&out.ewid, 1)) {
      err = "device allocation failed";
      return CUP2D_PS_ENOMEM;
    }
    ACHK(hipMemcpyAsync(out.etab, &zero, sizeof zero, hipMemcpyHostToDevice, st));
    ACHK(hipStreamSynchronize(st));
    return CUP2D_PS_OK;
  }

  // ---- explicit rows -------------------------------------------------------------
  unsigned *d_rlen, *d_rkey, *d_rkey_s;
  int *d_rrow, *d_rcol, *d_iota, *d_order;
  double *d_rval;
  long long *d_eoff;
  uint8_t *d_elen;
  ACHK(T.get(&d_rlen, n_exc));
  ACHK(T.get(&d_rkey, n_exc));
  ACHK(T.get(&d_rkey_s, n_exc));
  ACHK(T.get(&d_rrow, n_exc));
  ACHK(T.get(&d_rcol, (size_t)n_exc * RCAP));
  ACHK(T.get(&d_rval, (size_t)n_exc * RCAP));
  ACHK(T.get(&d_iota, n_exc));
  ACHK(T.get(&d_order, n_exc));
  ACHK(T.get(&d_eoff, K + 1));
  ACHK(T.get(&d_elen, n_exc));
  ka_rows<<<(n_exc + AWG - 1) / AWG, AWG, 0, st>>>(B, out.nbr, d_fid, out.mask, d_bptr, n_exc, d_rlen,
                                                 d_rkey, d_rrow, d_rcol, d_rval, d_g);
  ka_iota<<<(n_exc + AWG - 1) / AWG, AWG, 0, st>>>(n_exc, d_iota);
  ka_eoff<<<(K + AWG) / AWG, AWG, 0, st>>>(K, d_boff, d_bptr, d_eoff);
  ACHK(hipGetLastError());
  // stable: rows of one mesh and length keep their mesh order (length < 32: 5 bits)
  int kbits = 0;
  while ((1 << kbits) < K) kbits++;
  tb = 0;
  ACHK(hipcub::DeviceRadixSort::SortPairs(nullptr, tb, d_rkey, d_rkey_s, d_iota, d_order, n_exc, 0,
                                          5 + kbits, st));
  ACHK(T.get(&d_tmp, tb));
  ACHK(hipcub::DeviceRadixSort::SortPairs(d_tmp, tb, d_rkey, d_rkey_s, d_iota, d_order, n_exc, 0,
                                          5 + kbits, st));
  // distinct coefficients by bit pattern, ascending
  const long long nk = (long long)n_exc * RCAP + 1;
  unsigned long long *d_keys, *d_skeys, *d_table;
  int *d_ncodes;
  ACHK(T.get(&d_keys, nk));
  ACHK(T.get(&d_skeys, nk));
  ACHK(T.get(&d_table, nk));
  ACHK(T.get(&d_ncodes, 1));
  ka_keys<<<(int)((nk + AWG - 1) / AWG), AWG, 0, st>>>(n_exc, d_rlen, d_rval, d_keys);
  ACHK(hipGetLastError());
  tb = 0;
  ACHK(hipcub::DeviceRadixSort::SortKeys(nullptr, tb, d_keys, d_skeys, (int)nk, 0, 64, st));
  ACHK(T.get(&d_tmp, tb));
  ACHK(hipcub::DeviceRadixSort::SortKeys(d_tmp, tb, d_keys, d_skeys, (int)nk, 0, 64, st));
  tb = 0;
  ACHK(hipcub::DeviceSelect::Unique(nullptr, tb, d_skeys, d_table, d_ncodes, (int)nk, st));
  ACHK(T.get(&d_tmp, tb));
  ACHK(hipcub::DeviceSelect::Unique(d_tmp, tb, d_skeys, d_table, d_ncodes, (int)nk, st));
  Glob g;
  int n_codes = 0;
  ACHK(hipMemcpyAsync(&g, d_g, sizeof g, hipMemcpyDeviceToHost, st));
  ACHK(hipMemcpyAsync(&n_codes, d_ncodes, sizeof n_codes, hipMemcpyDeviceToHost, st));
  ACHK(hipMemcpyAsync(out.eoff.data(), d_eoff, (K + 1) * sizeof(long long), hipMemcpyDeviceToHost, st));
  ACHK(hipStreamSynchronize(st));
  if (g.wide != INT_MAX || g.maxlen > ELL_MAX) {
    err = "explicit row with more than " + std::to_string(ELL_MAX) + " entries";
    return CUP2D_PS_EINVAL;
  }
  if (n_codes > 0x10000) {
    err = "more than 65536 distinct coefficients in the explicit rows";
    return CUP2D_PS_EINVAL;
  }
  const int W = g.maxlen;
  out.ell_w = W;
  out.n_codes = n_codes;
  if (!give(&out.erow, n_exc) || !give(&out.ecol, (size_t)W * n_exc) ||
      !give(&out.ecode, (size_t)W * n_exc) || !give(&out.etab, n_codes) ||
      !give(&out.ewid, (n_exc + 63) / 64)) {
    err = "device allocation failed";
    return CUP2D_PS_ENOMEM;
  }
  ACHK(hipMemcpyAsync(out.etab, d_table, n_codes * sizeof(double), hipMemcpyDeviceToDevice, st));
  ka_ell<<<(n_exc + AWG - 1) / AWG, AWG, 0, st>>>(n_exc, W, d_order, d_rlen, d_rrow, d_rcol, d_rval,
                                                  d_table, n_codes, out.erow, out.ecol, out.ecode,
                                                  out.ewid, d_elen);
  ACHK(hipGetLastError());
  out.elen.resize(n_exc);
  ACHK(hipMemcpyAsync(out.elen.data(), d_elen, n_exc, hipMemcpyDeviceToHost, st));
  ACHK(hipStreamSynchronize(st));
  // nonzeros: the plain rows' (counted per block) and the explicit rows' lengths
  for (const Errs &e : E) out.nnz += (long long)e.nnz;
  for (uint8_t l : out.elen) out.nnz += l;
  return CUP2D_PS_OK;
}

}  // namespace cup2d
