// assembly.cpp -- builds the AMR Poisson operator (stencil form + CSR) from a
// block description.
//
// Operator definition (reference source/Poisson/ExpAMRSolver.cpp:180-313):
//   * faces of a cell inside its block, and block faces whose neighbour is a
//     same-level block, contribute +x_nb - x_self;
//   * non-periodic domain faces contribute nothing (zero-flux walls);
//   * a face whose neighbour block is one level coarser contributes
//     ghost - x_self, the ghost being 2/3 x_self - 1/5 x_inward + 8/15 times a
//     second-order Taylor expansion of the coarse cell along the face
//     (ExpAMRSolver.cpp:100-160, D1/D2 at ExpAMRSolver.h:268-300);
//   * a face whose neighbour is one level finer contributes, for each of the
//     two fine cells behind it, x_fine - ghost_fine (ExpAMRSolver.cpp:161-176),
//     which makes the coarse and fine fluxes cancel (zero column sums).
// Coefficients of a row are summed per column in the order the reference
// adds them (face order Xm, Xp, Ym, Yp), so the values are bit-identical to
// the reference's assembly.

#include "assembly.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <unordered_map>

namespace cup2d {
namespace {

enum FaceKind : int { SAME = 0, WALL = 1, COARSER = 2, FINER = 3 };

struct Face {
  int kind = WALL;
  long long same = -1;     // same-level neighbour block
  long long coarse = -1;   // coarser neighbour block
  long long fine[2] = {-1, -1};  // finer neighbour blocks along the tangential axis (low, high half)
};

struct BlockPos {
  int level;
  long long i, j;
};

uint64_t pos_key(int l, long long i, long long j) {
  return (uint64_t(l) << 58) | (uint64_t(i) << 29) | uint64_t(j);
}

// Coefficient accumulator for one explicit row: at most ~20 distinct columns.
struct RowAcc {
  long long c[48];
  double v[48];
  int n = 0;
  void add(long long col, double val) {
    for (int k = 0; k < n; k++)
      if (c[k] == col) {
        v[k] += val;
        return;
      }
    c[n] = col;
    v[n] = val;
    n++;
  }
};

inline long long gidx(long long b, int ix, int iy) { return b * BLEN + iy * BS + ix; }

// One coarse-fine ghost: 2/3 close - 1/5 far + 8/15 (c + sT*D1(c) + D2(c)),
// multiplied by sI.  `t` is the coarse cell's index along the face-tangential
// axis and `step(d)` the global index of the coarse cell d steps along it.
template <class Step>
void ghost(RowAcc &row, long long close, long long far_, long long c, int t, Step step, double sI,
           double sT) {
  row.add(close, sI * 2. / 3.);
  row.add(far_, -sI * 1. / 5.);
  const double w = sI * 8. / 15.;
  row.add(c, w);
  // first and second tangential differences of the coarse cell: one-sided at
  // the tangential indices 3,7 (backward) and 0,4 (forward), central
  // elsewhere (ExpAMRSolver.h:144-147, :206-209, :268-300)
  const bool bwd = (t == BS - 1) || (t == BS / 2 - 1);
  const bool fwd = (t == 0) || (t == BS / 2);
  if (bwd) {
    row.add(step(-2), sT * w * (1. / 8.));
    row.add(step(-1), sT * w * (-1. / 2.));
    row.add(c, sT * w * (3. / 8.));
    row.add(step(-2), w * (1. / 32.));
    row.add(step(-1), w * (-1. / 16.));
    row.add(c, w * (1. / 32.));
  } else if (fwd) {
    row.add(step(2), sT * w * (-1. / 8.));
    row.add(step(1), sT * w * (1. / 2.));
    row.add(c, sT * w * (-3. / 8.));
    row.add(step(2), w * (1. / 32.));
    row.add(step(1), w * (-1. / 16.));
    row.add(c, w * (1. / 32.));
  } else {
    row.add(step(-1), sT * w * (-1. / 8.));
    row.add(step(1), sT * w * (1. / 8.));
    row.add(c, sT * w * 0.);
    row.add(step(-1), w * (1. / 32.));
    row.add(step(1), w * (1. / 32.));
    row.add(c, w * (-1. / 16.));
  }
}

}  // namespace

bool build_operator(const MeshDesc &m, int mean_constraint, Operator &op, std::string &err,
                    bool with_csr) {
  const long long nb = m.n_blocks;
  if (nb <= 0 || m.bpdx <= 0 || m.bpdy <= 0 || m.levels <= 0 || m.levels > 24 || m.extent <= 0 ||
      !m.level || !m.bx || !m.by) {
    err = "invalid mesh description";
    return false;
  }
  if (nb * BLEN >= (1LL << 31) / 8) {
    err = "mesh too large for 32-bit column indices";
    return false;
  }
  op = Operator();
  op.nb = nb;
  op.n = nb * BLEN;

  // ---- block lookup --------------------------------------------------------
  std::vector<BlockPos> B(nb);
  std::unordered_map<uint64_t, long long> where;
  where.reserve(size_t(nb) * 2);
  for (long long b = 0; b < nb; b++) {
    const int l = m.level[b];
    if (l < 0 || l >= m.levels) {
      err = "block " + std::to_string(b) + ": level out of range";
      return false;
    }
    const long long nx = (long long)m.bpdx << l, ny = (long long)m.bpdy << l;
    if (m.bx[b] < 0 || m.bx[b] >= nx || m.by[b] < 0 || m.by[b] >= ny) {
      err = "block " + std::to_string(b) + ": index out of range";
      return false;
    }
    B[b] = BlockPos{l, m.bx[b], m.by[b]};
    if (!where.emplace(pos_key(l, m.bx[b], m.by[b]), b).second) {
      err = "block " + std::to_string(b) + ": duplicate";
      return false;
    }
  }
  auto find = [&](int l, long long i, long long j) -> long long {
    if (l < 0 || l >= m.levels) return -1;
    auto it = where.find(pos_key(l, i, j));
    return it == where.end() ? -1 : it->second;
  };

  // ---- face classification (ExpAMRSolver.cpp:221-249, :141-177) -------------
  std::vector<Face> F(size_t(nb) * 4);
  op.nbr.assign(size_t(nb) * 4, -1);
  op.fid.assign(size_t(nb) * 8, -1);
  static const int di[4] = {-1, 1, 0, 0};
  static const int dj[4] = {0, 0, -1, 1};
  for (long long b = 0; b < nb; b++) {
    const int l = B[b].level;
    const long long nx = (long long)m.bpdx << l, ny = (long long)m.bpdy << l;
    for (int f = 0; f < 4; f++) {
      Face &fc = F[4 * b + f];
      long long I = B[b].i + di[f], J = B[b].j + dj[f];
      const bool outx = I < 0 || I >= nx, outy = J < 0 || J >= ny;
      if ((outx && !m.periodic_x) || (outy && !m.periodic_y)) {
        fc.kind = WALL;
        continue;
      }
      I = (I + nx) % nx;
      J = (J + ny) % ny;
      if ((fc.same = find(l, I, J)) >= 0) {
        fc.kind = SAME;
        op.nbr[4 * b + f] = (int)fc.same;
      } else if ((fc.coarse = find(l - 1, I >> 1, J >> 1)) >= 0) {
        fc.kind = COARSER;
        op.nbr[4 * b + f] = -2;
        op.fid[8 * b + 2 * f] = (int)fc.coarse;
        op.fid[8 * b + 2 * f + 1] = -1 - (int)((f < 2 ? B[b].j : B[b].i) & 1);
      } else {
        // the two children of the neighbour position that touch this face
        for (int h = 0; h < 2; h++) {
          long long ci, cj;
          if (f < 2) {
            ci = 2 * I + (f == 0 ? 1 : 0);
            cj = 2 * J + h;
          } else {
            ci = 2 * I + h;
            cj = 2 * J + (f == 2 ? 1 : 0);
          }
          fc.fine[h] = find(l + 1, ci, cj);
        }
        if (fc.fine[0] < 0 || fc.fine[1] < 0) {
          err = "block " + std::to_string(b) + " face " + std::to_string(f) +
                ": neighbour is neither same-level, coarser nor finer (mesh not 2:1 balanced)";
          return false;
        }
        fc.kind = FINER;
        op.nbr[4 * b + f] = -2;
        op.fid[8 * b + 2 * f] = (int)fc.fine[0];
        op.fid[8 * b + 2 * f + 1] = (int)fc.fine[1];
      }
    }
  }

  // ---- rows -----------------------------------------------------------------
  if (with_csr) {
    op.rp.assign(op.n + 1, 0);
    op.col.reserve(op.n * 5 + op.n / 4);
    op.val.reserve(op.n * 5 + op.n / 4);
  }
  long long nnz = 0;
  op.exc_mask.assign(nb, 0);
  op.exc_bptr.assign(nb + 1, 0);
  op.exc_rptr.push_back(0);

  for (long long b = 0; b < nb; b++) {
    const Face *fb = &F[4 * b];
    for (int iy = 0; iy < BS; iy++)
      for (int ix = 0; ix < BS; ix++) {
        const long long self = gidx(b, ix, iy);
        const bool edge[4] = {ix == 0, ix == BS - 1, iy == 0, iy == BS - 1};
        bool cf = false;
        for (int f = 0; f < 4; f++) cf |= edge[f] && (fb[f].kind == COARSER || fb[f].kind == FINER);
        if (!with_csr && !cf) {  // a plain row only counts: self + its non-wall faces
          int cnt = 1;
          for (int f = 0; f < 4; f++) cnt += (!edge[f] || fb[f].kind == SAME) ? 1 : 0;
          nnz += cnt;
          continue;
        }
        RowAcc row;
        if (!edge[0] && !edge[1] && !edge[2] && !edge[3]) {
          // interior cell: ascending columns (ExpAMRSolver.cpp:270-277)
          row.add(gidx(b, ix, iy - 1), 1);
          row.add(gidx(b, ix - 1, iy), 1);
          row.add(self, -4);
          row.add(gidx(b, ix + 1, iy), 1);
          row.add(gidx(b, ix, iy + 1), 1);
        } else {
          for (int f = 0; f < 4; f++) {
            const Face &fc = fb[f];
            if (!edge[f]) {
              const long long inner = f == 0 ? gidx(b, ix - 1, iy) : f == 1 ? gidx(b, ix + 1, iy)
                                    : f == 2 ? gidx(b, ix, iy - 1) : gidx(b, ix, iy + 1);
              row.add(inner, 1);
              row.add(self, -1);
              continue;
            }
            if (fc.kind == WALL) continue;
            const bool xf = f < 2;
            if (fc.kind == SAME) {
              const long long c = f == 0 ? gidx(fc.same, BS - 1, iy) : f == 1 ? gidx(fc.same, 0, iy)
                                : f == 2 ? gidx(fc.same, ix, BS - 1) : gidx(fc.same, ix, 0);
              row.add(c, 1);
              row.add(self, -1);
            } else if (fc.kind == COARSER) {
              // coarse cell facing this fine cell, inward fine cell, Taylor sign
              int ixc, iyc;
              long long inward;
              if (xf) {
                ixc = f == 0 ? BS - 1 : 0;
                iyc = iy / 2 + ((B[b].j & 1) ? BS / 2 : 0);
                inward = gidx(b, f == 0 ? ix + 1 : ix - 1, iy);
              } else {
                iyc = f == 2 ? BS - 1 : 0;
                ixc = ix / 2 + ((B[b].i & 1) ? BS / 2 : 0);
                inward = gidx(b, ix, f == 2 ? iy + 1 : iy - 1);
              }
              const double sT = ((xf ? iy : ix) % 2 == 0) ? -1. : 1.;
              const long long cb = fc.coarse;
              auto step = [&](int d) { return xf ? gidx(cb, ixc, iyc + d) : gidx(cb, ixc + d, iyc); };
              ghost(row, self, inward, gidx(cb, ixc, iyc), xf ? iyc : ixc, step, 1., sT);
              row.add(self, -1);
            } else {  // FINER
              const long long fbk = fc.fine[(xf ? iy : ix) >= BS / 2];
              const int t0 = ((xf ? iy : ix) % (BS / 2)) * 2;
              auto step = [&](int d) { return xf ? gidx(b, ix, iy + d) : gidx(b, ix + d, iy); };
              for (int h = 0; h < 2; h++) {
                const int t = t0 + h;
                long long close, far_;
                switch (f) {
                  case 0: close = gidx(fbk, BS - 1, t); far_ = gidx(fbk, BS - 2, t); break;
                  case 1: close = gidx(fbk, 0, t); far_ = gidx(fbk, 1, t); break;
                  case 2: close = gidx(fbk, t, BS - 1); far_ = gidx(fbk, t, BS - 2); break;
                  default: close = gidx(fbk, t, 0); far_ = gidx(fbk, t, 1); break;
                }
                row.add(close, 1);
                ghost(row, close, far_, self, xf ? iy : ix, step, -1., h == 0 ? -1. : 1.);
              }
            }
          }
        }
        nnz += row.n;
        // sort columns ascending (the reference keeps rows in a std::map)
        int order[48];
        for (int k = 0; k < row.n; k++) order[k] = k;
        std::sort(order, order + row.n, [&](int a, int c) { return row.c[a] < row.c[c]; });
        if (with_csr) {
          for (int k = 0; k < row.n; k++) {
            op.col.push_back((int)row.c[order[k]]);
            op.val.push_back(row.v[order[k]]);
          }
          op.rp[self + 1] = (int)op.col.size();
        }
        if (cf) {
          op.exc_mask[b] |= uint64_t(1) << (iy * BS + ix);
          op.exc_cell.push_back(iy * BS + ix);
          for (int k = 0; k < row.n; k++) {
            op.exc_col.push_back((int)row.c[order[k]]);
            op.exc_val.push_back(row.v[order[k]]);
          }
          op.exc_rptr.push_back((int)op.exc_col.size());
        }
      }
    op.exc_bptr[b + 1] = (int)op.exc_cell.size();
  }
  op.nnz_rows = nnz;

  // ---- h^2 and the mean-constraint row ---------------------------------------
  op.h2.resize(nb);
  const double h0 = m.extent / (double(std::max(m.bpdx, m.bpdy)) * BS);
  for (long long b = 0; b < nb; b++) {
    const double h = h0 / double(1LL << B[b].level);
    op.h2[b] = h * h;
    if (B[b].i == 0 && B[b].j == 0) op.corner_block = b;
  }
  op.mean_row = (mean_constraint && op.corner_block >= 0) ? op.corner_block * BLEN : -1;
  return true;
}

// float -> bf16 bits, round to nearest even (finite inputs)
static uint16_t bf16_bits(float f) {
  uint32_t u;
  std::memcpy(&u, &f, 4);
  const uint32_t lsb = (u >> 16) & 1u;
  u += 0x7fffu + lsb;
  return uint16_t(u >> 16);
}


// -K^-1 for the 8x8 block (Cholesky-Crout of K then L^-1 by Gauss-Jordan and
// P = -(L^-T L^-1), the construction of ExpAMRSolver.cpp:37-95).
int ell_width(const Operator &op) {
  int w = 0;
  for (size_t e = 0; e + 1 < op.exc_rptr.size(); e++) w = std::max(w, op.exc_rptr[e + 1] - op.exc_rptr[e]);
  return w;
}

int EllRows::code(double v) {
  uint64_t bits;
  std::memcpy(&bits, &v, sizeof bits);
  auto it = std::lower_bound(codes_.begin(), codes_.end(), std::make_pair(bits, uint16_t(0)),
                             [](const auto &a, const auto &b) { return a.first < b.first; });
  if (it != codes_.end() && it->first == bits) return it->second;
  if (values.size() > 0xffff) return -1;
  const uint16_t c = (uint16_t)values.size();
  values.push_back(v);
  codes_.insert(it, {bits, c});
  return c;
}

bool EllRows::append(const Operator &op, long long row_off, std::string &err) {
  const long long n_exc = op.n_exc();
  std::vector<long long> order(n_exc);  // explicit rows of op in mesh order: block, then cell
  std::vector<int> rowof(n_exc);
  for (long long b = 0; b < op.nb; b++)
    for (int e = op.exc_bptr[b]; e < op.exc_bptr[b + 1]; e++) {
      order[e] = e;
      rowof[e] = (int)(b * BLEN + op.exc_cell[e] + row_off);
    }
  auto length = [&](long long e) { return op.exc_rptr[e + 1] - op.exc_rptr[e]; };
  std::stable_sort(order.begin(), order.end(), [&](long long a, long long b) { return length(a) < length(b); });
  for (long long e : order) {
    const int row = rowof[e], k0 = op.exc_rptr[e], k1 = op.exc_rptr[e + 1];
    if (k1 - k0 > W) {
      err = "explicit row wider than the ELL width";
      return false;
    }
    erow.push_back(row);
    len.push_back((uint8_t)(k1 - k0));
    for (int w = 0; w < W; w++) {
      const bool ok = k0 + w < k1;
      ecol_rm.push_back(ok ? (int)(op.exc_col[k0 + w] + row_off) : row);
      const int c = ok ? code(op.exc_val[k0 + w]) : 0;
      if (c < 0) {
        err = "more than 65536 distinct coefficients in the explicit rows";
        return false;
      }
      ecode_rm.push_back((uint16_t)c);
    }
  }
  return true;
}

int EllRows::width(long long e0, long long e1) const {
  int w = 0;
  for (long long e = e0; e < e1; e++) w = std::max(w, (int)len[e]);
  return w;
}

std::vector<double> make_pinv() {
  const int N = 64;
  auto K = [](int a, int b) {
    const int ya = a / 8, xa = a % 8, yb = b / 8, xb = b % 8;
    if (a == b) return 4.0;
    return (std::abs(xa - xb) + std::abs(ya - yb) == 1) ? -1.0 : 0.0;
  };
  std::vector<double> L(N * N, 0.0), Li(N * N, 0.0), P(N * N);
  for (int i = 0; i < N; i++) {
    double s1 = 0;
    for (int k = 0; k < i; k++) s1 += L[i * N + k] * L[i * N + k];
    L[i * N + i] = std::sqrt(K(i, i) - s1);
    for (int j = i + 1; j < N; j++) {
      double s2 = 0;
      for (int k = 0; k < i; k++) s2 += L[i * N + k] * L[j * N + k];
      L[j * N + i] = (K(j, i) - s2) / L[i * N + i];
    }
  }
  for (int i = 0; i < N; i++) Li[i * N + i] = 1.0;
  for (int br = 0; br < N; br++) {
    const double f = 1.0 / L[br * N + br];
    for (int c = 0; c <= br; c++) Li[br * N + c] *= f;
    for (int wr = br + 1; wr < N; wr++) {
      const double wf = L[wr * N + br];
      for (int c = 0; c <= br; c++) Li[wr * N + c] -= wf * Li[br * N + c];
    }
  }
  for (int i = 0; i < N; i++)
    for (int j = 0; j < N; j++) {
      double a = 0;
      for (int k = std::max(i, j); k < N; k++) a += Li[k * N + i] * Li[k * N + j];
      P[i * N + j] = -a;
    }
  return P;
}

// P in the A-operand layout of v_mfma_f32_16x16x32_bf16 (see kernels.h):
// fragment (mt, ks), lane l, element j = P[16 mt + (l & 15)][32 ks + 8 (l >> 4) + j].
std::vector<uint16_t> make_pfrag(const std::vector<double> &P) {
  std::vector<uint16_t> f(8 * 64 * 8);
  for (int mt = 0; mt < 4; mt++)
    for (int ks = 0; ks < 2; ks++)
      for (int l = 0; l < 64; l++)
        for (int j = 0; j < 8; j++) {
          const int row = 16 * mt + (l & 15), col = 32 * ks + 8 * (l >> 4) + j;
          f[((mt * 2 + ks) * 64 + l) * 8 + j] = bf16_bits((float)P[row * 64 + col]);
        }
  return f;
}

std::vector<double> make_pfrag64(const std::vector<double> &P) {
  std::vector<double> f(64 * 64);
  for (int mt = 0; mt < 4; mt++)
    for (int ks = 0; ks < 16; ks++)
      for (int l = 0; l < 64; l++) {
        const int m = l & 15, row = 16 * mt + 4 * (m & 3) + (m >> 2);
        f[(mt * 16 + ks) * 64 + l] = P[row * 64 + 16 * (l >> 4) + ks];
      }
  return f;
}

}  // namespace cup2d
