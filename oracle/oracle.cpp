// oracle.cpp -- CPU FP64 restatement of the CubismUP-2D pressure-Poisson path.
//
// TEST INFRASTRUCTURE ONLY.  This file is the checker for the HIP solver in
// cup2d_amd/: only tests/, __graft_entry__.smoke() and the cpu_baseline leg of
// bench.py may load it, and only to compare against or to time the reference
// algorithm on the host.  Nothing in the product links it.
//
// Every function cites the reference file:line it restates (paths relative to
// the reference root).  The reference itself cannot be compiled here (its
// Cubism submodule is empty and it needs nvcc/cuBLAS/cuSPARSE), so this is an
// independent C++ restatement written from reading those files.
//
// Parity pins (tests/test_oracle_*.py):
//   orc_pinv              DST-I closed form of K^-1, K*P = -I, symmetry.
//   orc_assemble          uniform grid == textbook 5-point Neumann Laplacian
//                         (scipy), LoggerAnalyser.ipynb:496-497 known-answer
//                         row, zero row sums and zero column sums on AMR
//                         meshes, O(h^2) consistency on smooth fields,
//                         exact rational coefficients by brute force.
//   orc_spmv              dense matrix-vector product from the CSR.
//   orc_bicgstab          dense direct solve (numpy) on tiny meshes; the
//                         independently recomputed true residual.
//   orc_pipelined_bicgstab  same solution as orc_bicgstab / dense solve.
//   orc_block_cholesky_apply  matches -K^-1 within the 1e-10 drop tolerance.
//   orc_mean_remove       closed form sum h^2 (p - pbar) = 0.

#include <cmath>
#include <cstdint>
#include <cstring>
#include <map>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <utility>
#include <vector>
#include <omp.h>

namespace {

constexpr int BS = 8;         // cells per block side (CUP2D_BLOCK_SIZE default)
constexpr int BLEN = BS * BS; // cells per block

std::string g_err;

// ---------------------------------------------------------------------------
// Block-quadtree topology: the subset of Cubism's Tree/BlockInfo semantics the
// assembly uses (ExpAMRSolver.cpp:141-163, :186-249):  Exists / CheckCoarser /
// CheckFiner of a same-level neighbour position, Zparent and Zchild.
// ---------------------------------------------------------------------------
struct Topo {
  int bpdx = 0, bpdy = 0;
  bool perx = false, pery = false;
  std::unordered_map<uint64_t, long long> leaf;  // (level,i,j) -> block id

  static uint64_t key(int l, long long i, long long j) {
    return (uint64_t(l) << 56) | (uint64_t(i & 0xFFFFFFF) << 28) | uint64_t(j & 0xFFFFFFF);
  }
  long long nx(int l) const { return (long long)bpdx << l; }
  long long ny(int l) const { return (long long)bpdy << l; }
  // Periodic wrap of a same-level position (Cubism Znei on periodic grids).
  void wrap(int l, long long &i, long long &j) const {
    if (perx) i = ((i % nx(l)) + nx(l)) % nx(l);
    if (pery) j = ((j % ny(l)) + ny(l)) % ny(l);
  }
  long long find(int l, long long i, long long j) const {
    if (l < 0) return -1;
    auto it = leaf.find(key(l, i, j));
    return it == leaf.end() ? -1 : it->second;
  }
};

// Row accumulator with the semantics of SpRowInfo::mapColVal
// (LocalSpMatDnVec.h:25-38): values of repeated columns are summed in call
// order; columns come out sorted (std::map).
struct Row {
  std::map<long long, double> cv;
  void add(long long c, double v) { cv[c] += v; }
};

struct Block {
  int level;
  long long I, J;
};

// CellIndexer::This (ExpAMRSolver.h:74-75): single rank, so the rank offset
// Nblocks_xcumsum_[rank] is 0.
inline long long cell(long long b, int ix, int iy) { return b * BLEN + iy * BS + ix; }

enum Face { XM = 0, XP = 1, YM = 2, YP = 3 };

// Tangential-direction helpers of the edge indexers (ExpAMRSolver.h:135-258).
inline bool isXface(int f) { return f == XM || f == XP; }

// taylorSign (ExpAMRSolver.h:142-143, :204-205)
inline double taylorSign(int f, int ix, int iy) {
  return isXface(f) ? (iy % 2 == 0 ? -1. : 1.) : (ix % 2 == 0 ? -1. : 1.);
}
// isBD / isFD (ExpAMRSolver.h:144-147, :206-209)
inline bool isBD(int f, int ix, int iy) {
  const int t = isXface(f) ? iy : ix;
  return t == BS - 1 || t == BS / 2 - 1;
}
inline bool isFD(int f, int ix, int iy) {
  const int t = isXface(f) ? iy : ix;
  return t == 0 || t == BS / 2;
}
// Nei (ExpAMRSolver.h:148-149, :210-211): step along the face-tangential axis.
inline long long nei(int f, long long b, int ix, int iy, int dist) {
  return isXface(f) ? cell(b, ix, iy + dist) : cell(b, ix + dist, iy);
}

// D1, D2 (ExpAMRSolver.h:268-300).
void D1(int f, long long b, int ix, int iy, long long col[3], double val[3]) {
  if (isBD(f, ix, iy)) {
    col[0] = nei(f, b, ix, iy, -2); val[0] = 1. / 8.;
    col[1] = nei(f, b, ix, iy, -1); val[1] = -1. / 2.;
    col[2] = cell(b, ix, iy);       val[2] = 3. / 8.;
  } else if (isFD(f, ix, iy)) {
    col[0] = nei(f, b, ix, iy, 2);  val[0] = -1. / 8.;
    col[1] = nei(f, b, ix, iy, 1);  val[1] = 1. / 2.;
    col[2] = cell(b, ix, iy);       val[2] = -3. / 8.;
  } else {
    col[0] = nei(f, b, ix, iy, -1); val[0] = -1. / 8.;
    col[1] = nei(f, b, ix, iy, 1);  val[1] = 1. / 8.;
    col[2] = cell(b, ix, iy);       val[2] = 0.;
  }
}
void D2(int f, long long b, int ix, int iy, long long col[3], double val[3]) {
  if (isBD(f, ix, iy)) {
    col[0] = nei(f, b, ix, iy, -2); val[0] = 1. / 32.;
    col[1] = nei(f, b, ix, iy, -1); val[1] = -1. / 16.;
    col[2] = cell(b, ix, iy);       val[2] = 1. / 32.;
  } else if (isFD(f, ix, iy)) {
    col[0] = nei(f, b, ix, iy, 2);  val[0] = 1. / 32.;
    col[1] = nei(f, b, ix, iy, 1);  val[1] = -1. / 16.;
    col[2] = cell(b, ix, iy);       val[2] = 1. / 32.;
  } else {
    col[0] = nei(f, b, ix, iy, -1); val[0] = 1. / 32.;
    col[1] = nei(f, b, ix, iy, 1);  val[1] = 1. / 32.;
    col[2] = cell(b, ix, iy);       val[2] = -1. / 16.;
  }
}

// ExpAMRSolver::interpolate (ExpAMRSolver.cpp:100-128).
void interpolate(int f, long long bc, int ixc, int iyc, long long closeIdx, long long farIdx,
                 double signI, double signT, Row &row) {
  row.add(closeIdx, signI * 2. / 3.);
  row.add(farIdx, -signI * 1. / 5.);
  const double tf = signI * 8. / 15.;
  row.add(cell(bc, ixc, iyc), tf);
  long long c[3];
  double v[3];
  D1(f, bc, ixc, iyc, c, v);
  for (int i = 0; i < 3; i++) row.add(c[i], signT * tf * v[i]);
  D2(f, bc, ixc, iyc, c, v);
  for (int i = 0; i < 3; i++) row.add(c[i], tf * v[i]);
}

// ExpAMRSolver::makeFlux (ExpAMRSolver.cpp:131-178) for face f of cell
// (ix,iy) of block b; (NI,NJ) is the same-level neighbour position (Znei).
void makeFlux(const Topo &T, const std::vector<Block> &B, long long b, int ix, int iy, int f,
              long long NI, long long NJ, Row &row) {
  const int l = B[b].level;
  const long long self = cell(b, ix, iy);
  T.wrap(l, NI, NJ);
  const long long nb = T.find(l, NI, NJ);
  if (nb >= 0) {
    // Same level: neiUnif (ExpAMRSolver.h:160-161, :183-184, :222-223, :245-246).
    long long c;
    switch (f) {
      case XM: c = cell(nb, BS - 1, iy); break;
      case XP: c = cell(nb, 0, iy); break;
      case YM: c = cell(nb, ix, BS - 1); break;
      default: c = cell(nb, ix, 0); break;
    }
    row.add(c, 1.);
    row.add(self, -1.);
    return;
  }
  const long long cb = T.find(l - 1, NI >> 1, NJ >> 1);  // Zparent leaf
  if (cb >= 0) {
    // Neighbour is coarser (ExpAMRSolver.cpp:150-160).
    int ixc, iyc;
    long long inward;
    if (isXface(f)) {
      ixc = (f == XM) ? BS - 1 : 0;                                     // ExpAMRSolver.h:166-167, :189-190
      iyc = (B[b].J % 2 == 0) ? iy / 2 : iy / 2 + BS / 2;               // ExpAMRSolver.h:119-120
      inward = (f == XM) ? cell(b, ix + 1, iy) : cell(b, ix - 1, iy);   // :163-164, :186-187
    } else {
      iyc = (f == YM) ? BS - 1 : 0;                                     // :228-229, :251-252
      ixc = (B[b].I % 2 == 0) ? ix / 2 : ix / 2 + BS / 2;               // :117-118
      inward = (f == YM) ? cell(b, ix, iy + 1) : cell(b, ix, iy - 1);   // :225-226, :248-249
    }
    interpolate(f, cb, ixc, iyc, self, inward, 1., taylorSign(f, ix, iy), row);
    row.add(self, -1.);
    return;
  }
  // Neighbour is finer (ExpAMRSolver.cpp:161-176).  Zchild choice:
  // ExpAMRSolver.h:171-172, :194-195, :233-234, :256-257.
  long long CI, CJ;
  switch (f) {
    case XM: CI = 2 * NI + 1; CJ = 2 * NJ + (iy >= BS / 2); break;
    case XP: CI = 2 * NI;     CJ = 2 * NJ + (iy >= BS / 2); break;
    case YM: CI = 2 * NI + (ix >= BS / 2); CJ = 2 * NJ + 1; break;
    default: CI = 2 * NI + (ix >= BS / 2); CJ = 2 * NJ; break;
  }
  const long long fb = T.find(l + 1, CI, CJ);
  if (fb < 0) throw std::runtime_error("Neighbour doesn't exist, isn't coarser, nor finer...");
  // neiFine1/neiFine2 with offset 0 (close) and 1 (far): ExpAMRSolver.h:169-170, :192-193, :231-232, :254-255.
  auto fine = [&](int which, int offset) -> long long {
    if (isXface(f)) {
      const int r = (iy % (BS / 2)) * 2 + which;
      return f == XM ? cell(fb, BS - 1 - offset, r) : cell(fb, offset, r);
    } else {
      const int c = (ix % (BS / 2)) * 2 + which;
      return f == YM ? cell(fb, c, BS - 1 - offset) : cell(fb, c, offset);
    }
  };
  long long fc = fine(0, 0), ff = fine(0, 1);
  row.add(fc, 1.);
  interpolate(f, b, ix, iy, fc, ff, -1., -1., row);
  fc = fine(1, 0);
  ff = fine(1, 1);
  row.add(fc, 1.);
  interpolate(f, b, ix, iy, fc, ff, -1., 1., row);
}

bool buildTopo(Topo &T, std::vector<Block> &B, const int *lev, const int *bi, const int *bj,
               long long nb, int bpdx, int bpdy, int perx, int pery) {
  T.bpdx = bpdx;
  T.bpdy = bpdy;
  T.perx = perx != 0;
  T.pery = pery != 0;
  B.resize(nb);
  T.leaf.reserve(nb * 2);
  for (long long b = 0; b < nb; b++) {
    B[b] = Block{lev[b], bi[b], bj[b]};
    if (lev[b] < 0 || lev[b] > 20 || bi[b] < 0 || bj[b] < 0 || bi[b] >= T.nx(lev[b]) ||
        bj[b] >= T.ny(lev[b])) {
      g_err = "block " + std::to_string(b) + " out of range";
      return false;
    }
    if (!T.leaf.emplace(Topo::key(lev[b], bi[b], bj[b]), b).second) {
      g_err = "duplicate block " + std::to_string(b);
      return false;
    }
  }
  return true;
}

double blockH(int level, int bpdx, int bpdy, double extent) {
  const int bpd = bpdx > bpdy ? bpdx : bpdy;
  return extent / (double(bpd) * BS) / double(1LL << level);
}

// Mean-constraint corner block: index == (0,0,0) (ExpAMRSolver.cpp:252-256,
// AMRSolver.h:69-74).
long long cornerBlock(const int *bi, const int *bj, long long nb) {
  for (long long b = 0; b < nb; b++)
    if (bi[b] == 0 && bj[b] == 0) return b;
  return -1;
}

// ---- vector helpers -------------------------------------------------------
double dot(const std::vector<double> &a, const std::vector<double> &b) {
  double s = 0;
  const long long n = (long long)a.size();
#pragma omp parallel for reduction(+ : s)
  for (long long i = 0; i < n; i++) s += a[i] * b[i];
  return s;
}
double amax(const std::vector<double> &a) {
  double m = 0;
  const long long n = (long long)a.size();
#pragma omp parallel for reduction(max : m)
  for (long long i = 0; i < n; i++) m = std::fabs(a[i]) > m ? std::fabs(a[i]) : m;
  return m;
}

struct Csr {
  long long n;
  const long long *rp;
  const int *col;
  const double *val;
};

// y = A x with the reference mean-constraint handling:
//   mode 1: y[mean_row] = sum_i h2(block(i)) x_i (BiCGSTAB.cu:393-409,
//           AMRSolver.h:83-109);
//   mode 2: y_i += h2(block(i)) * sum_j h2(block(j)) x_j (AMRSolver.h:110-118).
void spmv(const Csr &A, const double *x, double *y, int mode, long long mean_row, const double *h2) {
#pragma omp parallel for schedule(static)
  for (long long i = 0; i < A.n; i++) {
    double s = 0;
    for (long long k = A.rp[i]; k < A.rp[i + 1]; k++) s += A.val[k] * x[A.col[k]];
    y[i] = s;
  }
  if (mode > 0) {
    double m = 0;
    for (long long i = 0; i < A.n; i++) m += h2[i / BLEN] * x[i];
    if (mode == 1) {
      if (mean_row >= 0) y[mean_row] = m;
    } else {
#pragma omp parallel for schedule(static)
      for (long long i = 0; i < A.n; i++) y[i] += m * h2[i / BLEN];
    }
  }
}

// Dense block-Jacobi: z_b = P z_b for every block, P row-major 64x64
// (cublasDgemm(T,N,64,N_b,64) at BiCGSTAB.cu:520, :545).
void precDense(const double *P, const std::vector<double> &in, std::vector<double> &out) {
  const long long nb = (long long)in.size() / BLEN;
#pragma omp parallel for schedule(static)
  for (long long b = 0; b < nb; b++) {
    const double *r = &in[b * BLEN];
    double *z = &out[b * BLEN];
    for (int i = 0; i < BLEN; i++) {
      double s = 0;
      for (int k = 0; k < BLEN; k++) s += P[i * BLEN + k] * r[k];
      z[i] = s;
    }
  }
}

// ---- sparse Cholesky block preconditioner of AMRSolver ---------------------
struct Chol {
  std::vector<double> Ld;                                  // 1 / L_ii
  std::vector<std::vector<std::pair<int, double>>> Lrow;   // (J, L_IJ), J<I, |L|>1e-10
  std::vector<std::vector<std::pair<int, double>>> Lcol;   // (I, L_IJ), I>J, |L|>1e-10
};

// getA_local (ExpAMRSolver.cpp:11-23, AMRSolver.cpp:77-96)
double Klocal(int I1, int I2) {
  const int j1 = I1 / BS, i1 = I1 % BS, j2 = I2 / BS, i2 = I2 % BS;
  if (i1 == i2 && j1 == j2) return 4.0;
  if (std::abs(i1 - i2) + std::abs(j1 - j2) == 1) return -1.0;
  return 0.0;
}

// Cholesky-Crout of K (ExpAMRSolver.cpp:53-66, AMRSolver.cpp:105-122).
std::vector<std::vector<double>> cholK() {
  std::vector<std::vector<double>> L(BLEN);
  for (int i = 0; i < BLEN; i++) L[i].assign(i + 1, 0.0);
  for (int i = 0; i < BLEN; i++) {
    double s1 = 0;
    for (int k = 0; k <= i - 1; k++) s1 += L[i][k] * L[i][k];
    L[i][i] = std::sqrt(Klocal(i, i) - s1);
    for (int j = i + 1; j < BLEN; j++) {
      double s2 = 0;
      for (int k = 0; k <= i - 1; k++) s2 += L[i][k] * L[j][k];
      L[j][i] = (Klocal(j, i) - s2) / L[i][i];
    }
  }
  return L;
}

// Sparse factor lists (AMRSolver.cpp:124-146).
Chol makeChol() {
  auto L = cholK();
  Chol C;
  C.Lrow.resize(BLEN);
  C.Lcol.resize(BLEN);
  for (int i = 0; i < BLEN; i++) {
    C.Ld.push_back(1.0 / L[i][i]);
    for (int j = 0; j < i; j++)
      if (std::fabs(L[i][j]) > 1e-10) C.Lrow[i].push_back({j, L[i][j]});
  }
  for (int j = 0; j < BLEN; j++)
    for (int i = j + 1; i < BLEN; i++)
      if (std::fabs(L[i][j]) > 1e-10) C.Lcol[j].push_back({i, L[i][j]});
  return C;
}

// AMRSolver::getZ (AMRSolver.cpp:13-74): z = -(L L^T)^-1 z in place, with the
// traversal mirrored so the result keeps the block's symmetries.
void getZ(const Chol &C, double *in, int level, long long I, long long J, int bpdx, int bpdy) {
  const double c11 = in[(BS - 1) + BS * (BS - 1)];
  const double c10 = in[(BS - 1) + BS * 0];
  const double c01 = in[0 + BS * (BS - 1)];
  const double c00 = in[0];
  bool plus_x = I > (long long)bpdx * (1LL << level) / 2 - 1;
  bool plus_y = J > (long long)bpdy * (1LL << level) / 2 - 1;
  if (std::fabs(std::fabs(c11 + c10) - std::fabs(c01 + c00)) > 1e-14)
    plus_x = std::fabs(c11 + c10) > std::fabs(c01 + c00);
  if (std::fabs(std::fabs(c11 + c01) - std::fabs(c10 + c00)) > 1e-14)
    plus_y = std::fabs(c11 + c01) > std::fabs(c10 + c00);
  const int base_x = plus_x ? 0 : BS - 1;
  const int base_y = plus_y ? 0 : BS - 1;
  const int sx = plus_x ? 1 : -1;
  const int sy = plus_y ? 1 : -1;
  auto at = [&](int K) -> int {
    const int iy = base_y + (sy * K) / BS;
    const int ix = base_x + (sx * K) % BS;
    return iy * BS + ix;
  };
  for (int I1 = 0; I1 < BLEN; I1++) {
    double rhs = 0.0;
    for (const auto &e : C.Lrow[I1]) rhs += e.second * in[at(e.first)];
    const int k = at(I1);
    in[k] = (in[k] - rhs) * C.Ld[I1];
  }
  for (int I1 = BLEN - 1; I1 >= 0; I1--) {
    double rhs = 0.0;
    for (const auto &e : C.Lcol[I1]) rhs += -e.second * in[at(e.first)];
    const int k = at(I1);
    in[k] = (rhs - in[k]) * C.Ld[I1];
  }
}

}  // namespace

extern "C" {

const char *orc_last_error() { return g_err.c_str(); }

int orc_block_size() { return BS; }

// P_inv = -(L^-T L^-1) by Cholesky-Crout + Gauss-Jordan
// (ExpAMRSolver.cpp:37-95).  P is 64x64 row-major.
void orc_pinv(double *P) {
  auto L = cholK();
  std::vector<std::vector<double>> Li(BLEN);
  for (int i = 0; i < BLEN; i++) {
    Li[i].assign(i + 1, 0.0);
    Li[i][i] = 1.0;
  }
  for (int br = 0; br < BLEN; br++) {
    const double bsf = 1. / L[br][br];
    for (int c = 0; c <= br; c++) Li[br][c] *= bsf;
    for (int wr = br + 1; wr < BLEN; wr++) {
      const double wsf = L[wr][br];
      for (int c = 0; c <= br; c++) Li[wr][c] -= wsf * Li[br][c];
    }
  }
  for (int i = 0; i < BLEN; i++)
    for (int j = 0; j < BLEN; j++) {
      double aux = 0.;
      for (int k = 0; k < BLEN; k++) aux += (i <= k && j <= k) ? Li[k][i] * Li[k][j] : 0.;
      P[i * BLEN + j] = -aux;
    }
}

// Assemble the FP64 AMR Laplacian in CSR (ExpAMRSolver::getMat,
// ExpAMRSolver.cpp:180-313).  Rows in block order, cells iy-major.  Returns
// nnz, or -(needed nnz) when cap is too small (arrays untouched), or
// LLONG_MIN on error (see orc_last_error).  *mean_row receives the mean
// constraint row (ExpAMRSolver.cpp:252-256) or -1.
long long orc_assemble(const int *lev, const int *bi, const int *bj, long long nb, int bpdx,
                       int bpdy, int perx, int pery, int mean_constraint, long long *row_ptr,
                       int *col, double *val, long long cap, long long *mean_row) {
  try {
    Topo T;
    std::vector<Block> B;
    if (!buildTopo(T, B, lev, bi, bj, nb, bpdx, bpdy, perx, pery)) return INT64_MIN;
    std::vector<long long> rp(nb * BLEN + 1, 0);
    std::vector<int> cc;
    std::vector<double> vv;
    cc.reserve(6 * nb * BLEN);
    vv.reserve(6 * nb * BLEN);
    for (long long b = 0; b < nb; b++) {
      const int l = B[b].level;
      const long long I = B[b].I, J = B[b].J;
      const bool isBoundary[4] = {I == 0, I == T.nx(l) - 1, J == 0, J == T.ny(l) - 1};
      const bool isPeriodic[2] = {T.perx, T.pery};
      const long long NI[4] = {I - 1, I + 1, I, I};
      const long long NJ[4] = {J, J, J - 1, J + 1};
      for (int iy = 0; iy < BS; iy++)
        for (int ix = 0; ix < BS; ix++) {
          const long long self = cell(b, ix, iy);
          if (ix > 0 && ix < BS - 1 && iy > 0 && iy < BS - 1) {
            // Interior cells, ascending columns (ExpAMRSolver.cpp:270-277).
            const long long c[5] = {cell(b, ix, iy - 1), cell(b, ix - 1, iy), self,
                                    cell(b, ix + 1, iy), cell(b, ix, iy + 1)};
            const double v[5] = {1, 1, -4, 1, 1};
            for (int k = 0; k < 5; k++) {
              cc.push_back((int)c[k]);
              vv.push_back(v[k]);
            }
          } else {
            // Edge cells (ExpAMRSolver.cpp:279-305).
            const bool valid[4] = {ix > 0, ix < BS - 1, iy > 0, iy < BS - 1};
            const long long inner[4] = {cell(b, ix - 1, iy), cell(b, ix + 1, iy),
                                        cell(b, ix, iy - 1), cell(b, ix, iy + 1)};
            Row row;
            for (int f = 0; f < 4; f++) {
              if (valid[f]) {
                row.add(inner[f], 1);
                row.add(self, -1);
              } else if (!isBoundary[f] || isPeriodic[f / 2]) {
                makeFlux(T, B, b, ix, iy, f, NI[f], NJ[f], row);
              }
            }
            for (const auto &e : row.cv) {
              cc.push_back((int)e.first);
              vv.push_back(e.second);
            }
          }
          rp[self + 1] = (long long)cc.size();
        }
    }
    const long long nnz = (long long)cc.size();
    if (mean_row) {
      *mean_row = -1;
      if (mean_constraint) {
        const long long cb = cornerBlock(bi, bj, nb);
        if (cb >= 0) *mean_row = cell(cb, 0, 0);
      }
    }
    if (nnz > cap) return -nnz;
    std::memcpy(row_ptr, rp.data(), sizeof(long long) * rp.size());
    std::memcpy(col, cc.data(), sizeof(int) * nnz);
    std::memcpy(val, vv.data(), sizeof(double) * nnz);
    return nnz;
  } catch (const std::exception &e) {
    g_err = e.what();
    return INT64_MIN;
  }
}

// h^2 per block (ExpAMRSolver.cpp:334).
void orc_h2(const int *lev, long long nb, int bpdx, int bpdy, double extent, double *h2) {
  for (long long b = 0; b < nb; b++) {
    const double h = blockH(lev[b], bpdx, bpdy, extent);
    h2[b] = h * h;
  }
}

void orc_spmv(long long n, const long long *rp, const int *col, const double *val, const double *x,
              double *y, int mean_mode, long long mean_row, const double *h2) {
  Csr A{n, rp, col, val};
  spmv(A, x, y, mean_mode, mean_row, h2);
}

// r = b - A x accumulated in long double; returns ||r||_2 and ||r||_inf and
// writes ||b||_2, ||b||_inf.  mean handling as spmv().
void orc_true_residual(long long n, const long long *rp, const int *col, const double *val,
                       const double *b, const double *x, int mean_mode, long long mean_row,
                       const double *h2, double out[4]) {
  long double m = 0;
  if (mean_mode > 0)
    for (long long i = 0; i < n; i++) m += (long double)h2[i / BLEN] * x[i];
  long double r2 = 0, b2 = 0;
  double rinf = 0, binf = 0;
  for (long long i = 0; i < n; i++) {
    long double s = 0;
    for (long long k = rp[i]; k < rp[i + 1]; k++) s += (long double)val[k] * x[col[k]];
    if (mean_mode == 1 && i == mean_row) s = m;
    if (mean_mode == 2) s += m * h2[i / BLEN];
    const long double r = (long double)b[i] - s;
    r2 += r * r;
    b2 += (long double)b[i] * b[i];
    rinf = std::fmax(rinf, (double)fabsl(r));
    binf = std::fmax(binf, std::fabs(b[i]));
  }
  out[0] = (double)sqrtl(r2);
  out[1] = rinf;
  out[2] = (double)sqrtl(b2);
  out[3] = binf;
}

// Subtract the h^2-weighted mean in place (ExpAMRSolver.cpp:383-415,
// PressureSingle.cpp:853-879).  Returns the mean removed.
double orc_mean_remove(long long nb, const double *h2, double *x) {
  double avg = 0, avg1 = 0;
  for (long long b = 0; b < nb; b++)
    for (int c = 0; c < BLEN; c++) {
      avg += x[b * BLEN + c] * h2[b];
      avg1 += h2[b];
    }
  avg /= avg1;
  for (long long i = 0; i < nb * BLEN; i++) x[i] += -avg;
  return avg;
}

// Stats layout shared by both solvers (doubles):
// [0] iterations  [1] restarts  [2] converged  [3] err_init  [4] err_opt
// [5] err_last
enum { ST_ITERS, ST_RESTARTS, ST_CONV, ST_E0, ST_EOPT, ST_ELAST, ST_N };

// Plain right-preconditioned BiCGSTAB of the reference GPU path
// (BiCGSTABSolver::main, BiCGSTAB.cu:412-620) with the dense P_inv block
// preconditioner.  x holds x0 on entry and x_opt on exit.  The stopping norm
// is ||r||_inf as in the reference (norm_kind 0) or ||r||_2 (norm_kind 1).
// eps is the denominator guard (the reference uses 1e-21, BiCGSTAB.cu:427)
// and breakdown_tol the restart threshold on rho^2/(|r|^2 |rhat|^2) (the
// reference uses 1e-16, BiCGSTAB.cu:483).
void orc_bicgstab(long long n, const long long *rp, const int *col, const double *val,
                  const double *P, const double *b, double *x, int mean_mode, long long mean_row,
                  const double *h2, int max_iter, double max_error, double max_rel_error,
                  int max_restarts, int norm_kind, double eps, double breakdown_tol,
                  double *stats) {
  Csr A{n, rp, col, val};
  auto nrm = [&](const std::vector<double> &v) { return norm_kind ? std::sqrt(dot(v, v)) : amax(v); };
  std::vector<double> r(n), rhat(n), p(n, 0.0), nu(n, 0.0), z(n), t(n), xv(x, x + n), xopt(n);
  double alpha = 1, beta = 1, omega = 1, rho_prev = 1, rho = 1;
  // r = b - A x0 (BiCGSTAB.cu:431-433)
  spmv(A, xv.data(), nu.data(), mean_mode, mean_row, h2);
  for (long long i = 0; i < n; i++) r[i] = b[i] - nu[i];
  double error = nrm(r);
  const double error_init = error;
  double error_opt = error;
  xopt = xv;
  rhat = r;
  std::fill(nu.begin(), nu.end(), 0.0);
  int restarts = 0, k = 0;
  bool conv = false;
  for (k = 0; k < max_iter; k++) {
    rho = dot(rhat, r);
    const double n1 = dot(r, r), n2 = dot(rhat, rhat);
    const bool breakdown = rho * rho < breakdown_tol * n1 * n2;
    beta = (rho / (rho_prev + eps)) * (alpha / (omega + eps));
    if (breakdown && max_restarts > 0) {
      // Restart (BiCGSTAB.cu:488-509).
      restarts++;
      if (restarts >= max_restarts) break;
      rhat = r;
      rho = dot(rhat, rhat);
      std::fill(nu.begin(), nu.end(), 0.0);
      std::fill(p.begin(), p.end(), 0.0);
      rho_prev = 1.;
      alpha = 1.;
      omega = 1.;
      beta = (rho / (rho_prev + eps)) * (alpha / (omega + eps));
    }
    for (long long i = 0; i < n; i++) p[i] = r[i] + beta * (p[i] - omega * nu[i]);
    precDense(P, p, z);
    spmv(A, z.data(), nu.data(), mean_mode, mean_row, h2);
    alpha = rho / (dot(rhat, nu) + eps);
    for (long long i = 0; i < n; i++) {
      xv[i] += alpha * z[i];
      r[i] -= alpha * nu[i];
    }
    precDense(P, r, z);
    spmv(A, z.data(), t.data(), mean_mode, mean_row, h2);
    omega = dot(t, r) / (dot(t, t) + eps);
    for (long long i = 0; i < n; i++) {
      xv[i] += omega * z[i];
      r[i] -= omega * t[i];
    }
    error = nrm(r);
    if (error < error_opt) {
      error_opt = error;
      xopt = xv;
      if (error <= max_error || error / error_init <= max_rel_error) {
        conv = true;
        break;
      }
    }
    rho_prev = rho;
  }
  std::memcpy(x, xopt.data(), sizeof(double) * n);
  stats[ST_ITERS] = conv ? k + 1 : k;
  stats[ST_RESTARTS] = restarts;
  stats[ST_CONV] = conv;
  stats[ST_E0] = error_init;
  stats[ST_EOPT] = error_opt;
  stats[ST_ELAST] = error;
}

// Apply the AMRSolver block preconditioner (AMRSolver.h:141-153) to a whole
// vector: out = getZ(in) per block.
void orc_block_cholesky_apply(long long nb, const int *lev, const int *bi, const int *bj, int bpdx,
                              int bpdy, const double *in, double *out) {
  static const Chol C = makeChol();
  std::memcpy(out, in, sizeof(double) * nb * BLEN);
#pragma omp parallel for schedule(static)
  for (long long b = 0; b < nb; b++) getZ(C, out + b * BLEN, lev[b], bi[b], bj[b], bpdx, bpdy);
}

// Pipelined BiCGSTAB of the reference CPU path (AMRSolver::solve,
// AMRSolver.cpp:149-490; Cools & Vanroose Alg. 11) on the explicit operator.
// The matrix-free Cubism LHS of the reference is replaced by the CSR SpMV.
// Stops on ||r||_2 < tol or ||r||_2/||r0||_2 < rel_tol.  b is not modified
// (the reference overwrites its input grid; callers pass a copy if needed).
void orc_pipelined_bicgstab(long long n, const long long *rp, const int *col, const double *val,
                            const int *lev, const int *bi, const int *bj, int bpdx, int bpdy,
                            const double *bin, double *xio, int mean_mode, long long mean_row,
                            const double *h2, int max_iter, double max_error, double max_rel_error,
                            int max_restarts, double *stats) {
  Csr A{n, rp, col, val};
  const long long nb = n / BLEN;
  static const Chol C = makeChol();
  auto lhs = [&](const std::vector<double> &in, std::vector<double> &out) {
    spmv(A, in.data(), out.data(), mean_mode, mean_row, h2);
  };
  auto prec = [&](const std::vector<double> &in, std::vector<double> &out) {
#pragma omp parallel for schedule(static)
    for (long long i = 0; i < n; i++) out[i] = in[i];
#pragma omp parallel for schedule(static)
    for (long long b2 = 0; b2 < nb; b2++) getZ(C, &out[b2 * BLEN], lev[b2], bi[b2], bj[b2], bpdx, bpdy);
  };
  std::vector<double> phat(n), rhat(n), shat(n), what(n), zhat(n), qhat(n), s(n), w(n), z(n),
      t(n), v(n), q(n), r(n), y(n), x(n), r0(n), b(n), x_opt(n);
  const double eps = 1e-100;
  bool useXopt = false;
  int restarts = 0;
  double min_norm = 1e50, norm_1 = 0, norm_2 = 0;
  for (long long i = 0; i < n; i++) {
    b[i] = bin[i];
    r[i] = bin[i];
    x[i] = xio[i];
  }
  if (mean_mode == 1 && mean_row >= 0) {
    b[mean_row] = 0.0;  // AMRSolver.cpp:205-206
    r[mean_row] = 0.0;
  }
  lhs(x, r0);
#pragma omp parallel for
  for (long long i = 0; i < n; i++) {
    r0[i] = r[i] - r0[i];
    r[i] = r0[i];
  }
  prec(r0, rhat);
  lhs(rhat, w);
  prec(w, what);
  lhs(what, t);
  double alpha = 0, norm = 0, beta = 0, omega = 0, r0r_prev;
  {
    double t0 = 0, t1 = 0, t2 = 0;
#pragma omp parallel for reduction(+ : t0, t1, t2)
    for (long long j = 0; j < n; j++) {
      t0 += r0[j] * r0[j];
      t1 += r0[j] * w[j];
      t2 += r0[j] * r0[j];
    }
    alpha = t0 / (t1 + eps);
    r0r_prev = t0;
    norm = std::sqrt(t2);
  }
  const double init_norm = norm;
  int k;
  bool conv = false;
  for (k = 0; k < max_iter; k++) {
    double qy = 0, yy = 0;
    if (k % 50 != 0) {
#pragma omp parallel for reduction(+ : qy, yy)
      for (long long j = 0; j < n; j++) {
        phat[j] = rhat[j] + beta * (phat[j] - omega * shat[j]);
        s[j] = w[j] + beta * (s[j] - omega * z[j]);
        shat[j] = what[j] + beta * (shat[j] - omega * zhat[j]);
        z[j] = t[j] + beta * (z[j] - omega * v[j]);
        q[j] = r[j] - alpha * s[j];
        qhat[j] = rhat[j] - alpha * shat[j];
        y[j] = w[j] - alpha * z[j];
        qy += q[j] * y[j];
        yy += y[j] * y[j];
      }
    } else {
      // Residual replacement every 50 iterations (AMRSolver.cpp:282-303).
#pragma omp parallel for
      for (long long j = 0; j < n; j++) phat[j] = rhat[j] + beta * (phat[j] - omega * shat[j]);
      lhs(phat, s);
      prec(s, shat);
      lhs(shat, z);
#pragma omp parallel for reduction(+ : qy, yy)
      for (long long j = 0; j < n; j++) {
        q[j] = r[j] - alpha * s[j];
        qhat[j] = rhat[j] - alpha * shat[j];
        y[j] = w[j] - alpha * z[j];
        qy += q[j] * y[j];
        yy += y[j] * y[j];
      }
    }
    prec(z, zhat);
    lhs(zhat, v);
    omega = qy / (yy + eps);
    double r0r = 0, r0w = 0, r0s = 0, r0z = 0;
    norm = 0;
    norm_1 = 0;
    norm_2 = 0;
    if (k % 50 != 0) {
#pragma omp parallel for reduction(+ : r0r, r0w, r0s, r0z, norm_1, norm_2, norm)
      for (long long j = 0; j < n; j++) {
        x[j] = x[j] + alpha * phat[j] + omega * qhat[j];
        r[j] = q[j] - omega * y[j];
        rhat[j] = qhat[j] - omega * (what[j] - alpha * zhat[j]);
        w[j] = y[j] - omega * (t[j] - alpha * v[j]);
        r0r += r0[j] * r[j];
        r0w += r0[j] * w[j];
        r0s += r0[j] * s[j];
        r0z += r0[j] * z[j];
        norm += r[j] * r[j];
        norm_1 += r[j] * r[j];
        norm_2 += r0[j] * r0[j];
      }
    } else {
      // AMRSolver.cpp:352-380.
#pragma omp parallel for
      for (long long j = 0; j < n; j++) x[j] = x[j] + alpha * phat[j] + omega * qhat[j];
      lhs(x, r);
#pragma omp parallel for
      for (long long j = 0; j < n; j++) r[j] = b[j] - r[j];
      prec(r, rhat);
      lhs(rhat, w);
#pragma omp parallel for reduction(+ : r0r, r0w, r0s, r0z, norm_1, norm_2, norm)
      for (long long j = 0; j < n; j++) {
        r0r += r0[j] * r[j];
        r0w += r0[j] * w[j];
        r0s += r0[j] * s[j];
        r0z += r0[j] * z[j];
        norm += r[j] * r[j];
        norm_1 += r[j] * r[j];
        norm_2 += r0[j] * r0[j];
      }
    }
    prec(w, what);
    lhs(what, t);
    norm = std::sqrt(norm);
    beta = alpha / (omega + eps) * r0r / (r0r_prev + eps);
    alpha = r0r / (r0w + beta * r0s - beta * omega * r0z);
    double alphat = 1.0 / (omega + eps) + r0w / (r0r + eps) - beta * omega * r0z / (r0r + eps);
    alphat = 1.0 / (alphat + eps);
    if (std::fabs(alphat) < 10 * std::fabs(alpha)) alpha = alphat;
    r0r_prev = r0r;
    const bool breakdown = r0r * r0r < 1e-16 * norm_1 * norm_2;
    if (breakdown && restarts < max_restarts) {
      // AMRSolver.cpp:421-456.
      restarts++;
#pragma omp parallel for
      for (long long i = 0; i < n; i++) r0[i] = r[i];
      prec(r0, rhat);
      lhs(rhat, w);
      double t0 = 0, t1 = 0;
#pragma omp parallel for reduction(+ : t0, t1)
      for (long long j = 0; j < n; j++) {
        t0 += r0[j] * r0[j];
        t1 += r0[j] * w[j];
      }
      prec(w, what);
      lhs(what, t);
      alpha = t0 / (t1 + eps);
      r0r_prev = t0;
      beta = 0.0;
      omega = 0.0;
    }
    if (norm < min_norm) {
      useXopt = true;
      min_norm = norm;
#pragma omp parallel for
      for (long long i = 0; i < n; i++) x_opt[i] = x[i];
    }
    if (norm < max_error || norm / (init_norm + eps) < max_rel_error) {
      conv = true;
      break;
    }
  }
  const std::vector<double> &sol = useXopt ? x_opt : x;
  std::memcpy(xio, sol.data(), sizeof(double) * n);
  stats[ST_ITERS] = conv ? k + 1 : k;
  stats[ST_RESTARTS] = restarts;
  stats[ST_CONV] = conv;
  stats[ST_E0] = init_norm;
  stats[ST_EOPT] = min_norm;
  stats[ST_ELAST] = norm;
}

int orc_num_threads() { return omp_get_max_threads(); }
void orc_set_num_threads(int n) { omp_set_num_threads(n); }

}  // extern "C"
