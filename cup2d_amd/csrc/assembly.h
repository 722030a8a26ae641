// assembly.h -- host-side description of the AMR Poisson operator in the
// layouts the HIP kernels read.
//
// The operator is the h^2-scaled finite-volume Laplacian of the reference
// (source/Poisson/ExpAMRSolver.cpp:180-313).  It is stored two ways:
//
//  * stencil form (what the fused solver kernels read): every row whose cell
//    has no face on a coarse-fine block boundary is the plain 5-point stencil
//    sum_f (x_nb(f) - x_self) over the non-wall faces, described by one
//    same-level neighbour id per block face (nbr[4*b+f], -1 = wall, -2 =
//    coarse-fine).  Rows with a coarse-fine face ("explicit rows") are stored
//    as a small CSR (exc_*), grouped by block.
//  * plain CSR (int32 columns) of the whole matrix, for the CSR SpMV kernel
//    and for the host-side checks.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

namespace cup2d {

constexpr int BS = 8;
constexpr int BLEN = BS * BS;

struct MeshDesc {
  int bpdx, bpdy, levels;
  int periodic_x, periodic_y;
  double extent;
  long long n_blocks;
  const int *level, *bx, *by;
};

struct Operator {
  long long nb = 0;  // blocks
  long long n = 0;   // unknowns = 64 nb
  // stencil form
  std::vector<int> nbr;          // [4 nb] faces Xm, Xp, Ym, Yp
  // [8 nb] coarse-fine neighbours of each face, a pair per face: one level
  // coarser (coarse block, -1 - half), half = which half of the coarse
  // block's face this block touches (its index parity along the face);
  // one level finer (the two fine blocks, low then high along the face);
  // (-1, -1) otherwise
  std::vector<int> fid;
  std::vector<uint64_t> exc_mask;  // [nb] bit c set when cell c is an explicit row
  std::vector<int> exc_bptr;     // [nb+1] explicit rows of block b: exc_bptr[b]..exc_bptr[b+1]
  std::vector<int> exc_cell;     // [n_exc] cell index (0..63) of each explicit row
  std::vector<int> exc_rptr;     // [n_exc+1] entry range of each explicit row
  std::vector<int> exc_col;      // [nnz_exc] global column
  std::vector<double> exc_val;   // [nnz_exc]
  // full CSR
  std::vector<int> rp;           // [n+1]
  std::vector<int> col;          // [nnz]
  std::vector<double> val;       // [nnz]
  // h^2 per block and the mean-constraint row (-1 when unused)
  std::vector<double> h2;
  long long mean_row = -1;
  long long corner_block = -1;   // block with index (0,0), whose cell (0,0) is the mean row

  long long nnz_rows = 0;        // nonzeros of the whole matrix (also without the CSR copy)
  long long nnz() const { return nnz_rows; }
  long long n_exc() const { return (long long)exc_cell.size(); }
};

// Builds the operator for the mesh.  Returns false with a message on a mesh
// the operator cannot be built on (out-of-range or duplicate blocks, a face
// whose neighbour is neither same-level, one level coarser nor one level
// finer).
// with_csr = false skips the plain CSR copy (rp, col, val stay empty).
bool build_operator(const MeshDesc &m, int mean_constraint, Operator &op, std::string &err,
                    bool with_csr = true);

// Explicit rows in ELL form.  ell_width: most entries of an explicit row
// (0 when there are none).  EllRows::append adds op's explicit rows with
// global row / column offset row_off, W entries per row (row-major), padded
// with (own row, 0.0), shortest rows first (a stable sort, so rows of one
// length keep their mesh order): a run of rows then shares one width and
// the kernels stop at it.  A coefficient is stored as a 16-bit code into
// `values`, the distinct FP64 coefficients by bit pattern (code 0 is 0.0);
// the interface stencils have a few dozen.  append fails when op has a row
// wider than W or the codes run out.  ell_transpose: row-major [n][W] to
// the kernels' column-major [W][n].
int ell_width(const Operator &op);
struct EllRows {
  int W = 0;
  std::vector<int> erow;           // [n] global row
  std::vector<int> ecol_rm;        // [n][W] global column
  std::vector<uint16_t> ecode_rm;  // [n][W] coefficient code
  std::vector<uint8_t> len;        // [n] entries of the row
  std::vector<double> values{0.0};
  bool append(const Operator &op, long long row_off, std::string &err);
  // widest row in [e0, e1)
  int width(long long e0, long long e1) const;
 private:
  std::vector<std::pair<uint64_t, uint16_t>> codes_{{0, 0}};  // sorted by bit pattern
  int code(double v);  // -1 when the table is full
};
template <class T>
std::vector<T> ell_transpose(const std::vector<T> &rm, int W) {
  const size_t n = W > 0 ? rm.size() / W : 0;
  std::vector<T> cm(rm.size());
  for (size_t e = 0; e < n; e++)
    for (int w = 0; w < W; w++) cm[(size_t)w * n + e] = rm[e * W + w];
  return cm;
}

// P_inv = -K^-1 of the 8x8-block 5-point matrix K, row-major 64x64.
std::vector<double> make_pinv();

// P in the A-operand layout of v_mfma_f32_16x16x32_bf16 rounded to bf16
// (bit patterns): fragment (mt, ks), lane l, element j =
// P[16 mt + (l & 15)][32 ks + 8 (l >> 4) + j].
std::vector<uint16_t> make_pfrag(const std::vector<double> &P);

// P in the A-operand layout of v_mfma_f64_16x16x4f64 (FP64, 64 fragments of
// 64 lanes): fragment (mt, ks), lane l = P[16 mt + o(l & 15)][16 (l >> 4) + ks]
// with the row order o(m) = 4 (m & 3) + (m >> 2), which makes the four output
// registers of a lane four consecutive cells (see tile_prec64).
std::vector<double> make_pfrag64(const std::vector<double> &P);

}  // namespace cup2d
