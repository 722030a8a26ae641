// assemble_device.h -- builds the AMR Poisson operator on the GPU, in the
// layouts the solver kernels read, from the block descriptions alone: of one
// mesh (the single solver) or of K independent meshes concatenated (the
// batched solver).
//
// The result is the operator that build_operator + EllRows::append produce
// on the host, entry for entry: the same neighbour table, the same explicit
// rows with bit-identical coefficients (summed per column in the same order),
// the same ELL order (by mesh, shortest rows first, mesh order within a
// length) and lengths.  Only the numbering of the coefficient codes differs
// (sorted by bit pattern here, first-seen order on the host); a code decodes
// to the same value.  No plain CSR copy is built.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <functional>
#include <string>
#include <vector>

#include "assembly.h"

namespace cup2d {

constexpr int MAX_MESHES = 4096;

struct DeviceOperator {
  long long nb = 0, n = 0, nnz = 0;            // over all meshes
  long long mean_row = -1, corner_block = -1;  // of mesh 0
  long long n_exc = 0;  // explicit rows
  int ell_w = 0;        // widest explicit row
  int n_codes = 0;      // distinct coefficients (etab entries)
  // device arrays, allocated through the caller's allocator
  int4 *nbr = nullptr;                // [nb] faces Xm, Xp, Ym, Yp
  int2 *fid = nullptr;                // [4 nb] coarse-fine neighbours per face (Operator::fid)
  unsigned long long *mask = nullptr; // [nb] explicit cells
  double *h2 = nullptr;               // [nb]
  int *erow = nullptr;                // [n_exc]
  int *ecol = nullptr;                // [ell_w][n_exc]
  uint16_t *ecode = nullptr;          // [ell_w][n_exc]
  double *etab = nullptr;             // [n_codes], etab[0] = 0.0
  uint8_t *ewid = nullptr;            // [(n_exc + 63) / 64] width of each run of 64 rows
  // host copies
  std::vector<long long> boff;       // [K+1] first block of each mesh
  std::vector<long long> eoff;       // [K+1] first explicit row (ELL slot) of each mesh
  std::vector<long long> mean_rows;  // [K] global mean-constraint row, -1 when unused
  std::vector<uint8_t> elen;         // [n_exc] length of each explicit row, ELL order
};

// Device memory: returns nullptr on failure; the caller owns what it hands
// out.
using DeviceAlloc = std::function<void *(size_t bytes)>;

// Assembles the operator of the K meshes (host arrays level/bx/by) on
// `stream`, mesh k's blocks following mesh k-1's.  `alloc` provides the
// outputs, `temp` the work arrays (which the caller may reuse once the call
// has returned and the stream has drained).  Returns 0, or a CUP2D_PS_* code
// with a message in err: EINVAL for the meshes build_operator rejects (the
// same message, bad_mesh = the first such mesh) and for an explicit row
// wider than ELL_MAX or a batch beyond 32-bit indices (bad_mesh = -1),
// ENOMEM / EHIP for device failures.  Blocks until the outputs are complete.
int assemble_on_device(const MeshDesc *meshes, int K, int mean_constraint, hipStream_t stream,
                       const DeviceAlloc &alloc, const DeviceAlloc &temp, DeviceOperator &out,
                       int &bad_mesh, std::string &err);

}  // namespace cup2d
