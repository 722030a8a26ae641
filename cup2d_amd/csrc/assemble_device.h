// assemble_device.h -- builds the AMR Poisson operator on the GPU, in the
// layouts the solver kernels read, from the block description alone.
//
// The result is the operator that build_operator + EllRows::append produce
// on the host, entry for entry: the same neighbour table, the same explicit
// rows with bit-identical coefficients (summed per column in the same order),
// the same ELL order (shortest rows first, mesh order within a length) and
// the same run widths.  Only the numbering of the coefficient codes differs
// (sorted by bit pattern here, first-seen order on the host); a code decodes
// to the same value.  No plain CSR copy is built.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <functional>
#include <string>

#include "assembly.h"

namespace cup2d {

struct DeviceOperator {
  long long nb = 0, n = 0, nnz = 0;
  long long mean_row = -1, corner_block = -1;
  long long n_exc = 0;  // explicit rows
  int ell_w = 0;        // widest explicit row
  int n_codes = 0;      // distinct coefficients (etab entries)
  // device arrays, allocated through the caller's allocator
  int4 *nbr = nullptr;                // [nb] faces Xm, Xp, Ym, Yp
  unsigned long long *mask = nullptr; // [nb] explicit cells
  double *h2 = nullptr;               // [nb]
  int *erow = nullptr;                // [n_exc]
  int *ecol = nullptr;                // [ell_w][n_exc]
  uint16_t *ecode = nullptr;          // [ell_w][n_exc]
  double *etab = nullptr;             // [n_codes], etab[0] = 0.0
  uint8_t *ewid = nullptr;            // [(n_exc + 63) / 64] width of each run of 64 rows
};

// Device memory: returns nullptr on failure; the caller owns what it hands
// out.
using DeviceAlloc = std::function<void *(size_t bytes)>;

// Assembles the operator of mesh m (host arrays level/bx/by) on `stream`.
// `alloc` provides the outputs, `temp` the work arrays (which the caller may
// reuse once the call has returned and the stream has drained).
// Returns 0, or a CUP2D_PS_* code with a message in err: EINVAL for the
// meshes build_operator rejects (same messages) and for an explicit row
// wider than ELL_MAX, ENOMEM / EHIP for device failures.  Blocks until the
// outputs are complete.
int assemble_on_device(const MeshDesc &m, int mean_constraint, hipStream_t stream,
                       const DeviceAlloc &alloc, const DeviceAlloc &temp, DeviceOperator &out,
                       std::string &err);

}  // namespace cup2d
