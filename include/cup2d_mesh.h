/* cup2d_mesh.h -- C ABI of the block-quadtree generator (libcup2d_mesh.so).
 *
 * The reference solver takes its mesh from Cubism's block grid
 * (cubism::BlockInfo::{level,index}, read at source/Poisson/ExpAMRSolver.cpp:186-249);
 * Cubism is not part of this library, so this generator produces the same
 * description -- one (level, bx, by) triple per 8x8-cell block, 2:1 balanced
 * over faces and corners, ordered along a Hilbert curve like Cubism's
 * space-filling-curve block order -- for the synthetic workloads.
 */
#ifndef CUP2D_MESH_H
#define CUP2D_MESH_H

#ifdef __cplusplus
extern "C" {
#endif

/* Build the leaves of a quadtree over bpdx x bpdy level-0 blocks with
 * `levels` levels (0..levels-1) and domain extent `extent` (the longer side).
 * A block at level l is split when its distance to one of the `ndisks` filled
 * disks (disks[3k..3k+2] = cx, cy, r) is below bands[l] (bands has levels-1
 * entries).  The tree is then 2:1 balanced.
 * Returns the number of blocks N_b (or -1 on invalid arguments).  The arrays
 * level/bx/by are written only when cap >= N_b; call once with cap = 0 to
 * size them. */
long long cup2d_mesh_graded(int bpdx, int bpdy, int levels, double extent, int periodic_x,
                            int periodic_y, int ndisks, const double *disks, const double *bands,
                            int *level, int *bx, int *by, long long cap);

/* Message of the last failure in this library (thread-unsafe). */
const char *cup2d_mesh_last_error(void);

#ifdef __cplusplus
}
#endif

#endif
