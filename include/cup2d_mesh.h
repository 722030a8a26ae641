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

/* Adapt an existing mesh (the decision of Cubism's MeshAdaptation that
 * AdaptTheMesh drives, AdaptTheMesh.cpp:73-122): tag[b] = +1 refine, -1
 * compress, 0 leave, per old block.  Tagged blocks below the top level are
 * split; four sibling leaves all tagged -1 are merged when their parent
 * keeps the 2:1 balance; the tree is balanced again (which may split more).
 * Writes the new blocks in Hilbert order and, per new block b, where its
 * values come from: src_kind[b] = 0 the old block src[4b] unchanged, 1 the
 * old ancestor src[4b] (refinement by one level or more), 2 the four old
 * children src[4b..4b+3] (quadrant q = (x & 1) + 2 (y & 1)); unused src
 * entries are -1.  Returns the new block count (or -1); the arrays are
 * written only when cap >= it. */
long long cup2d_mesh_adapt(int bpdx, int bpdy, int levels, int periodic_x, int periodic_y,
                           long long nb_old, const int *level_old, const int *bx_old,
                           const int *by_old, const signed char *tag, int *level, int *bx, int *by,
                           int *src_kind, long long *src, long long cap);

/* Message of the last failure in this library (thread-unsafe). */
const char *cup2d_mesh_last_error(void);

#ifdef __cplusplus
}
#endif

#endif
