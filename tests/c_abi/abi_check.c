/* Compiled as C99 and as C++17 against include/cup2d_poisson.h and linked
 * with libcup2d_poisson.so by tests/test_abi.py: the header is valid in both
 * languages, every declared entry point links, and the host-only calls work
 * without a GPU.  Prints "ABI OK <nnz>" and exits 0 on success. */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "cup2d_poisson.h"

/* every entry point, by address: a missing symbol fails the link */
typedef void (*fn_t)(void);
static const fn_t entry_points[] = {
    (fn_t)&cup2d_ps_config_default, (fn_t)&cup2d_ps_create,
    (fn_t)&cup2d_ps_destroy, (fn_t)&cup2d_ps_set_mesh,
    (fn_t)&cup2d_ps_set_mesh_on_device, (fn_t)&cup2d_ps_solve,
    (fn_t)&cup2d_ps_solve_device, (fn_t)&cup2d_ps_solve_device_noalloc,
    (fn_t)&cup2d_ps_pressure_step_device, (fn_t)&cup2d_ps_advect_diffuse_device,
    (fn_t)&cup2d_ps_penalize_device, (fn_t)&cup2d_ps_chi_disk_device,
    (fn_t)&cup2d_ps_max_speed_device, (fn_t)&cup2d_ps_block_score_device,
    (fn_t)&cup2d_ps_transfer_device,
    (fn_t)&cup2d_ps_pressure_rhs_device, (fn_t)&cup2d_ps_pressure_correct_device,
    (fn_t)&cup2d_ps_spmv_device, (fn_t)&cup2d_ps_spmv_time,
    (fn_t)&cup2d_ps_kernel_time, (fn_t)&cup2d_ps_precond_device,
    (fn_t)&cup2d_ps_precond_bf16_device, (fn_t)&cup2d_ps_residual_device,
    (fn_t)&cup2d_ps_sizes, (fn_t)&cup2d_ps_get_csr, (fn_t)&cup2d_op_csr,
    (fn_t)&cup2d_ps_dmalloc, (fn_t)&cup2d_ps_dfree, (fn_t)&cup2d_ps_h2d,
    (fn_t)&cup2d_ps_d2h, (fn_t)&cup2d_ps_last_error,
    (fn_t)&cup2d_psb_create, (fn_t)&cup2d_psb_destroy,
    (fn_t)&cup2d_psb_set_meshes, (fn_t)&cup2d_psb_set_meshes_on_device,
    (fn_t)&cup2d_psb_pressure_step_device, (fn_t)&cup2d_psb_pressure_rhs_device,
    (fn_t)&cup2d_psb_pressure_correct_device, (fn_t)&cup2d_psb_offsets,
    (fn_t)&cup2d_psb_solve, (fn_t)&cup2d_psb_solve_device,
    (fn_t)&cup2d_psb_autotune, (fn_t)&cup2d_psb_item_blocks, (fn_t)&cup2d_psb_xcd_run,
    (fn_t)&cup2d_psb_spmv_device, (fn_t)&cup2d_psb_spmv_time,
    (fn_t)&cup2d_psb_kernel_time, (fn_t)&cup2d_psb_copy_time, (fn_t)&cup2d_psb_sizes,
    (fn_t)&cup2d_psb_last_error,
};

static int fail(const char *what) {
  printf("ABI FAIL %s\n", what);
  return 1;
}

int main(void) {
  size_t k;
  for (k = 0; k < sizeof(entry_points) / sizeof(entry_points[0]); k++)
    if (!entry_points[k]) return fail("null entry point");

  cup2d_ps_config cfg;
  memset(&cfg, 0xff, sizeof(cfg));
  cup2d_ps_config_default(&cfg);
  if (cfg.precision != CUP2D_PS_FP64 || cfg.mean_constraint != 0 || cfg.norm != CUP2D_PS_NORM_INF ||
      cfg.max_iter != 1000 || cfg.use_graph != 1 || cfg.eps != 1e-100 || cfg.breakdown_tol != 1e-16)
    return fail("config defaults");

  /* a uniform 2x1-block mesh: 128 unknowns, 5 * 128 - 2 * (2 + 1) * 8 entries */
  const int level[2] = {0, 0}, bx[2] = {0, 1}, by[2] = {0, 0};
  cup2d_mesh_desc m;
  m.bpdx = 2; m.bpdy = 1; m.levels = 1; m.periodic_x = 0; m.periodic_y = 0;
  m.extent = 2.0; m.n_blocks = 2; m.level = level; m.bx = bx; m.by = by;
  long long mean_row = -7;
  const long long need = cup2d_op_csr(&m, 0, NULL, NULL, NULL, 0, &mean_row);
  if (need != -592) return fail("size query");
  long long *rp = (long long *)malloc(129 * sizeof(long long));
  int *col = (int *)malloc(592 * sizeof(int));
  double *val = (double *)malloc(592 * sizeof(double));
  const long long nnz = cup2d_op_csr(&m, 1, rp, col, val, 592, &mean_row);
  if (nnz != 592 || rp[128] != 592 || mean_row != 0) return fail("assembly");
  int i;
  for (i = 0; i < 128; i++) {
    double s = 0.0;
    long long j;
    for (j = rp[i]; j < rp[i + 1]; j++) s += val[j];
    if (s != 0.0) return fail("row sum");
  }
  free(rp);
  free(col);
  free(val);

  /* an unbalanced mesh is rejected with a message */
  const int l2[2] = {0, 2}, x2[2] = {0, 4}, y2[2] = {0, 0};
  m.levels = 3; m.level = l2; m.bx = x2; m.by = y2;
  if (cup2d_op_csr(&m, 0, NULL, NULL, NULL, 0, &mean_row) != -1) return fail("invalid mesh accepted");
  const char *msg = cup2d_ps_last_error(NULL);
  if (!msg || !msg[0]) return fail("no message");

  /* no usable device here, or a GPU box: creation fails cleanly or succeeds */
  cup2d_ps *h = NULL;
  const int rc = cup2d_ps_create(&cfg, &h);
  if (rc == CUP2D_PS_OK) {
    if (!h || cup2d_ps_destroy(h) != CUP2D_PS_OK) return fail("destroy");
  } else if (rc != CUP2D_PS_ENODEV || h != NULL) {
    return fail("create");
  }
  if (cup2d_ps_solve(NULL, NULL, NULL, 0.0, 1e-10, 10, NULL) != CUP2D_PS_EINVAL) return fail("null handle");
  printf("ABI OK %lld\n", nnz);
  return 0;
}
