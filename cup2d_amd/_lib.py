"""ctypes binding of libcup2d_poisson.so (include/cup2d_poisson.h).

Loading fails loudly when the library has not been built: the solver has no
fallback path.
"""

from __future__ import annotations

import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libcup2d_poisson.so")

_P = ctypes.c_void_p
_I = ctypes.c_int
_D = ctypes.c_double
_LL = ctypes.c_longlong


class MeshDesc(ctypes.Structure):
    _fields_ = [("bpdx", _I), ("bpdy", _I), ("levels", _I), ("periodic_x", _I), ("periodic_y", _I),
                ("extent", _D), ("n_blocks", _LL), ("level", _P), ("bx", _P), ("by", _P)]


class Config(ctypes.Structure):
    _fields_ = [("device", _I), ("precision", _I), ("mean_constraint", _I), ("norm", _I),
                ("max_iter", _I), ("chunk", _I), ("use_graph", _I), ("profile", _I),
                ("item_blocks", _I), ("schedule", _I), ("xcd_run", _I), ("reserved_i", _I * 1),
                ("eps", _D), ("breakdown_tol", _D), ("inner_tol", _D), ("reserved_d", _D * 5)]


class Stats(ctypes.Structure):
    _fields_ = [("iterations", _I), ("restarts", _I), ("converged", _I), ("refinements", _I),
                ("err_init", _D), ("err_final", _D), ("ms", _D), ("true_rel2", _D),
                ("true_relinf", _D), ("ms_phase", _D * 5), ("reserved", _D * 2)]

    def as_dict(self):
        return {"iterations": self.iterations, "restarts": self.restarts,
                "converged": bool(self.converged), "refinements": self.refinements,
                "err_init": self.err_init, "err_final": self.err_final, "ms": self.ms,
                "true_rel2": self.true_rel2, "true_relinf": self.true_relinf,
                "ms_phase": dict(zip(PHASES, list(self.ms_phase)))}


PHASES = ("spmv", "prec", "update", "scalar", "setup")  # cup2d_ps_phase order


# every symbol of include/cup2d_poisson.h with (restype, argtypes)
SIGNATURES = {
    "cup2d_ps_config_default": (None, [_P]),
    "cup2d_ps_create": (_I, [_P, _P]),
    "cup2d_ps_destroy": (_I, [_P]),
    "cup2d_ps_set_mesh": (_I, [_P, _P]),
    "cup2d_ps_set_mesh_on_device": (_I, [_P, _P]),
    "cup2d_ps_solve": (_I, [_P, _P, _P, _D, _D, _I, _P]),
    "cup2d_ps_solve_device": (_I, [_P, _P, _P, _D, _D, _I, _P]),
    "cup2d_ps_solve_device_noalloc": (_I, [_P, _P, _P, _D, _D, _I, _P]),
    "cup2d_ps_advect_diffuse_device": (_I, [_P, _P, _P, _D, _D, _D, _D]),
    "cup2d_ps_penalize_device": (_I, [_P, _P, _D, _D, _D, _D, _P, _P]),
    "cup2d_ps_chi_disk_device": (_I, [_P, _D, _D, _D, _P]),
    "cup2d_ps_max_speed_device": (_I, [_P, _P, _P, _D, _D, _P]),
    "cup2d_ps_block_score_device": (_I, [_P, _P, _P, _P, _D, _P]),
    "cup2d_ps_transfer_device": (_I, [_P, _P, _P, _P, _P, _I, _P, _P]),
    "cup2d_ps_pressure_step_device": (_I, [_P, _P, _P, _D, _D, _I, _P]),
    "cup2d_ps_pressure_rhs_device": (_I, [_P, _P, _P, _P, _P, _P, _D, _P]),
    "cup2d_ps_pressure_correct_device": (_I, [_P, _P, _D, _P, _P]),
    "cup2d_ps_spmv_device": (_I, [_P, _P, _P, _I]),
    "cup2d_ps_spmv_time": (_I, [_P, _P, _P, _I, _I, _P]),
    "cup2d_ps_kernel_time": (_I, [_P, _I, _I, _P]),
    "cup2d_ps_precond_device": (_I, [_P, _P, _P]),
    "cup2d_ps_precond_bf16_device": (_I, [_P, _P, _P]),
    "cup2d_ps_residual_device": (_I, [_P, _P, _P, _P]),
    "cup2d_ps_sizes": (_I, [_P, _P, _P, _P, _P]),
    "cup2d_ps_get_csr": (_I, [_P, _P, _P, _P]),
    "cup2d_op_csr": (_LL, [_P, _I, _P, _P, _P, _LL, _P]),
    "cup2d_ps_dmalloc": (_I, [_P, ctypes.c_size_t, _P]),
    "cup2d_ps_dfree": (_I, [_P, _P]),
    "cup2d_ps_h2d": (_I, [_P, _P, _P, ctypes.c_size_t]),
    "cup2d_ps_d2h": (_I, [_P, _P, _P, ctypes.c_size_t]),
    "cup2d_ps_last_error": (ctypes.c_char_p, [_P]),
    "cup2d_psb_create": (_I, [_P, _P]),
    "cup2d_psb_destroy": (_I, [_P]),
    "cup2d_psb_set_meshes": (_I, [_P, _I, _P]),
    "cup2d_psb_set_meshes_on_device": (_I, [_P, _I, _P]),
    "cup2d_psb_pressure_step_device": (_I, [_P, _P, _P, _D, _D, _I, _P]),
    "cup2d_psb_pressure_rhs_device": (_I, [_P, _P, _P, _P, _P, _P, _D, _P]),
    "cup2d_psb_pressure_correct_device": (_I, [_P, _P, _D, _P, _P]),
    "cup2d_psb_offsets": (_I, [_P, _P]),
    "cup2d_psb_solve": (_I, [_P, _P, _P, _D, _D, _I, _P]),
    "cup2d_psb_solve_device": (_I, [_P, _P, _P, _D, _D, _I, _P]),
    "cup2d_psb_autotune": (_I, [_P, _P, _P, _P]),
    "cup2d_psb_item_blocks": (_I, [_P]),
    "cup2d_psb_xcd_run": (_I, [_P]),
    "cup2d_psb_spmv_device": (_I, [_P, _P, _P]),
    "cup2d_psb_spmv_time": (_I, [_P, _P, _P, _I, _P]),
    "cup2d_psb_kernel_time": (_I, [_P, _I, _I, _P]),
    "cup2d_psb_copy_time": (_I, [_P, _P, _P, _I, _P]),
    "cup2d_psb_sizes": (_I, [_P, _P, _P, _P]),
    "cup2d_psb_last_error": (ctypes.c_char_p, [_P]),
}

_lib = None


def load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with __graft_entry__.build() "
                               "(the Poisson solver has no non-HIP fallback)")
        lib = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
    return _lib


class SolverError(RuntimeError):
    pass
