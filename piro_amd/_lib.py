"""ctypes binding of libpiro_hip.so (the C ABI declared in include/piro_hip.h).

The library is built in-tree by ``make -C piro_amd/csrc`` (or
``__graft_entry__.build()``).  There is no fallback: if the library is
missing, importing the package raises.
"""
from __future__ import annotations

import ctypes as C
import os
import re

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libpiro_hip.so")
HEADER = os.path.join(os.path.dirname(HERE), "include", "piro_hip.h")


class PiroError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"piro error {code}: {msg}")
        self.code = code


if not os.path.exists(LIB_PATH):
    raise ImportError(
        f"{LIB_PATH} is missing: build it with `make -C piro_amd/csrc` "
        "(the HIP library is the only implementation; there is no CPU fallback)")

lib = C.CDLL(LIB_PATH)

vp = C.c_void_p
i32, i64, f32, f64, u64 = C.c_int, C.c_int64, C.c_float, C.c_double, C.c_uint64
P_i64 = C.POINTER(C.c_int64)
P_i32 = C.POINTER(C.c_int)
P_f32 = C.POINTER(C.c_float)
P_f64 = C.POINTER(C.c_double)

_SIGS = {
    "piro_ctx_create": [i32, C.POINTER(vp)],
    "piro_ctx_destroy": [vp],
    "piro_ctx_set_stream": [vp, vp],
    "piro_ctx_synchronize": [vp],
    "piro_device_count": [P_i32],
    "piro_timer_start": [vp],
    "piro_timer_stop": [vp, P_f32],
    "piro_malloc": [vp, C.c_size_t, C.POINTER(vp)],
    "piro_free": [vp, vp],
    "piro_memcpy_h2d": [vp, vp, vp, C.c_size_t],
    "piro_memcpy_d2h": [vp, vp, vp, C.c_size_t],
    "piro_memcpy_d2d": [vp, vp, vp, C.c_size_t],
    "piro_memset": [vp, vp, i32, C.c_size_t],
    "piro_lap_nnz": [i32, i32, i32, P_i64],
    "piro_assemble_laplacian_csr": [vp, i32, i32, i32, i32, f32, vp, vp, vp],
    "piro_assemble_laplacian_coo": [vp, i32, i32, i32, i32, f32, vp, vp, vp],
    "piro_grad_nnz": [i32, i32, i32, i32, P_i64],
    "piro_assemble_gradient_csr": [vp, i32, i32, i32, i32, i32, f32, vp, vp, vp],
    "piro_row_scale": [vp, i32, vp, vp, vp],
    "piro_assemble_laplacian2d_csr": [vp, i32, f32, vp, vp, vp],
    "piro_ht_table_size": [i32, i32, f32, f32, f32, f32, P_i32],
    "piro_ht_build": [vp, i32, vp, vp, vp, i32, vp, vp],
    "piro_ht_count": [vp, i32, vp, i32, P_i64],
    "piro_ht_to_csr": [vp, i32, vp, vp, i32, vp, vp, vp],
    "piro_coo_row_ptr": [vp, i32, i64, vp, vp],
    "piro_geam3_rowptr": [vp, i32, vp, vp, vp, vp, vp, vp, vp, P_i64],
    "piro_geam3_fill": [vp, i32, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp],
    "piro_geam2_rowptr": [vp, i32, vp, vp, vp, vp, vp, P_i64],
    "piro_geam2_fill": [vp, i32, f32, vp, vp, vp, f32, vp, vp, vp, vp, vp, vp],
    "piro_ic_gaussian": [vp, i32, i32, i32, P_f32, P_f32, P_f32, f32, vp],
    "piro_ic_tophat": [vp, i32, i32, i32, P_f32, P_f32, f32, vp],
    "piro_alpha_uniform": [vp, i64, u64, f32, f32, vp],
    "piro_apply_face_values": [vp, i32, i32, i32, vp, P_f32, vp],
    "piro_set_faces": [vp, i32, i32, i32, vp, vp],
    "piro_fe_step_csr": [vp, i32, i32, i32, vp, vp, vp, i64, f32, vp, vp, vp],
    "piro_fe_run_csr": [vp, i32, i32, i32, vp, vp, vp, i64, f32, i32, vp, vp, vp],
    "piro_fe_step_allrows": [vp, i32, vp, vp, vp, i64, f32, vp, vp],
    "piro_assemble_laplacian_dense": [vp, i32, i32, i32, i32, f32, vp, i64],
    "piro_fe_step_dense": [vp, i32, vp, i64, f32, vp, vp],
    "piro_spmv_csr": [vp, i32, vp, vp, vp, i64, vp, vp],
    "piro_fe_set_variant": [vp, i32],
    "piro_fe_get_variant": [vp, P_i32],
    "piro_fe_autotune": [vp, i32, i32, i32, vp, vp, vp, i64, f32, vp, vp, vp, P_i32, P_f32],
    "piro_reduce_sum": [vp, i64, vp, P_f64],
    "piro_reduce_sumsq": [vp, i64, vp, P_f64],
    "piro_reduce_dot": [vp, i64, vp, vp, P_f64],
    "piro_reduce_absmax": [vp, i64, vp, P_f64],
    "piro_reduce_diff_sumsq": [vp, i64, vp, vp, P_f64],
    "piro_reduce_diff_sumsq_pair_async": [vp, i64, vp, vp, vp],
    "piro_vec_axpby": [vp, i64, f32, vp, f32, vp, vp],
    "piro_vec_copy": [vp, i64, vp, vp],
    "piro_csr_to_dense": [vp, i32, i32, vp, vp, vp, vp, i32],
    "piro_lu_factor_f32": [vp, i32, vp, i32, vp, P_i32],
    "piro_lu_set_speculation": [vp, i32],
    "piro_lu_solve_f32": [vp, i32, vp, i32, vp, i32, vp, i32],
    "piro_lu_solve_refine": [vp, i32, vp, i32, vp, i32, vp, i32, vp, i32, vp, i32, f64, i32, P_i32, P_f64],
    "piro_jli_solve": [vp, i32, i32, i32, vp, vp, vp, i64, f32, vp, f32, f32, i32, vp, vp, P_i32, P_f32],
    "piro_jli_solve_noalloc": [vp, i32, i32, i32, vp, vp, vp, i64, f32, vp, f32, f32, i32, vp, vp, P_i32, P_f32],
    "piro_batch_create": [vp, i32, P_i32, vp, P_f32, C.POINTER(vp)],
    "piro_batch_create_mixed": [vp, i32, P_i32, P_i32, P_f32, vp, P_f32, C.POINTER(vp)],
    "piro_batch_set_schedule": [vp, i32, i32, i32],
    "piro_batch_get_schedule": [vp, P_i32, P_i32, P_i32],
    "piro_batch_set_xruns": [vp, i32],
    "piro_batch_get_xruns": [vp, P_i32],
    "piro_batch_destroy": [vp],
    "piro_batch_info": [vp, P_i64, P_i64, P_i64],
    "piro_batch_operator": [vp, C.POINTER(vp), C.POINTER(vp), C.POINTER(vp)],
    "piro_batch_set_tiling": [vp, i32, i32],
    "piro_batch_get_tiling": [vp, P_i32, P_i32],
    "piro_batch_autotune": [vp, vp, vp, P_i32, P_i32, P_f32],
    "piro_batch_step": [vp, vp, vp],
    "piro_batch_run": [vp, vp, vp, i32, i32],
    "piro_case_run": [C.c_char_p, C.c_char_p, C.c_char_p, i32, i32, P_i32, P_f32, i64],
    "piro_paraview_write": [C.c_char_p, C.c_char_p, i32, P_i32, P_i32, P_i32, i32, C.POINTER(C.c_char_p), P_i32,
                            C.POINTER(vp)],
    "piro_surrogate_create": [vp, u64, C.POINTER(vp)],
    "piro_surrogate_destroy": [vp],
    "piro_surrogate_get_params": [vp, P_f32],
    "piro_surrogate_set_params": [vp, P_f32],
    "piro_surrogate_forward": [vp, i32, i32, i32, i32, vp, vp, f32, vp],
    "piro_surrogate_forward_add": [vp, i32, i32, i32, i32, vp, vp, f32, vp, vp],
    "piro_surrogate_stencil": [vp, P_f64],
    "piro_surrogate_gradients": [vp, i32, i32, i32, i32, vp, vp, vp, f32, P_f32, P_f64],
    "piro_surrogate_train_step": [vp, i32, i32, i32, i32, vp, vp, vp, f32, f32, P_f64],
}

for _name, _args in _SIGS.items():
    _fn = getattr(lib, _name)
    _fn.argtypes = _args
    _fn.restype = C.c_int

lib.piro_last_error.restype = C.c_char_p
lib.piro_last_error.argtypes = []
lib.piro_version.restype = C.c_int
lib.piro_ctx_get_stream.restype = vp
lib.piro_ctx_get_stream.argtypes = [vp]
lib.piro_lap_norm.restype = C.c_float
lib.piro_lap_norm.argtypes = [C.c_float, C.c_int]
lib.piro_grad_norm.restype = C.c_float
lib.piro_grad_norm.argtypes = [C.c_float, C.c_int]
lib.piro_fe_variant_count.restype = C.c_int
lib.piro_fe_variant_count.argtypes = []
lib.piro_surrogate_num_params.restype = C.c_int
lib.piro_surrogate_num_params.argtypes = []
lib.piro_fe_variant_name.restype = C.c_char_p
lib.piro_fe_variant_name.argtypes = [C.c_int]


def check(rc: int) -> None:
    if rc != 0:
        raise PiroError(rc, lib.piro_last_error().decode(errors="replace"))


def declared_symbols() -> list[str]:
    """Function names declared in include/piro_hip.h."""
    with open(HEADER) as f:
        text = f.read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(piro_[a-z0-9_]+)\s*\(", text)))
