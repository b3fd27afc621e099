"""ctypes binding of librest_tensors_amd.so (the C ABI in include/*.h).

Loading never falls back to anything: if the library is missing or lacks a
symbol the import of the caller fails loudly.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "librest_tensors_amd.so")

i32, i64, u64, f64, vp = C.c_int, C.c_int64, C.c_uint64, C.c_double, C.c_void_p
cp = C.c_char_p
pi32 = C.POINTER(C.c_int)
pf64 = C.POINTER(C.c_double)
pi64 = C.POINTER(C.c_int64)
ch = C.c_char

# name -> (restype, argtypes) for the extended ABI (include/rest_tensors_amd.h)
RT_SIGNATURES = {
    "rt_last_error": (cp, []),
    "rt_clear_error": (None, []),
    "rt_version": (i32, []),
    "rt_device_count": (i32, [pi32]),
    "rt_set_device": (i32, [i32]),
    "rt_device_name": (i32, [cp, i32]),
    "rt_malloc": (i32, [C.POINTER(vp), C.c_size_t]),
    "rt_free": (i32, [vp]),
    "rt_memcpy_h2d": (i32, [vp, vp, C.c_size_t]),
    "rt_memcpy_d2h": (i32, [vp, vp, C.c_size_t]),
    "rt_memcpy_d2d": (i32, [vp, vp, C.c_size_t]),
    "rt_memset": (i32, [vp, i32, C.c_size_t]),
    "rt_synchronize": (i32, []),
    "rt_stream_create": (i32, [C.POINTER(vp)]),
    "rt_stream_destroy": (i32, [vp]),
    "rt_stream_synchronize": (i32, [vp]),
    "rt_event_create": (i32, [C.POINTER(vp)]),
    "rt_event_destroy": (i32, [vp]),
    "rt_event_record": (i32, [vp, vp]),
    "rt_event_elapsed_ms": (i32, [vp, vp, C.POINTER(C.c_float)]),
    "rt_kernel_timer_enable": (i32, [i32]),
    "rt_kernel_timer_reset": (i32, []),
    "rt_kernel_timer_get": (i32, [cp, pf64, pi64]),
    "rt_gen_ri3ao_dev": (i32, [vp, i64, i64, u64, f64, vp]),
    "rt_gen_normal_dev": (i32, [vp, i64, u64, vp]),
    "rt_dgemm_dev": (i32, [ch, ch, i64, i64, i64, f64, vp, i64, vp, i64, f64, vp, i64, vp]),
    "rt_dgemm_strided_batched_dev": (
        i32, [ch, ch, i64, i64, i64, f64, vp, i64, i64, vp, i64, i64, f64, vp, i64, i64, i64, vp]),
    "rt_general_dgemm_dev": (
        i32, [vp, i64, i64, i64, i64, i64, i64, ch,
              vp, i64, i64, i64, i64, i64, i64, ch,
              vp, i64, i64, i64, i64, i64, i64, f64, f64, vp]),
    "rt_ao2mo_dev": (i32, [vp, vp, vp, i64, i64, i64, vp]),
    "rt_ao2mo_occ_dev": (i32, [vp, i64, i64, i64, vp, i64, vp, i32, vp]),
    "rt_density_dev": (i32, [vp, i64, i64, vp, vp]),
    "rt_rij_dev": (i32, [vp, i64, i64, vp, vp, vp, pf64, vp]),
    "rt_rik_dev": (i32, [vp, i64, i64, i64, vp, pf64, i32, vp]),
    "rt_exchange_energy_dev": (i32, [vp, i64, i64, i64, pf64, vp]),
    "rt_ozaki_i8_group_size": (i64, [i64, i64, i64]),
    "rt_dgemv_dev": (i32, [ch, i64, i64, f64, vp, i64, vp, i64, f64, vp, i64, vp]),
    "rt_dsyrk_dev": (i32, [ch, ch, i64, i64, f64, vp, i64, f64, vp, i64, vp]),
    "rt_einsum_ip_ip_p_dev": (i32, [vp, i64, i64, vp, i64, i64, vp, vp]),
    "rt_dpotrf_dev": (i32, [ch, i64, vp, i64, pi32, vp]),
    "rt_cholesky_decompose_inverse_dev": (i32, [ch, i64, vp, i64, vp, i64, vp]),
    "rt_matrix_power_dev": (i32, [i64, vp, i64, f64, f64, vp, i64, vp]),
    "rt_dsyev_dev": (i32, [ch, i64, vp, i64, vp, pi32, vp]),
    "rt_dsteqr_host": (i32, [ch, i64, vp, vp, vp, i64, pi32]),
    "rt_dspgvx_dev": (i32, [i64, vp, vp, i64, vp, vp, pi32, vp]),
    "rt_dspevx_dev": (i32, [i64, vp, vp, vp, pi32, vp]),
    "rt_fock_rhf_packed_dev": (i32, [i64, vp, vp, vp, vp, vp]),
    "rt_upper_unpack_dev": (i32, [vp, i64, vp, vp]),
    "rt_upper_pack_dev": (i32, [vp, i64, vp, vp]),
    "rt_rifull_to_matfull_symm_dev": (i32, [vp, i64, i64, vp, vp]),
    "rt_upper_dim": (i64, [i64]),
    "rt_rifull_transpose_dev": (i32, [vp, i64, i64, i64, C.c_char_p, vp, vp]),
    "rt_special_dgemm_dev": (i32, [vp, i64, i64, i64, i64, i64, i64, i64, vp, i64, i64, i64,
                                   i64, i64, i64, f64, f64, vp]),
    "rt_copy_mm_dev": (i32, [i64, i64, vp, i64, i64, i64, i64, vp, i64, i64, i64, i64, vp]),
    "rt_copy_mr_dev": (i32, [i64, i64, vp, i64, i64, i64, i64, vp, i64, i64, i64, i64, i64,
                             i64, i32, vp]),
    "rt_copy_rm_dev": (i32, [i64, i64, vp, i64, i64, i64, i64, i64, i64, i32, vp, i64, i64,
                             i64, i64, vp]),
    "rt_copy_rr_dev": (i32, [i64, i64, i64, vp, i64, i64, i64, i64, i64, i64, vp, i64, i64,
                             i64, i64, i64, i64, vp]),
    # Davidson / RI operator / MLP
    "rt_davidson_solve_dev": (i32, [vp, vp, vp, vp, i64, vp, pi32, pf64, vp, vp, vp]),
    "rt_casida_op_create": (i32, [vp, i64, i64, i64, f64, vp, C.POINTER(vp)]),
    "rt_casida_op_destroy": (i32, [vp]),
    "rt_casida_matvec": (i32, [vp, vp, vp, i64, i64, vp]),
    "rt_casida_hdiag": (i32, [vp, vp, vp]),
    "rt_casida_si_create": (i32, [vp, f64, C.POINTER(vp), vp]),
    "rt_casida_si_destroy": (i32, [vp]),
    "rt_casida_si_apply": (i32, [vp, vp, vp, i64, i64, vp]),
    "rt_casida_si_info": (i32, [vp, pf64, pf64]),
    "rt_mlp_num_params": (i32, []),
    "rt_mlp_create": (i32, [u64, C.POINTER(vp)]),
    "rt_mlp_destroy": (i32, [vp]),
    "rt_mlp_get_params": (i32, [vp, C.POINTER(C.c_float)]),
    "rt_mlp_set_params": (i32, [vp, C.POINTER(C.c_float)]),
    "rt_mlp_apply_dev": (i32, [vp, vp, vp, vp, vp, pf64, pf64, i64, i32, vp, vp]),
    "rt_mlp_loss_grad_dev": (i32, [vp, vp, vp, vp, vp, pf64, pf64, vp, pf64, i64, i32,
                                   vp, pi64, pi32, pf64, pf64, vp]),
    "rt_mlp_train_dev": (i32, [vp, vp, vp, vp, vp, pf64, pf64, vp, pf64, i64, i32,
                               vp, pi64, pi32, i32, i32, f64, u64, pf64, vp]),
    "rt_mlp_targets_dev": (i32, [vp, pi32, i32, vp, i64, f64, vp, pi64, pi32, vp, pf64,
                                 vp]),
    # batched pipeline
    "rt_batch_create": (i32, [vp, i32, C.POINTER(vp)]),
    "rt_batch_autotune": (i32, [vp, i32, vp]),
    "rt_batch_set_stage": (i32, [vp, i32, i32, i32, i32, i32]),
    "rt_batch_reload": (i32, [vp, vp]),
    "rt_batch_run": (i32, [vp, vp]),
    "rt_batch_graph_capture": (i32, [vp, vp]),
    "rt_batch_graph_launch": (i32, [vp, vp]),
    "rt_batch_energies": (i32, [vp, pf64, pf64, vp]),
    "rt_batch_info": (i32, [vp, cp, i32]),
    "rt_batch_destroy": (i32, [vp]),
}

# the seven drop-in symbols (include/restmatr.h): every argument by reference
DROPIN_SIGNATURES = {
    "ri_ao2mo_f_": (None, [vp, vp, vp, pi32, pi32, pi32]),
    "general_dgemm_f_": (None, [vp, pi32, pi32, pi32, pi32, pi32, pi32, C.c_char_p,
                                vp, pi32, pi32, pi32, pi32, pi32, pi32, C.c_char_p,
                                vp, pi32, pi32, pi32, pi32, pi32, pi32, pf64, pf64]),
    "special_dgemm_f_01_": (None, [vp, pi32, pi32, pi32, pi32, pi32, pi32, pi32, pi32,
                                   vp, pi32, pi32, pi32, pi32, pi32, pi32, pf64, pf64]),
    "copy_mm_": (None, [pi32, pi32, vp, pi32, pi32, pi32, pi32, vp, pi32, pi32, pi32, pi32]),
    "copy_mr_": (None, [pi32, pi32, vp, pi32, pi32, pi32, pi32, vp, pi32, pi32, pi32, pi32,
                        pi32, pi32, pi32]),
    "copy_rm_": (None, [pi32, pi32, vp, pi32, pi32, pi32, pi32, pi32, pi32, pi32, vp, pi32,
                        pi32, pi32, pi32]),
    "copy_rr_": (None, [pi32, pi32, pi32, vp, pi32, pi32, pi32, pi32, pi32, pi32, vp, pi32,
                        pi32, pi32, pi32, pi32, pi32]),
}

EXTRA_SIGNATURES: dict = {}

_lib = None


def register(sigs: dict):
    """Adds signatures (used by modules that grow the ABI)."""
    EXTRA_SIGNATURES.update(sigs)
    if _lib is not None:
        _apply(_lib, sigs)


def _apply(lib, sigs):
    for name, (res, args) in sigs.items():
        fn = getattr(lib, name)  # AttributeError if the symbol is missing
        fn.restype = res
        fn.argtypes = args


def lib():
    """Returns the loaded library; raises if it is not built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"{LIB_PATH} is not built; run `python -m rest_tensors_amd.build`")
        l = C.CDLL(LIB_PATH, mode=C.RTLD_GLOBAL)
        _apply(l, RT_SIGNATURES)
        _apply(l, DROPIN_SIGNATURES)
        _apply(l, EXTRA_SIGNATURES)
        _lib = l
    return _lib


class RtError(RuntimeError):
    def __init__(self, status: int, msg: str):
        super().__init__(f"status {status}: {msg}")
        self.status = status


def check(status: int):
    if status != 0:
        raise RtError(status, lib().rt_last_error().decode(errors="replace"))
