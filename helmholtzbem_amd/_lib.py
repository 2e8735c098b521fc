"""Loader of the C-ABI library helmholtzbem_amd/libhbem.so (include/hbem.h).

The library must have been built (``python -m helmholtzbem_amd.build`` or
``__graft_entry__.build()``); there is no CPU fallback: every entry point runs
on the GPU and raises if the library or a device is missing.
"""
from __future__ import annotations

import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libhbem.so")

HBEM_EINVAL, HBEM_ENOTSUP, HBEM_EDEVICE, HBEM_ENOMEM, HBEM_ENUMERIC = -1, -2, -3, -4, -5

EXPORTED = [
    "hbem_last_error", "hbem_init", "hbem_geometry_create", "hbem_geometry_destroy", "hbem_geometry_info",
    "hbem_assemble_dense", "hbem_mass_matrix", "hbem_projection_factor", "hbem_gen_sol_op", "hbem_projection",
    "hbem_rsvd_sv", "hbem_rsvd_sv_der", "hbem_rsvd_sv_der2", "hbem_sigma_sweep", "hbem_profile_enable",
    "hbem_profile_query", "hbem_assemble_dense_batch", "hbem_tp_A_rhs", "hbem_tp_solve", "hbem_sweep_batch_size",
    "hbem_rsvd_benchmarking", "hbem_rsvd_timers", "hbem_rsvd_reset_timer", "hbem_rsvd_sv_noalloc",
    "hbem_sigma_sweep_der2", "hbem_sigma_sweep_mixed", "hbem_sigma_sweep_multi",
    "hbem_geometry_release_work",
]

_lib = None


class HbemError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(msg)
        self.code = code


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("libhbem.so is not built; run python -m helmholtzbem_amd.build")
        L = ctypes.CDLL(LIB_PATH)
        vp, dp, ip, i, d = ctypes.c_void_p, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_int), \
            ctypes.c_int, ctypes.c_double
        L.hbem_last_error.argtypes = [ctypes.c_char_p, ctypes.c_size_t]
        L.hbem_init.argtypes = [i]
        L.hbem_geometry_create.argtypes = [dp, i, dp, i, i, i, i, ctypes.POINTER(vp)]
        L.hbem_geometry_destroy.argtypes = [vp]
        L.hbem_geometry_info.argtypes = [vp, ip, ip, ip, ip]
        L.hbem_assemble_dense.argtypes = [vp, d, d, d, i, i, vp, vp, vp]
        L.hbem_assemble_dense_batch.argtypes = [vp, dp, i, d, i, i, vp, vp, vp]
        L.hbem_mass_matrix.argtypes = [vp, dp]
        L.hbem_projection_factor.argtypes = [vp, dp, ip]
        L.hbem_gen_sol_op.argtypes = [vp, d, d, d, d, i, vp, vp, vp]
        L.hbem_projection.argtypes = [vp, vp, vp]
        L.hbem_rsvd_sv.argtypes = [vp, i, vp, i, i, dp]
        L.hbem_rsvd_sv_der.argtypes = [vp, vp, i, vp, i, i, dp]
        L.hbem_rsvd_sv_der2.argtypes = [vp, vp, vp, i, vp, i, i, dp]
        L.hbem_sigma_sweep.argtypes = [vp, dp, i, d, d, vp, i, i, i, dp, dp]
        L.hbem_profile_enable.argtypes = [vp, i]
        L.hbem_profile_query.argtypes = [vp, i, dp, ctypes.POINTER(ctypes.c_longlong), dp]
        L.hbem_tp_A_rhs.argtypes = [vp, d, d, d, d, vp, i, vp, vp]
        L.hbem_tp_solve.argtypes = [vp, d, d, d, d, vp, i, vp]
        L.hbem_sweep_batch_size.argtypes = [i, i, i, i, ctypes.c_ulonglong, i, ip]
        L.hbem_rsvd_benchmarking.argtypes = [i]
        L.hbem_rsvd_timers.argtypes = [dp]
        L.hbem_rsvd_reset_timer.argtypes = []
        L.hbem_rsvd_sv_noalloc.argtypes = [vp, i, vp, i, i, vp, vp]
        L.hbem_sigma_sweep_der2.argtypes = [vp, dp, i, d, d, vp, i, i, dp, dp, dp]
        L.hbem_sigma_sweep_mixed.argtypes = [vp, dp, i, d, d, vp, i, i, i, i, dp, ip]
        L.hbem_geometry_release_work.argtypes = [vp]
        L.hbem_sigma_sweep_multi.argtypes = [ctypes.POINTER(vp), i, dp, i, d, dp, ctypes.POINTER(vp), i, i, i, dp]
        _lib = L
    return _lib


def check(rc: int):
    if rc != 0:
        buf = ctypes.create_string_buffer(1024)
        lib().hbem_last_error(buf, 1024)
        raise HbemError(rc, buf.value.decode())
