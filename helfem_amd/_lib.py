"""ctypes binding of libhfgpu.so (declarations in include/hfgpu.h).

The library is built in-tree (helfem_amd/libhfgpu.so).  There is no
fallback: if the shared object is missing the import fails loudly.
"""
from __future__ import annotations

import ctypes
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "libhfgpu.so"

c_int = ctypes.c_int
c_i64 = ctypes.c_int64
c_dbl = ctypes.c_double
c_vp = ctypes.c_void_p
c_ip = ctypes.POINTER(ctypes.c_int)
c_dp = ctypes.c_void_p

# name -> (restype, argtypes); mirrors include/hfgpu.h
SIGNATURES = {
    "hf_last_error": (ctypes.c_char_p, []),
    "hf_version": (ctypes.c_char_p, []),
    "hf_element_grid": (c_int, [c_dbl, c_int, c_int, c_dbl, c_dp]),
    "hf_ctx_create": (c_int, [c_int, ctypes.c_uint, ctypes.POINTER(c_vp)]),
    "hf_ctx_destroy": (c_int, [c_vp]),
    "hf_ctx_set_stream": (c_int, [c_vp, c_vp]),
    "hf_ctx_get_stream": (c_int, [c_vp, ctypes.POINTER(c_vp)]),
    "hf_ctx_synchronize": (c_int, [c_vp]),
    "hf_ctx_set_timing": (c_int, [c_vp, c_int]),
    "hf_ctx_kernel_stats": (c_int, [c_vp, ctypes.c_char_p, c_vp]),
    "hf_atomic_work": (c_int, [c_vp, c_vp]),
    "hf_atomic_create": (c_int, [c_vp, c_int, c_int, c_int, c_dp, c_int, c_vp, c_vp, c_int, ctypes.POINTER(c_vp)]),
    "hf_atomic_destroy": (c_int, [c_vp]),
    "hf_atomic_clone": (c_int, [c_vp, ctypes.POINTER(c_vp)]),
    "hf_atomic_dftgrid_create_rule": (c_int, [c_vp, c_int, c_int, c_int, ctypes.POINTER(c_vp)]),
    "hf_quadrature": (c_int, [c_int, c_int, c_dp, c_dp]),
    "hf_atomic_sizes": (c_int, [c_vp, c_vp]),
    "hf_atomic_elements": (c_int, [c_vp, c_vp, c_vp]),
    "hf_atomic_compute_tei": (c_int, [c_vp, c_int]),
    "hf_atomic_compute_tei_device": (c_int, [c_vp, c_vp]),
    "hf_atomic_set_cache": (c_int, [c_vp, c_int, c_int, c_int, c_int, c_int, c_dp]),
    "hf_atomic_commit_tei": (c_int, [c_vp]),
    "hf_atomic_get_cache": (c_int, [c_vp, c_int, c_int, c_int, c_vp, c_dp]),
    "hf_atomic_one_electron": (c_int, [c_vp, c_int, c_dp, c_i64]),
    "hf_atomic_one_electron_dev": (c_int, [c_vp, c_int, c_dp, c_i64]),
    "hf_atomic_one_electron_prepare": (c_int, [c_vp]),
    "hf_eigh_batched": (c_int, [c_vp, c_int, c_dp, c_i64, c_i64, c_dp, c_dp, c_vp, c_int]),
    "hf_atomic_coulomb": (c_int, [c_vp, c_dp, c_i64, c_dp, c_i64]),
    "hf_atomic_exchange": (c_int, [c_vp, c_dp, c_i64, c_dp, c_i64]),
    "hf_atomic_coulomb_dev": (c_int, [c_vp, c_dp, c_i64, c_dp, c_i64]),
    "hf_atomic_coulomb_dev_noalloc": (c_int, [c_vp, c_dp, c_i64, c_dp, c_i64]),
    "hf_atomic_batch_create": (c_int, [c_vp, c_vp, c_int, c_vp]),
    "hf_atomic_batch_destroy": (c_int, [c_vp]),
    "hf_atomic_batch_set_options": (c_int, [c_vp, c_int, c_int]),
    "hf_atomic_batch_tuning": (c_int, [c_vp, c_vp, c_vp]),
    "hf_atomic_batch_coulomb_dev": (c_int, [c_vp, c_vp, c_vp, c_vp, c_vp]),
    "hf_atomic_batch_exchange_dev": (c_int, [c_vp, c_vp, c_vp, c_vp, c_vp, c_vp]),
    "hf_atomic_exchange_dev": (c_int, [c_vp, c_dp, c_i64, c_dp, c_i64]),
    "hf_atomic_compute_yukawa": (c_int, [c_vp, c_dbl]),
    "hf_atomic_compute_erfc": (c_int, [c_vp, c_dbl]),
    "hf_atomic_rs_exchange": (c_int, [c_vp, c_dp, c_i64, c_dp, c_i64]),
    "hf_atomic_rs_exchange_dev": (c_int, [c_vp, c_dp, c_i64, c_dp, c_i64]),
    "hf_atomic_exchange_fp32_dev": (c_int, [c_vp, c_dp, c_i64, c_dp, c_i64]),
    "hf_atomic_dftgrid_create": (c_int, [c_vp, c_int, c_int, ctypes.POINTER(c_vp)]),
    "hf_diatomic_dftgrid_create": (c_int, [c_vp, c_int, ctypes.POINTER(c_vp)]),
    "hf_diatomic_dftgrid3d_create": (c_int, [c_vp, c_int, c_int, c_vp]),
    "hf_dftgrid_destroy": (c_int, [c_vp]),
    "hf_xc_exact_exchange": (c_int, [c_int, c_vp]),
    "hf_dftgrid_eval_fxc": (c_int, [c_vp, c_int, c_int, c_dp, c_i64, c_dp, c_i64, c_dbl, c_vp]),
    "hf_dftgrid_eval_fxc_dev": (c_int, [c_vp, c_int, c_int, c_dp, c_i64, c_dp, c_i64, c_dbl, c_vp]),
    "hf_dftgrid_eval_fxc_pol": (c_int, [c_vp, c_int, c_int, c_dp, c_i64, c_dp, c_i64, c_dp, c_i64, c_dp, c_i64,
                                        c_int, c_dbl, c_vp]),
    "hf_dftgrid_eval_fxc_pol_dev": (c_int, [c_vp, c_int, c_int, c_dp, c_i64, c_dp, c_i64, c_dp, c_i64, c_dp,
                                            c_i64, c_int, c_dbl, c_vp]),
    "hf_diatomic_create": (c_int, [c_vp, c_int, c_int, c_dbl, c_int, c_int, c_dp, c_int, c_vp, c_vp, c_int,
                                   ctypes.POINTER(c_vp)]),
    "hf_diatomic_destroy": (c_int, [c_vp]),
    "hf_diatomic_sizes": (c_int, [c_vp, c_vp]),
    "hf_diatomic_pure_index": (c_int, [c_vp, c_vp]),
    "hf_diatomic_maps": (c_int, [c_vp, c_vp, c_vp]),
    "hf_diatomic_set_absm_symmetric": (c_int, [c_vp, c_int]),
    "hf_diatomic_compute_tei": (c_int, [c_vp, c_int]),
    "hf_diatomic_compute_tei_device": (c_int, [c_vp, c_vp]),
    "hf_diatomic_get_cache": (c_int, [c_vp, c_int, c_int, c_int, c_vp, c_dp]),
    "hf_diatomic_one_electron": (c_int, [c_vp, c_int, c_dp, c_i64]),
    "hf_diatomic_one_electron_dev": (c_int, [c_vp, c_int, c_dp, c_i64]),
    "hf_diatomic_work": (c_int, [c_vp, c_vp]),
    "hf_diatomic_coulomb": (c_int, [c_vp, c_dp, c_i64, c_dp, c_i64]),
    "hf_diatomic_exchange": (c_int, [c_vp, c_dp, c_i64, c_dp, c_i64]),
    "hf_diatomic_coulomb_dev": (c_int, [c_vp, c_dp, c_i64, c_dp, c_i64]),
    "hf_diatomic_exchange_dev": (c_int, [c_vp, c_dp, c_i64, c_dp, c_i64]),
    "hf_diatomic_exchange_fp32_dev": (c_int, [c_vp, c_dp, c_i64, c_dp, c_i64]),
}

_LIB = None


class HfError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"hfgpu error {code}: {msg}")
        self.code = code


def lib():
    global _LIB
    if _LIB is None:
        if not LIB_PATH.exists():
            raise RuntimeError(f"{LIB_PATH} is missing: build it with `make -C helfem_amd/csrc` "
                               "(or __graft_entry__.build())")
        # PyTorch ships its own HIP runtime under the same soname as ROCm's.
        # Whichever loads first serves the whole process; if libhfgpu pulled
        # in ROCm's before torch initialises its own device layer, torch then
        # finds no GPU.  Loading torch first keeps one runtime for both.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        L = ctypes.CDLL(str(LIB_PATH))
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _LIB = L
    return _LIB


def check(rc):
    if rc != 0:
        raise HfError(rc, lib().hf_last_error().decode())
