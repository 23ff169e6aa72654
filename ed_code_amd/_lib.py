"""ctypes binding of libed_amd.so (include/ed_amd.h).

The library is the only compute path: if it is missing, or no HIP device is
visible, calls raise instead of falling back to anything on the CPU.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
# ED_AMD_LIB overrides the library path (for comparing builds); default in-tree.
LIB_PATH = os.environ.get("ED_AMD_LIB", os.path.join(HERE, "libed_amd.so"))

ED_OK, ED_EINVAL, ED_ENOMEM, ED_EDEVICE, ED_EEMPTY, ED_ENOTCONV = range(6)
ED_C64, ED_C128 = 0, 1
ED_LZ_REFERENCE, ED_LZ_FULL, ED_LZ_LEAN, ED_LZ_THICK = 0, 1, 2, 3
ED_LZ_SELECTIVE, ED_LZ_GRAPH = 1, 2


class EDError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"[ed status {code}] {msg}")
        self.code = code


class Terms(C.Structure):
    _fields_ = [
        ("lam", C.c_double), ("nbonds", C.c_int32), ("bond_xy", C.c_void_p), ("bond_J", C.c_void_p),
        ("bond_phase", C.c_void_p), ("Ch", C.c_double), ("nchiral", C.c_int32), ("chiral_ijk", C.c_void_p),
        ("chiral_phase", C.c_void_p), ("ChCh", C.c_double), ("nbowtie", C.c_int32), ("bowtie", C.c_void_p),
    ]


class LanczosParams(C.Structure):
    _fields_ = [("nev", C.c_int32), ("iterations", C.c_int32), ("ncycles", C.c_int32), ("precision", C.c_int32),
                ("mode", C.c_int32), ("flags", C.c_int32), ("tol", C.c_double), ("seed", C.c_uint64)]


class LanczosStats(C.Structure):
    _fields_ = [("hv_calls", C.c_int64), ("hv_seconds", C.c_double), ("total_seconds", C.c_double),
                ("residual", C.c_double * 16), ("rayleigh", C.c_double * 16), ("eval_error", C.c_double * 64)]


# Every symbol include/ed_amd.h declares, with (restype, argtypes).
_VP = C.c_void_p
SIGNATURES = {
    "ed_last_error": (C.c_char_p, []),
    "ed_init": (C.c_int, [C.c_int]),
    "ed_lattice_create": (C.c_int, [C.c_int32, _VP, _VP, C.POINTER(_VP)]),
    "ed_lattice_create3": (C.c_int, [C.c_int32, _VP, _VP, _VP, C.POINTER(_VP)]),
    "ed_lattice_order3": (C.c_int, [_VP, C.POINTER(C.c_int32)]),
    "ed_basis_build3": (C.c_int, [_VP, C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.POINTER(_VP)]),
    "ed_lattice_orders": (C.c_int, [_VP, C.POINTER(C.c_int32), C.POINTER(C.c_int32)]),
    "ed_lattice_destroy": (None, [_VP]),
    "ed_basis_build": (C.c_int, [_VP, C.c_int32, C.c_int32, C.c_int32, C.POINTER(_VP)]),
    "ed_basis_dim": (C.c_int, [_VP, C.POINTER(C.c_int64)]),
    "ed_basis_export": (C.c_int, [_VP, _VP, _VP, _VP]),
    "ed_basis_full_size": (C.c_int, [_VP, C.POINTER(C.c_int64), C.POINTER(C.c_int64)]),
    "ed_basis_export_full": (C.c_int, [_VP, _VP, _VP, _VP, _VP, C.POINTER(C.c_int32)]),
    "ed_basis_destroy": (None, [_VP]),
    "ed_op_create": (C.c_int, [_VP, C.POINTER(Terms), C.POINTER(_VP)]),
    "ed_op_destroy": (None, [_VP]),
    "ed_op_tune": (C.c_int, [_VP, C.c_int32, C.c_int32]),
    "ed_op_elist_stats": (C.c_int, [_VP, C.POINTER(C.c_int64), C.POINTER(C.c_int64)]),
    "ed_op_tuning": (C.c_int, [_VP, C.c_int32, C.POINTER(C.c_int32), C.POINTER(C.c_int32), C.POINTER(C.c_int32)]),
    "ed_tune_cache_clear": (C.c_int, []),
    "ed_op_count_targets": (C.c_int, [_VP, C.POINTER(C.c_int64), C.POINTER(C.c_int64)]),
    "ed_hv": (C.c_int, [_VP, C.c_int32, _VP, _VP, _VP]),
    "ed_hv_profile": (C.c_int, [_VP, C.c_int32, _VP, _VP, _VP, C.POINTER(C.c_float)]),
    "ed_hv_noalloc": (C.c_int, [_VP, C.c_int32, _VP, _VP, _VP]),
    "ed_hv_host": (C.c_int, [_VP, _VP, _VP]),
    "ed_make_h": (C.c_int, [_VP, _VP]),
    "ed_csr_hv_host": (C.c_int, [C.c_int64, C.c_int64, _VP, _VP, _VP, _VP, _VP]),
    "ed_pair_correlations": (C.c_int, [_VP, _VP, _VP, C.c_int32, _VP, _VP, _VP]),
    "ed_unwrap": (C.c_int, [_VP, _VP, C.c_int64, _VP, _VP, _VP]),
    "ed_density_matrix": (C.c_int, [_VP, _VP, _VP, C.c_int32, _VP, C.c_int32, _VP]),
    "ed_entanglement_entropies": (C.c_int, [_VP, C.c_int64, C.POINTER(C.c_double), C.POINTER(C.c_double)]),
    "ed_dense_eigh": (C.c_int, [_VP, _VP, _VP]),
    "ed_lanczos": (C.c_int, [_VP, C.POINTER(LanczosParams), _VP, _VP, C.POINTER(LanczosStats)]),
    "ed_block_create": (C.c_int, [_VP, C.c_int32, _VP, C.POINTER(_VP)]),
    "ed_block_destroy": (None, [_VP]),
    "ed_block_fast": (C.c_int, [_VP]),
    "ed_block_hv": (C.c_int, [_VP, C.c_int32, _VP, _VP, _VP]),
    "ed_block_lanczos": (C.c_int, [_VP, C.POINTER(LanczosParams), _VP, C.POINTER(LanczosStats)]),
    "ed_mblock_create": (C.c_int, [_VP, C.c_int32, C.c_int32, _VP, C.POINTER(_VP)]),
    "ed_mblock_destroy": (None, [_VP]),
    "ed_mblock_rows": (C.c_int, [_VP, C.c_int32, C.POINTER(C.c_int64), C.POINTER(C.c_int64)]),
    "ed_mblock_fast": (C.c_int, [_VP]),
    "ed_mblock_hv": (C.c_int, [_VP, C.c_int32, _VP, _VP, _VP]),
    "ed_mblock_lanczos": (C.c_int, [_VP, C.POINTER(LanczosParams), _VP, _VP, C.POINTER(LanczosStats)]),
    "ed_gather_calibrate": (C.c_int, [C.c_int64, C.c_int32, C.c_int64, C.c_int32, C.c_int32, C.c_int32,
                                      C.POINTER(C.c_float)]),
}

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built; run `python -m ed_code_amd.build` "
                               "(there is no CPU fallback)")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def check(rc: int):
    if rc != ED_OK:
        msg = lib().ed_last_error()
        raise EDError(rc, msg.decode() if msg else "")
