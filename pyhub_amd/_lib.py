"""ctypes binding of libphub.so (the C ABI declared in include/phub.h).

The library is the product path: there is no CPU fallback.  Loading fails with
a RuntimeError when the shared object is missing, and every entry point raises
when the HIP call reports an error.
"""
from __future__ import annotations

import ctypes as C
import os

import torch  # noqa: F401  (loads the HIP runtime the library binds to)

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libphub.so")

OK, ERR_ARG, ERR_UNSUPPORTED, ERR_OOM, ERR_NOCONV, ERR_BREAKDOWN, ERR_HIP, ERR_NORM_LOST = range(8)
LCZ_PLAIN64, LCZ_FULL64, LCZ_MIXED, LCZ_PLAIN32 = 0, 1, 2, 3
MAX_STATES = 16                     # PHUB_MAX_STATES


class HubbardDesc(C.Structure):
    _fields_ = [("n_sites", C.c_int), ("n_up", C.c_int), ("n_dn", C.c_int),
                ("t", C.POINTER(C.c_double)), ("u", C.POINTER(C.c_double)),
                ("t_dn", C.POINTER(C.c_double)), ("u4", C.POINTER(C.c_double))]


class HeisDesc(C.Structure):
    _fields_ = [("n_sites", C.c_int), ("j", C.POINTER(C.c_double)), ("sign_convention", C.c_int)]


class OpInfo(C.Structure):
    _fields_ = [("kind", C.c_int), ("n_sites", C.c_int), ("n_up", C.c_int), ("n_dn", C.c_int),
                ("dim_up", C.c_int64), ("dim_dn", C.c_int64), ("dim", C.c_int64),
                ("max_hops_up", C.c_int), ("max_hops_dn", C.c_int), ("table_bytes", C.c_double),
                ("dn_tile", C.c_int * 2)]


class LanczosOpts(C.Structure):
    _fields_ = [("mode", C.c_int), ("max_steps", C.c_int), ("tol", C.c_double),
                ("check_every", C.c_int), ("fixed_steps", C.c_int), ("refine_steps", C.c_int),
                ("seed", C.c_uint64), ("want_psi", C.c_int), ("n_states", C.c_int),
                ("workspace", C.c_void_p), ("workspace_bytes", C.c_size_t),
                ("psi_tol", C.c_double)]


class LanczosResult(C.Structure):
    _fields_ = [("e0", C.c_double), ("e0_ritz", C.c_double), ("residual", C.c_double),
                ("steps", C.c_int), ("converged", C.c_int), ("ndeg", C.c_int),
                ("seconds", C.c_double), ("seconds_lanczos", C.c_double),
                ("hv_ms_avg", C.c_double), ("e_states", C.c_double * MAX_STATES),
                ("norm_lost", C.c_int), ("corrections", C.c_int),
                ("correction_iters", C.c_int)]


class Cluster(C.Structure):
    _fields_ = [("n_sites", C.c_int), ("n_up", C.c_int), ("n_dn", C.c_int),
                ("t", C.POINTER(C.c_double)), ("u", C.POINTER(C.c_double))]


class BandOpts(C.Structure):
    _fields_ = [("max_steps", C.c_int), ("tol", C.c_double), ("workspace", C.c_void_p),
                ("workspace_bytes", C.c_size_t)]


class BatchOpts(C.Structure):
    _fields_ = [("gs_steps", C.c_int), ("gf_steps", C.c_int), ("seed", C.c_uint64),
                ("gs_tol", C.c_double)]


# every symbol declared in include/phub.h, with its ctypes signature
_vp = C.c_void_p
_dp = C.POINTER(C.c_double)
SIGNATURES = {
    "phub_ctx_create": (C.c_int, [C.c_int, _vp, C.POINTER(_vp)]),
    "phub_ctx_destroy": (None, [_vp]),
    "phub_last_error": (C.c_char_p, [_vp]),
    "phub_sync": (C.c_int, [_vp]),
    "phub_last_kernel_ms": (C.c_double, [_vp]),
    "phub_binom": (C.c_uint64, [C.c_int, C.c_int]),
    "phub_basis_build": (C.c_int, [_vp, C.c_int, C.c_int, _vp]),
    "phub_basis_rank": (C.c_int, [_vp, C.c_int, C.c_int, _vp, C.c_int64, _vp]),
    "phub_op_hubbard": (C.c_int, [_vp, C.POINTER(HubbardDesc), C.POINTER(_vp)]),
    "phub_op_heisenberg": (C.c_int, [_vp, C.POINTER(HeisDesc), C.POINTER(_vp)]),
    "phub_op_destroy": (None, [_vp]),
    "phub_op_get_info": (C.c_int, [_vp, C.POINTER(OpInfo)]),
    "phub_op_masks": (C.c_int, [_vp, _vp, _vp, _vp]),
    "phub_hv_f64": (C.c_int, [_vp, _vp, _vp, _vp]),
    "phub_hv_f32": (C.c_int, [_vp, _vp, _vp, _vp]),
    "phub_lanczos_workspace_bytes": (C.c_int, [_vp, C.POINTER(LanczosOpts), C.POINTER(C.c_size_t)]),
    "phub_lanczos": (C.c_int, [_vp, _vp, C.POINTER(LanczosOpts), _vp, _dp, _dp,
                               C.POINTER(LanczosResult), _vp]),
    "phub_rdm_hubbard": (C.c_int, [_vp, _vp, _vp, _vp, _dp, _dp, _dp, _dp]),
    "phub_rdm_heisenberg": (C.c_int, [_vp, _vp, _vp, _vp, _dp]),
    "phub_one_rdm_spin": (C.c_int, [_vp, _vp, _vp, _vp, _dp, _dp]),
    "phub_excite": (C.c_int, [_vp, _vp, _vp, C.c_int, C.c_int, C.c_int, _vp, _vp]),
    "phub_cf_chain": (C.c_int, [_vp, _vp, _vp, C.c_int, C.c_int, _dp, _dp, _dp, _vp, C.c_size_t]),
    "phub_cf_chain_workspace_bytes": (C.c_int, [_vp, C.c_int, C.POINTER(C.c_size_t)]),
    "phub_two_rdm_hubbard": (C.c_int, [_vp, _vp, _vp, _vp, _dp]),
    "phub_band_lanczos": (C.c_int, [_vp, _vp, _vp, C.c_int, C.POINTER(BandOpts), _dp, _dp,
                                    C.POINTER(C.c_int)]),
    "phub_band_lanczos_workspace_bytes": (C.c_int, [_vp, C.c_int, C.POINTER(C.c_size_t)]),
    "phub_batch_hubbard": (C.c_int, [_vp, C.POINTER(Cluster), C.c_int, C.POINTER(BatchOpts),
                                     _dp, _dp, _dp, _dp]),
    "phub_cf_eval": (C.c_int, [_vp, C.c_int, C.c_int, _dp, _dp, _dp, _dp, _dp, C.c_int, _dp,
                               C.c_double, _dp]),
    "phub_cf_eval_noalloc": (C.c_int, [_vp, C.c_int, C.c_int, _dp, _dp, _dp, _dp, _dp, C.c_int,
                                       _dp, C.c_double, _dp]),
    "phub_hf_grad_hubbard": (C.c_int, [_vp, _vp, _vp, _dp, _dp]),
    "phub_grad_hubbard": (C.c_int, [_vp, _vp, _vp, _vp, _dp, _dp]),
    "phub_hf_grad_heisenberg": (C.c_int, [_vp, _vp, _vp, _dp]),
    "phub_adjoint_minres": (C.c_int, [_vp, _vp, _vp, C.c_double, _vp, _vp, C.c_int, C.c_double,
                                      C.POINTER(C.c_int), _dp, _vp, C.c_size_t]),
    "phub_adjoint_minres_workspace_bytes": (C.c_int, [_vp, C.POINTER(C.c_size_t)]),
    "phub_alloc_count": (C.c_uint64, []),
    "phub_fock_create": (C.c_int, [_vp, C.c_int, C.POINTER(C.c_int), C.c_int, C.POINTER(C.c_int),
                                   C.c_int, C.POINTER(_vp)]),
    "phub_fock_destroy": (None, [_vp]),
    "phub_fock_dims": (C.c_int, [_vp, C.POINTER(C.c_int64), C.POINTER(C.c_int64)]),
    "phub_fock_masks": (C.c_int, [_vp, _vp, _vp, _vp]),
    "phub_opprog_create": (C.c_int, [_vp, C.c_int, _dp, C.POINTER(C.c_int), C.POINTER(C.c_int),
                                     C.POINTER(_vp)]),
    "phub_opprog_destroy": (None, [_vp]),
    "phub_opprog_terms": (C.c_int, [_vp]),
    "phub_operator_apply": (C.c_int, [_vp, _vp, _vp, _vp, _vp, C.c_int]),
    "phub_operator_diag": (C.c_int, [_vp, _vp, _vp, _vp]),
    "phub_operator_matrix": (C.c_int, [_vp, _vp, _vp, _vp]),
}

_lib = None


def load():
    """Load libphub.so (raises when it was not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"libphub.so not found at {LIB_PATH}: build it with `python -m pyhub_amd.build` "
                "(the HIP extension is required; there is no CPU fallback)")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


class PhubError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"phub error {code}: {msg}")
        self.code = code


def check(ctx, rc):
    if rc == OK:
        return
    msg = load().phub_last_error(ctx).decode(errors="replace") if ctx else "no context"
    if rc == ERR_ARG:
        raise ValueError(msg)
    if rc == ERR_NOCONV:
        raise StopIteration(msg)
    if rc == ERR_OOM:
        raise MemoryError(msg)
    raise PhubError(rc, msg)
