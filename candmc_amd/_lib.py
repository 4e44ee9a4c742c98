"""ctypes binding of libcandmc_qr.so (the C-ABI in include/candmc_qr.h).

The library is built in-tree by ``__graft_entry__.build()`` (or ``make`` in
``candmc_amd/csrc``).  Loading fails loudly when it is missing: there is no
CPU fallback anywhere in the product path.
"""
from __future__ import annotations

import ctypes
import os

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libcandmc_qr.so")

_c = ctypes
_P = _c.c_void_p
_I64 = _c.c_int64
_I = _c.c_int
_D = _c.c_double

# name -> (restype, argtypes); every symbol include/candmc_qr.h declares
SIGNATURES = {
    "cqr_create": (_I, [_c.POINTER(_P), _I]),
    "cqr_destroy": (_I, [_P]),
    "cqr_set_stream": (_I, [_P, _P]),
    "cqr_get_stream": (_P, [_P]),
    "cqr_sync": (_I, [_P]),
    "cqr_reserve": (_I, [_P, _c.c_size_t]),
    "cqr_workspace_bytes": (_c.c_size_t, [_P]),
    "cqr_set_check_finite": (_I, [_P, _I]),
    "cqr_set_leaf_rows": (_I, [_P, _I64]),
    "cqr_set_lookahead": (_I, [_P, _I]),
    "cqr_set_wait_limit": (_I, [_P, _I]),
    "cqr_status_string": (_c.c_char_p, [_I]),
    "cqr_profile_enable": (_I, [_P, _I]),
    "cqr_profile_read": (_I, [_P, _I, _c.POINTER(_D), _c.POINTER(_D), _c.POINTER(_D), _c.POINTER(_I64)]),
    "cqr_profile_reset": (_I, [_P]),
    "cqr_last_error": (_c.c_char_p, [_P]),
    "cqr_plan_tsqr": (_I, [_I64, _I64, _I64, _I64, _c.POINTER(_I64), _c.POINTER(_I64), _c.POINTER(_I64)]),
    "cqr_tree_data_doubles": (_I64, [_I64, _I64, _I64]),
    "cqr_local_tsqr": (_I, [_P, _P, _I64, _I64, _I64, _P, _P]),
    "cqr_tree_tsqr": (_I, [_P, _P, _I64, _P]),
    "cqr_bitree_tsqr": (_I, [_P, _P, _I64, _P, _P, _I64, _I64, _I64, _I, _P]),
    "cqr_construct_Q1": (_I, [_P, _P, _I64, _P, _P, _I64, _I64, _I64, _I64, _I64, _P]),
    "cqr_construct_Q1_id": (_I, [_P, _P, _I64, _P, _P, _I64, _I64, _I64, _I64, _I64, _P, _P, _I64]),
    "cqr_apply_tsqr_QT": (_I, [_P, _P, _I64, _P, _P, _I64, _I64, _I64, _I64, _I64, _P]),
    "cqr_apply_tsqr_Q": (_I, [_P, _P, _I64, _P, _P, _I64, _I64, _I64, _I64, _I64, _P]),
    "cqr_qr_tree_2d": (_I, [_P, _P, _I64, _I64, _I64, _I64, _I64, _I64]),
    "cqr_tree_data_export": (_I, [_P, _P, _I64, _I64, _I64, _I64, _P, _P]),
    "cqr_signed_nlu": (_I, [_P, _P, _I64, _P, _I64, _I64, _P]),
    "cqr_recursive_nlu": (_I, [_P, _P, _I64, _P, _I64, _I64, _P]),
    "cqr_hh_recon_qr": (_I, [_P, _P, _I64, _I64, _I64, _P, _I64]),
    "cqr_hh_recon_qr_ex": (_I, [_P, _P, _I64, _I64, _I64, _P, _P, _P, _I64]),
    "cqr_noalloc_scratch_bytes": (_c.c_size_t, []),
    "cqr_hh_recon_qr_noalloc": (_I, [_P, _P, _I64, _I64, _I64, _P, _P, _P, _I64]),
    "cqr_compute_invT_from_W": (_I, [_P, _P, _I64, _P]),
    "cqr_compute_invT_from_Y": (_I, [_P, _P, _I64, _P, _I64, _I64, _I64]),
    "cqr_compute_invT_from_Y2": (_I, [_P, _P, _I64, _P, _I64, _P, _I64, _I64, _I64, _I64]),
    "cqr_trim": (_I, [_P]),
    "cqr_set_lstsq_escalate": (_I, [_P, _I]),
    "cqr_lstsq_last_prec": (_I, [_P]),
    "cqr_upd_A": (_I, [_P, _P, _I64, _P, _I64, _I64, _I64, _I64, _P, _I]),
    "cqr_update_A": (_I, [_P, _P, _I64, _P, _I64, _I64, _I64, _I64, _P, _P, _I64, _I]),
    "cqr_qr_2d": (_I, [_P, _P, _I64, _I64, _I64, _I64, _P, _I64]),
    "cqr_qr_2d_2d": (_I, [_P, _P, _I64, _I64, _I64, _I64, _I64, _P, _I64]),
    "cqr_tau_recon": (_I, [_P, _c.c_char, _I64, _I64, _I64, _P, _P]),
    "cqr_geqrf2d": (_I, [_P, _P, _I64, _I64, _I64, _I64, _P, _I]),
    "cqr_lstsq_workspace_bytes": (_c.c_size_t, [_I64, _I64, _I64, _I64]),
    "cqr_server_create": (_I, [_c.POINTER(_P), _I, _I, _c.c_size_t]),
    "cqr_server_destroy": (_I, [_P]),
    "cqr_server_submit": (_I64, [_P, _P]),
    "cqr_server_wait": (_I, [_P, _I64, _P]),
    "cqr_server_drain": (_I, [_P, _P, _I, _c.POINTER(_I)]),
    "cqr_server_pending": (_I, [_P, _c.POINTER(_I64), _c.POINTER(_c.c_size_t)]),
    "cqr_server_set_batching": (_I, [_P, _I]),
    "cqr_lstsq_batched": (_I, [_P, _I64, _P, _I, _c.c_double, _P, _P, _P]),
    "cqr_lstsq_batched_workspace_bytes": (_c.c_size_t, [_I64, _P]),
    "cqr_lstsq_factored": (_I, [_P, _P, _I64, _I64, _I64, _P, _I64, _P, _I64, _P, _I64, _I64, _P, _I64, _I,
                                _c.c_double, _P, _P]),
    "cqr_lstsq_factored_workspace_bytes": (_c.c_size_t, [_I64, _I64, _I64, _I64]),
    "cqr_batch_create": (_I, [_P, _c.POINTER(_P), _I64, _P, _P]),
    "cqr_batch_destroy": (_I, [_P, _P]),
    "cqr_batch_offsets": (_I, [_P, _P, _P, _c.POINTER(_I64), _c.POINTER(_I64)]),
    "cqr_batch_groups": (_I, [_P]),
    "cqr_batch_group_info": (_I, [_P, _I, _c.POINTER(_I64), _c.POINTER(_I64), _c.POINTER(_I64),
                                  _c.POINTER(_I64), _c.POINTER(_I64)]),
    "cqr_batch_set_leaf_rows": (_I, [_P, _P, _I, _I64]),
    "cqr_batch_qr": (_I, [_P, _P, _P, _P, _I, _I]),
    "cqr_batch_qr_loop": (_I, [_P, _P, _P, _P]),
    "cqr_batch_tune": (_I, [_P, _P, _P, _I]),
    "cqr_apply_qt": (_I, [_P, _P, _I64, _P, _I64, _I64, _I64, _P, _I64, _I64, _I]),
    "cqr_trsm_r": (_I, [_P, _P, _I64, _I64, _P, _I64, _I64, _I]),
    "cqr_lstsq": (_I, [_P, _P, _I64, _I64, _I64, _P, _I64, _I64, _P, _I64, _I64, _I, _I, _D,
                       _c.POINTER(_I), _c.POINTER(_D)]),
    "cqr_pack_upper": (_I, [_P, _P, _P, _I64, _I64]),
    "cqr_unpack_upper": (_I, [_P, _P, _P, _I64, _I64]),
    "cqr_pack_lower": (_I, [_P, _P, _P, _I64, _I64, _I]),
    "cqr_unpack_lower": (_I, [_P, _P, _P, _I64, _I64, _I]),
}

STATUS = {0: "CQR_OK", 1: "CQR_ERR_INVALID", 2: "CQR_ERR_UNSUPPORTED", 3: "CQR_ERR_NOMEM",
          4: "CQR_ERR_DEVICE", 5: "CQR_ERR_NONFINITE", 6: "CQR_ERR_NOCONVERGE", 7: "CQR_ERR_CANCELLED"}

_lib = None


def load():
    """Load the HIP library once; raise if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: build it with __graft_entry__.build() "
                           "(there is no CPU fallback)")
    lib = _c.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


class CQRError(RuntimeError):
    def __init__(self, status, msg=""):
        self.status = status
        super().__init__(f"{STATUS.get(status, status)}: {msg}")
