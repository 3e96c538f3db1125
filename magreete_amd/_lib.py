"""ctypes binding of libmagreete_hip.so (the C ABI declared in include/magreete_hip.h).

The library is required: importing the product without it raises, there is
no fallback path.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libmagreete_hip.so")

# enums (include/magreete_hip.h)
MG_OK = 0
MG_2D_TM, MG_2D_TE, MG_3D_VECTOR, MG_3D_SCALAR = 0, 1, 2, 3
MG_C128, MG_C64, MG_BF16X3, MG_BF16X6 = 0, 1, 2, 3
MG_F_REGULARIZE, MG_F_ZERO_SELF, MG_F_DIAG, MG_F_NANFILL = 1, 2, 4, 8
MG_F_PERIODIC_X, MG_F_PERIODIC_Y, MG_F_PERIODIC_Z = 16, 32, 64
MG_F_BESSEL_EXACT = 128

_c = ctypes
_i64 = _c.c_int64
_dbl = _c.c_double
_vp = _c.c_void_p
_int = _c.c_int

# name -> argtypes (every symbol returns int unless listed in _RESTYPES)
SIGNATURES = {
    "mg_version": [],
    "mg_last_error": [],
    "mg_create": [_c.POINTER(_vp)],
    "mg_destroy": [_vp],
    "mg_assemble": [_int, _vp, _i64, _vp, _i64, _dbl, _dbl, _dbl, _dbl, _dbl, _dbl, _dbl, _int, _dbl,
                    _vp, _i64, _vp, _i64, _vp],
    "mg_cgemm": [_int, _i64, _i64, _i64, _dbl, _dbl, _vp, _i64, _vp, _i64, _dbl, _dbl, _vp, _i64, _vp],
    "mg_cgemm_work_elems": [_int, _i64, _i64, _i64],
    "mg_cgemm_ws": [_int, _i64, _i64, _i64, _dbl, _dbl, _vp, _i64, _vp, _i64, _dbl, _dbl, _vp, _i64, _vp, _i64,
                    _vp],
    "mg_getrf": [_vp, _int, _vp, _i64, _i64, _vp, _vp, _vp],
    "mg_getrs": [_vp, _int, _vp, _i64, _i64, _vp, _vp, _i64, _i64, _vp],
    "mg_check": [_vp, _vp],
    "mg_set_leaf_blocks": [_vp, _int],
    "mg_dinv_elems": [_i64],
    "mg_getrf_dinv": [_vp, _int, _vp, _i64, _i64, _vp, _vp],
    "mg_getrs_dinv": [_vp, _int, _vp, _i64, _i64, _vp, _vp, _vp, _i64, _i64, _vp, _vp],
    "mg_profile_enable": [_vp, _int],
    "mg_profile_read": [_vp, _vp, _c.POINTER(_c.c_longlong), _c.POINTER(_dbl), _c.POINTER(_dbl)],
    "mg_profile_read_classes": [_vp, _vp, _c.POINTER(_c.c_longlong), _c.POINTER(_dbl), _c.POINTER(_dbl)],
    "mg_convert": [_int, _vp, _i64, _vp, _i64, _i64, _i64, _vp],
    "mg_propagate": [_int, _vp, _i64, _vp, _i64, _dbl, _dbl, _dbl, _int, _dbl, _vp, _i64, _i64, _dbl, _dbl,
                     _vp, _i64, _vp],
    "mg_propagate_work_bytes": [_int, _i64, _i64, _i64],
    "mg_propagate_ws": [_int, _vp, _i64, _vp, _i64, _dbl, _dbl, _dbl, _int, _dbl, _vp, _i64, _i64, _dbl, _dbl,
                        _vp, _i64, _vp, _i64, _vp],
    "mg_hankel1": [_int, _int, _vp, _i64, _vp, _vp],
    "mg_coldot": [_vp, _i64, _vp, _i64, _i64, _i64, _vp, _vp],
    "mg_sources": [_int, _int, _vp, _i64, _vp, _i64, _vp, _i64, _dbl, _dbl, _vp, _vp],
    "mg_ldu_dinv_elems": [_i64],
    "mg_ldu_work_elems": [_i64],
    "mg_ldu_factor_elems": [_int, _i64],
    "mg_givens": [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _int, _int, _i64, _vp, _vp],
    "mg_ldu_solve_work_elems": [_int, _i64, _i64],
    "mg_getrf_ldu": [_vp, _int, _vp, _i64, _i64, _vp, _vp, _vp, _vp],
    "mg_getrs_ldu": [_vp, _int, _vp, _i64, _i64, _vp, _vp, _i64, _i64, _vp, _vp],
    "mg_mdot_work_elems": [_i64, _i64],
    "mg_mdot": [_vp, _i64, _i64, _int, _vp, _i64, _i64, _i64, _vp, _vp, _i64, _vp],
    "mg_maxpy": [_vp, _i64, _i64, _int, _vp, _i64, _dbl, _vp, _i64, _i64, _i64, _vp],
    "mg_maxpy_norm": [_vp, _i64, _i64, _int, _vp, _i64, _dbl, _vp, _i64, _i64, _i64, _vp, _vp, _vp],
    "mg_colnorm": [_vp, _i64, _i64, _i64, _vp, _vp, _vp],
    "mg_maxpy_mdot": [_vp, _i64, _i64, _int, _vp, _i64, _vp, _i64, _i64, _i64, _vp, _vp, _i64, _vp],
    "mg_hess_solve": [_vp, _vp, _int, _int, _i64, _vp, _vp],
    "mg_split_rows_bytes": [_i64, _i64],
    "mg_split_rows": [_int, _vp, _i64, _i64, _i64, _vp, _vp],
    "mg_relaxed_product_work_bytes": [_int, _i64, _i64, _i64],
    "mg_relaxed_product": [_int, _i64, _i64, _i64, _vp, _vp, _i64, _vp, _i64, _vp, _i64, _vp],
    "mg_propagate_src": [_int, _vp, _i64, _vp, _i64, _dbl, _dbl, _dbl, _int, _dbl, _vp, _i64, _i64, _int, _vp, _vp,
                         _dbl, _vp, _i64, _vp, _i64, _vp],
    "mg_group_create": [_int, _int, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _i64, _c.POINTER(_vp)],
    "mg_group_destroy": [_vp],
    "mg_group_configure": [_vp, _int, _int, _int, _int, _int],
    "mg_group_getrf": [_vp, _vp],
    "mg_group_getrs": [_vp, _vp],
    "mg_group_launches": [_vp],
    "mg_group_tune": [_vp, _vp, _vp, _vp],
    "mg_group_capture": [_vp, _int],
    "mg_group_replay": [_vp, _int, _vp],
    "mg_group_set_aux": [_vp, _vp, _vp],
    "mg_group_residual": [_vp, _vp],
    "mg_group_colnorm": [_vp, _vp, _vp],
    "mg_propagate_noalloc_max_bytes": [],
    "mg_propagate_noalloc": [_int, _vp, _i64, _vp, _i64, _dbl, _dbl, _dbl, _int, _dbl, _vp, _i64, _i64, _dbl, _dbl,
                             _vp, _i64, _vp],
    "mg_inverse_diag_work_elems": [_i64],
    "mg_inverse_diag_blocks": [_vp, _int, _vp, _i64, _i64, _vp, _vp, _int, _vp, _i64, _vp, _vp, _i64, _vp],
}
_RESTYPES = {"mg_last_error": _c.c_char_p, "mg_dinv_elems": _c.c_int64, "mg_cgemm_work_elems": _c.c_int64,
             "mg_mdot_work_elems": _c.c_int64, "mg_ldu_dinv_elems": _c.c_int64, "mg_ldu_work_elems": _c.c_int64,
             "mg_ldu_factor_elems": _c.c_int64, "mg_ldu_solve_work_elems": _c.c_int64,
             "mg_split_rows_bytes": _c.c_int64, "mg_relaxed_product_work_bytes": _c.c_int64,
             "mg_propagate_work_bytes": _c.c_int64, "mg_inverse_diag_work_elems": _c.c_int64,
             "mg_propagate_noalloc_max_bytes": _c.c_int64, "mg_group_launches": _c.c_int64}

_lib = None


def load():
    """Load the library once; raise if it is missing or lacks a symbol."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError("libmagreete_hip.so is not built (run magreete_amd/build.py or __graft_entry__.build())")
    lib = _c.CDLL(LIB_PATH)
    for name, args in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = _RESTYPES.get(name, _int)
    _lib = lib
    return lib


class MGError(RuntimeError):
    def __init__(self, status, where):
        msg = load().mg_last_error().decode(errors="replace")
        super().__init__("%s failed with status %d: %s" % (where, status, msg))
        self.status = status


def check(status, where):
    if status != MG_OK:
        raise MGError(status, where)
