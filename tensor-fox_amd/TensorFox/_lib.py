"""ctypes binding of libtfxhip.so (declared in include/tfxhip.h).

The library is loaded from this package directory.  Loading it needs no GPU; every compute
entry point needs one.  A missing library raises ImportError with the build command: there is
no CPU fallback on the product path.
"""

import ctypes
import os
import threading

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, 'libtfxhip.so')

TFX_F32, TFX_F64 = 0, 1
TFX_MAXL = 8          # largest tensor order of the library (include/tfxhip.h)
TFX_PREC_NATIVE, TFX_PREC_SPLIT16 = 0, 1
TFX_OK, TFX_E_ARG, TFX_E_HIP, TFX_E_NOMEM, TFX_E_UNSUPPORTED = 0, -1, -2, -3, -4

_c_int, _c_i64, _c_dbl, _c_vp, _c_uint = ctypes.c_int, ctypes.c_int64, ctypes.c_double, ctypes.c_void_p, ctypes.c_uint
_I64P = ctypes.POINTER(ctypes.c_int64)

# name -> argument types (all return int status except tfx_last_error)
SIGNATURES = {
    'tfx_ctx_create': [_c_int, ctypes.POINTER(_c_vp)],
    'tfx_ctx_destroy': [_c_vp],
    'tfx_device_info': [_c_vp, ctypes.POINTER(_c_int), ctypes.POINTER(_c_int), ctypes.POINTER(_c_int)],
    'tfx_mttkrp': [_c_vp, _c_vp, _c_int, _c_int, _I64P, _c_vp, _c_int, _c_uint, _c_vp, _c_int, _c_vp],
    'tfx_mttkrp_mixed': [_c_vp, _c_vp, _c_int, _c_int, _I64P, _c_vp, _c_int, _c_int, _c_uint, _c_vp, _c_int, _c_vp],
    'tfx_mttkrp_noalloc': [_c_vp, _c_vp, _c_int, _I64P, _c_vp, _c_int, _c_uint, _c_vp, _c_vp],
    'tfx_ttm_mixed': [_c_vp, _c_vp, _c_int, _c_int, _I64P, _c_int, _c_vp, _c_int, _c_vp, _c_int, _c_vp, _c_int, _c_vp],
    'tfx_cpd_residual': [_c_vp, _c_vp, _c_int, _c_int, _I64P, _c_vp, _c_int, _c_vp, _c_vp],
    'tfx_cpd_residual_mixed': [_c_vp, _c_vp, _c_int, _c_int, _I64P, _c_vp, _c_int, _c_int, _c_vp, _c_vp],
    'tfx_tensor_stats': [_c_vp, _c_vp, _c_int, _c_i64, _c_vp, _c_vp],
    'tfx_gramians': [_c_vp, _c_int, _c_int, _I64P, _c_int, _c_vp, _c_vp, _c_vp, _c_vp, _c_vp, _c_vp],
    'tfx_neg_grad': [_c_vp, _c_int, _c_int, _I64P, _c_int, _c_vp, _c_vp, _c_vp, _c_vp, _c_vp],
    'tfx_mttkrp_neg_grad': [_c_vp, _c_vp, _c_int, _c_int, _I64P, _c_vp, _c_int, _c_int, _c_vp, _c_vp,
                            _c_vp, _c_int, _c_vp],
    'tfx_cg': [_c_vp, _c_int, _c_int, _I64P, _c_int, _c_vp, _c_vp, _c_vp, _c_vp, _c_dbl, _c_vp,
               _c_int, _c_dbl, _c_vp, _c_vp, _c_vp, _c_vp],
    'tfx_dgn_update': [_c_vp, _c_int, _c_int, _I64P, _c_int, _c_vp, _c_vp, _c_dbl, _c_vp, _c_dbl,
                       _c_int, _c_int, _c_dbl, _c_vp, _c_vp, _c_vp, _c_vp],
    'tfx_dgn_direct': [_c_vp, _c_int, _I64P, _c_int, _c_vp, _c_vp, _c_vp, _c_vp, _c_dbl, _c_vp,
                       _c_vp, _c_vp, _c_vp],
    'tfx_als_solve': [_c_vp, _c_int, _c_int, _I64P, _c_int, _c_int, _c_vp, _c_vp, _c_vp, _c_vp, _c_vp],
    'tfx_dogleg_dots': [_c_vp, _c_int, _c_i64, _c_vp, _c_vp, _c_vp, _c_vp, _c_vp],
    'tfx_axpby': [_c_vp, _c_int, _c_i64, _c_dbl, _c_vp, _c_dbl, _c_vp, _c_vp, _c_vp],
    'tfx_gram_mode': [_c_vp, _c_vp, _c_int, _c_int, _I64P, _c_int, _c_vp, _c_int, _c_vp],
    'tfx_eig_top': [_c_vp, _c_vp, _c_int, _c_vp, _c_int, _c_int, _c_int, _c_vp, _c_vp, _c_vp],
    'tfx_ttm': [_c_vp, _c_vp, _c_int, _c_int, _I64P, _c_int, _c_vp, _c_int, _c_vp, _c_vp, _c_int, _c_vp],
    'tfx_gemm_small': [_c_vp, _c_int, _c_int, _c_int, _c_int, _c_int, _c_vp, _c_vp, _c_vp, _c_vp],
    'tfx_convert': [_c_vp, _c_i64, _c_vp, _c_int, _c_vp, _c_int, _c_vp],
    'tfx_sparse_pack': [_c_vp, _c_i64, _c_int, _c_vp, _c_vp, _I64P, _c_vp, _c_vp, _c_vp],
    'tfx_sparse_gram': [_c_vp, _c_i64, _c_int, _I64P, _c_vp, _c_vp, _c_int, _c_vp, _c_vp],
    'tfx_sparse_ttm0': [_c_vp, _c_i64, _c_int, _I64P, _c_vp, _c_vp, _c_vp, _c_int, _c_vp, _c_vp],
    'tfx_sparse_residual': [_c_vp, _c_i64, _c_int, _I64P, _c_vp, _c_vp, _c_vp, _c_int, _c_vp, _c_vp],
    'tfx_dgn_run': [_c_vp, _c_vp, _c_int, _c_int, _c_int, _I64P, _c_int, _c_vp, _c_vp, _c_vp, _c_vp,
                    _c_vp, _c_vp, _c_dbl, _c_vp, _c_i64, _c_vp, _c_vp, _c_vp, _c_vp],
    'tfx_batch_mlsvd': [_c_vp, _c_vp, _c_int, _c_int, _c_vp, _c_vp, _c_dbl, _c_vp, _c_vp, _c_vp,
                        _c_vp, _c_vp],
    'tfx_batch_dgn': [_c_vp, _c_int, _c_int, _c_vp, _c_vp, _c_vp, _c_vp, _c_vp, _c_vp, _c_vp,
                      _c_vp, _c_vp, _c_int, _c_int, _c_int, _c_vp],
    'tfx_batch_draws': [_c_int, _c_int, _c_vp, _c_vp, _c_vp, _c_vp, _c_int, _c_vp, _c_vp, _c_dbl,
                        _c_int, _c_int, _c_dbl, _c_vp, _c_vp, _c_vp, _c_vp, _c_int],
    'tfx_batch_finish': [_c_vp, _c_vp, _c_int, _c_int, _c_int, _c_vp, _c_vp, _c_vp, _c_vp, _c_vp,
                         _c_vp, _c_vp, _c_vp],
}


# host-only entry points (no context argument)
HOST_SIGNATURES = {
    'tfx_mttkrp_plan': [_c_int, _c_int, _c_int, _c_int, _I64P, _c_int, _c_uint],
}
# host-only entry points returning int64
HOST_I64_SIGNATURES = {
    'tfx_dgn_run_work': [_c_int, _I64P, _c_int],
}
TFX_PATH_SPLIT16_MFMA, TFX_PATH_FMA, TFX_PATH_F64_MFMA = 0, 1, 2


class DgnOpts(ctypes.Structure):
    """tfx_dgn_opts (include/tfxhip.h)."""
    _fields_ = [('maxiter', _c_int), ('tol', _c_dbl), ('tol_step', _c_dbl), ('tol_improv', _c_dbl),
                ('tol_grad', _c_dbl), ('tol_jump', _c_dbl), ('cg_tol', _c_dbl), ('cg_factor', _c_dbl),
                ('cg_maxiter', _c_int), ('damp0', _c_dbl), ('symm', _c_int), ('aliased', _c_int),
                ('factors_norm', _c_dbl)]


class BatchOpts(ctypes.Structure):
    """tfx_batch_opts (include/tfxhip.h)."""
    _fields_ = [('maxiter', _c_int), ('symm', _c_int), ('tol', _c_dbl), ('tol_step', _c_dbl),
                ('tol_improv', _c_dbl), ('tol_grad', _c_dbl), ('tol_jump', _c_dbl),
                ('init_damp', _c_dbl), ('cg_tol', _c_dbl), ('factors_norm', _c_dbl)]

_lib = None
_lock = threading.Lock()


def load():
    """Load libtfxhip.so (no GPU needed) and declare its prototypes."""
    global _lib
    with _lock:
        if _lib is not None:
            return _lib
        if not os.path.exists(LIB_PATH):
            raise ImportError('libtfxhip.so not found at %s; build it with '
                              '`python tensor-fox_amd/build_lib.py`' % LIB_PATH)
        lib = ctypes.CDLL(LIB_PATH)
        for name, args in list(SIGNATURES.items()) + list(HOST_SIGNATURES.items()):
            fn = getattr(lib, name)
            fn.argtypes = args
            fn.restype = ctypes.c_int
        for name, args in HOST_I64_SIGNATURES.items():
            fn = getattr(lib, name)
            fn.argtypes = args
            fn.restype = ctypes.c_int64
        lib.tfx_last_error.argtypes = [_c_vp]
        lib.tfx_last_error.restype = ctypes.c_char_p
        _lib = lib
        return lib


class TfxError(RuntimeError):
    pass


def check(rc, ctx=None):
    if rc != 0:
        msg = ''
        if ctx is not None:
            m = load().tfx_last_error(ctx)
            msg = m.decode() if m else ''
        raise TfxError('libtfxhip error %d: %s' % (rc, msg))


_contexts = {}


def context(device):
    """Per-device context (created on first use; needs the GPU)."""
    with _lock:
        c = _contexts.get(device)
    if c is not None:
        return c
    lib = load()
    out = _c_vp()
    check(lib.tfx_ctx_create(int(device), ctypes.byref(out)))
    with _lock:
        _contexts[device] = out
    return out


def dims_arg(dims):
    return (ctypes.c_int64 * len(dims))(*[int(d) for d in dims])
