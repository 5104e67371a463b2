"""ctypes binding of libemix3d_hip.so (the C ABI of include/emix3d.h).

The library is built in-tree by emix3d_amd.build; loading fails loudly when
it is missing -- there is no CPU fallback.
"""
import ctypes
import os
import re

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
LIB_PATH = os.path.join(HERE, "libemix3d_hip.so")
HEADER = os.path.join(ROOT, "include", "emix3d.h")

_P = ctypes.c_void_p
_I = ctypes.c_int
_D = ctypes.c_double

# argument types of every exported function
SIGNATURES = {
    "emx_last_error": ([], ctypes.c_char_p),
    "emx_version": ([], _I),
    "emx_create": ([ctypes.POINTER(_P), _I, _I, _I, _I, _P], _I),
    "emx_destroy": ([_P], _I),
    "emx_set_stream": ([_P, _P], _I),
    "emx_set_plan": ([_P, _I, _I, _I, _I], _I),
    "emx_get_plan": ([_P, _I, _I, ctypes.POINTER(_I), ctypes.POINTER(_I)], _I),
    "emx_timing_begin": ([_P], _I),
    "emx_timing_select": ([_P, ctypes.c_char_p], _I),
    "emx_timing_end": ([_P, ctypes.POINTER(_I)], _I),
    "emx_timing_entry": ([_P, _I, ctypes.POINTER(ctypes.c_char_p), ctypes.POINTER(_I),
                          ctypes.POINTER(_I), ctypes.POINTER(_D), ctypes.POINTER(_D)], _I),
    "emx_sync": ([_P], _I),
    "emx_deriv": ([_P, _I, _P, _P, _D], _I),
    "emx_grad": ([_P, _P, _P, _P, _P, _D, _D, _D], _I),
    "emx_metrics": ([_P, _P, _P], _I),
    "emx_sgs_geometry": ([_P, _P, _P, _P, ctypes.POINTER(_D)], _I),
    "emx_sgdm": ([_P, _P, _P, _P, _P, _D, _P, _P, _P], _I),
    "emx_flux": ([_P, _I, _P, _P, _P, _P, _P, _P, _P], _I),
    "emx_divergence": ([_P, _P, _P, _D, _D, _D, _P, _P], _I),
    "emx_flux_div": ([_P, _I, _P, _P, _P, _P, _P, _P, _D, _D, _D, _P, _P], _I),
    "emx_flux_div_all": ([_P, _P, _P, _P, _P, _P, _P, _D, _D, _D, _P, _P], _I),
    "emx_dynsmag": ([_P, _P, _P, _P, _P, _P, _D, _P], _I),
    "emx_dynsmag_noalloc": ([_I, _I, _I, _P, _P, _P, _P, _P, _D, _P, _P], _I),
    "emx_rk_update": ([_P, _I, _P, _P, _P, _P, _P, _D], _I),
    "emx_frontera_val": ([_P, _P], _I),
    "emx_io_write_field": ([ctypes.c_char_p, _I, _I, _I, _I, _P, _P, _P], _I),
    "emx_io_read_field": ([ctypes.c_char_p, _I, _I, _I, _P, _P, _P], _I),
    "emx_io_field_dims": ([ctypes.c_char_p] + [ctypes.POINTER(_I)] * 4, _I),
    "emx_io_write_grid": ([ctypes.c_char_p, _I, _I, _I, _I, _P], _I),
    "emx_io_read_grid": ([ctypes.c_char_p, _I, _I, _I, _P], _I),
    "emx_io_write_time": ([ctypes.c_char_p, _D, _I], _I),
    "emx_io_read_time": ([ctypes.c_char_p, ctypes.POINTER(_D), ctypes.POINTER(_I)], _I),
    "emx_io_write_scalar": ([ctypes.c_char_p, _I, _I, _I, _I, _P], _I),
    "emx_io_read_scalar": ([ctypes.c_char_p, _I, _I, _I, _I, _P], _I),
    "emx_ls_reinit": ([_P, _P, _P, _P, _D, _D, _D, _D], _I),
    "emx_heviside": ([_P, _P, _D, _P], _I),
    "emx_varc_to_varnc": ([_P, _P, _D, _P, _P, _P, _P], _I),
    "emx_varnc_to_varc": ([_P, _P, _P, _P, _D, _P, _P], _I),
    "emx_rk_clip": ([_P, _P, _P, _D, _P, _P, _P, _P], _I),
    "emx_rk_rebuild": ([_P, _P, _P, _P, _D], _I),
    "emx_filtrado": ([_P, _I, _P, _P, _D, _D, _P, _P, _P], _I),
    "emx_grad_div": ([_P, _P, _P, _P, _P, _D, _D, _D], _I),
    "emx_batch_create": ([ctypes.POINTER(_P), _I, ctypes.POINTER(_I), _I, _P], _I),
    "emx_batch_destroy": ([_P], _I),
    "emx_batch_set_stream": ([_P, _P], _I),
    "emx_batch_bind": ([_P, _I, _P, _P, _P, _P, _P, _D, _D, _D, _D, _D, _D, _D], _I),
    "emx_batch_set_plan": ([_P, _I, _I, _I], _I),
    "emx_batch_get_plan": ([_P, _I, ctypes.POINTER(_I), ctypes.POINTER(_I)], _I),
    "emx_batch_tune": ([_P, _I], _I),
    "emx_batch_stage": ([_P, _I], _I),
    "emx_mlp_nut": ([_P, _P, _P, _P, _P, _P, _P, _D, _P], _I),
    "emx_mlp_gather": ([_P, _P, _P, _P, _P, _I, _P], _I),
    "emx_mlp_forward_f32": ([_P, _P, _P, _I, _P, _P, _P], _I),
    "emx_mlp_backward_work": ([_I], ctypes.c_size_t),
    "emx_mlp_backward_f32": ([_P, _P, _P, _P, _P, _I, _P, _P, ctypes.c_size_t, _P], _I),
    "emx_adam_f32": ([_P, _P, _P, _P, _I, ctypes.c_float, ctypes.c_float, ctypes.c_float,
                      ctypes.c_float, _I, _P], _I),
    "emx_deriv_adj": ([_P, _I, _P, _P, _D], _I),
    "emx_grad_adj": ([_P, _P, _P, _P, _P, _D, _D, _D], _I),
    "emx_divergence_adj": ([_P, _P, _P, _P, _D, _D, _D, _P], _I),
    "emx_flux_div_all_adj": ([_P, _P, _P, _P, _P, _P, _P, _P, _D, _D, _D, _P, _P, _P, _P, _P],
                             _I),
    "emx_ls_extrapolate": ([_P, _P, _P, _P, _D, _D, _D, _D], _I),
    "emx_ls_advect_weno5": ([_P, _P, _P, _P, _D, _D, _D, _D], _I),
    "cudaDeriv": ([_P, _P, _P, _P, _P, _P, _P, _D, _D, _D, _I, _I, _I], None),
    "cuDivergence": ([_P, _P, _P, _P, _P, _P, _P, _P, _P, _D, _D, _D, _I, _I, _I], None),
    "advect_CUDA": ([_P, _P, _P, _D, _D, _D, _I, _I, _I, _D], None),
    "emx_host_flujos": ([_I, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _I, _I, _I], _I),
    "emx_host_sgdm": ([_P, _P, _P, _P, _D, _P, _P, _P, _I, _I, _I], _I),
}


def declared_symbols(header=HEADER):
    """Names of the functions declared in the public header."""
    text = open(header).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"^\s*(?:const\s+)?[A-Za-z_][A-Za-z0-9_]*\s*\**\s+\**([A-Za-z_][A-Za-z0-9_]*)\s*\(",
                       text, flags=re.M)
    return sorted(set(names))


class EmxError(RuntimeError):
    """A non-zero status from libemix3d_hip.so."""

    def __init__(self, fn, status, msg):
        super().__init__(f"{fn} failed with status {status}: {msg}")
        self.status = status


_lib = None


def lib():
    """Load the library (once).  Raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(
                f"{LIB_PATH} is missing: build it with `python -m emix3d_amd.build` "
                "(the operator has no CPU fallback)")
        L = ctypes.CDLL(LIB_PATH)
        for name, (args, res) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = res
        _lib = L
    return _lib


def check(fn_name, status):
    if status != 0:
        msg = lib().emx_last_error().decode(errors="replace")
        raise EmxError(fn_name, status, msg)
    return status


def call(fn_name, *args):
    return check(fn_name, getattr(lib(), fn_name)(*args))
