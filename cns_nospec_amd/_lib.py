"""Loader of libcns_hip.so (the HIP library behind include/cns.h).

There is no CPU fallback: if the library is missing or fails to load, every
entry point raises.  Build it with ``python -m cns_nospec_amd.build``.
"""
from __future__ import annotations

import ctypes as C
import os
import re

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "libcns_hip.so")
INCLUDE = os.path.join(os.path.dirname(PKG), "include")
HEADERS = [os.path.join(INCLUDE, h) for h in ("cns.h", "cns_batch.h", "cns_io.h", "cns_domain.h")]

CNS_OK, CNS_EINVAL, CNS_EHIP, CNS_ENOMEM, CNS_ENUMERIC, CNS_EBUDGET, CNS_EIO = 0, -1, -2, -3, -4, -5, -6
CNS_IMPL_TILED, CNS_IMPL_DIRECT = 0, 1

_lib = None


class CnsError(RuntimeError):
    """A C ABI call returned a negative CNS_E* code."""

    def __init__(self, code, message):
        super().__init__(f"{message} (code {code})")
        self.code = code


class CnsNumericError(CnsError):
    """The Courant number was not finite (the state is not physical)."""


class CnsBudgetError(CnsError):
    """A box does not fit in the batched engine's HBM budget."""


class CnsIOError(CnsError):
    """A reference-format file could not be read or written, or is malformed."""


class Config(C.Structure):
    """struct cns_config (include/cns_io.h): the fields of a general_input file."""
    _fields_ = [("ng", C.c_int), ("nc", C.c_int), ("ncells", C.c_int),
                ("lo", C.c_int * 3), ("hi", C.c_int * 3), ("dx", C.c_double * 3),
                ("cfl", C.c_double), ("eta", C.c_double), ("alam", C.c_double),
                ("nsteps", C.c_int), ("dt", C.c_double),
                ("dim", C.c_int * 3), ("dim_g", C.c_int * 3)]


def header_symbols(paths=None):
    """Every function name declared in the public headers (include/*.h)."""
    names = set()
    for path in paths or [os.path.join(INCLUDE, f) for f in sorted(os.listdir(INCLUDE))
                          if f.endswith(".h")]:
        text = re.sub(r"/\*.*?\*/", "", open(path).read(), flags=re.S)
        names |= set(re.findall(r"\b(cns_\w+)\s*\(", text))
    return sorted(names)


_I3 = C.c_int * 3
_D3 = C.c_double * 3
_VP = C.c_void_p
_DBL = C.c_double
_INT = C.c_int

_SIGS = {
    "cns_last_error_string": (C.c_char_p, []),
    "cns_set_stream": (_INT, [_VP]),
    "cns_get_stream": (_VP, []),
    "cns_set_scratch": (_INT, [_VP, C.c_size_t]),
    "cns_scratch_bytes": (C.c_size_t, [_I3, _I3, _INT]),
    "cns_set_stencil_impl": (_INT, [_INT]),
    "cns_get_stencil_impl": (_INT, []),
    "cns_fill_boundary": (_INT, [_VP, _I3, _I3]),
    "cns_ctoprim": (_INT, [_I3, _I3, _VP, _VP, _D3, _INT, C.POINTER(C.c_double)]),
    "cns_hypterm": (_INT, [_I3, _I3, _INT, _D3, _VP, _VP, _VP]),
    "cns_diffterm": (_INT, [_I3, _I3, _INT, _D3, _VP, _VP, _DBL, _DBL]),
    "cns_diffterm_noalloc": (_INT, [_I3, _I3, _INT, _D3, _VP, _VP, _DBL, _DBL]),
    "cns_rk_update": (_INT, [_INT, _I3, _I3, _INT, _VP, _VP, _VP, _VP, _DBL]),
    "cns_ctx_create": (_INT, [C.POINTER(_VP), _I3, _D3, _DBL, _DBL, _DBL]),
    "cns_ctx_destroy": (_INT, [_VP]),
    "cns_advance": (_INT, [_VP, _VP, _VP, _VP, _VP, _VP, C.POINTER(C.c_double)]),
    "cns_advance_async": (_INT, [_VP, _VP, _VP, _VP, _VP, _VP, _VP, _VP]),
    "cns_ctx_profile": (_INT, [_VP, _INT]),
    "cns_ctx_set_fused": (_INT, [_VP, _INT]),
    "cns_ctx_profile_read": (_INT, [_VP, C.POINTER(C.c_double), C.POINTER(C.c_longlong)]),
    "cns_batch_create": (_INT, [C.POINTER(_VP), C.c_size_t, _INT, _INT]),
    "cns_batch_destroy": (_INT, [_VP]),
    "cns_batch_box_bytes": (C.c_size_t, [_INT, _INT]),
    "cns_batch_add": (_INT, [_VP, _INT, _DBL, _DBL, _DBL, _DBL, _VP, C.POINTER(_INT)]),
    "cns_batch_remove": (_INT, [_VP, _INT]),
    "cns_batch_read": (_INT, [_VP, _INT, _VP]),
    "cns_batch_step": (_INT, [_VP, _INT]),
    "cns_batch_query": (_INT, [_VP, _INT, C.POINTER(_INT), C.POINTER(C.c_longlong),
                               C.POINTER(C.c_double), C.POINTER(C.c_double), C.POINTER(_INT)]),
    "cns_batch_set_tile": (_INT, [_VP, _INT, _INT, _INT]),
    "cns_batch_get_tile": (_INT, [_VP, _INT, C.POINTER(_INT), C.POINTER(_INT)]),
    "cns_batch_stats": (_INT, [_VP, C.POINTER(C.c_size_t), C.POINTER(_INT), C.POINTER(_INT)]),
    "cns_batch_set_target": (_INT, [_VP, _INT, C.c_longlong]),
    "cns_domain_create": (_INT, [C.POINTER(_VP), _I3, _INT, _D3, _DBL, _DBL, _DBL]),
    "cns_domain_destroy": (_INT, [_VP]),
    "cns_domain_boxes": (_INT, [_VP, _I3, C.POINTER(_INT)]),
    "cns_domain_box": (_INT, [_VP, _INT, _I3, _I3]),
    "cns_domain_array_len": (C.c_size_t, [_VP, _INT, _INT]),
    "cns_domain_fill_boundary": (_INT, [_VP, _VP]),
    "cns_domain_advance": (_INT, [_VP, _VP, _VP, _VP, _VP, _VP, C.POINTER(C.c_double)]),
    "cns_read_general_input": (_INT, [C.c_char_p, C.POINTER(Config)]),
    "cns_write_general_input": (_INT, [C.c_char_p, C.POINTER(Config)]),
    "cns_read_field": (_INT, [C.c_char_p, _I3, _INT, _INT, _VP]),
    "cns_write_field": (_INT, [C.c_char_p, _I3, _INT, _INT, _VP]),
    "cns_field_len": (C.c_size_t, [_I3, _INT]),
    "cns_write_plotfile": (_INT, [C.c_char_p, _INT, _INT, _VP, _VP, _INT, _VP, _D3, _D3, _D3,
                                  C.c_double]),
}

PROF_KINDS = ("fill", "ctoprim", "diffterm", "hypterm", "rk", "hypterm_rk", "fill_ctoprim", "other")


def load(path: str = LIB_PATH) -> C.CDLL:
    """Load the library once; raise if it is absent (no fallback exists)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise ImportError(f"{path} is not built: run `python -m cns_nospec_amd.build` "
                          "(the HIP library has no CPU fallback)")
    lib = C.CDLL(path)
    for name, (res, args) in _SIGS.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def register(name, restype, argtypes):
    """Declare the ctypes signature of a further entry point."""
    fn = getattr(load(), name)
    fn.restype = restype
    fn.argtypes = argtypes
    return fn


def check(rc: int):
    if rc != CNS_OK:
        msg = load().cns_last_error_string().decode(errors="replace")
        if rc == CNS_ENUMERIC:
            raise CnsNumericError(rc, msg)
        if rc == CNS_EBUDGET:
            raise CnsBudgetError(rc, msg)
        if rc == CNS_EIO:
            raise CnsIOError(rc, msg)
        raise CnsError(rc, msg)
    return rc
