"""Python mirror of the reference's operator interface, over the C ABI.

Names, argument order and meaning follow the reference's free functions
(cpu/header.h:97-156, cpu/helper_functions.h:6-19):

==========================  ==============================================
this module                 reference
==========================  ==============================================
``allocate_variables``      allocate_variables  cpu/helper_functions.cpp:9-22,
                            gpu/util.cu:149-181
``fill_boundary``           fill_boundary       cpu/helper_functions.cpp:226-262
``ctoprim``                 ctoprim (+courno)   cpu/ctoprim.cpp:11-105
``hypterm``                 hypterm             cpu/hypterm.cpp:11-209
``diffterm``                diffterm            cpu/diffterm.cpp:22-320
``rk_update``               gpu_Unew            gpu/advance.cu:81-105
``Stepper.new_advance``     new_advance         cpu/advance.cpp:13-154
==========================  ==============================================

Arrays are float64 CUDA tensors (torch on ROCm) in the ABI's padded SoA
layout (see :mod:`cns_nospec_amd.layout`); torch provides device memory and
the stream, every computation runs in libcns_hip.so.  ``double &courno``
becomes a return value.  Errors raise :class:`CnsError`; nothing exits.
"""
from __future__ import annotations

import ctypes as C

from . import _lib
from .layout import NC, NQ, BoxLayout
from .state import ALAM, CFL, ETA, box_dx

_I3 = C.c_int * 3
_D3 = C.c_double * 3


def _torch():
    import torch
    return torch


def _stream():
    torch = _torch()
    lib = _lib.load()
    lib.cns_set_stream(C.c_void_p(torch.cuda.current_stream().cuda_stream))
    return lib


def _ptr(t, n_expected: int, name: str):
    torch = _torch()
    if not isinstance(t, torch.Tensor):
        raise TypeError(f"{name}: expected a torch tensor")
    if not t.is_cuda:
        raise ValueError(f"{name}: expected a device tensor (the library has no CPU path)")
    if t.dtype != torch.float64:
        raise TypeError(f"{name}: expected float64, got {t.dtype}")
    if not t.is_contiguous():
        raise ValueError(f"{name}: expected a contiguous tensor")
    if t.numel() < n_expected:
        raise ValueError(f"{name}: {t.numel()} elements, the layout needs {n_expected}")
    return C.c_void_p(t.data_ptr())


def _layout(lo, hi, ng) -> BoxLayout:
    return BoxLayout(tuple(h - l + 1 for l, h in zip(lo, hi)), ng)


def allocate_variables(layout: BoxLayout, device="cuda"):
    """U, Unew (5 comps, ghosted), Q (6, ghosted), D, F (5, interior), zeroed."""
    torch = _torch()
    z = lambda n: torch.zeros(n, dtype=torch.float64, device=device)
    return (z(layout.ghosted_len(NC)), z(layout.ghosted_len(NC)), z(layout.ghosted_len(NQ)),
            z(layout.interior_len(NC)), z(layout.interior_len(NC)))


def set_stencil_impl(impl: int):
    _lib.check(_lib.load().cns_set_stencil_impl(impl))


def get_stencil_impl() -> int:
    return _lib.load().cns_get_stencil_impl()


def fill_boundary(U, dim, dim_g):
    lay = BoxLayout(tuple(dim))
    lib = _stream()
    _lib.check(lib.cns_fill_boundary(_ptr(U, lay.ghosted_len(NC), "U"), _I3(*dim), _I3(*dim_g)))


def ctoprim(lo, hi, u, q, dx, ng, courno=None):
    """Returns the updated Courant number when ``courno`` is given (the
    reference passes 1e-50, cpu/advance.cpp:52), else None."""
    lay = _layout(lo, hi, ng)
    lib = _stream()
    c = None if courno is None else C.c_double(courno)
    _lib.check(lib.cns_ctoprim(_I3(*lo), _I3(*hi), _ptr(u, lay.ghosted_len(NC), "u"),
                               _ptr(q, lay.ghosted_len(NQ), "q"), _D3(*dx), ng,
                               None if c is None else C.byref(c)))
    return None if c is None else c.value


def hypterm(lo, hi, ng, dx, cons, q, flux):
    lay = _layout(lo, hi, ng)
    lib = _stream()
    _lib.check(lib.cns_hypterm(_I3(*lo), _I3(*hi), ng, _D3(*dx),
                               _ptr(cons, lay.ghosted_len(NC), "cons"),
                               _ptr(q, lay.ghosted_len(NQ), "q"),
                               _ptr(flux, lay.interior_len(NC), "flux")))


def diffterm(lo, hi, ng, dx, q, difflux, eta, alam):
    lay = _layout(lo, hi, ng)
    lib = _stream()
    _lib.check(lib.cns_diffterm(_I3(*lo), _I3(*hi), ng, _D3(*dx),
                                _ptr(q, lay.ghosted_len(NQ), "q"),
                                _ptr(difflux, lay.interior_len(NC), "difflux"), eta, alam))


def rk_update(phase, lo, hi, ng, u, unew, d, f, dt):
    lay = _layout(lo, hi, ng)
    lib = _stream()
    _lib.check(lib.cns_rk_update(phase, _I3(*lo), _I3(*hi), ng,
                                 _ptr(u, lay.ghosted_len(NC), "u"),
                                 _ptr(unew, lay.ghosted_len(NC), "unew"),
                                 _ptr(d, lay.interior_len(NC), "d"),
                                 _ptr(f, lay.interior_len(NC), "f"), dt))


class Stepper:
    """The reference's h_const plus new_advance: one box's constants and the
    device scratch of its step (cns_ctx)."""

    def __init__(self, layout: BoxLayout, dx=None, cfl=CFL, eta=ETA, alam=ALAM):
        self.layout = layout
        self.dx = tuple(dx) if dx is not None else box_dx(layout.n)
        self.cfl, self.eta, self.alam = cfl, eta, alam
        lib = _lib.load()
        h = C.c_void_p()
        _lib.check(lib.cns_ctx_create(C.byref(h), _I3(*layout.n), _D3(*self.dx), cfl, eta, alam))
        self._h = h

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            _lib.load().cns_ctx_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _arrays(self, U, Unew, Q, D, F):
        lay = self.layout
        return (_ptr(U, lay.ghosted_len(NC), "U"), _ptr(Unew, lay.ghosted_len(NC), "Unew"),
                _ptr(Q, lay.ghosted_len(NQ), "Q"), _ptr(D, lay.interior_len(NC), "D"),
                _ptr(F, lay.interior_len(NC), "F"))

    def new_advance(self, U, Unew, Q, D, F) -> float:
        """One RK3 step (cpu/advance.cpp:13-154); returns dt."""
        lib = _stream()
        dt = C.c_double(0.0)
        _lib.check(lib.cns_advance(self._h, *self._arrays(U, Unew, Q, D, F), C.byref(dt)))
        return dt.value

    advance = new_advance

    def set_fused(self, fused: bool):
        """Fused stage (hypterm + update in one pass, default) or the reference's
        kernel-for-kernel composition; results are bit-identical."""
        _lib.check(_lib.load().cns_ctx_set_fused(self._h, int(fused)))

    def profile(self, enable: bool = True):
        """Turn live per-kernel timing on/off (clears the totals)."""
        _lib.check(_lib.load().cns_ctx_profile(self._h, int(enable)))

    def profile_read(self):
        """{kernel kind: (total ms, launches)} since the last profile() call."""
        ms = (C.c_double * 8)()
        n = (C.c_longlong * 8)()
        _lib.check(_lib.load().cns_ctx_profile_read(self._h, ms, n))
        return {k: (ms[i], n[i]) for i, k in enumerate(_lib.PROF_KINDS) if n[i]}

    def advance_async(self, U, Unew, Q, D, F, dt_dev, courno_dev=None):
        """The same step with dt (and the Courant number) left in device memory."""
        lib = _stream()
        _lib.check(lib.cns_advance_async(
            self._h, *self._arrays(U, Unew, Q, D, F), _ptr(dt_dev, 1, "dt_dev"),
            None if courno_dev is None else _ptr(courno_dev, 1, "courno_dev")))
