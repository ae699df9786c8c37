"""Multi-box periodic domains (include/cns_domain.h) from Python.

The reference's BoxLib driver cuts its n_cell^3 domain into boxes of at most
max_grid_size cells (fortran90/main.f90:108-116), fills every box's ghost cells
from its neighbours (multifab_fill_boundary, fortran90/advance.f90:81) and steps
all boxes with one dt reduced over them (:139).  :class:`Domain` does the same on
the device: box-stacked arrays, one launch per kernel for all boxes.  Its state
after any number of steps equals, bit for bit, one box of the whole domain
advanced by :meth:`cns_nospec_amd.ops.Stepper.new_advance`.

``scatter``/``gather`` move a whole-domain interior array (``[l][k][j][i]``) into
and out of the box-stacked layout; they run as torch copies (data movement only).
"""
from __future__ import annotations

import ctypes as C

from . import _lib
from .layout import NC, NQ, BoxLayout
from .ops import _ptr, _stream, _torch
from .state import ALAM, CFL, ETA, box_dx

_I3 = C.c_int * 3
_D3 = C.c_double * 3


class Domain:
    """A periodic domain of ``ncell`` cells cut into ``max_grid_size``^3 boxes."""

    def __init__(self, ncell, max_grid_size: int, dx=None, cfl=CFL, eta=ETA, alam=ALAM):
        ncell = (ncell,) * 3 if isinstance(ncell, int) else tuple(int(v) for v in ncell)
        self.ncell = ncell
        self.m = int(max_grid_size)
        self.dx = tuple(dx) if dx is not None else box_dx(ncell)
        lib = _lib.load()
        h = C.c_void_p()
        _lib.check(lib.cns_domain_create(C.byref(h), _I3(*ncell), self.m, _D3(*self.dx), cfl, eta, alam))
        self._h = h
        nb, nbox = _I3(), C.c_int()
        _lib.check(lib.cns_domain_boxes(h, nb, C.byref(nbox)))
        self.nb = tuple(nb)
        self.nbox = nbox.value
        self.box_layout = BoxLayout.cube(self.m)

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            _lib.load().cns_domain_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ------------------------------------------------------------- geometry
    def box(self, n: int):
        """(lo, hi) cell range of box n, 0-based and inclusive."""
        lo, hi = _I3(), _I3()
        _lib.check(_lib.load().cns_domain_box(self._h, n, lo, hi))
        return tuple(lo), tuple(hi)

    def array_len(self, ncomp: int, ghosted: bool) -> int:
        return int(_lib.load().cns_domain_array_len(self._h, ncomp, int(ghosted)))

    def allocate(self, device="cuda"):
        """Zeroed U, Unew, Q, D, F domain arrays (box-stacked)."""
        torch = _torch()
        f = lambda nc, gh: torch.zeros(self.array_len(nc, gh), dtype=torch.float64, device=device)
        return f(NC, True), f(NC, True), f(NQ, True), f(NC, False), f(NC, False)

    # ----------------------------------------------------- layout conversion
    def _boxes_view(self, flat, ncomp: int, ghosted: bool):
        lay, m = self.box_layout, self.m
        if ghosted:
            g = lay.dimg
            v = flat.view(self.nbox, ncomp, g[2], g[1], lay.xg)
            ng = lay.ng
            return v[:, :, ng:ng + m, ng:ng + m, ng:ng + m]
        return flat.view(self.nbox, ncomp, m, m, lay.x)[:, :, :, :, :m]

    def scatter(self, whole, ncomp: int, out, ghosted: bool = True):
        """Copy a whole-domain interior array ``whole`` ([ncomp][nz][ny][nx], device)
        into the boxes' interiors of the box-stacked array ``out``."""
        v = self._boxes_view(out, ncomp, ghosted)
        m = self.m
        for n in range(self.nbox):
            (x0, y0, z0), _ = self.box(n)
            v[n].copy_(whole[:, z0:z0 + m, y0:y0 + m, x0:x0 + m])
        return out

    def gather(self, flat, ncomp: int, ghosted: bool = True):
        """The whole-domain interior ([ncomp][nz][ny][nx]) of a box-stacked array."""
        torch = _torch()
        v = self._boxes_view(flat, ncomp, ghosted)
        m = self.m
        out = torch.empty((ncomp, self.ncell[2], self.ncell[1], self.ncell[0]), dtype=flat.dtype,
                          device=flat.device)
        for n in range(self.nbox):
            (x0, y0, z0), _ = self.box(n)
            out[:, z0:z0 + m, y0:y0 + m, x0:x0 + m].copy_(v[n])
        return out

    # ------------------------------------------------------------------ step
    def fill_boundary(self, U):
        """multifab_fill_boundary of a 5-component box-stacked array."""
        lib = _stream()
        _lib.check(lib.cns_domain_fill_boundary(self._h, _ptr(U, self.array_len(NC, True), "U")))

    def write_plotfile(self, dirname: str, U, time: float, ncomp: int = NC):
        """BoxLib plotfile of a ghosted domain array, one FAB per box, as the Fortran
        driver writes it (fortran90/write_plotfile.f90:11-53; include/cns_io.h)."""
        from .refio import write_plotfile
        host = U.detach().cpu().numpy() if hasattr(U, "detach") else U
        write_plotfile(dirname, host, [self.box(n) for n in range(self.nbox)], self.dx, time,
                       ncomp=ncomp, ng=self.box_layout.ng)

    def advance(self, U, Unew, Q, D, F) -> float:
        """One SSP-RK3 step of every box (fortran90/advance.f90:57-410); returns dt."""
        lib = _stream()
        dt = C.c_double(0.0)
        _lib.check(lib.cns_domain_advance(
            self._h, _ptr(U, self.array_len(NC, True), "U"), _ptr(Unew, self.array_len(NC, True), "Unew"),
            _ptr(Q, self.array_len(NQ, True), "Q"), _ptr(D, self.array_len(NC, False), "D"),
            _ptr(F, self.array_len(NC, False), "F"), C.byref(dt)))
        return dt.value
