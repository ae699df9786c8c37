"""Ghosted, padded SoA array layout of the CNS_NoSpec C ABI.

The layout is the one the reference's GPU main loop allocates (gpu/header.h:14,
gpu/util.cu:149-181, gpu/util_cpu.cpp:327-343):

* ghosted arrays (U, Unew: 5 components; Q: 6 components) hold the box plus
  ``ng = 4`` ghost cells on every side, component outermost, x unit-stride::

      idx = l*cg + k*pg + j*xg + i      xg = PAD16(n0 + 2ng), pg = xg*(n1 + 2ng), cg = pg*(n2 + 2ng)

* interior arrays (D, F: 5 components) have no ghosts::

      idx = l*c + k*p + j*x + i         x = PAD16(n0), p = x*n1, c = p*n2

``PAD16`` rounds x up to a multiple of 16 doubles; the padding lanes are never
read or written.  Index space is ghost-shifted as in cpu/advance.cpp:33-34:
``lo = ng`` and ``hi = n - 1 + ng`` per direction.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

NG = 4  # ghost cells (cpu/header.h:16)
NC = 5  # conserved components rho, rho*u, rho*v, rho*w, rho*E (cpu/header.h:15)
NQ = 6  # primitives rho, u, v, w, p, T (cpu/header.h:28-34)

IRHO, IMX, IMY, IMZ, IENE = range(5)
QRHO, QU, QV, QW, QPRES, QTEMP = range(6)
CONS_NAMES = ("rho", "mx", "my", "mz", "E")
PRIM_NAMES = ("rho", "u", "v", "w", "p", "T")


def pad16(x: int) -> int:
    """gpu/header.h:14 ``PAD(x)``: round up to a multiple of 16."""
    return (((x - 1) >> 4) + 1) << 4


@dataclass(frozen=True)
class BoxLayout:
    """Pitches and offsets of one box of interior extent ``n = (nx, ny, nz)``."""

    n: tuple
    ng: int = NG

    def __post_init__(self):
        n = tuple(int(v) for v in self.n)
        if len(n) != 3 or min(n) < 1:
            raise ValueError(f"bad box extent {self.n!r}")
        object.__setattr__(self, "n", n)

    @classmethod
    def cube(cls, edge: int) -> "BoxLayout":
        return cls((edge, edge, edge))

    @property
    def dimg(self):
        return tuple(v + 2 * self.ng for v in self.n)

    @property
    def lo(self):
        return (self.ng,) * 3

    @property
    def hi(self):
        return tuple(v - 1 + self.ng for v in self.n)

    @property
    def xg(self):
        return pad16(self.dimg[0])

    @property
    def pg(self):
        return self.xg * self.dimg[1]

    @property
    def cg(self):
        return self.pg * self.dimg[2]

    @property
    def x(self):
        return pad16(self.n[0])

    @property
    def p(self):
        return self.x * self.n[1]

    @property
    def c(self):
        return self.p * self.n[2]

    @property
    def cells(self) -> int:
        return self.n[0] * self.n[1] * self.n[2]

    def ghosted_len(self, ncomp: int) -> int:
        return ncomp * self.cg

    def interior_len(self, ncomp: int) -> int:
        return ncomp * self.c

    # ---- numpy views (host side; tests, I/O and state generation) ----
    def gview(self, flat, ncomp: int):
        """[l][k][j][i] view of a flat ghosted array, padding lanes cut off."""
        g = self.dimg
        return flat.reshape(ncomp, g[2], g[1], self.xg)[:, :, :, : g[0]]

    def iview(self, flat, ncomp: int):
        """[l][k][j][i] view of a flat interior array, padding lanes cut off."""
        n = self.n
        return flat.reshape(ncomp, n[2], n[1], self.x)[:, :, :, : n[0]]

    def interior_of(self, gv):
        """Interior cells of a ghosted [l][k][j][i] view."""
        ng, n = self.ng, self.n
        return gv[:, ng : ng + n[2], ng : ng + n[1], ng : ng + n[0]]

    def new_ghosted(self, ncomp: int, fill=0.0):
        return np.full(self.ghosted_len(ncomp), fill, dtype=np.float64)

    def new_interior(self, ncomp: int, fill=0.0):
        return np.full(self.interior_len(ncomp), fill, dtype=np.float64)

    # ---- conversions to the reference cpu/ AoS layout [i][j][k][l] (tests only) ----
    def ghosted_to_aos(self, flat, ncomp: int):
        return np.ascontiguousarray(self.gview(flat, ncomp).transpose(3, 2, 1, 0))

    def interior_to_aos(self, flat, ncomp: int):
        return np.ascontiguousarray(self.iview(flat, ncomp).transpose(3, 2, 1, 0))

    def aos_to_ghosted(self, aos, ncomp: int, out=None):
        out = self.new_ghosted(ncomp) if out is None else out
        self.gview(out, ncomp)[...] = aos.transpose(3, 2, 1, 0)
        return out

    def aos_to_interior(self, aos, ncomp: int, out=None):
        out = self.new_interior(ncomp) if out is None else out
        self.iview(out, ncomp)[...] = aos.transpose(3, 2, 1, 0)
        return out
