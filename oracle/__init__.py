"""The parity oracle (TEST INFRASTRUCTURE -- only tests/, the smoke check and
bench.py's cpu_baseline leg may import this package).

* :class:`Oracle` wraps ``liboracle.so``, the C restatement of the reference's
  cpu/ path in the ABI's padded SoA layout (cns_oracle.c).
* :class:`Reference` wraps ``_ref/libcnsref.so``: the unmodified reference
  cpu/ sources compiled by oracle/Makefile, driven in their own AoS layout
  ``[i][j][k][l]``.  It exists only where the reference tree was available
  at build time (and travels with the prebuilt files).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_ORACLE = os.path.join(HERE, "liboracle.so")
LIB_REF = os.path.join(HERE, "_ref", "libcnsref.so")
BIN_REF = os.path.join(HERE, "_ref", "cpu")
REFERENCE_CPU_DIR = "/root/reference/cpu"

_D = C.POINTER(C.c_double)
_I3 = C.c_int * 3
_D3 = C.c_double * 3


def build(ref: bool = True) -> None:
    """Compile liboracle.so (always) and _ref/ (when the reference tree is present)."""
    targets = ["oracle"] + (["ref"] if ref else [])
    subprocess.run(["make", "-s", "-C", HERE] + targets, check=True)


def _p(a):
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"], "need contiguous float64"
    return a.ctypes.data_as(_D)


class Oracle:
    """ctypes front end of liboracle.so (same argument order as the C ABI)."""

    def __init__(self, path: str = LIB_ORACLE):
        if not os.path.exists(path):
            build(ref=False)
        lib = C.CDLL(path)
        for name in ("orc_fill_boundary", "orc_ctoprim", "orc_hypterm", "orc_diffterm",
                     "orc_rk_update", "orc_advance", "orc_set_threads"):
            getattr(lib, name).restype = C.c_int
        self.lib = lib

    def set_threads(self, n: int = 0) -> int:
        """OpenMP threads of the per-cell loops (n > 0 sets it); returns the count in use.
        The results are the same for any count."""
        return int(self.lib.orc_set_threads(int(n)))

    def fill_boundary(self, layout, u):
        dim = _I3(*layout.n)
        dim_g = _I3(*layout.dimg)
        rc = self.lib.orc_fill_boundary(_p(u), dim, dim_g)
        assert rc == 0

    def ctoprim(self, layout, u, q, dx, courno=None):
        """Returns the updated courno when one is passed (cpu/ctoprim.cpp:11-70)."""
        c = None if courno is None else C.c_double(courno)
        rc = self.lib.orc_ctoprim(_I3(*layout.lo), _I3(*layout.hi), _p(u), _p(q), _D3(*dx),
                                  layout.ng, None if c is None else C.byref(c))
        assert rc == 0
        return None if c is None else c.value

    def hypterm(self, layout, dx, cons, q, flux):
        rc = self.lib.orc_hypterm(_I3(*layout.lo), _I3(*layout.hi), layout.ng, _D3(*dx),
                                  _p(cons), _p(q), _p(flux))
        assert rc == 0

    def diffterm(self, layout, dx, q, difflux, eta, alam):
        rc = self.lib.orc_diffterm(_I3(*layout.lo), _I3(*layout.hi), layout.ng, _D3(*dx),
                                   _p(q), _p(difflux), C.c_double(eta), C.c_double(alam))
        assert rc == 0

    def rk_update(self, phase, layout, u, unew, d, f, dt):
        rc = self.lib.orc_rk_update(phase, _I3(*layout.lo), _I3(*layout.hi), layout.ng,
                                    _p(u), _p(unew), _p(d), _p(f), C.c_double(dt))
        assert rc == 0

    def advance(self, layout, dx, cfl, eta, alam, U, Unew, Q, D, F):
        """One RK3 step (cpu/advance.cpp:13-154); returns dt."""
        assert layout.n[0] == layout.n[1] == layout.n[2]
        dt = C.c_double(0.0)
        rc = self.lib.orc_advance(layout.n[0], _D3(*dx), C.c_double(cfl), C.c_double(eta),
                                  C.c_double(alam), _p(U), _p(Unew), _p(Q), _p(D), _p(F),
                                  C.byref(dt))
        assert rc == 0
        return dt.value


def reference_available() -> bool:
    return os.path.exists(LIB_REF)


class Reference:
    """The unmodified reference cpu/ functions, on AoS buffers [i][j][k][l]."""

    def __init__(self, path: str = LIB_REF):
        if not os.path.exists(path):
            raise FileNotFoundError(f"{path} not built (needs the reference tree)")
        lib = C.CDLL(path)
        lib.ref_setup.restype = C.c_int
        lib.ref_setup.argtypes = [C.c_int] * 3 + [C.c_double] * 6
        self.lib = lib
        self.n = None

    def setup(self, n, dx, cfl=0.5, eta=1.8e-4, alam=1.5e2):
        self.n = tuple(n)
        self.lib.ref_setup(*self.n, *dx, cfl, eta, alam)

    def fill_boundary(self, U):
        self.lib.ref_fill_boundary(_p(U))

    def ctoprim(self, U, Q, courno=None):
        c = None if courno is None else C.c_double(courno)
        self.lib.ref_ctoprim(_p(U), _p(Q), None if c is None else C.byref(c))
        return None if c is None else c.value

    def hypterm(self, U, Q, F):
        self.lib.ref_hypterm(_p(U), _p(Q), _p(F))

    def diffterm(self, Q, D):
        self.lib.ref_diffterm(_p(Q), _p(D))

    def new_advance(self, U, Unew, Q, D, F):
        dt = C.c_double(0.0)
        self.lib.ref_new_advance(_p(U), _p(Unew), _p(Q), _p(D), _p(F), C.byref(dt))
        return dt.value


def reference_program_available() -> bool:
    return os.path.exists(BIN_REF)


def run_reference_program(workdir: str, timeout: float = 600):
    """Run the unmodified reference cpu/ program (oracle/_ref/cpu) on
    workdir/testcases/{general_input,multistep_input}.  It reads ../testcases/
    relative to its working directory (cpu/main.cpp:21, cpu/helper_functions.cpp:40)
    and writes ``output`` there, so it runs in workdir/run.  Returns
    (stdout, path of the output file)."""
    if not reference_program_available():
        raise FileNotFoundError(f"{BIN_REF} not built (needs the reference tree)")
    run = os.path.join(workdir, "run")
    os.makedirs(run, exist_ok=True)
    res = subprocess.run([BIN_REF], cwd=run, capture_output=True, text=True, timeout=timeout)
    if res.returncode != 0:
        raise RuntimeError(f"reference program failed ({res.returncode}): {res.stderr[-500:]}")
    return res.stdout, os.path.join(run, "output")
