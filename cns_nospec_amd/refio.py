"""The reference program's text formats and its main loop (include/cns_io.h).

Files of a testcase directory, as the reference reads and writes them:

* ``general_input`` -- ng, nc, ncells, lo[3], hi[3], dx[3], cfl, eta, alam, nsteps, dt
  (cpu/helper_functions.cpp:38-83, gpu/util_cpu.cpp:287-364);
* ``multistep_input`` -- the 5 conserved components over the ghosted box, component
  slowest and i fastest (read_3D, gpu/util_cpu.cpp:76-85);
* ``output`` -- "nc", the ghosted extents, then the same body as print_4D writes it
  (cpu/main.cpp:36-40, cpu/helper_functions.cpp:85-98).

Named testcases use ``<name>_general_input`` / ``<name>_multistep_input``
(gpu/util_cpu.cpp:292-301).  Parsing and formatting run in libcns_hip.so;
:func:`run_driver` runs ``cns_run``, the reference's main loop on the HIP path.
Host arrays here are flat float64 arrays in the padded ghosted layout of
:class:`.layout.BoxLayout` (what ``cns_field_len`` sizes).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

from . import _lib
from .layout import NC, NG

DRIVER = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cns_run")
Config = _lib.Config


def make_config(n, dx, cfl: float, eta: float, alam: float, nsteps: int, dt: float = 0.0,
                ng: int = NG, nc: int = NC) -> Config:
    """A config for an n[0] x n[1] x n[2] box (lo = 0, hi = n-1, as the reference's
    testcases write them); ncells is n[0] as in the reference's cubic cases."""
    n = (n, n, n) if np.isscalar(n) else tuple(n)
    c = Config()
    c.ng, c.nc, c.ncells = ng, nc, n[0]
    for d in range(3):
        c.lo[d], c.hi[d] = 0, n[d] - 1
        c.dx[d] = dx[d]
        c.dim[d] = n[d]
        c.dim_g[d] = n[d] + 2 * ng
    c.cfl, c.eta, c.alam, c.nsteps, c.dt = cfl, eta, alam, nsteps, dt
    return c


def read_general_input(path: str) -> Config:
    c = Config()
    _lib.check(_lib.load().cns_read_general_input(os.fsencode(path), C.byref(c)))
    return c


def write_general_input(path: str, cfg: Config):
    _lib.check(_lib.load().cns_write_general_input(os.fsencode(path), C.byref(cfg)))


def field_len(dim_g, ncomp: int) -> int:
    return int(_lib.load().cns_field_len((C.c_int * 3)(*dim_g), ncomp))


def read_field(path: str, dim_g, ncomp: int = NC, header: bool = False) -> np.ndarray:
    """A field file as a flat host array in the padded ghosted layout (padding = 0)."""
    out = np.zeros(field_len(dim_g, ncomp))
    _lib.check(_lib.load().cns_read_field(os.fsencode(path), (C.c_int * 3)(*dim_g), ncomp,
                                          int(header), out.ctypes.data))
    return out


def write_field(path: str, flat: np.ndarray, dim_g, ncomp: int = NC, header: bool = True):
    flat = np.ascontiguousarray(flat, dtype=np.float64)
    if flat.size < field_len(dim_g, ncomp):
        raise ValueError("array smaller than the padded field")
    _lib.check(_lib.load().cns_write_field(os.fsencode(path), (C.c_int * 3)(*dim_g), ncomp,
                                           int(header), flat.ctypes.data))


def testcase_paths(directory: str, name: str = None):
    pre = os.path.join(directory, f"{name}_" if name else "")
    return pre + "general_input", pre + "multistep_input"


def write_testcase(directory: str, cfg: Config, U: np.ndarray, name: str = None):
    """general_input + multistep_input of a state U (flat, padded ghosted layout)."""
    os.makedirs(directory, exist_ok=True)
    gpath, upath = testcase_paths(directory, name)
    write_general_input(gpath, cfg)
    write_field(upath, U, tuple(cfg.dim_g), NC, header=False)


def run_driver(directory: str, name: str = None, output: str = "output", nsteps: int = None,
               quiet: bool = False, cwd: str = None, timeout: float = 600):
    """Run cns_run on a testcase directory; returns the CompletedProcess (text)."""
    if not os.path.exists(DRIVER):
        raise FileNotFoundError(f"{DRIVER} is not built: run `python -m cns_nospec_amd.build`")
    cmd = [DRIVER, "-d", directory, "-o", output]
    if nsteps is not None:
        cmd += ["-n", str(nsteps)]
    if quiet:
        cmd.append("-q")
    if name:
        cmd.append(name)
    return subprocess.run(cmd, capture_output=True, text=True, cwd=cwd, timeout=timeout)


def step_lines(stdout: str):
    """The per-step "dt, courno = ..." lines of a run's standard output."""
    return [ln for ln in stdout.splitlines() if ln.startswith("dt, courno = ")]
