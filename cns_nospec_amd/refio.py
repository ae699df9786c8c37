"""The reference program's text formats and its main loop (include/cns_io.h).

Files of a testcase directory, as the reference reads and writes them:

* ``general_input`` -- ng, nc, ncells, lo[3], hi[3], dx[3], cfl, eta, alam, nsteps, dt
  (cpu/helper_functions.cpp:38-83, gpu/util_cpu.cpp:287-364);
* ``multistep_input`` -- the 5 conserved components over the ghosted box, component
  slowest and i fastest (read_3D, gpu/util_cpu.cpp:76-85);
* ``output`` -- "nc", the ghosted extents, then the same body as print_4D writes it
  (cpu/main.cpp:36-40, cpu/helper_functions.cpp:85-98).

Named testcases use ``<name>_general_input`` / ``<name>_multistep_input``
(gpu/util_cpu.cpp:292-301).  :func:`write_plotfile` / :func:`read_plotfile` handle the
BoxLib plotfiles of the Fortran driver (fortran90/write_plotfile.f90:11-53).  Parsing and formatting run in libcns_hip.so;
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
               quiet: bool = False, cwd: str = None, timeout: float = 600, plot_int: int = 0,
               plot_dir: str = None):
    """Run cns_run on a testcase directory; returns the CompletedProcess (text).
    plot_int > 0 also writes plotfiles plt%05d into plot_dir (fortran90/main.f90:122-168)."""
    if not os.path.exists(DRIVER):
        raise FileNotFoundError(f"{DRIVER} is not built: run `python -m cns_nospec_amd.build`")
    cmd = [DRIVER, "-d", directory, "-o", output]
    if nsteps is not None:
        cmd += ["-n", str(nsteps)]
    if quiet:
        cmd.append("-q")
    if plot_int:
        cmd += ["-p", str(plot_int)] + (["-P", plot_dir] if plot_dir else [])
    if name:
        cmd.append(name)
    return subprocess.run(cmd, capture_output=True, text=True, cwd=cwd, timeout=timeout)


def step_lines(stdout: str):
    """The per-step "dt, courno = ..." lines of a run's standard output."""
    return [ln for ln in stdout.splitlines() if ln.startswith("dt, courno = ")]


# ----------------------------------------------------------------------- plotfiles
def plotfile_name(istep: int) -> str:
    """The Fortran driver's plotfile directory name (fortran90/write_plotfile.f90:43)."""
    return f"plt{istep:05d}"


def write_plotfile(dirname: str, flat: np.ndarray, boxes, dx, time: float, ncomp: int = NC,
                   ng: int = NG, prob_lo=None, prob_hi=None):
    """A BoxLib plotfile of the valid cells of ``flat`` (include/cns_io.h:
    cns_write_plotfile).  ``boxes`` lists each box's (lo, hi) 0-based inclusive cell
    range; ``flat`` holds the boxes back to back, each a padded ghosted array with ng
    ghosts (one box: a BoxLayout array; a Domain's box-stacked array).  prob_lo/prob_hi
    default to the domain centred on the origin (-0.1 .. 0.1 for the reference's
    n_cell and dx, fortran90/main.f90:93-106)."""
    boxes = [(tuple(int(v) for v in lo), tuple(int(v) for v in hi)) for lo, hi in boxes]
    lo = np.array([b[0] for b in boxes], dtype=np.int32)
    hi = np.array([b[1] for b in boxes], dtype=np.int32)
    need = sum(field_len(tuple(h - l + 1 + 2 * ng for l, h in zip(*b)), ncomp) for b in boxes)
    flat = np.ascontiguousarray(flat, dtype=np.float64)
    if flat.size < need:
        raise ValueError(f"array of {flat.size} doubles; the boxes need {need}")
    ext = [(hi[:, d].max() - lo[:, d].min() + 1) * dx[d] for d in range(3)]
    if prob_hi is None:
        prob_hi = tuple(0.5 * e for e in ext)
    if prob_lo is None:
        prob_lo = tuple(h - e for h, e in zip(prob_hi, ext))
    _D3 = C.c_double * 3
    _lib.check(_lib.load().cns_write_plotfile(
        os.fsencode(dirname), ncomp, len(boxes), lo.ctypes.data, hi.ctypes.data, ng, flat.ctypes.data,
        _D3(*prob_lo), _D3(*prob_hi), _D3(*dx), float(time)))


def _box(tok: str):
    """((a,b,c) (d,e,f) (t,t,t)) -> ((a,b,c), (d,e,f))"""
    nums = [int(v) for v in tok.replace("(", " ").replace(")", " ").replace(",", " ").split()]
    return tuple(nums[0:3]), tuple(nums[3:6])


def read_plotfile(dirname: str) -> dict:
    """A one-level plotfile as written by :func:`write_plotfile`: the header fields and,
    per box, its (lo, hi) and its data as a [ncomp][nz][ny][nx] float64 array."""
    with open(os.path.join(dirname, "Header")) as f:
        lines = f.read().split("\n")
    it = iter(lines)
    version = next(it)
    nc = int(next(it))
    names = [next(it) for _ in range(nc)]
    dim = int(next(it))
    time = float(next(it))
    finest = int(next(it))
    prob_lo = tuple(float(v) for v in next(it).split())
    prob_hi = tuple(float(v) for v in next(it).split())
    next(it)  # refinement ratios (none for one level)
    domain = _box(next(it))
    next(it)  # level steps
    dx = tuple(float(v) for v in next(it).split())
    coord, bwidth = int(next(it)), int(next(it))
    lev, nbox, ltime = next(it).split()
    next(it)  # level step
    extents = [[tuple(float(v) for v in next(it).split()) for _ in range(dim)] for _ in range(int(nbox))]
    cell_path = next(it)
    with open(os.path.join(dirname, cell_path + "_H")) as f:
        h = f.read().split("\n")
    if h[2].strip() != str(nc) or h[3].strip() != "0":
        raise ValueError(f"{dirname}: Cell_H holds {h[2].strip()} components with {h[3].strip()} "
                         f"ghosts; the Header names {nc} components")
    nb = int(h[4].strip("( ").split()[0])
    box_list = [_box(h[5 + b]) for b in range(nb)]
    fabs = [ln.split() for ln in h if ln.startswith("FabOnDisk:")]
    data = []
    level_dir = os.path.dirname(os.path.join(dirname, cell_path))
    for (lo, hi), (_, fname, off) in zip(box_list, fabs):
        with open(os.path.join(level_dir, fname), "rb") as f:
            f.seek(int(off))
            head = f.readline().decode()
            if not head.startswith("FAB ((8, (64 11 52 0 1 12 0 1023)),(8, (8 7 6 5 4 3 2 1)))"):
                raise ValueError(f"{fname} at {off}: not a little-endian float64 FAB: {head[:80]!r}")
            n = [h_ - l_ + 1 for l_, h_ in zip(lo, hi)]
            vals = np.frombuffer(f.read(8 * nc * n[0] * n[1] * n[2]), dtype="<f8")
            data.append(vals.reshape(nc, n[2], n[1], n[0]).copy())
    return {"version": version, "names": names, "dim": dim, "time": time, "finest_level": finest,
            "prob_lo": prob_lo, "prob_hi": prob_hi, "domain": domain, "dx": dx, "coord_sys": coord,
            "bwidth": bwidth, "level": int(lev), "level_time": float(ltime), "box_extents": extents,
            "boxes": box_list, "data": data, "cell_h": h}

