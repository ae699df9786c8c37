"""The reference program's text formats (include/cns_io.h) on the CPU.

* The writer is pinned BYTE-FOR-BYTE to the unmodified reference program: run
  with nsteps = 0 it prints back the state it read (cpu/main.cpp:28-40), so its
  ``output`` must equal our write of the same state, and it must read our
  general_input/multistep_input (skipped where oracle/_ref is not built).
* Golden: our writer applied to the oracle's state after 3 steps reproduces the
  SHA-256 of the reference program's own ``output`` file (tests/golden/refio_16.npz),
  and the oracle's dt/courno lines are the program's.
* Round trips are bit-exact for every double, including signed zeros, subnormals
  and extremes; malformed, short or mismatched files raise CnsIOError.
* The driver binary rejects bad usage and missing testcases before any GPU call.
"""
import hashlib
import os
import subprocess

import numpy as np
import pytest

import oracle
from cns_nospec_amd import _lib, refio
from cns_nospec_amd.layout import NC, NQ, BoxLayout
from cns_nospec_amd.state import ALAM, CFL, ETA, box_dx
from conftest import golden

GOLD = golden("refio_16.npz")


def _state_from(fixture):
    z = golden(fixture)
    n = int(np.asarray(z["n"]).ravel()[0])
    lay = BoxLayout.cube(n)
    U = lay.new_ghosted(NC)
    lay.interior_of(lay.gview(U, NC))[...] = z["U0"]
    return lay, U


CASES = {"sin16": "steps_16.npz", "cgs16": "cgs_16.npz"}


def test_general_input_round_trip(tmp_path):
    cfg = refio.make_config((20, 12, 17), box_dx((20, 12, 17)), 0.45, 2.5e-4, 130.0, 7, dt=1e-3)
    p = str(tmp_path / "general_input")
    refio.write_general_input(p, cfg)
    back = refio.read_general_input(p)
    for f in ("ng", "nc", "ncells", "cfl", "eta", "alam", "nsteps", "dt"):
        assert getattr(back, f) == getattr(cfg, f), f
    assert list(back.lo) == [0, 0, 0] and list(back.hi) == [19, 11, 16]
    assert list(back.dx) == list(cfg.dx)
    assert list(back.dim) == [20, 12, 17] and list(back.dim_g) == [28, 20, 25]


def test_general_input_reference_layout(tmp_path):
    """A file as the reference's testcases lay it out (lo/hi on one line each,
    Fortran-style exponents) parses; dim = hi - lo + 1 (cpu/helper_functions.cpp:54-57)."""
    p = tmp_path / "general_input"
    p.write_text("4\n5\n32\n2 2 2\n33 33 33\n6.25E-3 6.25E-3 6.25e-03\n0.5\n1.8e-4\n1.5E2\n100\n0.0\n")
    c = refio.read_general_input(str(p))
    assert (c.ng, c.nc, c.ncells, c.nsteps) == (4, 5, 32, 100)
    assert list(c.dim) == [32, 32, 32] and list(c.dim_g) == [40, 40, 40]
    assert c.dx[0] == 6.25e-3 and c.alam == 150.0


@pytest.mark.parametrize("text", ["", "4 5 16\n0 0 0\n15 15\n", "4 5 16 0 0 0 15 15 15 x 1 1 0.5 1 1 3 0\n",
                                  "4 5 16 0 0 0 15 15 -3 1 1 1 0.5 1 1 3 0\n"])
def test_general_input_malformed(tmp_path, text):
    p = tmp_path / "general_input"
    p.write_text(text)
    with pytest.raises(_lib.CnsIOError):
        refio.read_general_input(str(p))


def test_missing_file(tmp_path):
    with pytest.raises(_lib.CnsIOError, match="cannot open"):
        refio.read_general_input(str(tmp_path / "nope"))
    with pytest.raises(_lib.CnsIOError, match="cannot open"):
        refio.read_field(str(tmp_path / "nope"), (12, 12, 12), NC)


def _special_field(lay, ncomp, seed):
    rng = np.random.default_rng(seed)
    U = lay.new_ghosted(ncomp)
    v = lay.gview(U, ncomp)
    v[...] = rng.standard_normal(v.shape) * 10.0 ** rng.integers(-300, 300, v.shape)
    flat = v.reshape(-1)
    specials = [0.0, -0.0, 5e-324, -5e-324, 2.2250738585072014e-308, 1.7976931348623157e308,
                -1.7976931348623157e308, 1.0, 0.1, 1.0 / 3.0, 2.0 ** -1074 * 3]
    flat[:len(specials)] = specials
    return U


@pytest.mark.parametrize("header", [False, True])
@pytest.mark.parametrize("n,ncomp", [((12, 12, 12), NC), ((9, 5, 7), NQ), ((4, 4, 4), 1)])
def test_field_round_trip_bit_exact(tmp_path, n, ncomp, header):
    lay = BoxLayout(n)
    U = _special_field(lay, ncomp, 5)
    p = str(tmp_path / "f")
    refio.write_field(p, U, lay.dimg, ncomp, header=header)
    back = refio.read_field(p, lay.dimg, ncomp, header=header)
    assert back.size == U.size
    a, b = lay.gview(back, ncomp), lay.gview(U, ncomp)
    assert np.array_equal(a.view(np.uint64), b.view(np.uint64))  # signed zeros included


def test_field_format_structure(tmp_path):
    """Header lines, one row per line, a blank line after each plane and after each
    component, "%.17e" tab-terminated values."""
    lay = BoxLayout((3, 2, 2))
    U = lay.new_ghosted(2)
    v = lay.gview(U, 2)
    v[...] = np.arange(v.size, dtype=np.float64).reshape(v.shape) * 0.5 - 7.0
    p = str(tmp_path / "f")
    refio.write_field(p, U, lay.dimg, 2, header=True)
    text = open(p).read()
    lines = text.split("\n")
    assert lines[0] == "2" and lines[1] == f"{lay.dimg[0]} {lay.dimg[1]} {lay.dimg[2]}"
    d0, d1, d2 = lay.dimg
    row0 = lines[2].split("\t")
    assert len(row0) == d0 + 1 and row0[-1] == ""
    assert row0[0] == "%.17e" % v[0, 0, 0, 0] and row0[1] == "%.17e" % v[0, 0, 0, 1]
    # 2 header lines + per component (d2 planes of (d1 rows + 1 blank) + 1 blank)
    assert text.count("\n") == 2 + 2 * (d2 * (d1 + 1) + 1)
    assert text.endswith("\t\n\n\n")


def test_field_errors(tmp_path):
    lay = BoxLayout((4, 4, 4))
    U = _special_field(lay, NC, 1)
    p = str(tmp_path / "f")
    refio.write_field(p, U, lay.dimg, NC, header=True)
    with pytest.raises(_lib.CnsIOError, match="header"):
        refio.read_field(p, lay.dimg, NQ, header=True)
    with pytest.raises(_lib.CnsIOError, match="header"):
        refio.read_field(p, (12, 12, 13), NC, header=True)
    with pytest.raises(_lib.CnsIOError, match="expected"):
        refio.read_field(p, lay.dimg, NC - 1, header=False)  # header ints read as values
    text = open(p).read()
    short = tmp_path / "short"
    short.write_text(text[: len(text) // 2].rsplit("\t", 1)[0] + "\t\n")
    with pytest.raises(_lib.CnsIOError, match="expected"):
        refio.read_field(str(short), lay.dimg, NC, header=True)
    bad = tmp_path / "bad"
    bad.write_text(text.replace("\t", "\tzz", 1))
    with pytest.raises(_lib.CnsIOError, match="not a number"):
        refio.read_field(str(bad), lay.dimg, NC, header=True)
    extra = tmp_path / "extra"
    extra.write_text(text + "1.0\n")
    with pytest.raises(_lib.CnsIOError, match="expected"):
        refio.read_field(str(extra), lay.dimg, NC, header=True)


def test_large_field_threads(tmp_path):
    """A field big enough to be split over many parse/format workers (chunk cuts
    inside rows and planes) still round-trips exactly."""
    lay = BoxLayout((60, 50, 40))
    rng = np.random.default_rng(3)
    U = lay.new_ghosted(NC)
    v = lay.gview(U, NC)
    v[...] = rng.standard_normal(v.shape)
    p = str(tmp_path / "big")
    refio.write_field(p, U, lay.dimg, NC, header=False)
    assert os.path.getsize(p) > 16 << 20
    back = refio.read_field(p, lay.dimg, NC, header=False)
    assert np.array_equal(lay.gview(back, NC), v)


@pytest.mark.skipif(not oracle.reference_program_available(), reason="oracle/_ref/cpu not built")
@pytest.mark.parametrize("case", list(CASES))
def test_writer_matches_reference_program(tmp_path, case):
    """nsteps = 0: the reference program prints back what it read."""
    lay, U = _state_from(CASES[case])
    cfg = refio.make_config(lay.n, box_dx(lay.n), CFL, ETA, ALAM, 0)
    refio.write_testcase(str(tmp_path / "testcases"), cfg, U)
    _, out = oracle.run_reference_program(str(tmp_path))
    ours = str(tmp_path / "ours")
    refio.write_field(ours, U, lay.dimg, NC, header=True)
    assert open(out, "rb").read() == open(ours, "rb").read()
    back = refio.read_field(out, lay.dimg, NC, header=True)
    assert np.array_equal(back, U)


@pytest.mark.parametrize("case", list(CASES))
def test_golden_output_via_oracle(tmp_path, orc, case):
    """The oracle's 3-step state, written by our writer, is byte-identical to the
    reference program's output file (by SHA-256); its dt/courno lines match."""
    lay, U = _state_from(CASES[case])
    Unew, Q = lay.new_ghosted(NC), lay.new_ghosted(NQ)
    D, F = lay.new_interior(NC), lay.new_interior(NC)
    steps = int(GOLD["steps"])
    lines = []
    dx = box_dx(lay.n)
    for _ in range(steps):
        dt = orc.advance(lay, dx, CFL, ETA, ALAM, U, Unew, Q, D, F)
        lines.append("dt, courno = %e, %e" % (dt, CFL / dt))
    p = str(tmp_path / "output")
    refio.write_field(p, U, lay.dimg, NC, header=True)
    data = open(p, "rb").read()
    assert len(data) == int(GOLD[f"{case}_size"])
    assert hashlib.sha256(data).hexdigest() == str(GOLD[f"{case}_sha256"])
    # courno is printed to 7 digits; cfl/dt recovers it to well within that
    assert [ln.split(",")[1] for ln in lines] == [ln.split(",")[1] for ln in GOLD[f"{case}_lines"]]
    assert [ln.split("=")[1].split(",")[0] for ln in lines] == \
        [ln.split("=")[1].split(",")[0] for ln in GOLD[f"{case}_lines"]]


def _driver():
    if not os.path.exists(refio.DRIVER):
        pytest.fail("cns_run is not built (python -m cns_nospec_amd.build)")
    return refio.DRIVER


def test_driver_usage_and_missing_input(tmp_path):
    drv = _driver()
    r = subprocess.run([drv, "-h"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 1 and "usage" in r.stderr
    r = subprocess.run([drv, "-d", str(tmp_path), "case"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 1 and "case_general_input" in r.stderr
    # wrong ng is refused before any device work
    cfg = refio.make_config(16, box_dx(16), CFL, ETA, ALAM, 1, ng=2)
    refio.write_general_input(str(tmp_path / "general_input"), cfg)
    r = subprocess.run([drv, "-d", str(tmp_path)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 1 and "ng = 2" in r.stderr
    # a short multistep_input is refused before any device work
    cfg = refio.make_config(16, box_dx(16), CFL, ETA, ALAM, 1)
    refio.write_general_input(str(tmp_path / "general_input"), cfg)
    (tmp_path / "multistep_input").write_text("1.0 2.0\n")
    r = subprocess.run([drv, "-d", str(tmp_path)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 1 and "expected" in r.stderr
