"""The reference-format driver (cns_run) on the GPU.

cns_run runs the reference program's main loop (gpu/main.cu:11-68) on the HIP
path.  Its ``output`` file must be BYTE-IDENTICAL to the one the unmodified
reference program writes for the same testcase files (SHA-256 from
tests/golden/refio_16.npz), and its per-step "dt, courno" lines identical to the
reference program's.  A 64^3 case is compared with the oracle's state written by
the (CPU-pinned) writer."""
import hashlib
import os
import sys

import numpy as np
import pytest

from cns_nospec_amd import refio
from cns_nospec_amd.layout import NC, NQ, BoxLayout
from cns_nospec_amd.state import ALAM, CFL, ETA, box_dx, sinusoid_state
from conftest import golden

pytestmark = pytest.mark.gpu

GOLD = golden("refio_16.npz")
CASES = {"sin16": "steps_16.npz", "cgs16": "cgs_16.npz"}


def _state_from(fixture):
    z = golden(fixture)
    n = int(np.asarray(z["n"]).ravel()[0])
    lay = BoxLayout.cube(n)
    U = lay.new_ghosted(NC)
    lay.interior_of(lay.gview(U, NC))[...] = z["U0"]
    return lay, U


def _run(tmp_path, **kw):
    print(f"cns_run on {tmp_path} {kw}", flush=True)
    r = refio.run_driver(str(tmp_path / "testcases"), output=str(tmp_path / "output"), timeout=120, **kw)
    sys.stdout.write(r.stdout[-2000:])
    assert r.returncode == 0, r.stderr
    return r


@pytest.mark.parametrize("case", list(CASES))
def test_driver_output_matches_reference_program(tmp_path, device, case):
    lay, U = _state_from(CASES[case])
    steps = int(GOLD["steps"])
    cfg = refio.make_config(lay.n, box_dx(lay.n), CFL, ETA, ALAM, steps)
    refio.write_testcase(str(tmp_path / "testcases"), cfg, U)
    r = _run(tmp_path)
    data = open(tmp_path / "output", "rb").read()
    assert len(data) == int(GOLD[f"{case}_size"])
    assert hashlib.sha256(data).hexdigest() == str(GOLD[f"{case}_sha256"])
    assert refio.step_lines(r.stdout) == [str(x) for x in GOLD[f"{case}_lines"]]
    assert "Total time: " in r.stdout


def test_driver_named_case_and_step_override(tmp_path, device, orc):
    """<name>_general_input / <name>_multistep_input and -n; a 64^3 box of seeded
    sinusoids (a named case of the reference's file format, not its cell64 CGS deck,
    which test_gpu_parity covers) against the oracle, byte for byte."""
    lay = BoxLayout.cube(64)
    U = sinusoid_state(lay, 11, kmax=2)
    cfg = refio.make_config(64, box_dx(64), CFL, ETA, ALAM, 100)
    refio.write_testcase(str(tmp_path / "testcases"), cfg, U, name="sin64")
    r = _run(tmp_path, name="sin64", nsteps=4, quiet=True)
    assert refio.step_lines(r.stdout) == []
    Unew, Q = lay.new_ghosted(NC), lay.new_ghosted(NQ)
    D, F = lay.new_interior(NC), lay.new_interior(NC)
    for _ in range(4):
        orc.advance(lay, box_dx(64), CFL, ETA, ALAM, U, Unew, Q, D, F)
    expect = str(tmp_path / "expect")
    refio.write_field(expect, U, lay.dimg, NC, header=True)
    assert open(tmp_path / "output", "rb").read() == open(expect, "rb").read()


def test_driver_nonphysical_state(tmp_path, device):
    """A state with negative pressure everywhere has no finite Courant number: the
    driver stops with exit code 3 and writes no output."""
    lay = BoxLayout.cube(16)
    U = lay.new_ghosted(NC)
    v = lay.gview(U, NC)
    v[0] = 1.0
    v[4] = -1.0
    cfg = refio.make_config(16, box_dx(16), CFL, ETA, ALAM, 2)
    refio.write_testcase(str(tmp_path / "testcases"), cfg, U)
    print("cns_run non-physical 16^3", flush=True)
    r = refio.run_driver(str(tmp_path / "testcases"), output=str(tmp_path / "output"), timeout=120)
    assert r.returncode == 3, (r.returncode, r.stderr)
    assert "not finite" in r.stderr
    assert not os.path.exists(tmp_path / "output")


def test_driver_plotfiles(tmp_path, device):
    """-p plot_int: plt00000 holds the input state, plt00002 / plt00004 and the last step
    (plt00005) are written as fortran90/main.f90:122-168 does; the last equals the
    output file's interior bit for bit, with time = the sum of the printed dts."""
    lay, U = _state_from(CASES["sin16"])
    cfg = refio.make_config(lay.n, box_dx(lay.n), CFL, ETA, ALAM, 5)
    refio.write_testcase(str(tmp_path / "testcases"), cfg, U)
    plots = tmp_path / "plots"
    plots.mkdir()
    r = _run(tmp_path, plot_int=2, plot_dir=str(plots))
    assert sorted(os.listdir(plots)) == ["plt00000", "plt00002", "plt00004", "plt00005"]
    first = refio.read_plotfile(str(plots / "plt00000"))
    assert first["time"] == 0.0
    assert np.array_equal(first["data"][0], lay.interior_of(lay.gview(U, NC)))
    last = refio.read_plotfile(str(plots / "plt00005"))
    out = refio.read_field(str(tmp_path / "output"), lay.dimg, NC, header=True)
    assert np.array_equal(last["data"][0], lay.interior_of(lay.gview(out, NC)))
    dts = [float(ln.split("=")[1].split(",")[0]) for ln in refio.step_lines(r.stdout)]
    assert abs(last["time"] - sum(dts)) <= 1e-6 * sum(dts)  # printed dts carry 7 digits
    assert last["boxes"] == [((0, 0, 0), tuple(v - 1 for v in lay.n))]
