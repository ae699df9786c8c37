"""The oracle against the golden vectors written by the compiled reference cpu/
path (tests/golden/make_golden.py).  Bit-exact: the oracle restates the
reference's operation order and is built without FP contraction."""
import hashlib

import numpy as np
import pytest

from cns_nospec_amd.layout import NC, NQ, BoxLayout
from cns_nospec_amd.state import ALAM, CFL, ETA, box_dx, pack_ghosted
from conftest import golden


def _interior(lay, flat, ncomp):
    return np.ascontiguousarray(lay.interior_of(lay.gview(flat, ncomp)))


@pytest.mark.parametrize("name", ["kernels_16.npz", "kernels_nc.npz"])
def test_kernels_bit_exact(orc, name):
    g = golden(name)
    lay = BoxLayout(tuple(int(v) for v in g["n"]))
    dx = box_dx(lay.n)
    U = pack_ghosted(lay, g["U0"])
    orc.fill_boundary(lay, U)
    Q = lay.new_ghosted(NQ)
    courno = orc.ctoprim(lay, U, Q, dx, 1.0e-50)
    assert courno == float(g["courno"])
    assert np.array_equal(_interior(lay, Q, NQ), g["Q"])
    F = lay.new_interior(NC)
    D = lay.new_interior(NC)
    orc.hypterm(lay, dx, U, Q, F)
    orc.diffterm(lay, dx, Q, D, ETA, ALAM)
    assert np.array_equal(lay.iview(F, NC), g["F"])
    assert np.array_equal(lay.iview(D, NC), g["D"])


def _run(lay, U, nsteps):
    import oracle
    orc = oracle.Oracle()
    dx = box_dx(lay.n)
    Unew, Q = lay.new_ghosted(NC), lay.new_ghosted(NQ)
    D, F = lay.new_interior(NC), lay.new_interior(NC)
    for _ in range(nsteps):
        yield orc.advance(lay, dx, CFL, ETA, ALAM, U, Unew, Q, D, F), _interior(lay, U, NC)


def test_steps_16_bit_exact():
    g = golden("steps_16.npz")
    lay = BoxLayout.cube(16)
    U = pack_ghosted(lay, g["U0"])
    for s, (dt, cur) in enumerate(_run(lay, U, 10), start=1):
        assert dt == g["dt"][s - 1]
        if s in (1, 2, 10):
            assert np.array_equal(cur, g[f"U{s}"]), f"step {s}"


def test_run_32_bit_exact():
    g = golden("run_32.npz")
    lay = BoxLayout.cube(32)
    U = pack_ghosted(lay, g["U0"])
    idx = g["sample_idx"]
    for s, (dt, cur) in enumerate(_run(lay, U, 10)):
        assert dt == g["dt"][s]
        assert hashlib.sha256(cur.tobytes()).hexdigest() == str(g["sha256"][s])
        assert np.array_equal(cur[idx[:, 0], idx[:, 1], idx[:, 2], idx[:, 3]], g["samples"][s])


def test_cgs_16_bit_exact():
    g = golden("cgs_16.npz")
    lay = BoxLayout.cube(16)
    U = pack_ghosted(lay, g["U0"])
    out = list(_run(lay, U, 3))
    assert [d for d, _ in out] == list(g["dt"])
    assert np.array_equal(out[-1][1], g["U3"])


@pytest.mark.parametrize("name", ["cgs_cell32.npz", "cgs_cell64.npz"])
def test_cgs_decks_100_steps_bit_exact(name):
    """The reference's cell32 / cell64 decks (fortran90/cell32, fortran90/cell64: one
    box, 100 steps) on the CGS initial condition (fortran90/init_data.f90:43-85): the
    oracle's dt and interior SHA-256 equal the reference's after every step."""
    from cns_nospec_amd.state import cgs_state
    g = golden(name)
    n = int(g["n"])
    lay = BoxLayout.cube(n)
    U = cgs_state(lay)
    assert hashlib.sha256(_interior(lay, U, NC).tobytes()).hexdigest() == str(g["u0_sha256"])
    idx = g["sample_idx"]
    for s, (dt, cur) in enumerate(_run(lay, U, int(g["steps"]))):
        assert dt == g["dt"][s], f"dt at step {s + 1}"
        assert hashlib.sha256(cur.tobytes()).hexdigest() == str(g["sha256"][s]), f"step {s + 1}"
        assert np.array_equal(cur[idx[:, 0], idx[:, 1], idx[:, 2], idx[:, 3]], g["samples"][s])
