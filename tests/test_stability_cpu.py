"""The viscous stability limit that bounds how long the 512^3 workload can run.

dt comes from the acoustic Courant number alone (cpu/advance.cpp:52-58), so the
viscous number eta*dt/dx^2 grows like n at fixed eta.  At 512^3 with the reference's
eta = 1.8e-4 the grid-scale modes sit outside SSP-RK3's stability region and round-off
in them grows ~4x per step.  The same ratio is reached on a small box by scaling eta
with 512/n: the oracle step (the reference's arithmetic) then turns non-finite within
a few dozen steps, yet stays finite through the 20 steps the benchmark times from each
fresh state; at the 256^3-equivalent eta it stays bounded."""
import numpy as np

from oracle import Oracle
from cns_nospec_amd.layout import NC, NQ, BoxLayout
from cns_nospec_amd.state import ALAM, CFL, ETA, box_dx, sinusoid_state

N = 24


def _run(eta_equiv_edge: int, nsteps: int):
    """Steps of the oracle on an N^3 box with eta scaled to the viscous number of an
    eta_equiv_edge^3 box; returns the step index (1-based) of the first non-finite dt
    or state, or None, and the final state."""
    orc = Oracle()
    lay = BoxLayout((N, N, N))
    dx = box_dx(N)
    eta = ETA * eta_equiv_edge / N
    U = sinusoid_state(lay, 512, mach=0.085, kmax=2)
    Unew = np.zeros_like(U)
    Q = np.zeros(lay.ghosted_len(NQ))
    D = np.zeros(lay.interior_len(NC))
    F = np.zeros_like(D)
    for s in range(1, nsteps + 1):
        dt = orc.advance(lay, dx, CFL, eta, ALAM, U, Unew, Q, D, F)
        if not (np.isfinite(dt) and np.isfinite(lay.interior_of(lay.gview(U, NC))).all()):
            return s, U
    return None, U


def test_512_equivalent_eta_is_finite_for_20_steps_then_blows_up():
    bad, _ = _run(512, 20)
    assert bad is None
    bad, _ = _run(512, 60)
    assert bad is not None and 20 < bad <= 60


def test_256_equivalent_eta_stays_bounded():
    bad, U = _run(256, 60)
    assert bad is None
    lay = BoxLayout((N, N, N))
    rho = lay.interior_of(lay.gview(U, NC))[0]
    assert 0.9 < rho.min() and rho.max() < 1.1
