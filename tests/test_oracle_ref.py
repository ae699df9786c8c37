"""The oracle against the compiled, unmodified reference cpu/ (oracle/_ref) on
fresh seeded states: bit-exact per kernel and per step.  Skipped where the
reference build is absent."""
import numpy as np
import pytest

from cns_nospec_amd.layout import NC, NQ, BoxLayout
from cns_nospec_amd.state import ALAM, CFL, ETA, box_dx, cgs_state, sinusoid_state


@pytest.mark.parametrize("n,seed,mach,kmax", [((16, 16, 16), 21, 0.085, 2),
                                              ((20, 12, 17), 22, 0.3, 4),
                                              ((4, 5, 6), 23, 0.1, 1)])
def test_kernels_vs_reference(orc, ref, n, seed, mach, kmax):
    lay = BoxLayout(n)
    dx = box_dx(n)
    ref.setup(n, dx, CFL, ETA, ALAM)
    U = sinusoid_state(lay, seed, mach=mach, kmax=kmax)
    gv = lay.gview(U, NC)
    gv[:, :4] = 9.0  # stale ghosts: the fill must overwrite them all
    Ua = lay.ghosted_to_aos(U, NC)
    orc.fill_boundary(lay, U)
    ref.fill_boundary(Ua)
    assert np.array_equal(lay.ghosted_to_aos(U, NC), Ua)
    Q = lay.new_ghosted(NQ)
    Qa = lay.ghosted_to_aos(Q, NQ)
    assert orc.ctoprim(lay, U, Q, dx, 1e-50) == ref.ctoprim(Ua, Qa, 1e-50)
    assert np.array_equal(lay.ghosted_to_aos(Q, NQ), Qa)
    F, D = lay.new_interior(NC), lay.new_interior(NC)
    Fa, Da = np.zeros(n + (NC,)), np.zeros(n + (NC,))
    orc.hypterm(lay, dx, U, Q, F)
    ref.hypterm(Ua, Qa, Fa)
    orc.diffterm(lay, dx, Q, D, ETA, ALAM)
    ref.diffterm(Qa, Da)
    assert np.array_equal(lay.interior_to_aos(F, NC), Fa)
    assert np.array_equal(lay.interior_to_aos(D, NC), Da)


@pytest.mark.parametrize("n,state", [(12, "sin"), (24, "sin"), (12, "cgs")])
def test_steps_vs_reference(orc, ref, n, state):
    lay = BoxLayout.cube(n)
    dx = box_dx(n)
    ref.setup(lay.n, dx, CFL, ETA, ALAM)
    U = sinusoid_state(lay, 31 + n, mach=0.2) if state == "sin" else cgs_state(lay)
    Ua = lay.ghosted_to_aos(U, NC)
    Unew, Q = lay.new_ghosted(NC), lay.new_ghosted(NQ)
    D, F = lay.new_interior(NC), lay.new_interior(NC)
    Unewa, Qa = lay.ghosted_to_aos(Unew, NC), lay.ghosted_to_aos(Q, NQ)
    Da, Fa = np.zeros(lay.n + (NC,)), np.zeros(lay.n + (NC,))
    for _ in range(4):
        dt1 = orc.advance(lay, dx, CFL, ETA, ALAM, U, Unew, Q, D, F)
        dt2 = ref.new_advance(Ua, Unewa, Qa, Da, Fa)
        assert dt1 == dt2
        assert np.array_equal(lay.ghosted_to_aos(U, NC), Ua)       # incl. ghosts
        assert np.array_equal(lay.ghosted_to_aos(Unew, NC), Unewa)
        assert np.array_equal(lay.ghosted_to_aos(Q, NQ), Qa)
        assert np.array_equal(lay.interior_to_aos(D, NC), Da)
        assert np.array_equal(lay.interior_to_aos(F, NC), Fa)
