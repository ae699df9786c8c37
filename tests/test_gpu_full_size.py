"""The 512^3 workload itself (one periodic 512^3 box, 29 GB of ABI arrays).

One RK3 step through cns_advance is compared with the oracle's step on the same state,
bit for bit: dt and every ABI array (U, Unew, Q, D, F) over its whole extent (the oracle
runs on the host's cores; it gives the same bits for any thread count).  Independently
of the oracle, two properties that hold at any size: the ghost cells of U and Unew are
the periodic images of interior cells (closed form), and the interior sums of rho and of
the momenta are conserved by the step to rounding (central differences telescope on a
periodic box)."""
import numpy as np
import pytest

from cns_nospec_amd import ops
from cns_nospec_amd.layout import NC, NG, NQ, BoxLayout
from cns_nospec_amd.state import ALAM, CFL, ETA, box_dx, device_sinusoid_state

pytestmark = pytest.mark.gpu

N = 512


def _ghosts_are_images(g, n):
    """[l][k][j][i] ghosted view (device): every ghost layer equals its periodic image."""
    import torch
    for axis in (1, 2, 3):
        lo = g.narrow(axis, 0, NG)
        lo_img = g.narrow(axis, n, NG)
        hi = g.narrow(axis, n + NG, NG)
        hi_img = g.narrow(axis, NG, NG)
        if not (torch.equal(lo, lo_img) and torch.equal(hi, hi_img)):
            return False
    return True


def test_512_step_equals_oracle_and_conserves(device, orc):
    import torch
    lay = BoxLayout.cube(N)
    U, Unew, Q, D, F = ops.allocate_variables(lay, device)
    device_sinusoid_state(lay, 2024, device, mach=0.085, kmax=4, out=U)
    U0 = U.cpu().numpy()
    g0 = lay.interior_of(lay.gview(U, NC))
    sums0 = [float(g0[l].sum()) for l in range(4)]
    print("512^3: GPU step", flush=True)
    st = ops.Stepper(lay)
    dt = st.new_advance(U, Unew, Q, D, F)
    st.close()
    torch.cuda.synchronize()

    g1 = lay.interior_of(lay.gview(U, NC))
    for l in range(4):  # rho, rho u, rho v, rho w
        s1 = float(g1[l].sum())
        scale = float(g1[l].abs().sum())
        assert abs(s1 - sums0[l]) <= 1e-12 * scale, f"component {l}: {sums0[l]} -> {s1}"
    assert _ghosts_are_images(lay.gview(Unew, NC), N)
    # U's ghosts hold the images of U^n (the stage-1 fill), its interior U^{n+1}

    print("512^3: oracle step on the host", flush=True)
    orc.set_threads(16)
    Uo = U0
    Unew_o, Qo = lay.new_ghosted(NC), lay.new_ghosted(NQ)
    Do, Fo = lay.new_interior(NC), lay.new_interior(NC)
    dto = orc.advance(lay, box_dx(N), CFL, ETA, ALAM, Uo, Unew_o, Qo, Do, Fo)
    assert dt == dto
    for name, got, exp in (("U", U, Uo), ("Unew", Unew, Unew_o), ("Q", Q, Qo), ("D", D, Do), ("F", F, Fo)):
        assert torch.equal(got, torch.from_numpy(exp).to(device)), name
        print(f"512^3: {name} bit-identical", flush=True)
