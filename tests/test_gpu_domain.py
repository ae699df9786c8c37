"""Multi-box domains (include/cns_domain.h) on the GPU.

* The box fill: every ghost cell of every box holds the domain cell at the periodic
  image of its global index (closed form, on index-encoded values), for a domain
  of boxes along all three axes and for a single box (where it is the plain fill).
* The step: a domain cut into boxes advances bit-identically to one box of the
  whole domain through cns_advance -- dt, and every box's U, Unew and Q (these two
  with their ghosts), D and F -- which the oracle and the reference program pin in
  test_gpu_parity.py.  Cubic and non-cubic domains, 1 to 24 boxes.
* A non-physical state is CNS_ENUMERIC; misaligned arrays are refused."""
import numpy as np
import pytest

from cns_nospec_amd import _lib, ops
from cns_nospec_amd.domain import Domain
from cns_nospec_amd.layout import NC, NQ, BoxLayout
from cns_nospec_amd.state import box_dx, sinusoid_state

pytestmark = pytest.mark.gpu


def _encoded(ncell):
    """[l][z][y][x] values that name their cell: l*1e7 + global linear index."""
    nx, ny, nz = ncell
    idx = np.arange(nx * ny * nz, dtype=np.float64).reshape(nz, ny, nx)
    return np.stack([idx + 1e7 * l for l in range(NC)])


@pytest.mark.parametrize("ncell,m", [((16, 24, 8), 8), ((12, 12, 12), 12), ((8, 8, 16), 4)])
def test_domain_fill_closed_form(device, ncell, m):
    import torch
    dom = Domain(ncell, m)
    U = torch.full((dom.array_len(NC, True),), -1.0, dtype=torch.float64, device=device)
    whole = _encoded(ncell)
    dom.scatter(torch.from_numpy(whole).to(device), NC, U)
    print(f"domain fill {ncell} m={m} boxes={dom.nbox}", flush=True)
    dom.fill_boundary(U)
    lay = dom.box_layout
    g = lay.dimg
    got = U.cpu().numpy().reshape(dom.nbox, NC, g[2], g[1], lay.xg)[..., : g[0]]
    ng = lay.ng
    a = np.arange(-ng, m + ng)
    for n in range(dom.nbox):
        lo, hi = dom.box(n)
        assert tuple(h - l + 1 for l, h in zip(lo, hi)) == (m, m, m)
        xs = (lo[0] + a) % ncell[0]
        ys = (lo[1] + a) % ncell[1]
        zs = (lo[2] + a) % ncell[2]
        expect = whole[:, zs][:, :, ys][:, :, :, xs]
        assert np.array_equal(got[n], expect), f"box {n} at {lo}"
    dom.close()


def _single_box_run(ncell, U0, nsteps, device):
    import torch
    lay = BoxLayout(ncell)
    U = torch.from_numpy(U0).to(device)
    _, Unew, Q, D, F = ops.allocate_variables(lay, device)
    st = ops.Stepper(lay, dx=box_dx(ncell))
    dts = [st.new_advance(U, Unew, Q, D, F) for _ in range(nsteps)]
    st.close()
    return lay, dts, [a.cpu().numpy() for a in (U, Unew, Q, D, F)]


@pytest.mark.parametrize("ncell,m", [((32, 32, 32), 16), ((48, 32, 16), 16), ((24, 16, 8), 8),
                                     ((20, 20, 20), 20)])
def test_domain_step_equals_single_box(device, ncell, m):
    import torch
    nsteps = 3
    lay = BoxLayout(ncell)
    U0 = sinusoid_state(lay, 7 + sum(ncell), mach=0.25, kmax=3)
    print(f"domain step {ncell} m={m}", flush=True)
    lay, dts_ref, (Ur, Unewr, Qr, Dr, Fr) = _single_box_run(ncell, U0, nsteps, device)

    dom = Domain(ncell, m)
    U, Unew, Q, D, F = dom.allocate(device)
    whole0 = torch.from_numpy(np.ascontiguousarray(lay.interior_of(lay.gview(U0, NC)))).to(device)
    dom.scatter(whole0, NC, U)
    dts = [dom.advance(U, Unew, Q, D, F) for _ in range(nsteps)]
    assert dts == dts_ref

    blay, ng = dom.box_layout, lay.ng
    g = blay.dimg
    views = {"U": (Ur, U, NC, True), "Unew": (Unewr, Unew, NC, True), "Q": (Qr, Q, NQ, True),
             "D": (Dr, D, NC, False), "F": (Fr, F, NC, False)}
    for name, (ref, got, nc, ghosted) in views.items():
        got = got.cpu().numpy()
        if ghosted:
            rv = lay.gview(ref, nc)
            gv = got.reshape(dom.nbox, nc, g[2], g[1], blay.xg)[..., : g[0]]
        else:
            rv = lay.iview(ref, nc)
            gv = got.reshape(dom.nbox, nc, m, m, blay.x)[..., :m]
        for n in range(dom.nbox):
            (x0, y0, z0), _ = dom.box(n)
            if ghosted and name != "U":
                # the box's ghosted block is the whole domain's ghosted window around it
                # (the single box's ghosts are periodic images too); U's ghosts hold U^n,
                # its interior U^{n+1}, so U is compared on the interior
                ext = m + 2 * ng
                ix = (np.arange(x0, x0 + ext) - ng) % ncell[0] + ng
                iy = (np.arange(y0, y0 + ext) - ng) % ncell[1] + ng
                iz = (np.arange(z0, z0 + ext) - ng) % ncell[2] + ng
                expect = rv[:, iz][:, :, iy][:, :, :, ix]
                assert np.array_equal(gv[n], expect), f"{name} box {n}"
            elif ghosted:
                inner = gv[n][:, ng:ng + m, ng:ng + m, ng:ng + m]
                assert np.array_equal(inner, rv[:, ng + z0:ng + z0 + m, ng + y0:ng + y0 + m,
                                                ng + x0:ng + x0 + m]), f"{name} box {n}"
            else:
                assert np.array_equal(gv[n], rv[:, z0:z0 + m, y0:y0 + m, x0:x0 + m]), f"{name} box {n}"
    dom.close()


def test_domain_plotfile(device, tmp_path):
    """Domain.write_plotfile: one FAB per box; reassembled, the boxes are the domain's
    state (gathered) bit for bit."""
    import torch
    from cns_nospec_amd import refio
    ncell, m = (24, 16, 8), 8
    dom = Domain(ncell, m)
    U, Unew, Q, D, F = dom.allocate(device)
    lay = BoxLayout(ncell)
    U0 = sinusoid_state(lay, 11, mach=0.2, kmax=2)
    dom.scatter(torch.from_numpy(np.ascontiguousarray(lay.interior_of(lay.gview(U0, NC)))).to(device), NC, U)
    dt = dom.advance(U, Unew, Q, D, F)
    d = str(tmp_path / "plt00001")
    dom.write_plotfile(d, U, dt)
    rd = refio.read_plotfile(d)
    whole = dom.gather(U, NC).cpu().numpy()
    assert rd["time"] == dt and len(rd["boxes"]) == dom.nbox
    got = np.full_like(whole, np.nan)
    for (lo, hi), data in zip(rd["boxes"], rd["data"]):
        got[:, lo[2]:hi[2] + 1, lo[1]:hi[1] + 1, lo[0]:hi[0] + 1] = data
    assert np.array_equal(got, whole)
    dom.close()


def test_domain_nonphysical_and_alignment(device):
    import torch
    dom = Domain((16, 16, 16), 8)
    U, Unew, Q, D, F = dom.allocate(device)
    whole = torch.ones((NC, 16, 16, 16), dtype=torch.float64, device=device)
    whole[4] = -1.0  # negative energy: negative pressure everywhere
    dom.scatter(whole, NC, U)
    with pytest.raises(_lib.CnsNumericError):
        dom.advance(U, Unew, Q, D, F)
    big = torch.zeros(dom.array_len(NC, True) + 1, dtype=torch.float64, device=device)
    with pytest.raises(_lib.CnsError, match="aligned"):
        dom.advance(big[1:], Unew, Q, D, F)
    dom.close()
