"""Parity of the HIP library (through its C ABI) with the oracle and with the
golden vectors of the reference cpu/ path.

The library keeps the reference's per-cell operation order and is built
without FP contraction, so the bar is BIT-EXACT equality for every array,
every ghost cell and every dt (np.array_equal on float64)."""
import hashlib

import numpy as np
import pytest

from cns_nospec_amd import _lib, ops
from cns_nospec_amd.layout import NC, NQ, BoxLayout
from cns_nospec_amd.state import ALAM, CFL, ETA, box_dx, cgs_state, pack_ghosted, sinusoid_state
from conftest import golden

pytestmark = pytest.mark.gpu

# "tiled" runs the DMA-pipelined kernels (the default for aligned arrays),
# "staged" the register-staged tiled kernels (CNS_PIPE=0), "direct" the
# one-thread-per-cell restatement.
IMPLS = {"tiled": _lib.CNS_IMPL_TILED, "staged": _lib.CNS_IMPL_TILED, "direct": _lib.CNS_IMPL_DIRECT}
SIZES = [(16, 16, 16), (20, 12, 17), (64, 64, 64), (70, 33, 40), (4, 4, 4), (130, 9, 5)]


def dev(a, device):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).to(device)


def host(t):
    import torch
    torch.cuda.synchronize()
    return t.cpu().numpy()


def _normwise(a, b):
    den = np.abs(b).max()
    return np.abs(a - b).max() / (den if den > 0 else 1.0)


@pytest.fixture(params=list(IMPLS))
def impl(request, device, monkeypatch):
    ops.set_stencil_impl(IMPLS[request.param])
    if request.param == "staged":
        monkeypatch.setenv("CNS_PIPE", "0")
    yield request.param
    ops.set_stencil_impl(_lib.CNS_IMPL_TILED)


@pytest.mark.parametrize("n", SIZES)
def test_kernels_bit_exact(orc, device, impl, n):
    lay = BoxLayout(n)
    dx = box_dx(n)
    U = sinusoid_state(lay, 100 + sum(n), mach=0.3, kmax=4)
    gv = lay.gview(U, NC)
    gv[:, :4] = -3.0
    gv[:, :, -4:] = 7.0  # stale ghosts the fill must replace
    Ud = dev(U, device)
    ops.fill_boundary(Ud, lay.n, lay.dimg)
    orc.fill_boundary(lay, U)
    assert np.array_equal(host(Ud), U), "fill_boundary"

    Q = lay.new_ghosted(NQ)
    Qd = dev(Q, device)
    c_gpu = ops.ctoprim(lay.lo, lay.hi, Ud, Qd, dx, lay.ng, 1.0e-50)
    c_orc = orc.ctoprim(lay, U, Q, dx, 1.0e-50)
    assert c_gpu == c_orc, "courno"
    assert np.array_equal(host(Qd), Q), "ctoprim"
    Qd2 = dev(lay.new_ghosted(NQ), device)
    assert ops.ctoprim(lay.lo, lay.hi, Ud, Qd2, dx, lay.ng) is None
    assert np.array_equal(host(Qd2), Q), "ctoprim without courno"

    F, D = lay.new_interior(NC), lay.new_interior(NC)
    Fd, Dd = dev(F, device), dev(D, device)
    ops.hypterm(lay.lo, lay.hi, lay.ng, dx, Ud, Qd, Fd)
    orc.hypterm(lay, dx, U, Q, F)
    got = host(Fd)
    assert np.array_equal(got, F), f"hypterm normwise {_normwise(got, F):.3e}"
    ops.diffterm(lay.lo, lay.hi, lay.ng, dx, Qd, Dd, ETA, ALAM)
    orc.diffterm(lay, dx, Q, D, ETA, ALAM)
    got = host(Dd)
    assert np.array_equal(got, D), f"diffterm normwise {_normwise(got, D):.3e}"

    Unew = sinusoid_state(lay, 7, mach=0.1)
    Unewd = dev(Unew, device)
    for phase, dt in ((1, 1.25e-3), (2, 1.25e-3), (3, 1.25e-3)):
        ops.rk_update(phase, lay.lo, lay.hi, lay.ng, Ud, Unewd, Dd, Fd, dt)
        orc.rk_update(phase, lay, U, Unew, D, F, dt)
        assert np.array_equal(host(Ud), U) and np.array_equal(host(Unewd), Unew), f"rk {phase}"


def test_kernels_vs_reference_golden(device, impl):
    for name in ("kernels_16.npz", "kernels_nc.npz"):
        g = golden(name)
        lay = BoxLayout(tuple(int(v) for v in g["n"]))
        dx = box_dx(lay.n)
        Ud = dev(pack_ghosted(lay, g["U0"]), device)
        Qd = dev(lay.new_ghosted(NQ), device)
        Fd, Dd = dev(lay.new_interior(NC), device), dev(lay.new_interior(NC), device)
        assert ops.ctoprim(lay.lo, lay.hi, Ud, Qd, dx, lay.ng, 1.0e-50) == float(g["courno"])
        ops.hypterm(lay.lo, lay.hi, lay.ng, dx, Ud, Qd, Fd)
        ops.diffterm(lay.lo, lay.hi, lay.ng, dx, Qd, Dd, ETA, ALAM)
        q = host(Qd)
        assert np.array_equal(lay.interior_of(lay.gview(q, NQ)), g["Q"])
        assert np.array_equal(lay.iview(host(Fd), NC), g["F"])
        assert np.array_equal(lay.iview(host(Dd), NC), g["D"])


def _stepper_run(lay, U0_host, nsteps, device):
    st = ops.Stepper(lay)
    U = dev(U0_host, device)
    _, Unew, Q, D, F = ops.allocate_variables(lay, device)
    for _ in range(nsteps):
        dt = st.new_advance(U, Unew, Q, D, F)
        yield dt, U, Unew, Q, D, F


def test_run_32_matches_reference_sha(device, impl):
    g = golden("run_32.npz")
    lay = BoxLayout.cube(32)
    idx = g["sample_idx"]
    for s, (dt, U, *_rest) in enumerate(_stepper_run(lay, pack_ghosted(lay, g["U0"]), 10, device)):
        assert dt == g["dt"][s]
        cur = np.ascontiguousarray(lay.interior_of(lay.gview(host(U), NC)))
        assert hashlib.sha256(cur.tobytes()).hexdigest() == str(g["sha256"][s]), f"step {s + 1}"
        assert np.array_equal(cur[idx[:, 0], idx[:, 1], idx[:, 2], idx[:, 3]], g["samples"][s])


def test_cgs_16_matches_reference(device):
    g = golden("cgs_16.npz")
    lay = BoxLayout.cube(16)
    out = list(_stepper_run(lay, pack_ghosted(lay, g["U0"]), 3, device))
    assert [d for d, *_ in out] == list(g["dt"])
    assert np.array_equal(lay.interior_of(lay.gview(host(out[-1][1]), NC)), g["U3"])


@pytest.mark.parametrize("name", ["cgs_cell32.npz", "cgs_cell64.npz"])
def test_cgs_decks_100_steps(orc, device, name):
    """The reference's cell32 / cell64 decks (fortran90/cell32, fortran90/cell64: one
    box, nsteps 100) on its CGS initial condition (fortran90/init_data.f90:43-85):
    after every step dt and the SHA-256 of the interior equal the reference cpu/
    path's (golden fixture); every ABI array equals the oracle's after the last
    step (and U, ghosts included, after every step at 32^3)."""
    g = golden(name)
    n = int(g["n"])
    lay = BoxLayout.cube(n)
    dx = box_dx(n)
    U = cgs_state(lay)
    Unew, Q = lay.new_ghosted(NC), lay.new_ghosted(NQ)
    D, F = lay.new_interior(NC), lay.new_interior(NC)
    idx = g["sample_idx"]
    nsteps = int(g["steps"])
    print(f"CGS deck {n}^3 x {nsteps} steps", flush=True)
    run = _stepper_run(lay, U.copy(), nsteps, device)
    for s, (dt, Ud, Unewd, Qd, Dd, Fd) in enumerate(run, 1):
        assert dt == g["dt"][s - 1], f"dt at step {s}"
        got = host(Ud)
        cur = np.ascontiguousarray(lay.interior_of(lay.gview(got, NC)))
        assert hashlib.sha256(cur.tobytes()).hexdigest() == str(g["sha256"][s - 1]), f"step {s}"
        assert np.array_equal(cur[idx[:, 0], idx[:, 1], idx[:, 2], idx[:, 3]], g["samples"][s - 1])
        if n == 32 or s == nsteps:
            dto = orc.advance(lay, dx, CFL, ETA, ALAM, U, Unew, Q, D, F) if n == 32 else None
            if n != 32:  # the oracle's 100 steps at once, then every array compared
                for _ in range(nsteps):
                    dto = orc.advance(lay, dx, CFL, ETA, ALAM, U, Unew, Q, D, F)
            assert dt == dto
            assert np.array_equal(got, U), f"U at step {s}"
            if s == nsteps:
                for a, b, nm in ((Unewd, Unew, "Unew"), (Qd, Q, "Q"), (Dd, D, "D"), (Fd, F, "F")):
                    assert np.array_equal(host(a), b), f"{nm} at step {s}"


def test_64_cube_100_steps_every_step(orc, device):
    """64^3, 100 RK3 steps, every conserved field, Unew, Q, D, F and dt compared
    with the oracle after every step."""
    lay = BoxLayout.cube(64)
    dx = box_dx(64)
    U = sinusoid_state(lay, 64)
    Unew, Q = lay.new_ghosted(NC), lay.new_ghosted(NQ)
    D, F = lay.new_interior(NC), lay.new_interior(NC)
    worst = 0.0
    for s, (dt, Ud, Unewd, Qd, Dd, Fd) in enumerate(_stepper_run(lay, U.copy(), 100, device), 1):
        dto = orc.advance(lay, dx, CFL, ETA, ALAM, U, Unew, Q, D, F)
        got = host(Ud)
        worst = max(worst, _normwise(got, U))
        assert dt == dto, f"dt at step {s}"
        assert np.array_equal(got, U), f"U at step {s}: normwise {_normwise(got, U):.3e}"
        if s % 25 == 0:
            for a, b, name in ((Unewd, Unew, "Unew"), (Qd, Q, "Q"), (Dd, D, "D"), (Fd, F, "F")):
                assert np.array_equal(host(a), b), f"{name} at step {s}"
    assert worst == 0.0


@pytest.mark.parametrize("n", [(70, 33, 40), (24, 24, 24), (20, 6, 5), (4, 4, 4)])
def test_step_compositions_agree(device, monkeypatch, n):
    """Every way the library composes a step leaves every ABI array bit-identical:
    pipelined kernels fused (hypterm + stage update in one pass, with the stage's
    primitives stored by that pass or by a ctoprim pass per stage; 32 x 8 or 64 x 8
    tiles) and unfused,
    register-staged kernels fused and unfused, and the direct kernels.
    Non-cubic boxes go beyond the reference cpu/ path (cubic only), so the direct
    one-thread-per-cell kernels are the cross-check here.  Extents below 8 give cells
    two periodic images along an axis (the pipelined step stores ghost images from
    its stage kernel instead of filling)."""
    lay = BoxLayout(n)
    U0 = sinusoid_state(lay, 9, mach=0.2, kmax=3)
    results = {}
    for name, pipe, fused, impl, stage_q, tx in (
            ("pipe_fused", "1", True, _lib.CNS_IMPL_TILED, "1", "32"),
            ("pipe_fused_wide_tiles", "1", True, _lib.CNS_IMPL_TILED, "1", "64"),
            ("pipe_fused_ctoprim", "1", True, _lib.CNS_IMPL_TILED, "0", "32"),
            ("pipe_unfused", "1", False, _lib.CNS_IMPL_TILED, "1", "32"),
            ("staged_fused", "0", True, _lib.CNS_IMPL_TILED, "1", "32"),
            ("staged_unfused", "0", False, _lib.CNS_IMPL_TILED, "1", "32"),
            ("direct", "1", False, _lib.CNS_IMPL_DIRECT, "1", "32")):
        monkeypatch.setenv("CNS_PIPE", pipe)
        monkeypatch.setenv("CNS_STAGE_Q", stage_q)
        monkeypatch.setenv("CNS_DIFF_TX", tx)
        monkeypatch.setenv("CNS_HYP_TX", tx)
        ops.set_stencil_impl(impl)
        try:
            st = ops.Stepper(lay)
            st.set_fused(fused)
            U = dev(U0, device)
            _, Unew, Q, D, F = ops.allocate_variables(lay, device)
            dts = [st.new_advance(U, Unew, Q, D, F) for _ in range(3)]
            results[name] = (dts, [host(a) for a in (U, Unew, Q, D, F)])
            st.close()
        finally:
            ops.set_stencil_impl(_lib.CNS_IMPL_TILED)
    ref_dts, ref_arrays = results.pop("direct")
    for name, (dts, arrays) in results.items():
        assert dts == ref_dts, name
        for a, b, what in zip(arrays, ref_arrays, ("U", "Unew", "Q", "D", "F")):
            assert np.array_equal(a, b), f"{name}: {what}"


def test_caller_workspace_gives_the_same_bits(device):
    """With a caller-owned workspace set (cns_set_scratch, sized by cns_scratch_bytes),
    cns_ctoprim's Courant reduction and cns_diffterm run in it and give the bits they
    give with the library's own per-thread buffers."""
    import ctypes as C
    import torch
    lay = BoxLayout((40, 24, 33))
    dx = box_dx(lay.n)
    U = dev(sinusoid_state(lay, 77, mach=0.25, kmax=3), device)
    lib = _lib.load()
    I3 = C.c_int * 3
    nbytes = lib.cns_scratch_bytes(I3(*lay.lo), I3(*lay.hi), lay.ng)
    assert nbytes == lay.cg * 8
    ws = torch.empty(nbytes // 8 + 32, dtype=torch.float64, device=device)
    base = (ws.data_ptr() + 255) // 256 * 256
    out = []
    for use_ws in (False, True):
        _lib.check(lib.cns_set_scratch(C.c_void_p(base if use_ws else None), nbytes if use_ws else 0))
        Q = dev(lay.new_ghosted(NQ), device)
        D = dev(lay.new_interior(NC, np.nan), device)
        c = ops.ctoprim(lay.lo, lay.hi, U, Q, dx, lay.ng, 1.0e-50)
        ops.diffterm(lay.lo, lay.hi, lay.ng, dx, Q, D, ETA, ALAM)
        out.append((c, host(Q), host(D)))
    _lib.check(lib.cns_set_scratch(None, 0))
    assert out[0][0] == out[1][0]
    assert np.array_equal(out[0][1], out[1][1])
    assert np.array_equal(lay.iview(out[0][2], NC), lay.iview(out[1][2], NC))
    assert not np.isnan(lay.iview(out[1][2], NC)).any()


def test_nonfinite_state_is_an_error(device):
    lay = BoxLayout.cube(16)
    U = sinusoid_state(lay, 3)
    lay.interior_of(lay.gview(U, NC))[0, 5, 6, 7] = -1.0  # negative density
    st = ops.Stepper(lay)
    Ud = dev(U, device)
    _, Unew, Q, D, F = ops.allocate_variables(lay, device)
    with pytest.raises(_lib.CnsNumericError):
        st.new_advance(Ud, Unew, Q, D, F)
    with pytest.raises(_lib.CnsError):
        ops.hypterm((3, 3, 3), (18, 18, 18), 3, box_dx(16), Ud, Q, D)


def test_side_stream(orc, device):
    import torch
    lay = BoxLayout((24, 20, 28))
    dx = box_dx(lay.n)
    U = sinusoid_state(lay, 9)
    s = torch.cuda.Stream()
    with torch.cuda.stream(s):
        Ud = dev(U, device)
        Qd = dev(lay.new_ghosted(NQ), device)
        Fd = dev(lay.new_interior(NC), device)
        c = ops.ctoprim(lay.lo, lay.hi, Ud, Qd, dx, lay.ng, 1e-50)
        ops.hypterm(lay.lo, lay.hi, lay.ng, dx, Ud, Qd, Fd)
    s.synchronize()
    Q, F = lay.new_ghosted(NQ), lay.new_interior(NC)
    assert c == orc.ctoprim(lay, U, Q, dx, 1e-50)
    orc.hypterm(lay, dx, U, Q, F)
    assert np.array_equal(Fd.cpu().numpy(), F)
