"""CPU-side checks of the C ABI: libcns_hip.so loads without a GPU and exports
every function include/cns.h declares; host logic of the layout and state."""
import ctypes
import os

import numpy as np
import pytest

from cns_nospec_amd import _lib
from cns_nospec_amd.layout import NC, BoxLayout, pad16
from cns_nospec_amd.state import box_dx, sinusoid_primitives, sinusoid_state


def test_library_exports_every_header_symbol():
    lib = _lib.load()
    syms = _lib.header_symbols()
    assert "cns_advance" in syms and "cns_ctoprim" in syms
    missing = [s for s in syms if not hasattr(lib, s)]
    assert not missing, f"declared in include/cns.h but not exported: {missing}"


def test_every_header_symbol_has_a_signature():
    for s in _lib.header_symbols():
        fn = getattr(_lib.load(), s)
        assert fn.restype is not None or s in ("cns_ctx_destroy",), s


def test_errors_are_reported_not_fatal():
    lib = _lib.load()
    I3, D3 = ctypes.c_int * 3, ctypes.c_double * 3
    # ng other than 4 is rejected before any device work
    rc = lib.cns_hypterm(I3(3, 3, 3), I3(10, 10, 10), 3, D3(1, 1, 1), None, None, None)
    assert rc == _lib.CNS_EINVAL
    assert b"ng" in lib.cns_last_error_string()
    rc = lib.cns_fill_boundary(None, I3(8, 8, 8), I3(16, 16, 16))
    assert rc == _lib.CNS_EINVAL
    rc = lib.cns_set_stencil_impl(7)
    assert rc == _lib.CNS_EINVAL
    with pytest.raises(_lib.CnsError):
        _lib.check(rc)


def test_layout_matches_reference_padding():
    # gpu/util_cpu.cpp:335-343 with PAD(x) = ((x-1)/16+1)*16
    lay = BoxLayout.cube(512)
    assert (lay.xg, lay.dimg) == (528, (520, 520, 520))
    assert lay.cg == 528 * 520 * 520 and lay.x == 512
    assert pad16(72) == 80 and pad16(40) == 48 and pad16(16) == 16
    lay = BoxLayout((20, 12, 17))
    flat = np.arange(lay.ghosted_len(NC), dtype=np.float64)
    assert lay.gview(flat, NC)[2, 3, 4, 5] == 2 * lay.cg + 3 * lay.pg + 4 * lay.xg + 5
    aos = lay.ghosted_to_aos(flat, NC)
    assert aos[5, 4, 3, 2] == flat[2 * lay.cg + 3 * lay.pg + 4 * lay.xg + 5]
    assert np.array_equal(lay.gview(lay.aos_to_ghosted(aos, NC), NC), lay.gview(flat, NC))


def test_seeded_state_is_deterministic_and_physical():
    a = sinusoid_primitives(16, 5)
    b = sinusoid_primitives(16, 5)
    for k in a:
        assert np.array_equal(a[k], b[k])
    assert abs(a["rho"].max() - 1.05) < 1e-12 and abs(a["rho"].min() - 0.95) < 0.02
    vmax = max(np.abs(a[k]).max() for k in "uvw")
    assert abs(vmax - 0.085 * np.sqrt(1.4)) < 1e-12
    lay = BoxLayout.cube(16)
    U = sinusoid_state(lay, 5)
    assert np.all(lay.gview(U, NC)[0] > 0)
    assert box_dx(16) == (0.0125, 0.0125, 0.0125)


def test_scratch_size_and_argument_checks():
    """cns_scratch_bytes is one ghosted component of doubles (the wz field; the Courant
    slot fits in it), 0 for bounds the library rejects; cns_set_scratch refuses a NULL
    buffer with a size and a misaligned one before touching the device, and NULL, 0
    (no workspace) is always accepted."""
    lib = _lib.load()
    I3 = ctypes.c_int * 3
    lay = BoxLayout((20, 12, 17))
    assert lib.cns_scratch_bytes(I3(*lay.lo), I3(*lay.hi), 4) == lay.cg * 8
    assert lib.cns_scratch_bytes(I3(*lay.lo), I3(*lay.hi), 3) == 0
    assert lib.cns_set_scratch(None, 0) == _lib.CNS_OK
    assert lib.cns_set_scratch(None, 64) == _lib.CNS_EINVAL
    assert lib.cns_set_scratch(ctypes.c_void_p(0x1008), 4096) == _lib.CNS_EINVAL
    assert b"aligned" in lib.cns_last_error_string()
