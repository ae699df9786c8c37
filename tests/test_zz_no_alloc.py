"""cns_diffterm_noalloc (include/cns.h) through the C ABI: the same bits as
cns_diffterm on two calls in a row, the second on a smaller box (the static
scratch keeps the first call's contents), and boxes above its scratch size are
rejected with CNS_EINVAL."""
import ctypes as C

import numpy as np
import pytest

from cns_nospec_amd import _lib
from cns_nospec_amd.layout import NC, NQ, BoxLayout
from cns_nospec_amd.state import ALAM, ETA, box_dx, sinusoid_state

pytestmark = pytest.mark.gpu

_I3 = C.c_int * 3
_D3 = C.c_double * 3


def _lib_on_stream():
    import torch
    lib = _lib.load()
    lib.cns_set_stream(C.c_void_p(torch.cuda.current_stream().cuda_stream))
    return lib


def _diffterm_both(lay, seed, device):
    """D from cns_diffterm and from cns_diffterm_noalloc on the same filled state."""
    import torch
    from cns_nospec_amd import ops
    U = torch.from_numpy(sinusoid_state(lay, seed, mach=0.25, kmax=3)).to(device)
    Q = torch.zeros(lay.ghosted_len(NQ), dtype=torch.float64, device=device)
    ops.ctoprim(lay.lo, lay.hi, U, Q, box_dx(lay.n), lay.ng)
    out = {}
    lib = _lib_on_stream()
    for name in ("cns_diffterm", "cns_diffterm_noalloc"):
        D = torch.full((lay.interior_len(NC),), np.nan, dtype=torch.float64, device=device)
        rc = getattr(lib, name)(_I3(*lay.lo), _I3(*lay.hi), lay.ng, _D3(*box_dx(lay.n)),
                                C.c_void_p(Q.data_ptr()), C.c_void_p(D.data_ptr()), ETA, ALAM)
        _lib.check(rc)
        torch.cuda.synchronize()
        out[name] = D.cpu().numpy()
    return out


@pytest.mark.parametrize("impl", ["pipelined", "staged"])
def test_noalloc_diffterm_equals_diffterm(device, monkeypatch, impl):
    if impl == "staged":  # the register-staged kernels need no scratch
        monkeypatch.setenv("CNS_PIPE", "0")
    print(f"no-alloc diffterm, {impl}: 96^3 then 40x24x33", flush=True)
    for n, seed in (((96, 96, 96), 11), ((40, 24, 33), 12)):
        lay = BoxLayout(n)
        d = _diffterm_both(lay, seed, device)
        a, b = d["cns_diffterm_noalloc"], d["cns_diffterm"]
        assert not np.isnan(lay.iview(b, NC)).any()
        assert np.array_equal(lay.iview(a, NC), lay.iview(b, NC)), \
            f"{impl} {n}: no-alloc diffterm differs from cns_diffterm"


def test_noalloc_rejects_boxes_above_scratch(device):
    import torch
    lib = _lib_on_stream()
    n = 193  # PAD16(201) * 201 * 201 values per ghosted component > 8 Mi
    lay = BoxLayout.cube(n)
    assert lay.cg > 8 << 20
    dummy = torch.zeros(16, dtype=torch.float64, device=device)
    p = C.c_void_p(dummy.data_ptr())
    rc = lib.cns_diffterm_noalloc(_I3(*lay.lo), _I3(*lay.hi), lay.ng, _D3(*box_dx(n)), p, p, ETA, ALAM)
    assert rc == _lib.CNS_EINVAL
    assert b"scratch" in lib.cns_last_error_string()
    ok = BoxLayout.cube(192)
    assert ok.cg <= 8 << 20
