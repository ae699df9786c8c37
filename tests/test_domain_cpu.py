"""Multi-box domains (include/cns_domain.h) on the CPU: the argument checks that
run before any device work, and the exported symbols."""
import ctypes as C

import pytest

from cns_nospec_amd import _lib


def _create(ncell, m):
    lib = _lib.load()
    h = C.c_void_p()
    rc = lib.cns_domain_create(C.byref(h), (C.c_int * 3)(*ncell), m, (C.c_double * 3)(0.01, 0.01, 0.01),
                               0.5, 1.8e-4, 150.0)
    return rc, h, lib.cns_last_error_string().decode()


@pytest.mark.parametrize("ncell,m,msg", [((32, 32, 30), 16, "multiple of max_grid_size"),
                                         ((12, 12, 12), 3, "below 4"),
                                         ((0, 8, 8), 8, "multiple"),
                                         ((4096, 4096, 64), 4, "boxes")])
def test_domain_create_rejects(ncell, m, msg):
    rc, h, err = _create(ncell, m)
    assert rc == _lib.CNS_EINVAL and not h.value
    assert msg in err


def test_domain_null_arguments():
    lib = _lib.load()
    assert lib.cns_domain_destroy(None) == _lib.CNS_OK
    assert lib.cns_domain_array_len(None, 5, 1) == 0
    assert lib.cns_domain_fill_boundary(None, None) == _lib.CNS_EINVAL
    assert lib.cns_domain_advance(None, None, None, None, None, None, None) == _lib.CNS_EINVAL
