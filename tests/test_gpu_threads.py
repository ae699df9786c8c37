"""Two host threads, each with its own context and stream, stepping different boxes at
the same time (ctypes releases the GIL during the library calls).  The stream and the
last error are per thread and contexts share nothing, so each thread's results are
bit-identical to the same steps run alone on the null stream."""
import threading

import pytest

from cns_nospec_amd import ops
from cns_nospec_amd.layout import BoxLayout
from cns_nospec_amd.state import sinusoid_state

pytestmark = pytest.mark.gpu

CASES = [(24, 5), (40, 6)]  # (edge, seed)
NSTEPS = 3


def _alone(device, n, seed):
    import torch
    lay = BoxLayout.cube(n)
    arrs = ops.allocate_variables(lay, device)
    arrs[0].copy_(torch.from_numpy(sinusoid_state(lay, seed, mach=0.2, kmax=3)).to(device))
    st = ops.Stepper(lay)
    dts = [st.new_advance(*arrs) for _ in range(NSTEPS)]
    st.close()
    return dts, [a.clone() for a in arrs]


def test_two_threads_two_contexts(device):
    import torch
    ref = {n: _alone(device, n, seed) for n, seed in CASES}
    out, errors = {}, []

    def work(n, seed):
        try:
            lay = BoxLayout.cube(n)
            s = torch.cuda.Stream()
            with torch.cuda.stream(s):
                arrs = ops.allocate_variables(lay, device)
                arrs[0].copy_(torch.from_numpy(sinusoid_state(lay, seed, mach=0.2, kmax=3)).to(device))
                st = ops.Stepper(lay)
                dts = [st.new_advance(*arrs) for _ in range(NSTEPS)]
                s.synchronize()
                out[n] = (dts, [a.clone() for a in arrs])
                st.close()
            s.synchronize()
        except Exception as e:  # reported by the main thread
            errors.append(repr(e))

    print("two threads", CASES, flush=True)
    th = [threading.Thread(target=work, args=c) for c in CASES]
    for t in th:
        t.start()
    for t in th:
        t.join(timeout=120)
    assert not errors, errors
    for n, _ in CASES:
        dts, arrs = out[n]
        rdts, rarrs = ref[n]
        assert dts == rdts
        for a, b in zip(arrs, rarrs):
            assert torch.equal(a, b)
