"""Continuous-batching server on the GPU: requests admitted while others run
give the same result as when served alone, and the FP64 server agrees with the
plain (cns_advance) path bit for bit."""
import numpy as np
import pytest

from cns_nospec_amd.batch import BatchEngine
from cns_nospec_amd.serve import BatchServer, PlainServer, make_stream, request_state

pytestmark = pytest.mark.gpu


def _alone(req, fp32):
    eng = BatchEngine(1 << 30, fp32=fp32, max_boxes=2)
    b = eng.add(request_state(req).reshape(-1), req.edge)
    eng.step(req.steps)
    out = eng.read(b).cpu().numpy()
    eng.close()
    return out


@pytest.mark.parametrize("fp32", [True, False])
def test_served_equals_alone(device, fp32):
    reqs = make_stream(12, 200.0, 99, edge_range=(16, 40), steps_range=(5, 20))
    keep = [0, 5, 11]
    srv = BatchServer(3 << 30, fp32=fp32, max_boxes=64, max_chunk=4, keep_states=keep)
    res, summ = srv.run(reqs)
    srv.close()
    assert summ["failed"] == 0 and summ["requests"] == 12
    assert all(r.done >= r.admitted >= r.arrival for r in res)
    for rid in keep:
        got = res[rid].state.cpu().numpy()
        assert np.array_equal(got, _alone(reqs[rid], fp32)), f"request {rid}"
    if not fp32:
        plain, psum = PlainServer(keep_states=keep).run([reqs[i] for i in keep])
        for r in plain:
            assert np.array_equal(r.state.cpu().numpy(), res[r.rid].state.cpu().numpy())


def test_budget_queues_instead_of_failing(device):
    from cns_nospec_amd.batch import box_bytes
    reqs = make_stream(6, 1000.0, 3, edge_range=(32, 32), steps_range=(3, 6))
    # room for two boxes at a time: the rest must wait, not fail
    srv = BatchServer(2 * box_bytes(32, True) + 4096, fp32=True, max_boxes=8, max_chunk=2)
    res, summ = srv.run(reqs)
    srv.close()
    assert summ["failed"] == 0 and summ["requests"] == 6
    admitted = sorted(r.admitted for r in res)
    assert admitted[2] >= min(r.done for r in res)  # the third waited for a slot
