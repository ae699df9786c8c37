"""The batched and served workloads at their configured sizes, on the GPU.

* Batched: 64 independent boxes, edges uniform over {16,24,32,48,64,96} (seed 4),
  Mach U[0.05,0.3], 50 RK3 steps.  The FP64 engine gives, per box, the bits of
  the plain path (cns_advance one box at a time, which the oracle and the
  reference pin) on all 64 boxes, and the oracle's bits directly on four of them
  (the smallest and largest edges among them); the FP32-flux engine stays within
  1e-6 relative L2 per field of FP64.
* Served: the 2000-request stream (edge 24-128, 50-500 steps, Mach U[0.05,0.3],
  Poisson arrivals at a seeded rate) under the 200 GB budget with FP32 flux: no
  failures, and the smallest, largest and longest requests give the bits they
  give when run alone; on those boxes FP32 flux stays within 1e-6 relative L2 of
  FP64 after 100 steps.
* Edges of 476 and above: a box's five components pass 4 GiB, so the engine's
  component offsets must be formed in 64 bits; one 512^3 step (the largest edge
  the engine admits) equals the plain path's."""
import numpy as np
import pytest

from cns_nospec_amd import ops
from cns_nospec_amd.batch import BatchEngine, box_bytes
from cns_nospec_amd.layout import NC, NQ, BoxLayout
from cns_nospec_amd.serve import BatchServer, batch_mix, make_stream, request_state
from cns_nospec_amd.state import ALAM, CFL, ETA, box_dx, pack_ghosted

pytestmark = pytest.mark.gpu


def _plain_box(req, steps):
    """The request's box through cns_advance (ghosted layout), `steps` steps; the
    interior state as a (5, e, e, e) device tensor."""
    lay = BoxLayout.cube(req.edge)
    U, Unew, Q, D, F = ops.allocate_variables(lay, "cuda")
    g, ng, e = lay.dimg, lay.ng, req.edge
    view = U.view(NC, g[2], g[1], lay.xg)[:, ng:ng + e, ng:ng + e, ng:ng + e]
    view.copy_(request_state(req))
    st = ops.Stepper(lay)
    for _ in range(steps):
        st.new_advance(U, Unew, Q, D, F)
    out = view.clone()
    st.close()
    return out


def _engine_run(reqs, fp32, steps):
    eng = BatchEngine(sum(box_bytes(r.edge, fp32) for r in reqs) + (256 << 20), fp32=fp32,
                      max_boxes=len(reqs))
    ids = [eng.add(request_state(r).reshape(-1), r.edge) for r in reqs]
    eng.step(steps)
    q = eng.query(ids)
    assert all(q[i]["steps"] == steps and q[i]["err"] == 0 for i in ids)
    out = [eng.read(i) for i in ids]
    eng.close()
    return out


def _rel_l2(a, b):
    return [float(np.linalg.norm(a[l] - b[l]) / np.linalg.norm(b[l])) for l in range(NC)]


def test_batched_mix_64_boxes_50_steps(orc, device):
    import torch
    reqs = batch_mix()
    assert len(reqs) == 64 and {r.edge for r in reqs} <= {16, 24, 32, 48, 64, 96}
    steps = reqs[0].steps
    print(f"batched mix: 64 boxes, edges {sorted({r.edge for r in reqs})}, {steps} steps", flush=True)
    s64 = _engine_run(reqs, False, steps)
    for r, got in zip(reqs, s64):
        ref = _plain_box(r, steps)
        assert torch.equal(got, ref), f"box {r.rid} (edge {r.edge}): FP64 engine differs from cns_advance"
    # the oracle itself on four boxes, the largest edge among them
    by_edge = sorted(reqs, key=lambda r: (r.edge, r.rid))
    picks = {by_edge[0].rid, by_edge[-1].rid, by_edge[len(by_edge) // 2].rid, reqs[1].rid}
    assert max(reqs[i].edge for i in picks) == 96
    for i in sorted(picks):
        r = reqs[i]
        lay = BoxLayout.cube(r.edge)
        U = pack_ghosted(lay, request_state(r).cpu().numpy())
        Unew, Q = lay.new_ghosted(NC), lay.new_ghosted(NQ)
        D, F = lay.new_interior(NC), lay.new_interior(NC)
        for _ in range(steps):
            orc.advance(lay, box_dx(r.edge), CFL, ETA, ALAM, U, Unew, Q, D, F)
        ref = np.ascontiguousarray(lay.interior_of(lay.gview(U, NC)))
        assert np.array_equal(s64[i].cpu().numpy(), ref), f"box {i} (edge {r.edge}) vs the oracle"
    s32 = _engine_run(reqs, True, steps)
    worst = np.zeros(NC)
    for a, b in zip(s32, s64):
        worst = np.maximum(worst, _rel_l2(a.cpu().numpy(), b.cpu().numpy()))
    print("FP32 flux, worst relative L2 per field after 50 steps:", worst.tolist(), flush=True)
    assert (worst <= 1e-6).all()


@pytest.mark.timeout(900)
def test_served_stream_2000_requests(device):
    reqs = make_stream(2000, 40.0, 2024)
    smallest = min(reqs, key=lambda r: (r.edge, r.rid))
    largest = max(reqs, key=lambda r: (r.edge, -r.rid))
    longest = max(reqs, key=lambda r: (r.steps, -r.rid))
    keep = (smallest.rid, largest.rid, longest.rid)
    print(f"served stream: 2000 requests, Poisson 40/s, kept {keep}", flush=True)
    srv = BatchServer(200 << 30, fp32=True, keep_states=keep, progress_every=250)
    res, summ = srv.run(reqs)
    srv.close()
    print({k: summ[k] for k in ("requests", "failed", "cell_steps_per_s", "p50_latency_s",
                                "p99_latency_s")}, flush=True)
    assert summ["requests"] == 2000 and summ["failed"] == 0
    assert all(r.done >= r.admitted >= r.arrival for r in res)
    for rid in keep:
        r = reqs[rid]
        eng = BatchEngine(box_bytes(r.edge, True) + (64 << 20), fp32=True, max_boxes=1)
        b = eng.add(request_state(r).reshape(-1), r.edge)
        eng.step(r.steps)
        alone = eng.read(b).cpu().numpy()
        eng.close()
        assert np.array_equal(res[rid].state.cpu().numpy(), alone), f"request {rid} alone vs served"
    picks = [reqs[i] for i in keep]
    s32 = _engine_run(picks, True, 100)
    s64 = _engine_run(picks, False, 100)
    for r, a, b in zip(picks, s32, s64):
        err = _rel_l2(a.cpu().numpy(), b.cpu().numpy())
        print(f"request {r.rid} edge {r.edge} mach {r.mach:.3f}: FP32 rel L2 after 100 steps {err}",
              flush=True)
        assert max(err) <= 1e-6, f"request {r.rid}"


def test_engine_edge_512_equals_plain(device):
    """The largest edge the engine admits: one FP64 step of a 512^3 box equals one
    cns_advance step of the same box, bit for bit (the box's components lie 1 GiB
    apart, so the fourth and fifth pass 4 GiB from its base)."""
    import torch
    from cns_nospec_amd.serve import Request
    r = Request(0, 0.0, 512, 1, 77, 0.2)
    print("engine edge 512: one step, FP64 engine vs cns_advance", flush=True)
    eng = BatchEngine(box_bytes(512, False) + (64 << 20), fp32=False, max_boxes=1)
    b = eng.add(request_state(r).reshape(-1), 512)
    eng.step(1)
    assert eng.query([b])[b]["err"] == 0
    got = eng.read(b)
    eng.close()
    ref = _plain_box(r, 1)
    assert torch.equal(got, ref)
    del got, ref
    torch.cuda.empty_cache()
