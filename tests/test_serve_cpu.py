"""Host logic of the serving layer (no GPU): the seeded request stream, the
latency/throughput summary, and the batch server's admission loop over a host
stand-in engine."""
import math

from cns_nospec_amd.serve import Result, make_stream, summarize


def test_stream_is_seeded_and_in_range():
    a = make_stream(500, 20.0, 7)
    b = make_stream(500, 20.0, 7)
    assert [(r.arrival, r.edge, r.steps, r.seed) for r in a] == [(r.arrival, r.edge, r.steps, r.seed) for r in b]
    assert all(24 <= r.edge <= 128 and 50 <= r.steps <= 500 and 0.05 <= r.mach <= 0.3 for r in a)
    arr = [r.arrival for r in a]
    assert arr == sorted(arr)
    # Poisson: mean inter-arrival 1/rate within a few percent for 500 samples
    assert abs(arr[-1] / len(arr) - 1 / 20.0) < 0.2 / 20.0
    assert make_stream(10, 0, 1)[-1].arrival == 0.0  # rate 0: everything at t = 0


def test_summary_counts_only_successful_work():
    rs = [Result(0, 10, 5, 0.0, 0.0, 2.0), Result(1, 20, 2, 1.0, 1.0, 4.0),
          Result(2, 30, 9, 1.0, 1.0, 3.0, failed=True)]
    s = summarize(rs, 4.0)
    assert s["failed"] == 1 and s["cell_steps"] == 10 ** 3 * 5 + 20 ** 3 * 2
    assert math.isclose(s["cell_steps_per_s"], s["cell_steps"] / 4.0)
    assert s["p50_latency_s"] == 2.5


class _FakeEngine:
    """Host stand-in for BatchEngine: boxes advance by whole steps up to their target;
    the pool holds `budget` bytes at edge^3 bytes per box."""
    log = []

    def __init__(self, hbm_budget, fp32=True, max_boxes=4096, tuning=None):
        self.budget, self.used, self.boxes, self.next = hbm_budget, 0, {}, 0
        self.resident_peak = 0
        _FakeEngine.log = []

    def add(self, u0, edge):
        from cns_nospec_amd import _lib
        if self.used + edge ** 3 > self.budget:
            raise _lib.CnsBudgetError("no room (fake)", _lib.CNS_EBUDGET)
        self.used += edge ** 3
        bid = self.next
        self.next += 1
        self.boxes[bid] = {"edge": edge, "steps": 0, "target": 0}
        _FakeEngine.log.append(("add", bid, edge))
        self.resident_peak = max(self.resident_peak, sum(b["edge"] ** 3 for b in self.boxes.values()))
        return bid

    def set_target(self, bid, steps):
        self.boxes[bid]["target"] = steps

    def step(self, k):
        for b in self.boxes.values():
            b["steps"] = min(b["target"], b["steps"] + k) if b["target"] else b["steps"] + k

    def query(self, ids):
        return {i: {"steps": self.boxes[i]["steps"], "dt": 1.0, "courno": 1.0, "err": 0} for i in ids}

    def read(self, bid):
        raise AssertionError("no state kept in this test")

    def remove(self, bid):
        self.used -= self.boxes.pop(bid)["edge"] ** 3

    def close(self):
        pass


class _Flat:
    def reshape(self, *a):
        return self


def test_batch_server_admission_targets_and_cap(monkeypatch):
    """FIFO admission under the HBM budget and the resident-cell cap (requests that arrive
    together taken longest first), each box run to exactly its own step target, every
    request served once."""
    from cns_nospec_amd import batch, serve
    monkeypatch.setattr(batch, "BatchEngine", _FakeEngine)
    monkeypatch.setattr(serve, "request_state", lambda req: _Flat())
    reqs = make_stream(40, 0.0, 11)
    cap = 3 * 100 ** 3
    srv = serve.BatchServer(10 * 128 ** 3, fp32=True, max_chunk=7, max_resident_cells=cap)
    eng = srv.engine
    out, summ = srv.run(reqs)
    assert summ["requests"] == 40 and summ["failed"] == 0
    assert [r.rid for r in out] == list(range(40))
    order = sorted(reqs, key=lambda r: (r.arrival, -r.steps, r.rid))  # all arrive at t = 0
    assert [e[1] for e in _FakeEngine.log] == list(range(40))  # one admission each
    assert [e[2] for e in _FakeEngine.log] == [r.edge for r in order]
    assert cap // 2 < eng.resident_peak <= cap  # the cap binds, and holds
    assert all(r.done >= r.admitted >= r.arrival for r in out)
    assert summ["cell_steps"] == sum(r.edge ** 3 * r.steps for r in reqs)


def test_batch_server_fifo_across_arrival_times(monkeypatch):
    """Distinct arrival times keep first-come-first-served order whatever the step counts;
    only requests that arrive together are reordered (longest first)."""
    from cns_nospec_amd import batch, serve
    monkeypatch.setattr(batch, "BatchEngine", _FakeEngine)
    monkeypatch.setattr(serve, "request_state", lambda req: _Flat())
    reqs = [serve.Request(i, 0.01 * (i // 2), 20, 10 + 7 * (i % 2) + i, i, 0.1) for i in range(12)]
    srv = serve.BatchServer(10 ** 9, fp32=True, max_chunk=3, max_resident_cells=3 * 20 ** 3)
    out, summ = srv.run(reqs)
    assert summ["failed"] == 0
    admitted = [e[1] for e in _FakeEngine.log]
    edges_order = [reqs[i] for i in sorted(range(12), key=lambda i: (reqs[i].arrival, -reqs[i].steps, i))]
    assert len(admitted) == 12
    assert [r.arrival for r in sorted(out, key=lambda r: r.admitted)] == sorted(r.arrival for r in reqs)
    assert [r.rid for r in edges_order][:2] == [1, 0]  # the pair arriving first: longer one first


def test_batch_server_burst_ramps_up_in_waves(monkeypatch):
    """With every request waiting at t = 0, the server does not build all initial states
    before the first chunk: at most `prepare_ahead` are built while no chunk runs, the
    rest while chunks run, and every request is still served once, in order."""
    from cns_nospec_amd import batch, serve
    monkeypatch.setattr(batch, "BatchEngine", _FakeEngine)

    def build(req):
        _FakeEngine.log.append(("build", req.rid, req.edge))
        return _Flat()

    monkeypatch.setattr(serve, "request_state", build)
    step = _FakeEngine.step

    def logged_step(self, k):
        _FakeEngine.log.append(("step", k, 0))
        step(self, k)

    monkeypatch.setattr(_FakeEngine, "step", logged_step)
    reqs = make_stream(30, 0.0, 5)
    srv = serve.BatchServer(10 ** 9, fp32=True, max_chunk=5, max_resident_cells=0, prepare_ahead=4)
    out, summ = srv.run(reqs)
    log = _FakeEngine.log
    first_step = next(i for i, e in enumerate(log) if e[0] == "step")
    assert sum(1 for e in log[:first_step] if e[0] == "build") == 4
    assert sum(1 for e in log if e[0] == "build") == 30
    assert summ["failed"] == 0 and summ["cell_steps"] == sum(r.edge ** 3 * r.steps for r in reqs)
    order = [r.rid for r in sorted(reqs, key=lambda r: (r.arrival, -r.steps, r.rid))]
    assert [e[1] for e in log if e[0] == "add"] == list(range(30))
    assert [e[2] for e in log if e[0] == "add"] == [reqs[i].edge for i in order]


def test_request_state_is_the_seeded_sinusoid():
    """A request's initial state is the seeded sinusoidal state of the oracle's
    generator (modes drawn field by field from the request's seed, wavenumbers 1-2,
    amplitude Mach * c0 on the velocities): the server's device construction, run here
    on the host, agrees with the NumPy one to rounding."""
    import numpy as np
    from cns_nospec_amd.serve import request_state
    from cns_nospec_amd.state import conserved_from_primitives, sinusoid_primitives
    for req in make_stream(4, 0.0, 11):
        e = min(req.edge, 32)
        req.edge = e
        got = request_state(req, device="cpu").numpy()
        prim = sinusoid_primitives(e, req.seed, mach=req.mach, kmax=2, nmodes=4)
        ref = conserved_from_primitives(prim["rho"], prim["u"], prim["v"], prim["w"], prim["p"])
        assert got.shape == (5, e, e, e)
        for l in range(5):
            scale = np.abs(ref[l]).max()
            assert np.abs(got[l] - ref[l]).max() <= 1e-13 * scale, f"component {l}"
