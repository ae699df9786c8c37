"""Serving: a request loop that accepts CNS solves as they arrive.

A request is one periodic box (edge, its own seeded sinusoidal state and Mach
number) to be advanced a number of RK3 steps.  :class:`BatchServer` does
continuous batching on the batched engine (:mod:`.batch`): requests that arrive
while others are on the GPU are admitted as soon as the engine's HBM budget has
room, every resident box advances in the same graph-captured launches, and a box
leaves as soon as its own step count is reached.  :class:`PlainServer` serves
the same stream through the plain path (one box at a time via ``cns_advance``,
dt to the host every step, as the reference's main loop), first come first served.

Both report per-request latency (arrival to result), p50/p99, aggregate
cell-steps/s, and failures (a request whose state becomes non-physical fails
alone).
"""
from __future__ import annotations

import math
import sys
import time
from collections import deque
from dataclasses import dataclass, field

import numpy as np

from . import _lib
from .layout import NC, BoxLayout


@dataclass
class Request:
    rid: int
    arrival: float  # seconds after the stream starts
    edge: int
    steps: int
    seed: int
    mach: float


@dataclass
class Result:
    rid: int
    edge: int
    steps: int
    arrival: float
    admitted: float = math.nan
    done: float = math.nan
    failed: bool = False
    state: object = None  # final U (device tensor) when kept

    @property
    def latency(self):
        return self.done - self.arrival


def make_stream(n: int, rate: float, seed: int, edge_range=(24, 128), steps_range=(50, 500),
                mach_range=(0.05, 0.3)):
    """n requests with Poisson arrivals at `rate` per second, edge ~ U{edge_range},
    steps ~ U{steps_range}, Mach ~ U(mach_range), seeded."""
    rng = np.random.default_rng(seed)
    t = np.cumsum(rng.exponential(1.0 / rate, n)) if rate > 0 else np.zeros(n)
    edges = rng.integers(edge_range[0], edge_range[1] + 1, n)
    steps = rng.integers(steps_range[0], steps_range[1] + 1, n)
    machs = rng.uniform(mach_range[0], mach_range[1], n)
    seeds = rng.integers(0, 2 ** 31 - 1, n)
    return [Request(i, float(t[i]), int(edges[i]), int(steps[i]), int(seeds[i]), float(machs[i]))
            for i in range(n)]


BATCH_MIX_EDGES = (16, 24, 32, 48, 64, 96)


def batch_mix(nbox: int = 64, steps: int = 50, seed: int = 4, edges=BATCH_MIX_EDGES,
              mach_range=(0.05, 0.3)):
    """The batched workload: nbox independent boxes, edges uniform over `edges`,
    Mach ~ U(mach_range), each with its own seeded sinusoidal state, all advanced
    `steps` RK3 steps (64 boxes of {16,...,96} x 50 steps by default)."""
    rng = np.random.default_rng(seed)
    e = rng.choice(edges, nbox)
    machs = rng.uniform(mach_range[0], mach_range[1], nbox)
    seeds = rng.integers(0, 2 ** 31 - 1, nbox)
    return [Request(i, 0.0, int(e[i]), steps, int(seeds[i]), float(machs[i])) for i in range(nbox)]


def request_state(req: Request, device="cuda"):
    """The request's initial ghost-free U (5, e, e, e) on the device.

    The five primitive fields are evaluated together (one set of launches for all
    of them), mode by mode in the order the modes are drawn.  Nothing in it waits for
    the device: the per-field parameters reach it in one asynchronous copy from pinned
    host memory (a copy from pageable memory would wait for all the work queued before
    it, e.g. a chunk of steps that a server is running meanwhile)."""
    from .state import GAMMA, _draw_modes
    import torch
    e = req.edge
    rng = np.random.default_rng(req.seed)
    crd = (torch.arange(e, dtype=torch.float64, device=device) + 0.5) / e
    c0 = math.sqrt(GAMMA)
    modes = [_draw_modes(rng, 2, 4) for _ in range(5)]  # rho, u, v, w, p
    # rows: f0, amplitude, then per mode k (3 rows), phase (3 rows), mode amplitude
    par = np.empty((2 + 4 * 7, 5))
    par[0] = (1.0, 0.0, 0.0, 0.0, 1.0)
    par[1] = (0.05, req.mach * c0, req.mach * c0, req.mach * c0, 0.05)
    for m in range(4):
        for f in range(5):
            k, ph, a = modes[f][m]
            par[2 + 7 * m:5 + 7 * m, f] = [int(k[d]) for d in range(3)]
            par[5 + 7 * m:8 + 7 * m, f] = [float(ph[d]) for d in range(3)]
            par[8 + 7 * m, f] = float(a)
    if torch.device(device).type == "cuda":
        host = torch.empty(par.shape, dtype=torch.float64, pin_memory=True)
        host.copy_(torch.from_numpy(par))
        dev = host.to(device, non_blocking=True)
    else:
        dev = torch.from_numpy(par).to(device)
    f0, amp = dev[0], dev[1]
    s = torch.zeros((5, e, e, e), dtype=torch.float64, device=device)
    for m in range(4):
        k = dev[2 + 7 * m:5 + 7 * m].t()
        ph = dev[5 + 7 * m:8 + 7 * m].t()
        a = dev[8 + 7 * m]
        fx = torch.sin(2 * math.pi * k[:, 0:1] * crd + ph[:, 0:1])
        fy = torch.sin(2 * math.pi * k[:, 1:2] * crd + ph[:, 1:2])
        fz = torch.sin(2 * math.pi * k[:, 2:3] * crd + ph[:, 2:3])
        s.add_(a[:, None, None, None] * fz[:, :, None, None] * fy[:, None, :, None] * fx[:, None, None, :])
    smax = s.abs().amax(dim=(1, 2, 3))
    s.mul_((amp / smax)[:, None, None, None]).add_(f0[:, None, None, None])
    rho, u, v, w, p = s[0], s[1], s[2], s[3], s[4]
    return torch.stack([rho, rho * u, rho * v, rho * w,
                        p / (GAMMA - 1.0) + 0.5 * rho * (u * u + v * v + w * w)]).contiguous()


def summarize(results, wall: float):
    ok = [r for r in results if not r.failed]
    lat = np.array([r.latency for r in ok]) if ok else np.array([math.nan])
    work = sum(r.edge ** 3 * r.steps for r in ok)
    return {"requests": len(results), "failed": len(results) - len(ok),
            "cell_steps": int(work), "wall_s": wall,
            "cell_steps_per_s": work / wall if wall > 0 else math.nan,
            "p50_latency_s": float(np.percentile(lat, 50)),
            "p99_latency_s": float(np.percentile(lat, 99)),
            "mean_latency_s": float(lat.mean())}


def _progress(every: int, done: int, total: int, t: float):
    """One line on stderr per `every` finished requests (long runs stay visibly alive
    without touching stdout)."""
    if every and done % every == 0:
        print(f"served {done}/{total} requests at t={t:.1f} s", file=sys.stderr, flush=True)


class _Clock:
    """Stream time: real time since start, fast-forwarded over idle gaps when the
    server has nothing to do (so a sparse stream does not waste the GPU's time
    budget; latency is still measured in stream time)."""

    def __init__(self):
        self.t0 = time.perf_counter()
        self.skip = 0.0

    def now(self):
        return time.perf_counter() - self.t0 + self.skip

    def advance_to(self, t):
        cur = self.now()
        if t > cur:
            self.skip += t - cur


class BatchServer:
    """Continuous batching on one BatchEngine."""

    def __init__(self, hbm_budget: int, fp32: bool = True, max_boxes: int = 4096,
                 max_chunk: int = 8, keep_states=(), progress_every: int = 0,
                 max_resident_cells: int = 128 << 20, prepare_ahead: int = 16):
        from .batch import BatchEngine
        self.engine = BatchEngine(hbm_budget, fp32=fp32, max_boxes=max_boxes)
        self.max_chunk = max_chunk
        # Admission also stops once the resident boxes hold this many cells (0: only
        # the HBM budget limits it).  The engine's rate falls as its resident footprint
        # grows (every step touches every resident box): on the request mix 3.5 G
        # cell-steps/s at 113 M cells, 3.2 G at 490 M, 3.0 G at 960 M; a 2000-request
        # burst runs 10 % faster with 128 M cells resident than with all that fit.
        self.max_resident_cells = max_resident_cells
        self.keep = set(keep_states)
        self.progress_every = progress_every
        # Initial states of up to this many waiting requests are built while a chunk runs
        # (their host work then overlaps the GPU's), so admission itself is a copy.
        self.prepare_ahead = prepare_ahead
        # per chunk: (stream time at launch, steps, resident boxes, resident cells,
        # seconds from launch to the end of its synchronising query)
        self.chunks = []

    def run(self, requests):
        import torch
        eng = self.engine
        clock = _Clock()
        # FIFO by arrival; requests that arrive together are taken longest first (every
        # resident box advances one step per engine step, so a box's steps set when it
        # finishes, and the last boxes admitted then run the shortest jobs: no tail of a
        # few long boxes running alone on an idle GPU)
        pending = deque(sorted(requests, key=lambda r: (r.arrival, -r.steps, r.rid)))
        waiting = deque()
        prepared = {}  # request id -> initial state, for requests at the head of `waiting`
        active = {}  # box id -> (request, steps left, result)
        results = {}
        resident = 0  # cells of the active boxes

        def arrivals():
            now = clock.now()
            while pending and pending[0].arrival <= now:
                waiting.append(pending.popleft())

        while pending or waiting or active:
            arrivals()
            # admission: FIFO while the budget has room; each box stops by itself on the
            # device once it has taken its requested steps.  Initial states not built
            # during the last chunk are built here with the GPU idle, so once boxes are
            # resident at most `prepare_ahead` of them are built before the next chunk
            # starts (a burst then ramps up in waves instead of waiting for all of them).
            built = 0
            while waiting:
                req = waiting[0]
                if (self.max_resident_cells and active
                        and resident + req.edge ** 3 > self.max_resident_cells):
                    break
                u0 = prepared.get(req.rid)
                if u0 is None and active and built >= self.prepare_ahead:
                    break
                if u0 is None:
                    built += 1
                try:
                    bid = eng.add(u0 if u0 is not None else request_state(req).reshape(-1), req.edge)
                except _lib.CnsBudgetError:
                    break
                prepared.pop(req.rid, None)
                eng.set_target(bid, req.steps)
                waiting.popleft()
                res = Result(req.rid, req.edge, req.steps, req.arrival, admitted=clock.now())
                results[req.rid] = res
                active[bid] = [req, req.steps, res]
                resident += req.edge ** 3
            if not active:
                if pending:
                    clock.advance_to(pending[0].arrival)
                    continue
                if waiting:
                    raise RuntimeError("a single request does not fit in the engine's budget")
                break
            # a chunk of steps: boxes that reach their target mid-chunk stop there, so the
            # chunk need not end at the first box's end (that would cost a sync and a
            # work-list rebuild per finishing box)
            k = min(self.max_chunk, max(v[1] for v in active.values()))
            t_launch = clock.now()
            eng.step(k)
            # while the chunk runs: the next admissions' initial states (their kernels queue
            # behind the chunk on the same stream; the host work overlaps it)
            arrivals()
            for req in list(waiting)[:self.prepare_ahead]:
                if req.rid not in prepared:
                    prepared[req.rid] = request_state(req).reshape(-1)
            q = eng.query(list(active))  # synchronises
            self.chunks.append((t_launch, k, len(active), resident, clock.now() - t_launch))
            for bid, v in active.items():
                v[1] = v[0].steps - q[bid]["steps"]
            finished = [bid for bid, v in active.items() if v[1] <= 0 or q[bid]["err"]]
            if finished:
                t = clock.now()
                for bid in finished:
                    req, _, res = active.pop(bid)
                    res.done = t
                    res.failed = bool(q[bid]["err"])
                    if req.rid in self.keep:
                        res.state = eng.read(bid).clone()
                    eng.remove(bid)
                    resident -= req.edge ** 3
                    _progress(self.progress_every, sum(1 for r in results.values() if r.done == r.done),
                              len(requests), t)
        wall = max((r.done for r in results.values()), default=0.0) - min(
            (r.arrival for r in results.values()), default=0.0)
        out = [results[r.rid] for r in sorted(requests, key=lambda r: r.rid)]
        return out, summarize(out, wall)

    def close(self):
        self.engine.close()


class PlainServer:
    """The plain path on the same stream: one box at a time through cns_advance
    (ghosted layout, dt to the host every step), first come first served."""

    def __init__(self, keep_states=(), progress_every: int = 0):
        self.keep = set(keep_states)
        self.progress_every = progress_every

    def run(self, requests):
        import torch
        from .ops import Stepper, allocate_variables
        clock = _Clock()
        results = []
        for req in sorted(requests, key=lambda r: r.arrival):
            clock.advance_to(req.arrival)
            res = Result(req.rid, req.edge, req.steps, req.arrival, admitted=clock.now())
            lay = BoxLayout.cube(req.edge)
            U, Unew, Q, D, F = allocate_variables(lay, "cuda")
            g, ng, e = lay.dimg, lay.ng, req.edge
            U.view(NC, g[2], g[1], lay.xg)[:, ng:ng + e, ng:ng + e, ng:ng + e].copy_(request_state(req))
            st = Stepper(lay)
            try:
                for _ in range(req.steps):
                    st.new_advance(U, Unew, Q, D, F)
            except _lib.CnsNumericError:
                res.failed = True
            torch.cuda.synchronize()
            res.done = clock.now()
            if req.rid in self.keep:
                res.state = U.view(NC, g[2], g[1], lay.xg)[:, ng:ng + e, ng:ng + e, ng:ng + e].clone()
            st.close()
            del U, Unew, Q, D, F
            results.append(res)
            _progress(self.progress_every, len(results), len(requests), res.done)
        wall = max(r.done for r in results) - min(r.arrival for r in results)
        out = sorted(results, key=lambda r: r.rid)
        return out, summarize(out, wall)
