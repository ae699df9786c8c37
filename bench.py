#!/usr/bin/env python3
"""Benchmark of the CNS_NoSpec RK3 time step on one MI355X.

    python bench.py [--gpus 1] [--steps 20] [--warmup 3] [--edge 512]

Headline (``value``): the plain FP64 path through the C ABI (``cns_advance``:
ghost fill, ctoprim + Courant reduction, diffterm, hypterm, RK update, three
stages, dt synchronised to the host every step as in the reference) on one
periodic 512^3 box (520^3 ghosted, ng = 4), seeded sinusoidal state with
random phases on wavenumbers 1-4.  A step is one RK3 step of the whole box;
five windows of exactly ``--steps`` steps are timed and the median window's
rate is reported: value = interior cells x steps / wall time of that window
(inputs resident in HBM).  The other workloads (the cell64 CGS deck, the
64-box batch, the 2000-request served stream) are reported under
``secondary``.

``roofline``: the dominant kernel's algorithmic HBM bytes per launch (each
array element it needs read once over its needed extent, each output written
once) divided by its average duration, measured with HIP events around every
launch inside the timed windows on the launch stream.  ``traffic`` is the
measured HBM bytes per launch from rocprofv3 PMC counters
(profiles/hbm_traffic.txt).  ``step_bytes`` sets the step's measured traffic
beside the reference ABI's bytes and the fused composition's own minimum.
``cpu_baseline``: the reference cpu/ path itself (oracle/_ref, compiled from
the reference sources) timed on one host core at 32^3, 64^3 and 128^3, or the
oracle restatement where that build is absent.
"""
from __future__ import annotations

import argparse
import contextlib
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "cell-steps/s (plain FP64 vs batched/FP32-flux) + p50/p99 solve latency"
HBM_PEAK_GBS = 8000.0
C2050_64_CELL_STEPS = 64 ** 3 * 100 / 2.7245  # cns_nospec.xls!Sheet1!A7:E7
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "hbm_traffic.txt")


def algorithmic_bytes(kind: str, n: int, stage: int = 2) -> float:
    """Per-launch algorithmic HBM bytes of one kernel on an n^3 box (f64): each input
    element read once over its needed extent, each output written once."""
    g3, n3 = (n + 8) ** 3, n ** 3
    rk = 20 * n3 if stage == 1 else 25 * n3
    table = {
        "fill": 10 * (g3 - n3),
        "ctoprim": 11 * g3,
        "diffterm": 4 * g3 + 5 * n3,
        "hypterm": 8 * g3 + 5 * n3,
        "rk": rk,
        # the fused hypterm + RK pass, averaged over a step's three stages: the five
        # conserved fields over their halos (5g; u, v, w, p are derived from them in
        # the tile), D without its zero rho row (4c), the stage base in stages 2-3
        # (10c / 3), the result (5c), F in stage 3 only (5c / 3), and the primitives
        # of the result stored for the next stage (u, v, w, T after stage 1, all six
        # after stage 2: 10c / 3)
        "hypterm_rk": 5 * g3 + 9 * n3 + 25 * n3 / 3,
        "fill_ctoprim": 10 * (g3 - n3) + 11 * g3,
    }
    return 8.0 * table[kind]


def step_algorithmic_bytes(n: int) -> float:
    """The reference's ABI traffic per step: every array of the unfused composition
    read and written once per use."""
    return 8.0 * (99 * (n + 8) ** 3 + 70 * n ** 3)


def fused_step_min_bytes(n: int) -> float:
    """What the fused composition itself must move per step (f64): fills 50(g-c)
    (state shell in three stages, plus u, v, w, T and then all six primitives in
    stages 2 and 3), ctoprim + wz 10g (stage 1), wz 2g (stages 2 and 3), diffterm
    5g + 4c/4c/5c, fused hypterm + RK 3 x (5g + 17.33c): 94g + 15c in all."""
    g3, n3 = (n + 8) ** 3, n ** 3
    return 8.0 * (94 * g3 + 15 * n3)


# Minimum of the batched / fused path per interior cell-step: read U, write U1; read
# U1 and U, write U2; read U2 and U, write U (40 doubles).
BATCHED_MIN_BYTES_PER_CELL_STEP = 320.0


WINDOWS = 5  # timed windows of the headline; the median is reported


def step_traffic(kern: dict, edge: int):
    """Measured HBM bytes of one step: per kernel kind, the calibrated bytes per launch
    (profiles/hbm_traffic.txt) times the launches per step; None when a kind has no
    measurement at this edge."""
    steps_kinds = {k: v for k, v in kern.items() if v[1] > 0}
    if not steps_kinds:
        return None
    launches_per_step = {"fill": 3, "ctoprim": 1, "diffterm": 3, "hypterm_rk": 3, "wz": 2}
    total = 0.0
    for kind, per_step in launches_per_step.items():
        b = read_traffic(kind, edge)
        if b is None:
            return None
        total += b * per_step
    return total


def read_traffic(kind: str, edge: int):
    """Measured HBM bytes per launch for (kind, edge) from profiles/hbm_traffic.txt."""
    if not os.path.exists(TRAFFIC_FILE):
        return None
    for line in open(TRAFFIC_FILE):
        if line.startswith("#") or "=" not in line:
            continue
        kv = dict(tok.split("=", 1) for tok in line.split() if "=" in tok)
        if kv.get("kernel") == kind and kv.get("edge") == str(edge) and "bytes" in kv:
            return float(kv["bytes"])
    return None


# Steps run from one seeded state before it is seeded again.  At 512^3 with the
# reference's eta (dx = 0.2/512, dt from cfl 0.5) the explicit viscous term of the
# grid-scale modes lies outside SSP-RK3's stability region (dt * eta * |D2| ~ 3.9 against
# 2.51), so round-off in those modes grows ~4x per step and the state turns non-finite
# after ~25 steps -- the reference program does the same.  Longer runs are therefore
# timed in chunks of CHUNK steps, each from a freshly seeded state (the seeding is not
# timed and every timed step is a full cns_advance).
CHUNK = 20


def run_plain(edge: int, steps: int, warmup: int, seed: int = 2024, kmax: int = 4,
              profile: bool = True, windows: int = 1):
    """`windows` timed windows of exactly `steps` steps each (after `warmup` untimed
    steps); the rate reported is the median window's.  Kernel times are HIP events
    around every launch of every window."""
    import torch
    from cns_nospec_amd import ops
    from cns_nospec_amd.layout import BoxLayout
    from cns_nospec_amd.state import device_sinusoid_state

    dev = torch.device("cuda:0")
    lay = BoxLayout.cube(edge)
    U, Unew, Q, D, F = ops.allocate_variables(lay, dev)

    def seed_state():
        device_sinusoid_state(lay, seed, dev, mach=0.085, kmax=kmax, out=U)

    st = ops.Stepper(lay)
    for i in range(warmup):
        if i % CHUNK == 0:
            seed_state()
        st.new_advance(U, Unew, Q, D, F)
    if profile:
        st.profile(True)
    window_s, dts = [], []
    for _ in range(windows):
        el, done = 0.0, 0
        while done < steps:
            n = min(CHUNK, steps - done)
            seed_state()
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            dts += [st.new_advance(U, Unew, Q, D, F) for _ in range(n)]
            torch.cuda.synchronize()
            el += time.perf_counter() - t0
            done += n
        window_s.append(el)
    kern = st.profile_read() if profile else {}
    st.profile(False)
    st.close()
    del U, Unew, Q, D, F
    torch.cuda.empty_cache()
    el = sorted(window_s)[len(window_s) // 2]
    return {"edge": edge, "elapsed_s": el, "cell_steps_per_s": edge ** 3 * steps / el,
            "ms_per_step": 1e3 * el / steps, "dt_last": dts[-1], "kernels": kern,
            "window_ms_per_step": [round(1e3 * w / steps, 3) for w in window_s]}


def run_cgs_deck(edge: int = 64, steps: int = 100, warmup: int = 5, reps: int = 3):
    """The reference's cell64 deck (fortran90/cell64: one 64^3 box, 100 steps, CGS
    initial condition of fortran90/init_data.f90:43-85) through cns_advance, the
    configuration of the reference's published C2050 figure (cns_nospec.xls, 64^3 x 100
    steps in 2.7245 s).  A 64^3 step is launch- and sync-bound (a fraction of a
    millisecond), so the deck is run `reps` times from its initial state and the median
    run is reported."""
    import torch
    from cns_nospec_amd import ops
    from cns_nospec_amd.layout import BoxLayout
    from cns_nospec_amd.state import cgs_state

    lay = BoxLayout.cube(edge)
    U0 = torch.from_numpy(cgs_state(lay)).to("cuda")
    U, Unew, Q, D, F = ops.allocate_variables(lay, "cuda")
    st = ops.Stepper(lay)
    U.copy_(U0)
    for _ in range(warmup):
        st.new_advance(U, Unew, Q, D, F)
    runs = []
    for _ in range(reps):
        U.copy_(U0)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            st.new_advance(U, Unew, Q, D, F)
        torch.cuda.synchronize()
        runs.append(time.perf_counter() - t0)
    st.close()
    el = sorted(runs)[len(runs) // 2]
    rate = edge ** 3 * steps / el
    return {"workload": f"cell{edge} deck: one periodic {edge}^3 box, CGS initial condition, "
                        f"{steps} RK3 steps through cns_advance, dt to the host every step",
            "value": rate, "unit": "cell-steps/s", "ms_per_step": 1e3 * el / steps,
            "run_ms_per_step": [round(1e3 * r / steps, 4) for r in runs],
            "timing": f"median of {reps} runs of the deck",
            "steps": steps, "vs_c2050_published": rate / C2050_64_CELL_STEPS}


@contextlib.contextmanager
def quiet_stdout():
    """The reference prints `dt, courno` each step; keep it off our one JSON line."""
    import ctypes
    libc = ctypes.CDLL(None)
    sys.stdout.flush()
    libc.fflush(None)
    saved = os.dup(1)
    null = os.open(os.devnull, os.O_WRONLY)
    os.dup2(null, 1)
    try:
        yield
    finally:
        libc.fflush(None)  # C stdio's buffered lines go to /dev/null, not to our stdout
        os.dup2(saved, 1)
        os.close(null)
        os.close(saved)


def cpu_model() -> str:
    """The host CPU as /proc/cpuinfo names it (some hosts report only the brand, so the
    family/model numbers and clock are added), and the logical CPUs this process may use."""
    info = {}
    try:
        for line in open("/proc/cpuinfo"):
            if ":" in line:
                k, v = (x.strip() for x in line.split(":", 1))
                info.setdefault(k, v)
    except OSError:
        return "unknown"
    name = info.get("model name", "unknown")
    extra = []
    if "cpu family" in info and "model" in info:
        extra.append(f"family {info['cpu family']} model {info['model']}")
    if "cpu MHz" in info:
        extra.append(f"{float(info['cpu MHz']) / 1e3:.2f} GHz")
    try:
        extra.append(f"{len(os.sched_getaffinity(0))} logical CPUs usable")
    except (AttributeError, OSError):
        pass
    return f"{name} ({', '.join(extra)})" if extra else name


def cpu_baseline_one(edge: int, budget_s: float):
    """Reference cpu/ new_advance on one core: as many steps of an edge^3 box as fit
    the time budget (at least 2).  Returns (cell-steps/s, steps, seconds, kind)."""
    import numpy as np
    import oracle
    from cns_nospec_amd.layout import NC, NQ, BoxLayout
    from cns_nospec_amd.state import ALAM, CFL, ETA, box_dx, sinusoid_state

    lay = BoxLayout.cube(edge)
    U = sinusoid_state(lay, 7, kmax=4)
    if oracle.reference_available():
        kind = "reference"
        ref = oracle.Reference()
        ref.setup(lay.n, box_dx(edge), CFL, ETA, ALAM)
        Ua = lay.ghosted_to_aos(U, NC)
        Unew, Q = np.zeros_like(Ua), np.zeros(lay.dimg + (NQ,))
        D, F = np.zeros(lay.n + (NC,)), np.zeros(lay.n + (NC,))
        step = lambda: ref.new_advance(Ua, Unew, Q, D, F)
    else:
        kind = "port"
        orc = oracle.Oracle()
        orc.set_threads(1)
        Unew, Q = lay.new_ghosted(NC), lay.new_ghosted(NQ)
        D, F = lay.new_interior(NC), lay.new_interior(NC)
        dx = box_dx(edge)
        step = lambda: orc.advance(lay, dx, CFL, ETA, ALAM, U, Unew, Q, D, F)
    n = 0
    with quiet_stdout():
        t0 = time.perf_counter()
        while n < 2 or (time.perf_counter() - t0) < budget_s:
            step()
            n += 1
        el = time.perf_counter() - t0
    return edge ** 3 * n / el, n, el, kind


def cpu_baseline(budgets=((32, 4.0), (64, 6.0), (128, 12.0))):
    """The reference cpu/ path (oracle/_ref: the unmodified sources, g++ -O2 as
    cpu/Makefile:2) on one host core at 32^3, 64^3 and 128^3 (cpu/main.cpp:27-34 runs
    whatever box its input names); the value is the 128^3 rate, the size closest to
    the 512^3 GPU workload that a bounded CPU sample can step (one 512^3 step takes
    about a minute on one core)."""
    by_edge, parts, kind = {}, [], "port"
    for edge, budget in budgets:
        rate, n, el, kind = cpu_baseline_one(edge, budget)
        by_edge[str(edge)] = rate
        parts.append(f"{n} steps of {edge}^3 in {el:.1f} s")
    return {"value": by_edge[str(budgets[-1][0])], "unit": "cell-steps/s", "cores": 1, "kind": kind,
            "cpu_model": cpu_model(), "by_edge": by_edge,
            "sample": "RK3 steps of one box (seeded sinusoids on wavenumbers 1-4), single thread "
                      "(the path has no threading: its OpenMP is commented out): " + "; ".join(parts)
                      + f"; value = the {budgets[-1][0]}^3 rate"}


def cpu_baseline_all_cores(edge: int = 128, budget_s: float = 10.0):
    """The oracle's CPU restatement of new_advance (same operation order, gcc -O2) on
    every host core the process may use (at most 16, the host share of one GPU), with
    OpenMP over its per-cell loops: the CPU path at full host width, beside the
    reference's single-threaded one."""
    import oracle
    from cns_nospec_amd.layout import NC, NQ, BoxLayout
    from cns_nospec_amd.state import ALAM, CFL, ETA, box_dx, sinusoid_state

    orc = oracle.Oracle()
    cores = min(16, len(os.sched_getaffinity(0)))
    threads = orc.set_threads(cores)
    lay = BoxLayout.cube(edge)
    U = sinusoid_state(lay, 7, kmax=4)
    Unew, Q = lay.new_ghosted(NC), lay.new_ghosted(NQ)
    D, F = lay.new_interior(NC), lay.new_interior(NC)
    dx = box_dx(edge)
    n = 0
    t0 = time.perf_counter()
    while n < 2 or (time.perf_counter() - t0) < budget_s:
        orc.advance(lay, dx, CFL, ETA, ALAM, U, Unew, Q, D, F)
        n += 1
    el = time.perf_counter() - t0
    return {"value": edge ** 3 * n / el, "unit": "cell-steps/s", "cores": threads, "kind": "port",
            "cpu_model": cpu_model(), "sample": f"{n} RK3 steps of one {edge}^3 box (seeded sinusoids), {threads} OpenMP "
                      f"threads, {el:.1f} s"}


def field_errors(a, b):
    """Per-field relative L2 and normwise max error of a vs b, (5, ...) arrays."""
    import numpy as np
    out = {}
    for l, name in enumerate(("rho", "mx", "my", "mz", "E")):
        x, y = a[l].ravel(), b[l].ravel()
        out[name] = {"rel_l2": float(np.linalg.norm(x - y) / np.linalg.norm(y)),
                     "max_rel": float(np.abs(x - y).max() / np.abs(y).max())}
    return out


def worst_errors(errs):
    return {f: {k: max(e[f][k] for e in errs) for k in ("rel_l2", "max_rel")} for f in errs[0]}


def run_batched(reqs, fp32: bool):
    """All boxes resident in one engine; returns rate and final states (host)."""
    import torch
    from cns_nospec_amd.batch import BatchEngine, box_bytes
    from cns_nospec_amd.serve import request_state
    budget = sum(box_bytes(r.edge, fp32) for r in reqs) + (256 << 20)
    eng = BatchEngine(budget, fp32=fp32, max_boxes=len(reqs))
    ids = [eng.add(request_state(r).reshape(-1), r.edge) for r in reqs]
    eng.step(0)  # captures the step graph
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    eng.step(reqs[0].steps)
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    q = eng.query(ids)
    states = [eng.read(i).cpu().numpy() for i in ids]
    eng.close()
    torch.cuda.empty_cache()
    work = sum(r.edge ** 3 * r.steps for r in reqs)
    return {"cell_steps_per_s": work / el, "elapsed_s": el,
            "failed": sum(1 for i in ids if q[i]["err"])}, states


def run_plain_boxes(reqs):
    """The same boxes one after another through cns_advance (dt to the host each step)."""
    import torch
    from cns_nospec_amd import ops
    from cns_nospec_amd.layout import NC, BoxLayout
    from cns_nospec_amd.serve import request_state
    states, el = [], 0.0
    for r in reqs:
        lay = BoxLayout.cube(r.edge)
        U, Unew, Q, D, F = ops.allocate_variables(lay, "cuda")
        g, ng, e = lay.dimg, lay.ng, r.edge
        view = U.view(NC, g[2], g[1], lay.xg)[:, ng:ng + e, ng:ng + e, ng:ng + e]
        view.copy_(request_state(r))
        st = ops.Stepper(lay)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(r.steps):
            st.new_advance(U, Unew, Q, D, F)
        torch.cuda.synchronize()
        el += time.perf_counter() - t0
        states.append(view.cpu().numpy())
        st.close()
        del U, Unew, Q, D, F
    work = sum(r.edge ** 3 * r.steps for r in reqs)
    return {"cell_steps_per_s": work / el, "elapsed_s": el}, states


def batched_secondary():
    from cns_nospec_amd.serve import batch_mix
    reqs = batch_mix()
    plain, s64 = run_plain_boxes(reqs)
    b64, sb64 = run_batched(reqs, False)
    b32, sb32 = run_batched(reqs, True)
    import numpy as np
    exact = all(np.array_equal(a, b) for a, b in zip(sb64, s64))
    errs = worst_errors([field_errors(a, b) for a, b in zip(sb32, s64)])
    work = sum(r.edge ** 3 * r.steps for r in reqs)
    return {"workload": "64 periodic boxes, edges uniform over {16,24,32,48,64,96} (seed 4), "
                        "Mach U[0.05,0.3], own seeded sinusoidal states, 50 RK3 steps each",
            "cell_steps": work, "unit": "cell-steps/s",
            "plain_fp64": plain, "batched_fp64": b64, "batched_fp32_flux": b32,
            "batched_fp64_bit_exact_vs_plain": exact,
            "ratio_batched_fp32_vs_plain_fp64": b32["cell_steps_per_s"] / plain["cell_steps_per_s"],
            "ratio_batched_fp64_vs_plain_fp64": b64["cell_steps_per_s"] / plain["cell_steps_per_s"],
            "fp32_flux_error_vs_fp64_after_50_steps": errs}


def fp32_error_secondary(nbox: int = 8, steps: int = 100, seed: int = 2024):
    """FP32-flux error after exactly 100 steps on the first boxes of the request stream
    (edge 24-128, Mach up to 0.3): FP32-flux engine against the FP64 engine, which is
    bit-identical to the plain path and the oracle."""
    from cns_nospec_amd.serve import make_stream
    reqs = [r for r in make_stream(nbox, 0.0, seed)]
    for r in reqs:
        r.steps = steps
    _, s64 = run_batched(reqs, False)
    _, s32 = run_batched(reqs, True)
    return {"boxes": [r.edge for r in reqs], "machs": [round(r.mach, 3) for r in reqs], "steps": steps,
            "per_field_worst": worst_errors([field_errors(a, b) for a, b in zip(s32, s64)])}


def served_secondary(nreq: int, seed: int = 2024, poisson_nreq: int = 200):
    """The mixed-size request stream (edge 24-128, 50-500 steps) served twice: as a
    burst of all `nreq` requests (every request waiting at t = 0: both servers
    saturated, so the aggregate rates compare capacities), and as a Poisson stream of
    the first `poisson_nreq` requests offering 1.2x the plain path's measured capacity
    (latency under load).  Progress goes to stderr."""
    import torch
    from cns_nospec_amd.serve import BatchServer, PlainServer, make_stream
    budget = 200 << 30
    keep = tuple(range(min(4, nreq)))
    every = max(1, nreq // 10)
    burst = make_stream(nreq, 0.0, seed)
    work = sum(r.edge ** 3 * r.steps for r in burst)
    res_p, plain = PlainServer(keep_states=keep, progress_every=every).run(burst)
    torch.cuda.empty_cache()
    srv = BatchServer(budget, fp32=True, keep_states=keep, progress_every=every)
    res_b, bat = srv.run(burst)
    srv.close()
    torch.cuda.empty_cache()
    errs = worst_errors([field_errors(res_b[i].state.cpu().numpy(), res_p[i].state.cpu().numpy())
                         for i in keep])
    npo = min(nreq, poisson_nreq)
    lam = 1.2 * plain["cell_steps_per_s"] / (work / nreq)
    stream = make_stream(npo, lam, seed)
    _, plain_l = PlainServer().run(stream)
    torch.cuda.empty_cache()
    srv = BatchServer(budget, fp32=True)
    _, bat_l = srv.run(stream)
    srv.close()
    torch.cuda.empty_cache()
    ach = BATCHED_MIN_BYTES_PER_CELL_STEP * bat["cell_steps_per_s"] / 1e9
    measured = read_traffic("batched_fp32_step", 0)
    return {"workload": f"{nreq} requests (seed {seed}), edge U{{24..128}}, steps U{{50..500}}, "
                        "Mach U[0.05,0.3], 200 GB budget, FP32-flux continuous batching vs "
                        "plain FP64 one box at a time (cns_advance, dt to the host each step)",
            "cell_steps": int(work), "unit": "cell-steps/s",
            "burst": {"requests": nreq, "plain_fp64": plain, "batched_fp32_flux": bat,
                      "ratio": bat["cell_steps_per_s"] / plain["cell_steps_per_s"]},
            "poisson": {"requests": npo,
                        "workload": f"the first {npo} requests of the same seeded stream, Poisson "
                                    f"arrivals at 1.2x the plain path's burst capacity",
                        "rate_per_s": lam, "offered_vs_plain_capacity": 1.2,
                        "plain_fp64": plain_l, "batched_fp32_flux": bat_l},
            "roofline": {"bound": "hbm", "basis": "320 B per interior cell-step (read U, write U1; "
                                                  "read U1 and U, write U2; read U2 and U, write U)",
                         "achieved": round(ach, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": round(ach / HBM_PEAK_GBS, 4),
                         "traffic_bytes_per_cell_step": measured,
                         "traffic_vs_minimum": round(measured / BATCHED_MIN_BYTES_PER_CELL_STEP, 2)
                         if measured else None,
                         "traffic_GBps_at_served_rate": round(measured * bat["cell_steps_per_s"] / 1e9, 1)
                         if measured else None},
            "fp32_flux_error_vs_fp64": errs,
            "error_sample": f"requests {list(keep)} of the burst run, final state (50-500 steps)"}


def domain_secondary(edge: int = 256, m: int = 64, steps: int = 10, warmup: int = 2):
    """A periodic edge^3 domain cut into m^3 boxes (the reference Fortran driver's
    max_grid_size split), stepped as one multi-box domain and as one box; the dt
    sequences must agree exactly (the states are bit-identical, as the tests check)."""
    import torch
    from cns_nospec_amd import ops
    from cns_nospec_amd.domain import Domain
    from cns_nospec_amd.layout import NC, BoxLayout
    from cns_nospec_amd.state import device_sinusoid_state

    dev = torch.device("cuda:0")
    lay = BoxLayout.cube(edge)
    U, Unew, Q, D, F = ops.allocate_variables(lay, dev)
    device_sinusoid_state(lay, 2024, dev, kmax=4, out=U)
    g, ng = lay.dimg, lay.ng
    whole = U.view(NC, g[2], g[1], lay.xg)[:, ng:ng + edge, ng:ng + edge, ng:ng + edge]
    dom = Domain(edge, m)
    dU, dUnew, dQ, dD, dF = dom.allocate(dev)
    dom.scatter(whole, NC, dU)
    st = ops.Stepper(lay)
    res = {}
    for name, step in (("domain", lambda: dom.advance(dU, dUnew, dQ, dD, dF)),
                       ("single_box", lambda: st.new_advance(U, Unew, Q, D, F))):
        dts = [step() for _ in range(warmup)]
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        dts += [step() for _ in range(steps)]
        torch.cuda.synchronize()
        el = time.perf_counter() - t0
        res[name] = {"cell_steps_per_s": edge ** 3 * steps / el, "ms_per_step": el / steps * 1e3, "dts": dts}
    same = res["domain"].pop("dts") == res["single_box"].pop("dts")
    st.close()
    dom.close()
    return {"workload": f"periodic {edge}^3 domain, max_grid_size {m} ({dom.nbox} boxes), seeded "
                        f"sinusoids, {steps} RK3 steps", "unit": "cell-steps/s", **res,
            "dt_sequence_identical": same}


def kernels_once(edge: int, reps: int = 3):
    """One warm-up and `reps` launches of every kernel of the fused plain step on an
    edge^3 box, plus a plain device copy of known size (for calibrating rocprofv3
    byte counters).  Used as the target of rocprofv3 runs; prints nothing timed."""
    import torch
    from cns_nospec_amd import ops
    from cns_nospec_amd.layout import BoxLayout
    from cns_nospec_amd.state import device_sinusoid_state

    dev = torch.device("cuda:0")
    lay = BoxLayout.cube(edge)
    U, Unew, Q, D, F = ops.allocate_variables(lay, dev)
    device_sinusoid_state(lay, 2024, dev, kmax=4, out=U)
    st = ops.Stepper(lay)
    for _ in range(reps + 1):
        st.new_advance(U, Unew, Q, D, F)
    src = torch.empty(1 << 27, dtype=torch.float64, device=dev).fill_(1.0)  # 1 GiB
    dst = torch.empty_like(src)
    for _ in range(reps):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"kernels_once": edge, "reps": reps, "copy_bytes_each_way": src.numel() * 8}))


def batched_once(nreq: int = 200, steps: int = 4, seed: int = 2024):
    """The served request mix (edge 24-128) all resident in the batched engine, FP32
    flux then FP64: one captured warm-up step and `steps` timed steps each.  Used as
    the target of rocprofv3 runs of the batched kernels; ends with three copies of a
    1 GiB buffer for calibrating byte counters."""
    import torch
    from cns_nospec_amd.batch import BatchEngine, box_bytes
    from cns_nospec_amd.serve import make_stream, request_state
    reqs = make_stream(nreq, 0.0, seed)
    cells = sum(r.edge ** 3 for r in reqs)
    for fp32 in (True, False):
        eng = BatchEngine(sum(box_bytes(r.edge, fp32) for r in reqs) + (256 << 20), fp32=fp32,
                          max_boxes=nreq)
        for r in reqs:
            eng.add(request_state(r).reshape(-1), r.edge)
        eng.step(1)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        eng.step(steps)
        torch.cuda.synchronize()
        el = time.perf_counter() - t0
        print(json.dumps({"batched_once": nreq, "fp32_flux": fp32, "cells": cells, "steps": steps,
                          "cell_steps_per_s": cells * steps / el}), flush=True)
        eng.close()
        torch.cuda.empty_cache()
    src = torch.empty(1 << 27, dtype=torch.float64, device="cuda").fill_(1.0)  # 1 GiB
    dst = torch.empty_like(src)
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--edge", type=int, default=512)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-secondary", action="store_true")
    ap.add_argument("--serve-requests", type=int, default=2000,
                    help="requests of the served burst (0 skips it); the Poisson leg serves "
                         "the first 200 of them")
    ap.add_argument("--kernels-once", action="store_true",
                    help="run each step kernel a few times and a calibration copy (for rocprofv3)")
    ap.add_argument("--batched-once", action="store_true",
                    help="a few steps of the batched engines on the request mix (for rocprofv3)")
    ap.add_argument("--once-requests", type=int, default=200,
                    help="boxes of the request mix resident in --batched-once")
    args = ap.parse_args()
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}, WORLD_SIZE {world}): "
              "independent boxes need no communication and a single box is not sharded here.",
              file=sys.stderr)
        sys.exit(2)

    import torch
    from cns_nospec_amd import _lib
    if not torch.cuda.is_available():
        print("bench.py needs a GPU", file=sys.stderr)
        sys.exit(1)
    _lib.load()
    if args.kernels_once:
        kernels_once(args.edge)
        return
    if args.batched_once:
        batched_once(args.once_requests)
        return

    main_run = run_plain(args.edge, args.steps, args.warmup, windows=WINDOWS)
    kern = main_run["kernels"]
    dom = max(kern, key=lambda k: kern[k][0]) if kern else None
    roof = None
    if dom:
        ms_tot, cnt = kern[dom]
        avg_s = ms_tot / cnt / 1e3
        nbytes = algorithmic_bytes(dom, args.edge)
        ach = nbytes / avg_s / 1e9
        traffic = read_traffic(dom, args.edge)
        roof = {"bound": "hbm", "kernel": dom, "achieved": round(ach, 1), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(ach / HBM_PEAK_GBS, 4),
                "traffic": traffic, "algorithmic_bytes": nbytes, "avg_ms": round(avg_s * 1e3, 4),
                "traffic_vs_algorithmic": round(traffic / nbytes, 3) if traffic else None}
    step_bytes = step_algorithmic_bytes(args.edge)
    fused_min = fused_step_min_bytes(args.edge)
    measured_step = step_traffic(kern, args.edge)
    out = {
        "metric": METRIC,
        "value": main_run["cell_steps_per_s"],
        "unit": "cell-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": main_run["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"plain FP64 RK3 step through the C ABI, one periodic "
                               f"{args.edge}^3 box ({args.edge + 8}^3 ghosted, ng=4), seeded "
                               f"sinusoids on wavenumbers 1-4, dt to host every step",
                   "edge": args.edge, "ng": 4, "cfl": 0.5, "eta": 1.8e-4, "alam": 150.0},
        "timing": f"median of {WINDOWS} windows of exactly {args.steps} steps each (window "
                  f"ms/step {main_run['window_ms_per_step']})",
        "step_hbm_fraction": round(step_bytes / (main_run["ms_per_step"] / 1e3) / 1e9
                                   / HBM_PEAK_GBS, 4),
        "step_bytes": {"abi_unfused": step_bytes, "fused_minimum": fused_min,
                       "measured": measured_step,
                       "measured_vs_fused_minimum": round(measured_step / fused_min, 3)
                       if measured_step else None,
                       "fused_minimum_hbm_fraction": round(fused_min / (main_run["ms_per_step"] / 1e3)
                                                           / 1e9 / HBM_PEAK_GBS, 4)},
        "roofline": roof,
        "kernels_ms_per_launch": {k: round(v[0] / v[1], 4) for k, v in kern.items()},
    }
    if not args.no_secondary:
        sec = {}
        sec["cgs_cell64"] = run_cgs_deck(64, 100, 5)
        sec["batched"] = batched_secondary()
        sec["fp32_flux_error_100_steps"] = fp32_error_secondary()
        sec["domain"] = domain_secondary()
        if args.serve_requests > 0:
            sec["served"] = served_secondary(args.serve_requests)
        out["secondary"] = sec
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline()
        out["cpu_baseline_all_cores"] = cpu_baseline_all_cores()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
