"""Batched engine: many independent periodic boxes resident in HBM and advanced
together by libcns_hip.so (include/cns_batch.h).

Each box is the reference's problem (cpu/advance.cpp:13-154) with its own dt;
the engine advances all resident boxes one RK3 step per replay of a captured
hipGraph.  Box state is ghost-free ([5][e][e][e] float64, component outermost).

``fp32=True`` selects the FP32-flux mode (stencil arithmetic of hypterm and
diffterm in FP32; state, RK3 accumulation, p/T recovery and dt in FP64).

The tile shape of each box edge comes from ``tuning_gfx950.json``, which the
autotuner below (:func:`autotune`, ``python -m cns_nospec_amd.batch``) writes: it
times whole tile tables on the served request mix resident in one engine (one width
for every edge, so each RK stage stays one launch per kernel, and a z run per edge
bucket).  Edges the table does not list get 32-wide tiles and z runs of min(edge, 16).
"""
from __future__ import annotations

import ctypes as C
import json
import os
import time

import numpy as np

from . import _lib
from .layout import NC, BoxLayout
from .state import ALAM, CFL, ETA

TUNING_FILE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tuning_gfx950.json")


def _torch():
    import torch
    return torch


def _lib_stream():
    torch = _torch()
    lib = _lib.load()
    lib.cns_set_stream(C.c_void_p(torch.cuda.current_stream().cuda_stream))
    return lib


def box_bytes(edge: int, fp32: bool) -> int:
    return int(_lib.load().cns_batch_box_bytes(edge, int(fp32)))


class BatchEngine:
    """Python front end of cns_batch (one engine = one pool of hbm_budget bytes)."""

    def __init__(self, hbm_budget: int, fp32: bool = False, max_boxes: int = 4096, tuning=None):
        self.fp32 = bool(fp32)
        lib = _lib.load()
        h = C.c_void_p()
        _lib.check(lib.cns_batch_create(C.byref(h), C.c_size_t(int(hbm_budget)), int(fp32), max_boxes))
        self._h = h
        self.edges = {}
        if tuning is None:
            tuning = load_tuning_for(self.fp32)
        for edge, (w, kz) in tuning.items():
            self.set_tile(edge, w, kz)

    # ------------------------------------------------------------------ boxes
    def add(self, u0, edge: int, dx: float = None, cfl: float = CFL, eta: float = ETA,
            alam: float = ALAM) -> int:
        """Admit a box from a ghost-free float64 device tensor of 5*e^3 values.
        Raises CnsBudgetError when the pool cannot hold it."""
        torch = _torch()
        if not (isinstance(u0, torch.Tensor) and u0.is_cuda and u0.dtype == torch.float64
                and u0.is_contiguous() and u0.numel() == NC * edge ** 3):
            raise ValueError("u0 must be a contiguous float64 device tensor of 5*e^3 values")
        dx = 0.2 / edge if dx is None else dx
        lib = _lib_stream()
        bid = C.c_int(-1)
        _lib.check(lib.cns_batch_add(self._h, edge, dx, cfl, eta, alam, C.c_void_p(u0.data_ptr()),
                                     C.byref(bid)))
        self.edges[bid.value] = edge
        return bid.value

    def remove(self, box_id: int):
        lib = _lib_stream()
        _lib.check(lib.cns_batch_remove(self._h, box_id))
        self.edges.pop(box_id, None)

    def read(self, box_id: int, out=None):
        """Current U of a box as a (5, e, e, e) float64 device tensor."""
        torch = _torch()
        e = self.edges[box_id]
        if out is None:
            out = torch.empty((NC, e, e, e), dtype=torch.float64, device="cuda")
        lib = _lib_stream()
        _lib.check(lib.cns_batch_read(self._h, box_id, C.c_void_p(out.data_ptr())))
        return out

    def set_target(self, box_id: int, steps: int):
        """Stop advancing the box once it has taken `steps` steps (0: no limit)."""
        lib = _lib_stream()
        _lib.check(lib.cns_batch_set_target(self._h, box_id, int(steps)))

    def step(self, nsteps: int = 1):
        lib = _lib_stream()
        _lib.check(lib.cns_batch_step(self._h, int(nsteps)))

    def query(self, ids):
        ids = list(ids)
        n = len(ids)
        a_ids = (C.c_int * max(n, 1))(*ids)
        steps = (C.c_longlong * max(n, 1))()
        dt = (C.c_double * max(n, 1))()
        courno = (C.c_double * max(n, 1))()
        err = (C.c_int * max(n, 1))()
        lib = _lib_stream()
        _lib.check(lib.cns_batch_query(self._h, n, a_ids, steps, dt, courno, err))
        return {i: {"steps": steps[k], "dt": dt[k], "courno": courno[k], "err": err[k]}
                for k, i in enumerate(ids)}

    # ----------------------------------------------------------------- tuning
    def set_tile(self, edge: int, width: int, kz: int):
        _lib.check(_lib.load().cns_batch_set_tile(self._h, edge, width, kz))

    def get_tile(self, edge: int):
        w, kz = C.c_int(), C.c_int()
        _lib.check(_lib.load().cns_batch_get_tile(self._h, edge, C.byref(w), C.byref(kz)))
        return w.value, kz.value

    def stats(self):
        used, boxes, tiles = C.c_size_t(), C.c_int(), C.c_int()
        lib = _lib_stream()
        _lib.check(lib.cns_batch_stats(self._h, C.byref(used), C.byref(boxes), C.byref(tiles)))
        return {"used_bytes": used.value, "boxes": boxes.value, "tiles": tiles.value}

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            _lib.load().cns_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def load_tuning_for(fp32: bool, path: str = TUNING_FILE):
    """{edge: (width, kz)} of one precision from a tuning file; a key is one edge or an
    inclusive range "lo-hi"."""
    if not os.path.exists(path):
        return {}
    raw = json.load(open(path))
    out = {}
    for k, v in raw.get("fp32" if fp32 else "fp64", {}).items():
        lo, _, hi = k.partition("-")
        for e in range(int(lo), int(hi or lo) + 1):
            out[e] = (int(v[0]), int(v[1]))
    return out


# ------------------------------------------------------------------ layouts
def compact_from_ghosted(layout: BoxLayout, flat):
    """Interior of a ghosted ABI array as a ghost-free (5, nz, ny, nx) array."""
    return layout.interior_of(layout.gview(flat, NC))


def compact_sinusoid(edge: int, seed: int, mach: float = 0.085, kmax: int = 2, device=None):
    """The seeded sinusoidal state of :mod:`.state`, ghost-free, as a device tensor
    (evaluated on the host with numpy, so it is bit-identical to the state the
    plain path and the oracle use)."""
    from .state import conserved_from_primitives, sinusoid_primitives
    torch = _torch()
    p = sinusoid_primitives(edge, seed, mach=mach, kmax=kmax)
    cons = np.ascontiguousarray(conserved_from_primitives(p["rho"], p["u"], p["v"], p["w"], p["p"]))
    return torch.from_numpy(cons).to(device or "cuda")


# ---------------------------------------------------------------- autotuner
WIDTHS = (16, 32, 64)
KZS = (8, 16, 32, 64)
# edge ranges whose z run the tuner chooses separately (the engine admits edges 8..512)
EDGE_BUCKETS = ((8, 31), (32, 47), (48, 63), (64, 95), (96, 512))


def _rate(eng, cells: int, steps: int, reps: int) -> float:
    """Best of `reps` timed windows of `steps` engine steps (after one untimed step that
    rebuilds and captures the step graph), in cell-steps/s."""
    torch = _torch()
    eng.step(1)
    torch.cuda.synchronize()
    best = 0.0
    for _ in range(reps):
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record()
        eng.step(steps)
        b.record()
        torch.cuda.synchronize()
        best = max(best, cells * steps / (a.elapsed_time(b) / 1e3))
    return best


def autotune(requests, fp32: bool, widths=WIDTHS, kzs=KZS, buckets=EDGE_BUCKETS, steps: int = 4,
             reps: int = 2, verbose: bool = False):
    """Tile table for a mixed workload, timed on the mix itself.

    All `requests` (objects with .edge, and what serve.request_state needs) are resident in
    one engine, which is what the server runs.  Stage 1: every (width, kz) applied to every
    edge (one width class, so each RK stage stays one launch per kernel); the fastest is the
    base.  Stage 2: for each edge bucket in turn, every kz with the base width for the
    bucket's edges, kept when the whole mix steps faster.  Tables are judged by the mix's
    aggregate cell-steps/s; tiling never changes results (each cell's arithmetic is the
    same whichever lane computes it).

    Returns (table {"lo-hi": (width, kz)} per edge bucket, log of the trials)."""
    from .serve import request_state
    torch = _torch()
    eng = BatchEngine(sum(box_bytes(r.edge, fp32) for r in requests) + (256 << 20), fp32=fp32,
                      max_boxes=len(requests), tuning={})
    for r in requests:
        eng.add(request_state(r).reshape(-1), r.edge)
    present = sorted({r.edge for r in requests})
    cells = sum(r.edge ** 3 for r in requests)
    log = []

    def trial(table, label):
        for e in present:
            w, kz = table(e)
            eng.set_tile(e, w, kz)
        rate = _rate(eng, cells, steps, reps)
        log.append({"trial": label, "cell_steps_per_s": rate})
        if verbose:
            print(json.dumps({"autotune": label, "fp32": fp32, "cell_steps_per_s": rate}), flush=True)
        return rate

    best, base = 0.0, None
    for w in widths:
        for kz in kzs:
            rate = trial(lambda e, w=w, kz=kz: (w, kz), f"all edges {w} x {kz}")
            if rate > best:
                best, base = rate, (w, kz)
    choice = {bk: base[1] for bk in buckets}

    def bucket_of(e):
        return next((bk for bk in buckets if bk[0] <= e <= bk[1]), None)

    for bk in buckets:
        if not any(bk[0] <= e <= bk[1] for e in present):
            continue
        for kz in kzs:
            if kz == choice[bk]:
                continue
            trial_choice = dict(choice)
            trial_choice[bk] = kz
            rate = trial(lambda e, c=trial_choice: (base[0], c.get(bucket_of(e), base[1])),
                         f"edges {bk[0]}-{bk[1]} kz {kz}")
            if rate > best:
                best, choice = rate, trial_choice
    eng.close()
    del eng
    torch.cuda.empty_cache()
    table = {f"{bk[0]}-{bk[1]}": (base[0], choice[bk]) for bk in buckets}
    log.append({"chosen": {k: list(v) for k, v in table.items()}, "cell_steps_per_s": best})
    return table, log


def save_tuning(table_fp32: dict, table_fp64: dict, method: str, logs: dict, path: str = TUNING_FILE):
    doc = {"device": "gfx950 (MI355X)", "generated": time.strftime("%Y-%m-%d"),
           "written_by": "cns_nospec_amd.batch autotune", "method": method,
           "fp32": {str(k): [v[0], v[1]] for k, v in table_fp32.items()},
           "fp64": {str(k): [v[0], v[1]] for k, v in table_fp64.items()},
           "trials": logs}
    with open(path, "w") as f:
        json.dump(doc, f, indent=1)


if __name__ == "__main__":
    import argparse
    ap = argparse.ArgumentParser(description="autotune the batched engine's tile table on the "
                                             "served request mix")
    ap.add_argument("--requests", type=int, default=200, help="boxes of the request mix resident")
    ap.add_argument("--seed", type=int, default=2024)
    ap.add_argument("--out", default=TUNING_FILE)
    args = ap.parse_args()
    from .serve import make_stream
    reqs = make_stream(args.requests, 0.0, args.seed)
    t32, l32 = autotune(reqs, True, verbose=True)
    t64, l64 = autotune(reqs, False, verbose=True)
    method = (f"{args.requests} requests of the served stream (seed {args.seed}: edges U{{24..128}}) "
              "resident in one engine; every (width, kz) for all edges, then kz per edge bucket "
              "with the best width; aggregate cell-steps/s of the mix decides")
    save_tuning(t32, t64, method, {"fp32": l32, "fp64": l64}, args.out)
    print(json.dumps({"saved": args.out}))
