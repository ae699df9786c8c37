"""Batched engine: many independent periodic boxes resident in HBM and advanced
together by libcns_hip.so (include/cns_batch.h).

Each box is the reference's problem (cpu/advance.cpp:13-154) with its own dt;
the engine advances all resident boxes one RK3 step per replay of a captured
hipGraph.  Box state is ghost-free ([5][e][e][e] float64, component outermost).

``fp32=True`` selects the FP32-flux mode (stencil arithmetic of hypterm and
diffterm in FP32; state, RK3 accumulation, p/T recovery and dt in FP64).

The tile shape of each box edge comes from ``tuning_gfx950.json``: 32-wide tiles
and z runs of min(edge, 16) for every edge, so that all boxes share one width class
(measured on a mixed request stream; edges the table does not list get the same).
The autotuner below (:func:`autotune`) times every shape per edge on batches of one
edge, for workloads that are not mixed; its table can replace the file.
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
    if not os.path.exists(path):
        return {}
    raw = json.load(open(path))
    table = raw.get("fp32" if fp32 else "fp64", {})
    return {int(k): (int(v[0]), int(v[1])) for k, v in table.items()}


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


def candidate_tiles(edge: int):
    kzs = sorted({k for k in (8, 16, 32, 64, edge) if k <= edge})
    return [(w, kz) for w in WIDTHS for kz in kzs if w <= max(16, 2 * edge)]


def autotune(edges, fp32: bool, cells_per_trial: int = 24_000_000, steps: int = 3,
             budget: int = 96 << 30, verbose: bool = False):
    """Measure every candidate tile shape per box edge and keep the fastest.

    For each edge a batch of identical-edge boxes (about cells_per_trial cells)
    is stepped with each candidate (width, kz); the rate in cell-steps/s decides.
    Returns {edge: (width, kz, cell_steps_per_s)}."""
    torch = _torch()
    out = {}
    for edge in edges:
        nbox = max(1, min(512, cells_per_trial // edge ** 3))
        eng = BatchEngine(min(budget, nbox * box_bytes(edge, fp32) + (64 << 20)), fp32=fp32,
                          max_boxes=nbox, tuning={})
        u0 = compact_sinusoid(edge, 7, kmax=2)
        for _ in range(nbox):
            eng.add(u0.reshape(-1), edge)
        best = None
        for w, kz in candidate_tiles(edge):
            eng.set_tile(edge, w, kz)
            eng.step(1)
            torch.cuda.synchronize()
            a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            a.record()
            eng.step(steps)
            b.record()
            torch.cuda.synchronize()
            rate = nbox * edge ** 3 * steps / (a.elapsed_time(b) / 1e3)
            if verbose:
                print(json.dumps({"autotune": edge, "fp32": fp32, "width": w, "kz": kz,
                                  "cell_steps_per_s": rate}), flush=True)
            if best is None or rate > best[2]:
                best = (w, kz, rate)
        out[edge] = best
        eng.close()
        del u0
        torch.cuda.empty_cache()
    return out


def save_tuning(results_fp32: dict, results_fp64: dict, path: str = TUNING_FILE):
    doc = {"device": "gfx950 (MI355X)", "generated": time.strftime("%Y-%m-%d"),
           "fp32": {str(k): [v[0], v[1]] for k, v in sorted(results_fp32.items())},
           "fp64": {str(k): [v[0], v[1]] for k, v in sorted(results_fp64.items())},
           "rates": {"fp32": {str(k): v[2] for k, v in results_fp32.items()},
                     "fp64": {str(k): v[2] for k, v in results_fp64.items()}}}
    with open(path, "w") as f:
        json.dump(doc, f, indent=1)


if __name__ == "__main__":
    import argparse
    ap = argparse.ArgumentParser(description="autotune the batched engine's tile table")
    ap.add_argument("--edges", default="16,24,32,48,64,96,128")
    ap.add_argument("--out", default=TUNING_FILE)
    args = ap.parse_args()
    edges = [int(x) for x in args.edges.split(",")]
    r32 = autotune(edges, True, verbose=True)
    r64 = autotune(edges, False, verbose=True)
    save_tuning(r32, r64, args.out)
    print(json.dumps({"saved": args.out}))
