"""The batched engine (include/cns_batch.h) on the GPU.

* FP64 engine: bit-identical to the oracle (and so to the reference) per box,
  for uneven edges and partial tiles, whatever the tile shape.
* A box gives the same bits alone and in a mix (both precisions).
* FP32-flux mode: relative L2 / max error per field against FP64 after 100 steps.
* Admission by budget, and per-box fault isolation (a non-physical box is
  flagged alone)."""
import numpy as np
import pytest

from cns_nospec_amd import _lib
from cns_nospec_amd.batch import BatchEngine, box_bytes, compact_sinusoid
from cns_nospec_amd.layout import NC, NQ, BoxLayout
from cns_nospec_amd.state import ALAM, CFL, ETA, box_dx, pack_ghosted

pytestmark = pytest.mark.gpu


def oracle_run(orc, edge, u0_compact, nsteps):
    lay = BoxLayout.cube(edge)
    U = pack_ghosted(lay, u0_compact)
    Unew, Q = lay.new_ghosted(NC), lay.new_ghosted(NQ)
    D, F = lay.new_interior(NC), lay.new_interior(NC)
    dts = [orc.advance(lay, box_dx(edge), CFL, ETA, ALAM, U, Unew, Q, D, F) for _ in range(nsteps)]
    return np.ascontiguousarray(lay.interior_of(lay.gview(U, NC))), dts


@pytest.mark.parametrize("tiles", [None, {8: (32, 8), 13: (32, 5), 16: (64, 5), 20: (16, 7), 24: (32, 24),
                                           36: (16, 8)}])
def test_fp64_engine_bit_exact_vs_oracle(orc, device, tiles):
    """The FP64 engine equals the oracle bit for bit on every box, down to the smallest
    edge (8, where a 32-wide tile holds each cell four times over) and odd edges."""
    edges = (8, 13, 16, 20, 24, 36)
    eng = BatchEngine(2 << 30, fp32=False, max_boxes=8, tuning=tiles or {})
    init = {e: compact_sinusoid(e, 50 + e, mach=0.2, kmax=3) for e in edges}
    ids = {e: eng.add(init[e].reshape(-1), e) for e in edges}
    nsteps = 4
    eng.step(nsteps)
    q = eng.query(list(ids.values()))
    for e in edges:
        ref, dts = oracle_run(orc, e, init[e].cpu().numpy(), nsteps)
        got = eng.read(ids[e]).cpu().numpy()
        assert q[ids[e]]["steps"] == nsteps and q[ids[e]]["err"] == 0
        assert np.array_equal(got, ref), f"edge {e}: max |d| {np.abs(got - ref).max():.3e}"
        # dt of the next step = cfl/courno of the state reached
        assert q[ids[e]]["dt"] > 0
    eng.close()


@pytest.mark.parametrize("fp32", [False, True])
def test_alone_equals_in_mix(device, fp32):
    u = compact_sinusoid(24, 3, mach=0.25)
    alone = BatchEngine(1 << 30, fp32=fp32, max_boxes=4)
    a = alone.add(u.reshape(-1), 24)
    alone.step(10)
    ua = alone.read(a).cpu().numpy()
    mix = BatchEngine(4 << 30, fp32=fp32, max_boxes=16)
    others = [compact_sinusoid(e, 11 + e, mach=0.1) for e in (16, 40, 32)]
    mix.add(others[0].reshape(-1), 16)
    m = mix.add(u.reshape(-1), 24)
    mix.add(others[1].reshape(-1), 40)
    mix.step(4)
    mix.add(others[2].reshape(-1), 32)   # admitted mid-run: must not disturb the others
    mix.step(6)
    um = mix.read(m).cpu().numpy()
    assert np.array_equal(ua, um)
    alone.close()
    mix.close()


@pytest.mark.parametrize("fp32", [False, True])
def test_step_target_stops_a_box_exactly(device, fp32):
    """A box with a step target stops there: its state equals the same box advanced
    exactly that many steps, while the other boxes keep going; a target of 0 lifts
    the limit again."""
    u = compact_sinusoid(24, 8, mach=0.2)
    v = compact_sinusoid(32, 9, mach=0.1)
    ref = BatchEngine(1 << 30, fp32=fp32, max_boxes=2)
    r = ref.add(u.reshape(-1), 24)
    ref.step(3)
    u3 = ref.read(r).cpu().numpy()
    ref.step(2)
    u5 = ref.read(r).cpu().numpy()
    eng = BatchEngine(1 << 30, fp32=fp32, max_boxes=4)
    a = eng.add(u.reshape(-1), 24)
    b = eng.add(v.reshape(-1), 32)
    eng.set_target(a, 3)
    eng.step(7)
    q = eng.query([a, b])
    assert q[a]["steps"] == 3 and q[b]["steps"] == 7
    assert np.array_equal(eng.read(a).cpu().numpy(), u3)
    eng.set_target(a, 0)
    eng.step(2)
    assert eng.query([a])[a]["steps"] == 5
    assert np.array_equal(eng.read(a).cpu().numpy(), u5)
    ref.close()
    eng.close()


def test_fp32_flux_error_after_100_steps(device):
    edge, nsteps = 32, 100
    u = compact_sinusoid(edge, 5, mach=0.085)
    e64 = BatchEngine(1 << 30, fp32=False, max_boxes=2)
    e32 = BatchEngine(1 << 30, fp32=True, max_boxes=2)
    i64, i32 = e64.add(u.reshape(-1), edge), e32.add(u.reshape(-1), edge)
    e64.step(nsteps)
    e32.step(nsteps)
    a, b = e32.read(i32).cpu().numpy(), e64.read(i64).cpu().numpy()
    for l in range(NC):
        rel_l2 = np.linalg.norm(a[l] - b[l]) / np.linalg.norm(b[l])
        rel_max = np.abs(a[l] - b[l]).max() / np.abs(b[l]).max()
        print(f"field {l}: rel L2 {rel_l2:.3e}  rel max {rel_max:.3e}")
        assert rel_l2 <= 1e-6, f"field {l}: relative L2 {rel_l2:.3e} above 1e-6"
    e64.close()
    e32.close()


def test_fp32_flux_error_request_boxes_100_steps(device):
    """The FP32-flux target on the served workload's own boxes: the first 8 requests of
    the request stream (edges 24-128, Mach up to 0.3), 100 steps each, FP32-flux engine
    against the FP64 engine (bit-identical to the oracle): per box and field, relative
    L2 <= 1e-6."""
    from cns_nospec_amd.batch import box_bytes
    from cns_nospec_amd.serve import make_stream, request_state
    reqs = make_stream(8, 0.0, 2024)
    out = {}
    for fp32 in (False, True):
        eng = BatchEngine(sum(box_bytes(r.edge, fp32) for r in reqs) + (64 << 20), fp32=fp32, max_boxes=8)
        ids = [eng.add(request_state(r).reshape(-1), r.edge) for r in reqs]
        eng.step(100)
        q = eng.query(ids)
        assert all(q[i]["steps"] == 100 and q[i]["err"] == 0 for i in ids)
        out[fp32] = [eng.read(i).cpu().numpy() for i in ids]
        eng.close()
    for r, a, b in zip(reqs, out[True], out[False]):
        for l in range(NC):
            rel_l2 = np.linalg.norm(a[l] - b[l]) / np.linalg.norm(b[l])
            print(f"edge {r.edge} mach {r.mach:.3f} field {l}: rel L2 {rel_l2:.3e}")
            assert rel_l2 <= 1e-6, f"edge {r.edge} field {l}: relative L2 {rel_l2:.3e}"


def test_fp32_tile_widths_agree(device):
    """The tile shape only decides which lane computes which cell, so the FP32 engine's
    16-, 32- and 64-wide classes (and any z run) give the same bits as the default
    32-wide tiles."""
    edges = (16, 24, 40)
    init = {e: compact_sinusoid(e, 70 + e, mach=0.2, kmax=3) for e in edges}
    results = []
    for tiles in ({}, {16: (16, 16), 24: (64, 7), 40: (16, 40)}, {16: (64, 4), 24: (16, 24), 40: (64, 9)}):
        eng = BatchEngine(2 << 30, fp32=True, max_boxes=4, tuning=tiles)
        ids = {e: eng.add(init[e].reshape(-1), e) for e in edges}
        eng.step(5)
        results.append({e: eng.read(ids[e]).cpu().numpy() for e in edges})
        eng.close()
    for r in results[1:]:
        for e in edges:
            assert np.array_equal(r[e], results[0][e]), f"edge {e}"


def test_budget_admission(device):
    edge = 32
    need = box_bytes(edge, True)
    eng = BatchEngine(2 * need + 4096, fp32=True, max_boxes=8)
    u = compact_sinusoid(edge, 1).reshape(-1)
    a = eng.add(u, edge)
    eng.add(u, edge)
    with pytest.raises(_lib.CnsBudgetError):
        eng.add(u, edge)
    eng.step(2)
    eng.remove(a)
    c = eng.add(u, edge)  # memory freed by the removal is reused
    eng.step(1)
    st = eng.stats()
    assert st["boxes"] == 2 and st["used_bytes"] <= 2 * need + 4096
    q = eng.query([c])
    assert q[c]["steps"] == 1
    eng.close()


def test_nonphysical_box_fails_alone(device):
    good = compact_sinusoid(24, 8)
    bad = good.clone()
    bad[0, 3, 4, 5] = -1.0  # negative density
    ref = BatchEngine(1 << 30, fp32=False, max_boxes=2)
    r = ref.add(good.reshape(-1), 24)
    ref.step(3)
    mix = BatchEngine(1 << 30, fp32=False, max_boxes=4)
    g = mix.add(good.reshape(-1), 24)
    b = mix.add(bad.reshape(-1), 24)
    mix.step(3)
    q = mix.query([g, b])
    assert q[b]["err"] == 1 and q[g]["err"] == 0
    assert np.array_equal(mix.read(g).cpu().numpy(), ref.read(r).cpu().numpy())
    ref.close()
    mix.close()


@pytest.mark.parametrize("fp32", [True, False])
def test_engine_variants_agree(device, monkeypatch, fp32):
    """The engine's kernel variants compute the same expressions on the same values, so
    they give the same bits, for every tile width and partial tiles: the diffterm tile
    whose planes are staged by direct-to-LDS loads (with a wz field) against the one that
    gathers its halos through registers, the hypterm tile staged by direct-to-LDS loads
    (FP32 flux and FP64) against the one that gathers its halos, and the FP32 hypterm
    tile that keeps its z window's FP64 leading values for the centre plane against the
    one that loads them again."""
    edges = (8, 13, 16, 20, 24, 40)
    init = {e: compact_sinusoid(e, 90 + e, mach=0.25, kmax=3) for e in edges}
    tiles = {8: (32, 8), 13: (32, 5), 16: (16, 5), 20: (32, 7), 24: (64, 24), 40: (32, 16)}
    out = []
    for diff, keep, hyp in (("1", "0", "1"), ("1", "0", "0"), ("0", "0", "0"), ("1", "1", "0")):
        monkeypatch.setenv("CNS_BATCH_DIFF", diff)
        monkeypatch.setenv("CNS_BATCH_KEEP", keep)
        monkeypatch.setenv("CNS_BATCH_HYP", hyp)
        monkeypatch.setenv("CNS_BATCH_HYP64", hyp)
        eng = BatchEngine(2 << 30, fp32=fp32, max_boxes=8, tuning=tiles)
        ids = {e: eng.add(init[e].reshape(-1), e) for e in edges}
        eng.step(6)
        out.append({e: eng.read(ids[e]).cpu().numpy() for e in edges})
        eng.close()
    for variant in out[1:]:
        for e in edges:
            assert np.array_equal(variant[e], out[0][e]), f"edge {e}"


def test_autotune_table_loads_and_keeps_bits(device, tmp_path):
    """The autotuner on a two-edge mix: its table uses one width for every bucket and
    kz values from the candidates, survives the file round trip (every edge of every
    bucket), is what the engine then runs, and gives the same bits as the default
    tiles in both precisions."""
    from cns_nospec_amd.batch import autotune, load_tuning_for, save_tuning
    from cns_nospec_amd.serve import Request, request_state
    reqs = [Request(i, 0.0, e, 5, 300 + i, 0.2) for i, e in enumerate((16, 24, 16, 24))]
    buckets = ((8, 20), (21, 40))
    tables, logs = {}, {}
    for fp32 in (True, False):
        tables[fp32], logs[fp32] = autotune(reqs, fp32, widths=(16, 32), kzs=(8, 16),
                                            buckets=buckets, steps=2, reps=1)
        t = tables[fp32]
        assert set(t) == {"8-20", "21-40"}
        assert len({w for w, _ in t.values()}) == 1 and all(w in (16, 32) for w, _ in t.values())
        assert all(kz in (8, 16) for _, kz in t.values())
        assert len(logs[fp32]) == 2 * 2 + 2 * 1 + 1  # uniform trials, bucket trials, choice
    path = str(tmp_path / "tuning.json")
    save_tuning(tables[True], tables[False], "two-edge test mix", {"fp32": logs[True], "fp64": logs[False]},
                path)
    for fp32 in (True, False):
        loaded = load_tuning_for(fp32, path)
        assert set(loaded) == set(range(8, 41))
        assert loaded[16] == tuple(tables[fp32]["8-20"]) and loaded[24] == tuple(tables[fp32]["21-40"])
        out = []
        for tuning in (loaded, {}):
            eng = BatchEngine(1 << 30, fp32=fp32, max_boxes=8, tuning=tuning)
            if tuning:
                assert eng.get_tile(24) == loaded[24]
            ids = [eng.add(request_state(r).reshape(-1), r.edge) for r in reqs]
            eng.step(5)
            out.append([eng.read(i).cpu().numpy() for i in ids])
            eng.close()
        for a, b in zip(*out):
            assert np.array_equal(a, b)
