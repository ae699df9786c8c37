"""Host logic of the serving layer (no GPU): the seeded request stream and the
latency/throughput summary."""
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
