"""The benchmark's arithmetic (no GPU): the algorithmic bytes behind `roofline` and the
step's HBM fraction, and the error summaries.

* Per step, the unfused ABI composition moves 8 (99 (n+8)^3 + 70 n^3) bytes: 2107 B per
  cell-step at 32^3, 1688 at 64^3, 1389.7 at 512^3 (the per-kernel terms summed over a
  step's three stages must give the same total).
* The fused hypterm + RK kernel's per-launch figure is the mean of its three stages:
  5g + 13c (stage 1: the five conserved fields over halos, D without its rho row, the
  result, u v w T), 5g + 20c (stage 2: plus the base, all six primitives), 5g + 19c
  (stage 3: the base, the result and F).
* The fused step's own minimum is the sum of its launches: 94g + 15c doubles, 121.8 GB
  at 512^3.
* profiles/hbm_traffic.txt parses, and the measured bytes of the fused kernel exceed its
  algorithmic bytes (halo re-reads), as a measurement of real traffic must."""
import numpy as np
import pytest

import bench


@pytest.mark.parametrize("n,per_cell", [(32, 2107.0), (64, 1688.0), (512, 1389.7)])
def test_step_bytes_per_cell_step(n, per_cell):
    assert bench.step_algorithmic_bytes(n) / n ** 3 == pytest.approx(per_cell, abs=0.5)


@pytest.mark.parametrize("n", [16, 64, 512])
def test_kernel_terms_sum_to_the_step(n):
    a = bench.algorithmic_bytes
    total = sum(a("fill", n) + a("ctoprim", n) + a("diffterm", n) + a("hypterm", n) for _ in range(3))
    total += a("rk", n, stage=1) + 2 * a("rk", n, stage=2)
    assert total == pytest.approx(bench.step_algorithmic_bytes(n), rel=1e-12)


def test_fused_kernel_bytes_are_the_stage_mean():
    n = 512
    g, c = (n + 8) ** 3, n ** 3
    stages = [5 * g + 13 * c, 5 * g + 20 * c, 5 * g + 19 * c]
    assert bench.algorithmic_bytes("hypterm_rk", n) == pytest.approx(8 * sum(stages) / 3, rel=1e-12)


@pytest.mark.parametrize("n", [32, 512])
def test_fused_step_minimum_is_the_sum_of_its_launches(n):
    g, c = (n + 8) ** 3, n ** 3
    shell = g - c
    fills = 10 * shell + 2 * (5 + 4) * shell + 2 * (5 + 6) * shell  # U; T + uvwT; Unew + Q
    ctoprim_wz = 5 * g + 4 * g + g  # U in, u v w T and wz out
    wz = 2 * (g + g)  # w in, wz out, stages 2 and 3
    diff = 3 * (4 * g + g) + (4 + 4 + 5) * c  # u v w T and wz in; D out (rho row in stage 3)
    hyp = (5 * g + 13 * c) + (5 * g + 20 * c) + (5 * g + 19 * c)
    total = 8 * (fills + ctoprim_wz + wz + diff + hyp)
    assert bench.fused_step_min_bytes(n) == pytest.approx(total, rel=1e-12)
    if n == 512:
        assert total / 1e9 == pytest.approx(121.8, abs=0.05)


def test_measured_traffic_file():
    t = bench.read_traffic("hypterm_rk", 512)
    assert t is not None
    assert t > bench.algorithmic_bytes("hypterm_rk", 512)
    assert bench.read_traffic("hypterm_rk", 7) is None


def test_error_summaries():
    b = np.ones((5, 4, 4, 4))
    a = b.copy()
    a[1] *= 1 + 1e-7
    e = bench.field_errors(a, b)
    assert e["rho"]["rel_l2"] == 0.0
    assert e["mx"]["rel_l2"] == pytest.approx(1e-7, rel=1e-6)
    w = bench.worst_errors([e, bench.field_errors(b, b)])
    assert w["mx"]["rel_l2"] == pytest.approx(1e-7, rel=1e-6)
