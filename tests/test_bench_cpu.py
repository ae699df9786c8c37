"""The benchmark's arithmetic (no GPU): the algorithmic bytes behind `roofline` and the
step's HBM fraction, and the error summaries.

* Per step, the unfused ABI composition moves 8 (99 (n+8)^3 + 70 n^3) bytes: 2107 B per
  cell-step at 32^3, 1688 at 64^3, 1389.7 at 512^3 (the per-kernel terms summed over a
  step's three stages must give the same total).
* The fused hypterm + RK kernel's per-launch figure is the mean of its three stages:
  8g + 14c (stage 1: cons over halos, rho and D, result, u v w T), 8g + 21c (stage 2:
  plus the base, all six primitives), 8g + 20c (stage 3: the base, the result and F).
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
    stages = [8 * g + 14 * c, 8 * g + 21 * c, 8 * g + 20 * c]
    assert bench.algorithmic_bytes("hypterm_rk", n) == pytest.approx(8 * sum(stages) / 3, rel=1e-12)


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
