"""The launch-occupancy analysis of a wavefront trace (blockvecchia_amd.trace)
on hand-built records whose gaps, occupancy and drain are known."""
import numpy as np
import pytest

from blockvecchia_amd import trace


def _rec(rows):
    """rows of (start_us, end_us, kernel, step, workgroup)."""
    return np.array(rows, dtype=float)


def test_two_launches_known_gap_and_occupancy():
    # one CU of capacity 2 * 4 = 8 wavefronts (paired kernel): launch A holds
    # 8 wavefronts over [0, 10); launch B starts at 12 (gap 2) with 8
    # wavefronts over [12, 20) and 4 more over [20, 30) (drain: half the slots idle)
    a = [(0.0, 10.0, 3, 0, i) for i in range(8)]
    b = [(12.0, 20.0, 3, 1, i) for i in range(8)] + [(20.0, 30.0, 3, 1, 8 + i) for i in range(4)]
    rows, tot = trace.launch_table(_rec(a + b), cus=1)
    assert [r["step"] for r in rows] == [0, 1]
    assert rows[0]["mean_occupancy"] == pytest.approx(1.0)
    assert rows[1]["gap_before_us"] == pytest.approx(2.0)
    assert rows[1]["duration_us"] == pytest.approx(18.0)
    assert rows[1]["mean_occupancy"] == pytest.approx((8 * 8 + 4 * 10) / (8 * 18.0))
    assert rows[1]["drain_us"] == pytest.approx(10.0)
    assert rows[1]["drain_idle_share"] == pytest.approx(0.5, abs=0.01)
    assert tot["span_us"] == pytest.approx(30.0)
    assert tot["launch_gaps_us"] == pytest.approx(2.0)
    assert tot["slot_utilisation"] == pytest.approx((80 + 64 + 40) / 8 / 30.0)
    assert tot["drain_idle_us"] == pytest.approx(5.0, abs=0.1)


def test_kernel_capacity_differs():
    # the lookahead kernel holds 3 wavefronts per SIMD: 12 on one CU
    r = [(0.0, 5.0, 2, 0, i) for i in range(12)]
    rows, tot = trace.launch_table(_rec(r), cus=1)
    assert rows[0]["kernel"] == "lookahead"
    assert rows[0]["mean_occupancy"] == pytest.approx(1.0)
    assert tot["slot_utilisation"] == pytest.approx(1.0)
