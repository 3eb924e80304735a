"""Failure handling on the device, forced from the host.

Capacity overflows are forced by shrinking the neighbour-list storage or the
PME spread bins (every write stays within the storage; an overflowing build
truncates and raises a status bit).  The step then restores the state its run
of steps started from, grows the storage, rebuilds the lists that state had
and replays, ending on the trajectory of a run that never overflowed.

Fixed-point overflow and non-finite results are forced with two atoms placed
0.01 nm apart (an LJ force beyond the fixed-point range) and two coincident
atoms (a 0/0 displacement); the kernels turn both into status bits, which the
host reports as OMM_ERR_OVERFLOW and OMM_ERR_NONFINITE, the reference's panics
(neighbors.rs:108-111) as error codes.
"""
import numpy as np
import pytest

import ommhip
from helpers import water

pytestmark = pytest.mark.gpu


def _pair(graph_steps, s):
    runs = []
    for _ in range(2):
        c = ommhip.Context(s, temperature=300.0, friction=1.0, seed=5)
        c.set_graph_steps(graph_steps)
        c.set_velocities_to_temperature(300.0, 9)
        runs.append(c)
    return runs


@pytest.mark.parametrize("graph_steps", [10, 0])
def test_list_overflow_replays_bit_for_bit(graph_steps):
    s = water(1000, 3.1, seed=51)
    ref, ctx = _pair(graph_steps, s)
    cap0, _, _ = ctx.capacity()
    ctx.set_capacity(list_cap=64)  # one tile per cluster: every list truncated
    assert ctx.capacity()[0] == 64
    ref.step(30)
    ctx.step(30)
    cap, _, replays = ctx.capacity()
    print(f"graph_steps={graph_steps} capacity {cap0} -> 64 -> {cap}, replays {replays}")
    assert replays >= 1 and cap > 64
    a, b = ref.get_state(), ctx.get_state()
    assert np.array_equal(a["positions"], b["positions"])
    assert np.array_equal(a["velocities"], b["velocities"])
    assert ref.capacity()[2] == 0


def test_bin_overflow_replays_bit_for_bit():
    s = water(1000, 3.1, seed=52)
    ref, ctx = _pair(10, s)
    # The bins belong to the tile-owned spread (the atom-owned one has none).
    t = ctx.launch_tuning()
    ctx.set_launch_tuning(8, t["fft_lines_per_workgroup"], t["fft_plane"], t["pair_scale"])
    ctx.set_capacity(bin_cap=8)
    ref.step(20)
    ctx.step(20)
    _, bins, replays = ctx.capacity()
    print(f"bin capacity 8 -> {bins}, replays {replays}")
    assert replays >= 1 and bins > 8
    a, b = ref.get_state(), ctx.get_state()
    assert np.array_equal(a["positions"], b["positions"])
    assert np.array_equal(a["velocities"], b["velocities"])


def _moved(offset):
    s = water(300, 2.1, seed=53)
    pos = s.positions.copy()
    pos[3:6] += pos[0] + offset - pos[3]  # molecule 1 rigidly onto molecule 0's oxygen + offset
    return s, pos


@pytest.mark.parametrize("precision", [ommhip.PRECISION_MIXED, ommhip.PRECISION_DOUBLE])
def test_fixed_point_overflow_is_reported(precision):
    s, pos = _moved(np.array([0.01, 0.0, 0.0]))
    op = ommhip.GpuNonbondedForce(s.nonbonded, precision)
    op.initialize(s.box)
    if precision == ommhip.PRECISION_MIXED:
        with pytest.raises(ommhip.OmmError) as e:
            op.calc_forces_and_energy(pos, s.box)
        assert e.value.status == ommhip.STATUS_CODES["OVERFLOW"]
    else:  # FP64 accumulation has the range: the force is huge but finite
        F, E = op.calc_forces_and_energy(pos, s.box)
        assert np.isfinite(F).all() and np.abs(F).max() > 1e15
    # The operator stays usable: the next evaluation at sane positions succeeds.
    F, E = op.calc_forces_and_energy(s.positions, s.box)
    assert np.isfinite(F).all() and np.isfinite(E)


def test_nonfinite_result_is_reported():
    s, pos = _moved(np.zeros(3))
    op = ommhip.GpuNonbondedForce(s.nonbonded)
    op.initialize(s.box)
    with pytest.raises(ommhip.OmmError) as e:
        op.calc_forces_and_energy(pos, s.box)
    assert e.value.status == ommhip.STATUS_CODES["NONFINITE"]
    ctx = ommhip.Context(s, temperature=300.0, friction=1.0)
    ctx.set_positions(pos)
    with pytest.raises(ommhip.OmmError) as e:
        ctx.get_state()
    assert e.value.status == ommhip.STATUS_CODES["NONFINITE"]


def test_host_rejects_nonfinite_input():
    s = water(300, 2.1, seed=54)
    ctx = ommhip.Context(s, temperature=300.0, friction=1.0)
    bad = s.positions.copy()
    bad[7, 2] = np.inf
    with pytest.raises(ommhip.OmmError) as e:
        ctx.set_positions(bad)
    assert e.value.status == ommhip.STATUS_CODES["NONFINITE"]
    op = ommhip.GpuNonbondedForce(s.nonbonded)
    op.initialize(s.box)
    with pytest.raises(ommhip.OmmError) as e:
        op.calc_forces_and_energy(bad, s.box)
    assert e.value.status == ommhip.STATUS_CODES["NONFINITE"]
