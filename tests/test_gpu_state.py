"""Context::set_state (context.rs:274-293) on the device: a context restored
from a saved state continues the saved trajectory."""
import numpy as np
import pytest

import ommhip
import oracle as orc
from helpers import energy_rel, force_rms_rel, water

pytestmark = pytest.mark.gpu


def _run(s, steps, graph_steps, state=None, seed=5):
    ctx = ommhip.Context(s, temperature=300.0, friction=1.0, seed=seed)
    ctx.set_graph_steps(graph_steps)
    if state is None:
        ctx.set_velocities_to_temperature(300.0, 9)
    else:
        ctx.set_state(state)
    ctx.step(steps)
    return ctx


@pytest.mark.parametrize("graph_steps", [10, 0])
def test_restored_context_continues_bit_for_bit(graph_steps):
    # 40 steps in one context against 20 steps, a checkpoint (get_state with
    # the list positions), a new context restored from it and 20 more steps.
    s = water(1000, 3.1, seed=31)
    whole = _run(s, 40, graph_steps).get_state()
    first = _run(s, 20, graph_steps)
    saved = first.get_state()
    assert saved["step"] == 20
    resumed = _run(s, 20, graph_steps, state=saved)
    end = resumed.get_state()
    assert end["step"] == 40
    assert np.array_equal(end["positions"], whole["positions"])
    assert np.array_equal(end["velocities"], whole["velocities"])
    assert end["potential_energy"] == whole["potential_energy"]


def test_restore_without_list_positions_stays_close():
    # Without the list positions the list is built from the restored
    # positions: FP32 pair sums are taken around other block centres, so the
    # trajectory differs in the last bits only.
    s = water(1000, 3.1, seed=32)
    whole = _run(s, 40, 10).get_state()
    saved = _run(s, 20, 10).get_state()
    saved.pop("list_positions")
    end = _run(s, 20, 10, state=saved).get_state()
    assert end["step"] == 40
    assert np.abs(end["positions"] - whole["positions"]).max() < 1e-5


def test_batch_replica_restore_bit_for_bit():
    a, b = water(500, 2.5, seed=33), water(800, 2.9, seed=34)
    whole = ommhip.Batch([a, b], temperature=300.0, friction=1.0, seeds=[3, 4])
    whole.set_velocities_to_temperature(300.0, 21)
    whole.step(30)
    half = ommhip.Batch([a, b], temperature=300.0, friction=1.0, seeds=[3, 4])
    half.set_velocities_to_temperature(300.0, 21)
    half.step(15)
    saved = [half.get_state(r) for r in range(2)]
    resumed = ommhip.Batch([a, b], temperature=300.0, friction=1.0, seeds=[3, 4])
    for r in range(2):
        resumed.set_state(r, saved[r])
    resumed.step(15)
    for r in range(2):
        x, y = resumed.get_state(r), whole.get_state(r)
        assert x["step"] == 30
        assert np.array_equal(x["positions"], y["positions"])
        assert np.array_equal(x["velocities"], y["velocities"])


def test_set_state_box_matches_oracle():
    # A restored box (molecules scaled with it) is used by every term; alpha
    # and the grid stay frozen at their initial values (nonbonded.rs:1055-1107).
    s = water(1000, 3.1, seed=35)
    ctx = ommhip.Context(s, temperature=0.0, friction=0.0)
    scale = 1.02
    box = s.box * scale
    cent = s.positions.reshape(-1, 3, 3).mean(axis=1, keepdims=True)
    pos = (s.positions.reshape(-1, 3, 3) - cent + cent * scale).reshape(-1, 3)
    ctx.set_state(dict(step=7, box=box, positions=pos))
    st = ctx.get_state()
    assert st["step"] == 7
    assert np.array_equal(st["box"], box)
    s2 = water(1000, 3.1, seed=35)
    s2.positions = pos
    s2.box = box
    F_ref, E_ref = orc.water_energy_forces(s2)
    assert force_rms_rel(st["forces"], F_ref) <= 2e-6
    assert energy_rel(st["potential_energy"], E_ref) <= 1e-6


def test_set_state_rejects_bad_input():
    s = water(300, 2.1, seed=36)
    ctx = ommhip.Context(s, temperature=300.0, friction=1.0)
    bad = s.positions.copy()
    bad[3, 1] = np.nan
    with pytest.raises(ommhip.OmmError):
        ctx.set_state(dict(positions=bad))
    with pytest.raises(ommhip.OmmError):
        ctx.set_state(dict(step=-1))
    with pytest.raises(ommhip.OmmError):  # cutoff 0.9 nm > half of 1.5 nm
        ctx.set_state(dict(box=np.eye(3) * 1.5))
    with pytest.raises(ommhip.OmmError):
        ctx.set_state(dict(list_positions=s.positions))
