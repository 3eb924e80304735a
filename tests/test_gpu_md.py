"""The device MD step (LangevinMiddle + bonds + angles + PME) against the oracle."""
import copy

import numpy as np
import pytest

import ommhip
import oracle as orc
from helpers import energy_rel, force_rms_rel, water

pytestmark = pytest.mark.gpu


def test_state_forces_and_energy_match_oracle():
    s = water(1000, 3.1, seed=21)
    ctx = ommhip.Context(s, temperature=0.0, friction=0.0)
    st = ctx.get_state()
    F_ref, E_ref = orc.water_energy_forces(s)
    assert force_rms_rel(st["forces"], F_ref) <= 2e-6
    assert energy_rel(st["potential_energy"], E_ref) <= 1e-6
    assert np.array_equal(st["positions"], s.positions)


def test_zero_temperature_step_is_leapfrog():
    # T = 0 and friction 0 make LFMiddle deterministic (langevin.rs:103-142).
    s = water(1000, 3.1, seed=22)
    rng = np.random.default_rng(1)
    v0 = rng.normal(scale=0.3, size=s.positions.shape)
    ctx = ommhip.Context(s, temperature=0.0, friction=0.0, step_size=0.001)
    ctx.set_velocities(v0)
    ctx.step(1)
    st = ctx.get_state()
    F_ref, _ = orc.water_energy_forces(s)
    x1, v1 = orc.lfmiddle_step(s.positions, v0, F_ref, s.masses, 0.001, 0.0, 0.0)
    assert np.abs(st["velocities"] - v1).max() < 1e-5
    assert np.abs(st["positions"] - x1).max() < 1e-8


def test_graph_steps_match_eager_steps():
    # Graph replays and eager launches run the same kernels with the same
    # list-rebuild decisions (the integrator's skin flag), so the trajectories
    # are identical bit for bit.
    s = water(500, 2.5, seed=23)
    a = ommhip.Context(s, temperature=300.0, friction=1.0, seed=5)
    b = ommhip.Context(s, temperature=300.0, friction=1.0, seed=5)
    a.set_velocities_to_temperature(300.0, 9)
    b.set_velocities_to_temperature(300.0, 9)
    a.step(20)  # two graph replays of 10 steps
    for _ in range(20):
        b.step(1)  # eager launches
    sa, sb = a.get_state(), b.get_state()
    assert np.array_equal(sa["positions"], sb["positions"])
    assert np.array_equal(sa["velocities"], sb["velocities"])


def test_autotuned_launch_shapes():
    # The autotuner picks among exact regroupings: its choices are reported
    # and a context tuned alone matches a batch (tuned for the whole batch)
    # bit for bit.
    s = water(500, 2.5, seed=27)
    c = ommhip.Context(s, temperature=300.0, friction=1.0, seed=5)
    t = c.launch_tuning()
    assert t["spread_tile_edge"] in (0, 6, 8)
    assert t["fft_lines_per_workgroup"] in (2, 4, 8, 16)
    assert t["pair_grid"] >= 64
    assert t["spread_ms"] > 0 and t["fft_ms"] > 0 and t["pair_ms"] > 0
    other = water(1500, 3.6, seed=28)
    b = ommhip.Batch([s, other], temperature=300.0, friction=1.0, seeds=[5, 6])
    print("single", t, "batch", b.launch_tuning())
    c.set_velocities_to_temperature(300.0, 11)
    st = c.get_state()
    b.set_velocities(0, st["velocities"])
    c.step(20)
    b.step(20)
    sa, sb = c.get_state(), b.get_state(0)
    assert np.array_equal(sa["positions"], sb["positions"])
    assert np.array_equal(sa["velocities"], sb["velocities"])


def test_graph_trajectory_is_bit_reproducible():
    s = water(500, 2.5, seed=25)
    runs = []
    for _ in range(2):
        c = ommhip.Context(s, temperature=300.0, friction=1.0, seed=6)
        c.set_velocities_to_temperature(300.0, 10)
        c.step(30)
        runs.append(c.get_state())
    assert np.array_equal(runs[0]["positions"], runs[1]["positions"])
    assert np.array_equal(runs[0]["velocities"], runs[1]["velocities"])
    assert runs[0]["potential_energy"] == runs[1]["potential_energy"]


def test_thermostat_holds_temperature():
    s = water(500, 2.5, seed=24)
    ctx = ommhip.Context(s, temperature=300.0, friction=5.0, step_size=0.001, seed=3)
    ctx.set_velocities_to_temperature(300.0, 4)
    ctx.step(300)
    ke = np.mean([ctx.kinetic_energy() for _ in range(1) if ctx.step(50) is None])
    ndof = 3 * len(s.masses)
    T = 2 * ke / (ndof * orc.boltzmann())
    st = ctx.get_state()
    assert np.isfinite(st["potential_energy"])
    assert 150.0 < T < 600.0


def test_langevin_noise_is_standard_normal():
    # One step from rest at 300 K and at 0 K with the same seed: the velocity
    # difference is the noise term sqrt(kT (1 - a^2) / m) xi alone
    # (langevin.rs:103-142), so xi can be read back and its moments checked.
    s = water(1000, 3.1, seed=29)
    gamma, dt, temp = 1.0, 0.001, 300.0
    hot = ommhip.Context(s, temperature=temp, friction=gamma, step_size=dt, seed=41)
    cold = ommhip.Context(s, temperature=0.0, friction=gamma, step_size=dt, seed=41)
    hot.step(1)
    cold.step(1)
    dv = hot.get_state()["velocities"] - cold.get_state()["velocities"]
    kt = orc.boltzmann() * temp
    scale = np.sqrt(kt * -np.expm1(-2.0 * gamma * dt) / np.asarray(s.masses))[:, None]
    xi = (dv / scale).ravel()
    n = xi.size
    assert abs(xi.mean()) < 4.0 / np.sqrt(n)
    assert abs(xi.var() - 1.0) < 4.0 * np.sqrt(2.0 / n)
    kurt = np.mean((xi - xi.mean()) ** 4) / xi.var() ** 2
    assert abs(kurt - 3.0) < 4.0 * np.sqrt(24.0 / n)
    # Components are uncorrelated.
    x = xi.reshape(-1, 3)
    assert abs(np.corrcoef(x[:, 0], x[:, 1])[0, 1]) < 4.0 / np.sqrt(len(x))


def test_larger_box_multi_chunk_paths():
    # 3,200 waters in 4.6 nm: 13^3 = 2,197 list cells (the sort's scan spans
    # several workgroup chunks) and ~300 list blocks; forces against the
    # oracle, then graph steps against eager steps.
    s = water(3200, 4.6, seed=27)
    a = ommhip.Context(s, temperature=300.0, friction=1.0, seed=8)
    st = a.get_state()
    F_ref, E_ref = orc.water_energy_forces(s)
    assert force_rms_rel(st["forces"], F_ref) <= 2e-6
    assert energy_rel(st["potential_energy"], E_ref) <= 1e-6
    b = ommhip.Context(s, temperature=300.0, friction=1.0, seed=8)
    a.set_velocities_to_temperature(300.0, 12)
    b.set_velocities_to_temperature(300.0, 12)
    a.step(20)
    for _ in range(20):
        b.step(1)
    sa, sb = a.get_state(), b.get_state()
    assert np.array_equal(sa["positions"], sb["positions"])
    assert np.isfinite(sa["potential_energy"])


def test_superblock_search_paths():
    # 12,000 waters in 7.1 nm: 1,125 list blocks, above the 1,024 from which
    # clusters search 64-block superblocks first.  Forces against the oracle,
    # graph steps against eager steps, and a batch against lone contexts: the
    # batch builds superblock bounds for all systems, but the path is chosen
    # per system, so the 300-water replica scans its blocks flat beside the
    # large one's superblock search in the same launch.
    s = water(12000, 7.1, seed=41)
    small = water(300, 2.1, seed=42)
    a = ommhip.Context(s, temperature=300.0, friction=1.0, seed=8)
    st = a.get_state()
    F_ref, E_ref = orc.water_energy_forces(s)
    assert force_rms_rel(st["forces"], F_ref) <= 2e-6
    assert energy_rel(st["potential_energy"], E_ref) <= 1e-6
    batch = ommhip.Batch([s, small], temperature=300.0, friction=1.0, seeds=[8, 9])
    lone_small = ommhip.Context(small, temperature=300.0, friction=1.0, seed=9)
    assert np.array_equal(batch.get_state(0)["forces"], st["forces"])
    assert np.array_equal(batch.get_state(1)["forces"], lone_small.get_state()["forces"])
    b = ommhip.Context(s, temperature=300.0, friction=1.0, seed=8)
    a.set_velocities_to_temperature(300.0, 12)
    b.set_velocities_to_temperature(300.0, 12)
    a.step(20)
    for _ in range(20):
        b.step(1)
    assert np.array_equal(a.get_state()["positions"], b.get_state()["positions"])


@pytest.mark.parametrize("graph_steps", [10, 0])
def test_fast_atom_keeps_the_list_exact(graph_steps):
    # Atoms launched at ~48 nm/ps (0.05 nm per step, about half the list
    # skin) cross the skin/2 rebuild threshold within two steps, in graph
    # replays and in eager steps alike; the list must follow them, so the
    # forces after the run still match the oracle.
    s = water(500, 2.5, seed=31)
    ctx = ommhip.Context(s, temperature=300.0, friction=1.0, seed=7)
    ctx.set_graph_steps(graph_steps)
    ctx.set_velocities_to_temperature(300.0, 8)
    v = ctx.get_state()["velocities"]
    for k in range(5):
        v[3 * 40 * k] = [40.0, -25.0, 10.0]
    ctx.set_velocities(v)
    ctx.step(30)
    st = ctx.get_state()
    moved = copy.deepcopy(s)
    moved.positions = st["positions"]
    F_ref, E_ref = orc.water_energy_forces(moved)
    frel, erel = force_rms_rel(st["forces"], F_ref), energy_rel(st["potential_energy"], E_ref)
    print(f"graph_steps={graph_steps}: force_rms_rel={frel:.3e} energy_rel={erel:.3e}")
    assert frel <= 2e-6
    assert erel <= 1e-6


@pytest.mark.parametrize("method", ["PME", "LJPME"])
def test_odd_graph_replays_match_eager_steps(method):
    # The mesh reads one half of its own slot-order copy while the pair fold
    # writes the list's order into the other; a graph of 7 steps starts each
    # replay on the half its capture did not end on (an older order of the
    # same layout).  Replays, eager steps and the oracle agree.
    s = water(500, 2.5, seed=33, method=method)
    a = ommhip.Context(s, temperature=300.0, friction=1.0, seed=9)
    b = ommhip.Context(s, temperature=300.0, friction=1.0, seed=9)
    a.set_graph_steps(7)
    b.set_graph_steps(0)
    a.set_velocities_to_temperature(300.0, 13)
    b.set_velocities_to_temperature(300.0, 13)
    for n in (7, 7, 3, 14):
        a.step(n)
        b.step(n)
    sa, sb = a.get_state(), b.get_state()
    assert np.array_equal(sa["positions"], sb["positions"])
    assert np.array_equal(sa["velocities"], sb["velocities"])
    assert np.array_equal(sa["forces"], sb["forces"])
    if method == "PME":
        moved = copy.deepcopy(s)
        moved.positions = sa["positions"]
        F_ref, E_ref = orc.water_energy_forces(moved)
        assert force_rms_rel(sa["forces"], F_ref) <= 2e-6
        assert energy_rel(sa["potential_energy"], E_ref) <= 1e-6
