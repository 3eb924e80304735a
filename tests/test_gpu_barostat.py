"""Monte Carlo barostat on the device (barostat.rs:129-218, program.rs:5-45).

One attempt at a time is checked against the oracle: the energies before and
after the move (same frozen alpha and grid, FP64 oracle), the rigid-molecule
scaling of the positions and box (context.rs:976-1029), the work of the
Metropolis test (barostat.rs:160-163), and the outcome (accepted moves keep
the scaled state, rejected ones restore the old one bit for bit).  The
random stream itself is the device's own (Philox), so which moves are drawn
is not compared with the reference.
"""
import dataclasses

import numpy as np
import pytest

import ommhip
import oracle as orc
from helpers import water

pytestmark = pytest.mark.gpu


def oracle_energy(system, pos, box):
    s = dataclasses.replace(system, positions=pos, box=box)
    return orc.water_energy_forces(s)[1]


@pytest.mark.parametrize("pressure", [1.0, 2.0e4, -2.0e4])
def test_attempts_match_oracle(pressure):
    s = water(500, 2.5, seed=31)
    ctx = ommhip.Context(s, temperature=300.0, friction=1.0, seed=3)
    ctx.set_barostat(pressure, 300.0, 25, seed=77)
    mol, nmol = orc.molecules(len(s.masses), s.bonds)
    outcomes = set()
    pos0, box0 = s.positions.copy(), s.box.copy()
    for attempt in range(6):
        ctx.barostat_attempt()
        info = ctx.barostat_info()
        st = ctx.get_state()
        box1 = ctx.box()
        v0 = box0[0, 0] * box0[1, 1] * box0[2, 2]
        assert info["volume_old"] == pytest.approx(v0, rel=1e-14)
        scale = np.cbrt(info["volume_new"] / info["volume_old"])
        pos_scaled = orc.scale_positions(pos0, scale, True, mol, nmol)
        box_scaled = box0 * scale
        e_old = oracle_energy(s, pos0, box0)
        e_new = oracle_energy(s, pos_scaled, box_scaled)
        print(f"P={pressure} attempt {attempt}: dV={info['volume_new'] - info['volume_old']:+.4f} "
              f"dE={info['energy_new'] - info['energy_old']:+.4f} (oracle {e_new - e_old:+.4f}) "
              f"w={info['work']:+.3f} accepted={info['accepted_last']}")
        assert abs(info["energy_old"] - e_old) <= 1e-6 * abs(e_old)
        assert abs(info["energy_new"] - e_new) <= 1e-6 * abs(e_new)
        assert abs((info["energy_new"] - info["energy_old"]) - (e_new - e_old)) < 2e-3
        w_ref = orc.barostat_work(info["energy_new"] - info["energy_old"], pressure, info["volume_old"],
                                  info["volume_new"], nmol, 300.0)
        assert info["work"] == pytest.approx(w_ref, rel=1e-12, abs=1e-9)
        if info["work"] <= 0.0:
            assert info["accepted_last"] == 1
        if info["accepted_last"]:
            assert np.abs(st["positions"] - pos_scaled).max() < 1e-12
            assert np.allclose(box1, box_scaled, rtol=1e-15, atol=0)
        else:
            assert np.array_equal(st["positions"], pos0)
            assert np.array_equal(box1, box0)
        outcomes.add(info["accepted_last"])
        pos0, box0 = st["positions"], box1
    assert info["total_attempted"] == 6
    if pressure != 1.0:
        assert outcomes == {0, 1}  # a strong pressure accepts one direction and rejects the other


def test_batch_barostat_matches_single_contexts():
    reps = [(216, 1.86), (350, 2.19), (500, 2.465)]
    sys_ = [water(m, e, seed=200 + k) for k, (m, e) in enumerate(reps)]
    seeds = [31, 32, 33]
    batch = ommhip.Batch(sys_, temperature=300.0, friction=1.0, seeds=seeds)
    batch.set_barostat(1.0, 300.0, 5, seed=50)
    batch.set_velocities_to_temperature(300.0, 60)
    batch.step(27)  # five attempts (the graph spans one barostat period)
    for r, s in enumerate(sys_):
        single = ommhip.Context(s, temperature=300.0, friction=1.0, seed=seeds[r])
        single.set_barostat(1.0, 300.0, 5, seed=50 + r)
        single.set_velocities_to_temperature(300.0, 60 + r)
        single.step(27)
        a, b = batch.get_state(r), single.get_state()
        ia, ib = batch.barostat_info(r), single.barostat_info()
        print(f"replica {r}: accepted {ia['total_accepted']}/{ia['total_attempted']}, box {batch.box(r)[0, 0]:.6f}")
        assert ia["total_attempted"] == ib["total_attempted"] == 5
        assert ia["total_accepted"] == ib["total_accepted"]
        assert np.array_equal(batch.box(r), single.box())
        assert np.array_equal(a["positions"], b["positions"])


def test_npt_run_stays_physical():
    s = water(500, 2.5, seed=41)
    ctx = ommhip.Context(s, temperature=300.0, friction=5.0, seed=9)
    ctx.set_barostat(1.0, 300.0, 10, seed=5)
    ctx.set_velocities_to_temperature(300.0, 11)
    ctx.step(200)
    info = ctx.barostat_info()
    box = ctx.box()
    st = ctx.get_state()
    assert info["total_attempted"] == 20
    assert np.isfinite(st["potential_energy"])
    v0 = 2.5 ** 3
    v = box[0, 0] * box[1, 1] * box[2, 2]
    assert 0.8 * v0 < v < 1.2 * v0
    assert box[0, 0] == box[1, 1] == box[2, 2]  # isotropic


@pytest.mark.parametrize("pressure,axes", [((1.0, 1.0, 1.0), (True, True, True)),
                                           ((2.0e4, -2.0e4, 1.0), (True, True, False))])
def test_anisotropic_attempts_match_oracle(pressure, axes):
    """MonteCarloAnisotropicBarostat (anisotropic_barostat.rs:122-211): each
    attempt scales one enabled axis by V'/V; energies, rigid scaling, box,
    work (with that axis' pressure) and outcome against the oracle."""
    s = water(500, 2.5, seed=32)
    ctx = ommhip.Context(s, temperature=300.0, friction=1.0, seed=4)
    ctx.set_anisotropic_barostat(pressure, 300.0, axes, 25, seed=78)
    mol, nmol = orc.molecules(len(s.masses), s.bonds)
    pos0, box0 = s.positions.copy(), s.box.copy()
    seen = set()
    for attempt in range(8):
        ctx.barostat_attempt()
        info = ctx.barostat_info()
        st = ctx.get_state()
        box1 = ctx.box()
        a = info["axis"]
        assert axes[a]
        seen.add(a)
        scale = np.ones(3)
        scale[a] = info["volume_new"] / info["volume_old"]
        pos_scaled = orc.scale_positions_axes(pos0, scale, mol, nmol).reshape(-1, 3)
        box_scaled = box0 * scale[None, :]
        e_old = oracle_energy(s, pos0, box0)
        e_new = oracle_energy(s, pos_scaled, box_scaled)
        print(f"P={pressure} attempt {attempt}: axis {a} dV={info['volume_new'] - info['volume_old']:+.4f} "
              f"w={info['work']:+.3f} accepted={info['accepted_last']}")
        assert abs(info["energy_old"] - e_old) <= 1e-6 * abs(e_old)
        assert abs(info["energy_new"] - e_new) <= 1e-6 * abs(e_new)
        w_ref = orc.barostat_work(info["energy_new"] - info["energy_old"], pressure[a], info["volume_old"],
                                  info["volume_new"], nmol, 300.0)
        assert info["work"] == pytest.approx(w_ref, rel=1e-12, abs=1e-9)
        if info["accepted_last"]:
            assert np.abs(st["positions"] - pos_scaled).max() < 1e-12
            assert np.allclose(box1, box_scaled, rtol=1e-15, atol=0)
        else:
            assert np.array_equal(st["positions"], pos0)
            assert np.array_equal(box1, box0)
        pos0, box0 = st["positions"], box1
    assert info["total_attempted"] == 8
    assert seen == {k for k in range(3) if axes[k]}


def test_anisotropic_npt_keeps_disabled_axis_and_batch_matches():
    reps = [(216, 1.86), (350, 2.19)]
    sys_ = [water(m, e, seed=210 + k) for k, (m, e) in enumerate(reps)]
    batch = ommhip.Batch(sys_, temperature=300.0, friction=1.0, seeds=[41, 42])
    batch.set_anisotropic_barostat((1.0, 1.0, 1.0), 300.0, (True, False, True), 5, seed=60)
    batch.set_velocities_to_temperature(300.0, 70)
    batch.step(25)
    for r, s in enumerate(sys_):
        single = ommhip.Context(s, temperature=300.0, friction=1.0, seed=41 + r)
        single.set_anisotropic_barostat((1.0, 1.0, 1.0), 300.0, (True, False, True), 5, seed=60 + r)
        single.set_velocities_to_temperature(300.0, 70 + r)
        single.step(25)
        assert batch.box(r)[1, 1] == s.box[1, 1]  # y never scaled
        assert np.array_equal(batch.box(r), single.box())
        assert np.array_equal(batch.get_state(r)["positions"], single.get_state()["positions"])
        assert batch.barostat_info(r)["total_attempted"] == 5


@pytest.mark.parametrize("xy_mode,z_mode,tension", [(0, 0, 0.0), (1, 2, 500.0), (1, 1, -300.0)])
def test_membrane_attempts_match_oracle(xy_mode, z_mode, tension):
    """MonteCarloMembraneBarostat (membrane_barostat.rs:167-292): coupled or
    independent x/y, z free / fixed / at constant volume, and the surface
    tension term of the work; each attempt against the oracle."""
    s = water(500, 2.5, seed=33)
    ctx = ommhip.Context(s, temperature=300.0, friction=1.0, seed=5)
    ctx.set_membrane_barostat(1.0, tension, 300.0, xy_mode, z_mode, 25, seed=79)
    mol, nmol = orc.molecules(len(s.masses), s.bonds)
    pos0, box0 = s.positions.copy(), s.box.copy()
    axes = set()
    for attempt in range(8):
        ctx.barostat_attempt()
        info = ctx.barostat_info()
        st = ctx.get_state()
        box1 = ctx.box()
        a, sc = info["axis"], np.array(info["scale"])
        axes.add(a)
        if xy_mode == 0 and a < 2:
            assert a == 0 and sc[0] == sc[1]
        if z_mode == 1:
            assert a < 2 and sc[2] == 1.0
        if z_mode == 2:
            assert a < 2 and sc[2] == 1.0 / (sc[0] * sc[1]) and info["volume_new"] == info["volume_old"]
        area0 = box0[0, 0] * box0[1, 1]
        assert info["area_change"] == pytest.approx(area0 * (sc[0] * sc[1] - 1.0), rel=1e-12, abs=1e-15)
        pos_scaled = orc.scale_positions_axes(pos0, sc, mol, nmol).reshape(-1, 3)
        box_scaled = box0 * sc[None, :]
        e_old = oracle_energy(s, pos0, box0)
        e_new = oracle_energy(s, pos_scaled, box_scaled)
        assert abs(info["energy_old"] - e_old) <= 1e-6 * abs(e_old)
        assert abs(info["energy_new"] - e_new) <= 1e-6 * abs(e_new)
        w_ref = orc.barostat_work(info["energy_new"] - info["energy_old"], 1.0, info["volume_old"],
                                  info["volume_new"], nmol, 300.0) - tension * 0.0602214076 * info["area_change"]
        print(f"modes {xy_mode}/{z_mode}: axis {a} scale {sc} w={info['work']:+.3f} accepted={info['accepted_last']}")
        assert info["work"] == pytest.approx(w_ref, rel=1e-12, abs=1e-9)
        if info["accepted_last"]:
            assert np.abs(st["positions"] - pos_scaled).max() < 1e-12
            assert np.allclose(box1, box_scaled, rtol=1e-15, atol=0)
        else:
            assert np.array_equal(st["positions"], pos0)
            assert np.array_equal(box1, box0)
        pos0, box0 = st["positions"], box1
    assert info["total_attempted"] == 8
    if z_mode == 0 and xy_mode == 0:
        assert axes <= {0, 2}


def test_flexible_attempts_match_oracle():
    """MonteCarloFlexibleBarostat (flexible_barostat.rs:132-238): each lower-
    triangle box element moves by up to the length step, the trial box is
    reduced, molecules scale by the diagonal ratios, and the work carries
    (N - 2) kT ln(V'/V) and kT (2 ln s_x + ln s_y)."""
    s = water(500, 2.5, seed=34)
    ctx = ommhip.Context(s, temperature=300.0, friction=1.0, seed=6)
    ctx.set_flexible_barostat(1.0, 300.0, 25, seed=80)
    mol, nmol = orc.molecules(len(s.masses), s.bonds)
    pos0, box0 = s.positions.copy(), s.box.copy()
    kt = orc.boltzmann() * 300.0
    outcomes = set()
    step = ctx.barostat_info()["volume_scale"]
    assert step == pytest.approx(0.01 * np.cbrt(np.prod(np.diag(box0))), rel=1e-14)
    for attempt in range(8):
        ctx.barostat_attempt()
        info = ctx.barostat_info()
        st = ctx.get_state()
        box1 = ctx.box()
        bn = info["box_new"]
        lower = np.tril(np.ones((3, 3), bool))
        assert np.all(bn[~lower] == 0.0)
        assert np.abs(bn - box0)[lower].max() <= step * (1 + 1e-12)
        assert np.array_equal(orc.reduced_box(bn), bn)
        sc = np.diag(bn) / np.diag(box0)
        assert np.allclose(info["scale"], sc, rtol=1e-15, atol=0)
        assert info["log_shape"] == pytest.approx(2 * np.log(sc[0]) + np.log(sc[1]), rel=1e-12, abs=1e-15)
        assert info["volume_new"] == pytest.approx(np.prod(np.diag(bn)), rel=1e-15)
        pos_scaled = orc.scale_positions_axes(pos0, sc, mol, nmol).reshape(-1, 3)
        e_old = oracle_energy(s, pos0, box0)
        e_new = oracle_energy(s, pos_scaled, bn)
        assert abs(info["energy_old"] - e_old) <= 1e-6 * abs(e_old)
        assert abs(info["energy_new"] - e_new) <= 1e-6 * abs(e_new)
        de = info["energy_new"] - info["energy_old"]
        w_ref = (de + 1.0 * 0.0602214076 * (info["volume_new"] - info["volume_old"])
                 - (nmol - 2) * kt * np.log(info["volume_new"] / info["volume_old"]) - kt * info["log_shape"])
        print(f"attempt {attempt}: box_new diag {np.diag(bn)} tilt {bn[1, 0]:+.4f} w={info['work']:+.3f} "
              f"accepted={info['accepted_last']}")
        assert info["work"] == pytest.approx(w_ref, rel=1e-12, abs=1e-9)
        if info["accepted_last"]:
            assert np.abs(st["positions"] - pos_scaled).max() < 1e-12
            assert np.array_equal(box1, bn)
        else:
            assert np.array_equal(st["positions"], pos0)
            assert np.array_equal(box1, box0)
        outcomes.add(info["accepted_last"])
        pos0, box0 = st["positions"], box1
        step = info["volume_scale"]
    assert info["total_attempted"] == 8


def test_flexible_npt_run_stays_physical():
    s = water(500, 2.5, seed=42)
    ctx = ommhip.Context(s, temperature=300.0, friction=5.0, seed=10)
    ctx.set_flexible_barostat(1.0, 300.0, 10, seed=6)
    ctx.set_velocities_to_temperature(300.0, 12)
    ctx.step(200)
    box = ctx.box()
    st = ctx.get_state()
    assert ctx.barostat_info()["total_attempted"] == 20
    assert np.isfinite(st["potential_energy"])
    assert np.array_equal(orc.reduced_box(box), box)
    assert 0.8 * 2.5 ** 3 < np.prod(np.diag(box)) < 1.2 * 2.5 ** 3
