"""Distance constraints in the device MD step (SHAKE/RATTLE inside
LangevinMiddle, langevin.rs:215-277; constraints.rs) against the oracle."""
import copy
import math

import numpy as np
import pytest

import ommhip
import oracle as orc
from helpers import water

pytestmark = pytest.mark.gpu

R_OH = 0.09572
R_HH = 2 * R_OH * math.sin(math.radians(104.52) / 2)


def rigid(system):
    """The water box with each molecule held by three constraints instead of
    its bonds and angle."""
    s = copy.deepcopy(system)
    n = len(s.masses) // 3
    s.constraints = [c for m in range(n) for c in ((3 * m, 3 * m + 1, R_OH), (3 * m, 3 * m + 2, R_OH),
                                                    (3 * m + 1, 3 * m + 2, R_HH))]
    s.bonds = np.zeros((0, 2), np.int32)
    s.bond_params = np.zeros((0, 2))
    s.angles = np.zeros((0, 3), np.int32)
    s.angle_params = np.zeros((0, 2))
    return s


def two_particles():
    f = ommhip.NonbondedForce()
    f.set_nonbonded_method("CutoffPeriodic")
    f.set_cutoff_distance(1.0)
    for _ in range(2):
        f.add_particle(0.0, 0.1, 0.0)
    pos = np.array([[0.0, 0, 0], [1.2, 0, 0]])
    s = ommhip.WaterSystem(f, pos, np.diag([5.0] * 3), np.array([1.0, 2.0]), np.zeros((0, 2), np.int32),
                           np.zeros((0, 2)), np.zeros((0, 3), np.int32), np.zeros((0, 2)))
    s.constraints = [(0, 1, 1.0)]
    return s


@pytest.mark.parametrize("precision", [ommhip.PRECISION_MIXED, ommhip.PRECISION_DOUBLE])
def test_projection_known_answer(precision):
    # tests/constraints.rs:15-38
    s = two_particles()
    ctx = ommhip.Context(s, temperature=0.0, friction=0.0, precision=precision)
    ctx.set_velocities(np.array([[1.0, 2.0, 0.0], [-1.0, -1.0, 0.0]]))
    ctx.apply_constraints(1e-6)
    ctx.apply_velocity_constraints(1e-6)
    st = ctx.get_state()
    p, v = st["positions"], st["velocities"]
    assert p[1, 0] - p[0, 0] == pytest.approx(1.0, abs=2e-6)
    assert p[0, 0] + 2.0 * p[1, 0] == pytest.approx(2.4, abs=2e-6)
    assert v[1, 0] - v[0, 0] == pytest.approx(0.0, abs=2e-6)
    assert v[0, 0] + 2.0 * v[1, 0] == pytest.approx(-1.0, abs=2e-6)
    assert v[0, 1] == 2.0 and v[1, 1] == -1.0
    # The same sweeps as the oracle's solver.
    p_ref = orc.constraints_solve(s.constraints, s.masses, s.positions, s.positions, False, 1e-6).reshape(-1, 3)
    assert np.abs(p - p_ref).max() < 1e-12


def test_singular_projection_is_an_error():
    # tests/constraints.rs:77-97: coincident particles.
    s = two_particles()
    s.positions = np.zeros((2, 3))
    ctx = ommhip.Context(s, temperature=0.0, friction=0.0)
    with pytest.raises(ommhip.OmmError):
        ctx.apply_constraints(1e-5)


def test_massless_partner_is_rejected():
    # tests/constraints.rs:100-110
    s = two_particles()
    s.masses = np.array([1.0, 0.0])
    with pytest.raises(ommhip.OmmError):
        ommhip.Context(s, temperature=0.0, friction=0.0)


@pytest.mark.parametrize("precision,tol", [(ommhip.PRECISION_DOUBLE, 1e-11), (ommhip.PRECISION_MIXED, 1e-9)])
def test_rigid_water_step_matches_oracle(precision, tol):
    # T = 0, no friction: the constrained step is deterministic; one step from
    # random velocities against the oracle's constrained LFMiddle with the
    # oracle's forces.
    s = rigid(water(500, 2.5, seed=71))
    rng = np.random.default_rng(4)
    v0 = rng.normal(scale=0.5, size=s.positions.shape)
    ctx = ommhip.Context(s, temperature=0.0, friction=0.0, step_size=0.001, precision=precision)
    ctx.set_velocities(v0)
    ctx.step(1)
    st = ctx.get_state()
    F, _ = orc.water_energy_forces(s)
    x1, v1 = orc.lfmiddle_step_constrained(s.positions, v0, F, s.masses, 0.001, 0.0, 0.0, s.constraints, 1e-5)
    dx = np.abs(st["positions"] - x1.reshape(-1, 3)).max()
    dv = np.abs(st["velocities"] - v1.reshape(-1, 3)).max()
    print(f"max |dx| = {dx:.3e} nm, max |dv| = {dv:.3e} nm/ps")
    assert dx < tol
    assert dv < 1e3 * tol


def test_rigid_water_nvt_keeps_lengths_and_temperature():
    # tests/constraints.rs:41-74 with 500 waters: graph-replayed steps at
    # 300 K keep every constraint within 2e-5 relative; after the lattice
    # start has relaxed (400 steps at 10/ps), the kinetic temperature over
    # 3N - constraints degrees of freedom averages near 300 K.
    s = rigid(water(500, 2.5, seed=72))
    ctx = ommhip.Context(s, temperature=300.0, friction=10.0, step_size=0.001, seed=9)
    ctx.set_velocities_to_temperature(300.0, 10)
    ctx.apply_velocity_constraints(1e-5)
    c = np.array(s.constraints)
    i, j, d = c[:, 0].astype(int), c[:, 1].astype(int), c[:, 2]
    ndof = 3 * len(s.masses) - len(s.constraints)
    ctx.step(400)
    temps, worst = [], 0.0
    for _ in range(10):
        ctx.step(20)
        p = ctx.get_state()["positions"]
        worst = max(worst, np.abs(np.linalg.norm(p[i] - p[j], axis=1) / d - 1.0).max())
        temps.append(2 * ctx.kinetic_energy() / (ndof * orc.boltzmann()))
    print(f"max relative constraint error {worst:.2e}, T = {np.mean(temps):.1f} K")
    assert worst <= 2e-5
    assert 250.0 < np.mean(temps) < 350.0


def test_constrained_batch_matches_single_contexts():
    systems = [rigid(water(m, e, seed=73 + k)) for k, (m, e) in enumerate([(300, 2.1), (500, 2.5)])]
    b = ommhip.Batch(systems, temperature=300.0, friction=1.0, seeds=[5, 6])
    b.set_velocities_to_temperature(300.0, 20)
    b.step(30)
    for r, s in enumerate(systems):
        c = ommhip.Context(s, temperature=300.0, friction=1.0, seed=5 + r)
        c.set_velocities_to_temperature(300.0, 20 + r)
        c.step(30)
        assert np.array_equal(b.get_state(r)["positions"], c.get_state()["positions"])
