"""Pressure queries on the device (context/pressure.rs) against the
reference's known answer, closed forms and the oracle.

The device evaluates the potential energy of every replica at trial boxes
deformed by +-0.1 % with the molecules moved rigidly (context.rs:942-1029)
and adds the kinetic term of the molecular momenta (pressure.rs:107-131).
Both precisions evaluate these energies on the FP64 energy path (FP64 pair
energies, FP64 mesh on the fixed-point charge grid), so the bars are tight:
water pressures within 1e-3 bar of the oracle (measured: 2e-7 bar FP64,
5e-6 bar mixed, on pressures of ~4000 bar), the Lennard-Jones pair within
1e-9 relative (FP64) or 1e-6 bar absolute (mixed) of its closed forms.
"""
import dataclasses

import numpy as np
import pytest

import ommhip
import oracle as orc
from helpers import LJ_BOX, lj_pair_closed_forms, lj_pair_system, water

pytestmark = pytest.mark.gpu

PRECISIONS = [ommhip.PRECISION_DOUBLE, ommhip.PRECISION_MIXED]
BAR = orc.BAR_NM3_TO_KJ_PER_MOL


def no_terms(k):
    return np.zeros((0, k), np.int32), np.zeros((0, 2))


def reference_barostat_system():
    """tests/barostat.rs:17-41: masses 1, 3, 0, a zero-stiffness bond 0-1 and a
    triclinic box of 60 nm^3; the NonbondedForce (zero charges and depths)
    adds no energy."""
    f = ommhip.NonbondedForce()
    f.set_nonbonded_method("CutoffPeriodic")
    f.set_cutoff_distance(1.0)
    for _ in range(3):
        f.add_particle(0.0, 0.1, 0.0)
    box = np.array([[3.0, 0, 0], [0.4, 4.0, 0], [-0.2, 0.3, 5.0]])
    pos = np.array([[0.2, 0.3, 0.4], [0.9, 0.3, 0.4], [1.4, 1.1, 0.8]])
    angles, angle_params = no_terms(3)
    return ommhip.WaterSystem(f, pos, box, np.array([1.0, 3.0, 0.0]), np.array([[0, 1]], np.int32),
                              np.array([[0.7, 0.0]]), angles, angle_params)


@pytest.mark.parametrize("precision", PRECISIONS)
def test_kinetic_known_answer_and_state_kept(precision):
    """pressure_uses_raw_molecular_momentum_and_does_not_change_state
    (tests/barostat.rs:112-149), at barostat frequency 0."""
    s = reference_barostat_system()
    ctx = ommhip.Context(s, temperature=0.0, friction=0.0, precision=precision)
    with pytest.raises(ommhip.OmmError) as err:
        ctx.compute_current_pressure()
    assert err.value.status == 6
    ctx.set_barostat(0.0, 0.0, 0, seed=47)
    vel = np.array([[1.0, 2.0, 3.0], [-1.0, 0.0, 1.0], [10.0, 20.0, 30.0]])
    ctx.set_velocities(vel)
    ctx.step(3)
    ctx.set_velocities(vel)
    before, box = ctx.get_state(), ctx.box()
    expected = 2.0 * (0.5 * (4.0 + 4.0 + 36.0) / 4.0) / (3.0 * 60.0 * BAR)
    assert ctx.compute_current_pressure() == pytest.approx(expected, rel=1e-14)
    after = ctx.get_state()
    assert np.array_equal(after["positions"], before["positions"])
    assert np.array_equal(after["velocities"], vel)
    assert np.array_equal(ctx.box(), box)
    assert after["step"] == before["step"] == 3
    for query in (ctx.compute_current_anisotropic_pressure, ctx.compute_current_membrane_pressure,
                  ctx.compute_current_flexible_pressure, lambda: ctx.compute_stress_tensor(True)):
        with pytest.raises(ommhip.OmmError) as err:
            query()
        assert err.value.status == 6


def lj_context(precision, kind):
    f, pos, box = lj_pair_system(LJ_BOX)
    angles, angle_params = no_terms(3)
    bonds, bond_params = no_terms(2)
    masses = np.array([2.0, 3.0])
    s = ommhip.WaterSystem(f, pos, box, masses, bonds, bond_params, angles, angle_params)
    ctx = ommhip.Context(s, temperature=0.0, friction=0.0, precision=precision)
    if kind == "iso":
        ctx.set_barostat(1.0, 300.0, 0)
    elif kind == "aniso":
        ctx.set_anisotropic_barostat([1.0, 1.0, 1.0], 300.0, frequency=0)
    elif kind == "membrane":
        ctx.set_membrane_barostat(1.0, 0.0, 300.0, frequency=0)
    else:
        ctx.set_flexible_barostat(1.0, 300.0, 0)
    vel = np.array([[0.3, -0.2, 0.5], [-0.1, 0.4, 0.2]])
    ctx.set_velocities(vel)
    return ctx, pos, vel, masses, box


@pytest.mark.parametrize("precision", PRECISIONS)
def test_lj_pair_closed_forms(precision):
    """Every query on two one-atom molecules in a triclinic box equals the
    central difference of the textbook pair energy of the deformed separation."""
    closed = None
    for kind in ("iso", "aniso", "membrane", "flex"):
        ctx, pos, vel, masses, box = lj_context(precision, kind)
        if closed is None:
            closed = lj_pair_closed_forms(pos, vel, masses, box)
        iso, axis, flex, stress = closed
        double = precision == ommhip.PRECISION_DOUBLE
        tol_iso = 1e-9 * abs(iso) if double else 1e-6
        tol_ax = 1e-9 * np.abs(axis).max() if double else 1e-6
        if kind == "iso":
            got = ctx.compute_current_pressure()
            print(f"iso {got:.12g} closed {iso:.12g}")
            assert abs(got - iso) <= tol_iso
        elif kind in ("aniso", "membrane"):
            got = (ctx.compute_current_anisotropic_pressure() if kind == "aniso"
                   else ctx.compute_current_membrane_pressure())
            print(f"{kind} {got} closed {axis}")
            assert np.abs(got - axis).max() <= tol_ax
        else:
            got = ctx.compute_current_flexible_pressure()
            print(f"flex {got} closed {flex}")
            assert np.abs(got - flex).max() <= tol_ax
            for kinetic in (False, True):
                got = ctx.compute_stress_tensor(kinetic)
                ref = stress if kinetic else lj_pair_closed_forms(pos, 0 * vel, masses, box)[3]
                print(f"stress kinetic={kinetic} {got} closed {ref}")
                assert np.abs(got - ref).max() <= tol_ax
        assert np.array_equal(ctx.get_state()["positions"], pos)
        assert np.array_equal(ctx.box(), box)


def oracle_energy(system, pos, box):
    s = dataclasses.replace(system, positions=np.asarray(pos).reshape(-1, 3), box=np.asarray(box))
    return orc.water_energy_forces(s)[1]


def water_context(kind, precision):
    s = water(500, 2.5, seed=71)
    ctx = ommhip.Context(s, temperature=300.0, friction=1.0, seed=9, precision=precision)
    if kind == "iso":
        ctx.set_barostat(1.0, 300.0, 25, seed=3)
    elif kind == "aniso":
        ctx.set_anisotropic_barostat([1.0, 2.0, 3.0], 300.0, frequency=25, seed=3)
    elif kind == "membrane":
        ctx.set_membrane_barostat(1.0, 10.0, 300.0, 1, 0, frequency=25, seed=3)
    else:
        ctx.set_flexible_barostat(1.0, 300.0, 25, seed=3)
    ctx.set_velocities_to_temperature(300.0, 5)
    return s, ctx


@pytest.mark.parametrize("precision", PRECISIONS)
@pytest.mark.parametrize("kind", ["iso", "aniso", "membrane", "flex"])
def test_water_pressures_match_oracle(kind, precision):
    """500 flexible waters after a short NPT run (a triclinic box for the
    flexible barostat): each query against the oracle's restatement over its
    FP64 energies at the device's positions, velocities and box."""
    s, ctx = water_context(kind, precision)
    ctx.step(50)
    st, box = ctx.get_state(), ctx.box()
    mol, nmol = orc.molecules(len(s.masses), s.bonds)
    args = (lambda q, b: oracle_energy(s, q, b), st["positions"], st["velocities"], s.masses, box, mol, nmol)
    tol = 1e-3  # bar
    if kind == "iso":
        got, ref = ctx.compute_current_pressure(), orc.current_pressure(*args)
    else:
        if kind == "aniso":
            got, ref = ctx.compute_current_anisotropic_pressure(), orc.axis_pressures(*args)
        elif kind == "membrane":
            got, ref = ctx.compute_current_membrane_pressure(), orc.axis_pressures(*args)
        else:
            got, ref = ctx.compute_current_flexible_pressure(), orc.flexible_pressure(*args)
    print(f"{kind}: device {got} oracle {ref} max diff {np.max(np.abs(np.subtract(got, ref))):.3g} "
          f"(tolerance {tol:.3g} bar)")
    assert np.max(np.abs(np.subtract(got, ref))) <= tol
    if kind == "flex":
        got = ctx.compute_stress_tensor(True)
        ref = orc.stress_tensor(*args, True)
        print(f"stress: device {got} oracle {ref} max diff {np.abs(got - ref).max():.3g}")
        assert np.abs(got - ref).max() <= tol
    after = ctx.get_state()
    assert np.array_equal(after["positions"], st["positions"])
    assert np.array_equal(after["velocities"], st["velocities"])
    assert np.array_equal(ctx.box(), box)
    assert after["potential_energy"] == pytest.approx(st["potential_energy"], rel=1e-9)


@pytest.mark.parametrize("kind", ["iso", "flex"])
def test_batch_pressures_equal_single_contexts(kind):
    """The same trial evaluations serve every replica: replica r's pressure is
    that of its own context."""
    reps = [(216, 1.86), (350, 2.19), (500, 2.465)]
    sys_ = [water(m, e, seed=300 + k) for k, (m, e) in enumerate(reps)]
    batch = ommhip.Batch(sys_, temperature=300.0, friction=1.0, seeds=[5, 6, 7])
    if kind == "iso":
        batch.set_barostat(1.0, 300.0, 0)
    else:
        batch.set_flexible_barostat(1.0, 300.0, 0)
    batch.set_velocities_to_temperature(300.0, 8)
    got = batch.compute_current_pressure() if kind == "iso" else batch.compute_stress_tensor(True)
    for r, s in enumerate(sys_):
        st = batch.get_state(r)
        ctx = ommhip.Context(dataclasses.replace(s, positions=st["positions"]), temperature=300.0, friction=1.0)
        ctx.set_velocities(st["velocities"])
        if kind == "iso":
            ctx.set_barostat(1.0, 300.0, 0)
            one = ctx.compute_current_pressure()
        else:
            ctx.set_flexible_barostat(1.0, 300.0, 0)
            one = ctx.compute_stress_tensor(True)
        print(f"replica {r}: batch {got[r]} single {one}")
        assert np.array_equal(got[r], one)
