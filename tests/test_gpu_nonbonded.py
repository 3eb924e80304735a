"""Parity of the gfx950 NonbondedForce/PME operator with the FP64 oracle.

Bars (mixed precision, against the FP64 oracle on the same PME grid and
alpha): per-atom force RMS relative error <= 2e-6, total-energy relative error
<= 1e-6, direct-space pair set bit-exact (count and order-independent hash).
The plain FP64 form must agree to 1e-9.
"""
import math

import numpy as np
import pytest

import ommhip
import oracle as orc
from helpers import energy_rel, force_max_atom_rel, force_rms_rel, water

pytestmark = pytest.mark.gpu

FORCE_TOL = 2e-6
ENERGY_TOL = 1e-6


def device_op(system, precision=ommhip.PRECISION_MIXED):
    op = ommhip.GpuNonbondedForce(system.nonbonded, precision)
    op.initialize(system.box)
    return op


def check_against_oracle(system, precision=ommhip.PRECISION_MIXED, ftol=FORCE_TOL, etol=ENERGY_TOL):
    desc = system.nonbonded.desc()
    F_ref, E_ref, parts_ref = orc.nb_evaluate(desc, system.positions, system.box)
    op = device_op(system, precision)
    F, E = op.calc_forces_and_energy(system.positions, system.box)
    parts = op.energy_parts(system.positions, system.box)
    frel = force_rms_rel(F, F_ref)
    erel = energy_rel(E, E_ref)
    print(f"N={len(F)} force_rms_rel={frel:.3e} force_max_atom_rel={force_max_atom_rel(F, F_ref):.3e} "
          f"energy_rel={erel:.3e} E={E:.6f} E_ref={E_ref:.6f}")
    print("parts", parts, "ref", parts_ref)
    assert frel <= ftol
    assert erel <= etol
    # The force-only path (what the MD step uses) meets the same force bar.
    F_only = op.calc_forces(system.positions, system.box)
    frel_only = force_rms_rel(F_only, F_ref)
    print(f"force-only path force_rms_rel={frel_only:.3e}")
    assert frel_only <= ftol
    # Every energy term individually, relative to the largest term.
    scale = np.abs(parts_ref).max()
    assert np.abs(parts - parts_ref).max() <= etol * scale
    count_ref, hash_ref = orc.nb_pairs(desc, system.positions, system.box)
    count, h = op.pair_stats(system.positions, system.box)
    assert count == count_ref
    assert h == hash_ref
    return op


@pytest.mark.parametrize("molecules,edge", [(300, 2.1), (1000, 3.1), (3000, 4.48)])
def test_pme_water_parity(molecules, edge):
    check_against_oracle(water(molecules, edge, seed=molecules))


def test_pme_water_parity_plain_fp64():
    check_against_oracle(water(1000, 3.1, seed=3), ommhip.PRECISION_DOUBLE, ftol=1e-9, etol=1e-9)


def test_pair_set_on_cutoff_lattice():
    # Lattice spacing a with cutoff 3a: many pairs sit exactly on the cutoff,
    # so the set is decided by the last bits of the FP64 r^2 comparison.
    a = 0.3
    m = 8
    g = np.arange(m) * a
    pos = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3) + 0.013
    n = len(pos)
    f = ommhip.NonbondedForce()
    f.set_nonbonded_method("PME")
    f.set_cutoff_distance(3 * a)
    for i in range(n):
        f.add_particle(0.5 if (i % 2) else -0.5, 0.3, 0.2)
    for i in range(0, n - 1, 2):
        f.add_exception(i, i + 1, 0.0, 0.3, 0.0)
    box = np.diag([m * a] * 3)
    alpha, grid = orc.effective_pme_parameters(5e-4, 3 * a, box)
    f.set_pme_parameters(alpha, *grid)
    sys_ = ommhip.WaterSystem(f, pos, box, np.ones(n), np.zeros((0, 2), np.int32), np.zeros((0, 2)),
                              np.zeros((0, 3), np.int32), np.zeros((0, 2)))
    check_against_oracle(sys_)


def test_triclinic_box_parity():
    s = water(600, 2.6, seed=11)
    s.box = np.array([[2.6, 0, 0], [0.5, 2.6, 0], [-0.4, 0.6, 2.6]])
    alpha, grid = orc.effective_pme_parameters(5e-4, 0.9, s.box)
    s.nonbonded.set_pme_parameters(alpha, *grid)
    check_against_oracle(s)


def test_reaction_field_parity():
    s = water(1000, 3.1, seed=5, method="CutoffPeriodic")
    check_against_oracle(s)


@pytest.mark.parametrize("method", ["PME", "CutoffPeriodic"])
def test_switching_function_parity(method):
    """LJ switching from 0.75 nm to the 0.9 nm cutoff (nonbonded.rs:887-899),
    including the switched dispersion correction."""
    s = water(1000, 3.1, seed=13, method=method)
    s.nonbonded.set_use_switching_function(True)
    s.nonbonded.set_switching_distance(0.75)
    check_against_oracle(s)


def test_groups_split_and_none():
    s = water(500, 2.5, seed=2)
    s.nonbonded.set_force_group(1)
    s.nonbonded.set_reciprocal_space_force_group(2)
    op = device_op(s)
    F, E = op.calc_forces_and_energy(s.positions, s.box)
    Fd, Ed = op.calc_forces_and_energy_groups(s.positions, s.box, 1 << 1)
    Fr, Er = op.calc_forces_and_energy_groups(s.positions, s.box, 1 << 2)
    assert op.calc_forces_and_energy_groups(s.positions, s.box, 1 << 5) is None
    assert abs(Ed + Er - E) <= 1e-9 * abs(E) + 1e-9
    assert np.abs(Fd + Fr - F).max() <= 1e-6
    desc = s.nonbonded.desc()
    Fr_ref, Er_ref, _ = orc.nb_evaluate(desc, s.positions, s.box, include_direct=False)
    assert force_rms_rel(Fr, Fr_ref) <= 1e-5
    assert energy_rel(Er, Er_ref) <= 1e-6


def test_bit_reproducible():
    s = water(1000, 3.1, seed=9)
    op = device_op(s)
    F1, E1 = op.calc_forces_and_energy(s.positions, s.box)
    F2, E2 = op.calc_forces_and_energy(s.positions + 0.0, s.box)
    assert E1 == E2
    assert np.array_equal(F1, F2)


def test_moved_positions_rebuild_list():
    s = water(1000, 3.1, seed=4)
    op = device_op(s)
    op.calc_forces_and_energy(s.positions, s.box)
    rng = np.random.default_rng(0)
    moved = s.positions + rng.normal(scale=0.05, size=s.positions.shape)
    F, E = op.calc_forces_and_energy(moved, s.box)
    F_ref, E_ref, _ = orc.nb_evaluate(s.nonbonded.desc(), moved, s.box)
    assert force_rms_rel(F, F_ref) <= FORCE_TOL
    assert energy_rel(E, E_ref) <= ENERGY_TOL
    count_ref, hash_ref = orc.nb_pairs(s.nonbonded.desc(), moved, s.box)
    assert op.pair_stats(moved, s.box) == (count_ref, hash_ref)


def test_cutoff_larger_than_half_box_is_rejected():
    s = water(100, 1.7, seed=1)
    op = ommhip.GpuNonbondedForce(s.nonbonded)
    with pytest.raises(ommhip.OmmError) as e:
        op.initialize(s.box)
    assert e.value.status == 2


def _offset_water(value, explicit):
    """300 waters with a global parameter 'lambda' offsetting the charge and
    sigma of atom 0 and the charge product / sigma of exception 1 (an O-H
    pair); explicit=True folds the offsets into the base parameters instead."""
    s = water(300, 2.1, seed=51)
    f = s.nonbonded
    q, sg, ep = f.particles[0]
    p1, p2, eq, es, ee = f.exceptions[1]
    if explicit:
        f.particles[0] = (q + value * 0.1, sg + value * 0.01, ep)
        f.exceptions[1] = (p1, p2, eq + value * 0.05, es + value * 0.02, ee)
    else:
        f.add_global_parameter("lambda", 0.0)
        f.add_particle_parameter_offset("lambda", 0, 0.1, 0.01, 0.0)
        f.add_exception_parameter_offset("lambda", 1, 0.05, 0.02, 0.0)
    return s


def test_parameter_offsets_on_device():
    # set_global_parameter moves the device parameters exactly as explicit ones would.
    s = _offset_water(0.0, explicit=False)
    op = device_op(s)
    assert op.set_global_parameter("lambda", 0.0)
    assert not op.set_global_parameter("unknown", 1.0)
    for value in (0.0, 0.6, -0.35):
        assert op.set_global_parameter("lambda", value)
        assert op.get_global_parameter("lambda") == value
        F, E = op.calc_forces_and_energy(s.positions, s.box)
        ref = _offset_water(value, explicit=True)
        F2, E2 = device_op(ref).calc_forces_and_energy(ref.positions, ref.box)
        print(f"lambda={value}: E={E:.9f} explicit={E2:.9f}")
        assert np.array_equal(F, F2)
        assert E == E2
        F_ref, E_ref, _ = orc.nb_evaluate(ref.nonbonded.desc(), ref.positions, ref.box)
        assert force_rms_rel(F, F_ref) <= FORCE_TOL
        assert energy_rel(E, E_ref) <= ENERGY_TOL


def test_update_parameters_from():
    s = water(300, 2.1, seed=52)
    op = device_op(s)
    f = s.nonbonded
    q, sg, ep = f.particles[3]
    f.particles[3] = (q * 0.5, sg, ep)
    op.update_parameters_from(f)
    F, E = op.calc_forces_and_energy(s.positions, s.box)
    F2, E2 = device_op(s).calc_forces_and_energy(s.positions, s.box)
    assert np.array_equal(F, F2) and E == E2
    bad = water(300, 2.1, seed=52).nonbonded
    p1, p2, eq, es, ee = bad.exceptions[0]
    bad.exceptions[0] = (p1, p2, 0.1, es, ee)  # an exclusion cannot become an interaction
    with pytest.raises(ommhip.OmmError):
        op.update_parameters_from(bad)


def test_context_set_parameter():
    s = _offset_water(0.0, explicit=False)
    ctx = ommhip.Context(s, temperature=0.0, friction=0.0)
    assert ctx.set_parameter("lambda", 0.4)
    ref = _offset_water(0.4, explicit=True)
    a, b = ctx.get_state(), ommhip.Context(ref, temperature=0.0, friction=0.0).get_state()
    assert np.array_equal(a["forces"], b["forces"])
    assert abs(a["potential_energy"] - b["potential_energy"]) < 1e-7
