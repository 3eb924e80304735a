"""Batched replicas: one set of launches for many independent water boxes.

Replica r of a batch must reproduce, bit for bit, the forces and the
trajectory of a one-system context built from the same descriptor and seed
(forces are fixed-point sums, so the grouping of systems into launches cannot
change them; the RNG is keyed by the replica's own seed and atom index), and
each replica must meet the oracle bars on its own.
"""
import numpy as np
import pytest

import ommhip
import oracle as orc
from helpers import energy_rel, force_rms_rel, water

pytestmark = pytest.mark.gpu

# Uneven sizes and boxes (33.4 molecules/nm^3 as in the batched workloads),
# each with its own PME grid.
REPLICAS = [(216, 1.86), (500, 2.465), (350, 2.19), (800, 2.88)]


def systems():
    return [water(m, edge, seed=100 + k) for k, (m, edge) in enumerate(REPLICAS)]


def test_batch_forces_match_single_contexts_bitwise():
    sys_ = systems()
    batch = ommhip.Batch(sys_, temperature=0.0, friction=0.0, seeds=[11, 12, 13, 14])
    for r, s in enumerate(sys_):
        single = ommhip.Context(s, temperature=0.0, friction=0.0, seed=11 + r)
        a, b = batch.get_state(r), single.get_state()
        print(f"replica {r}: N={len(s.masses)} E_batch={a['potential_energy']:.9f} E_single={b['potential_energy']:.9f}")
        assert np.array_equal(a["forces"], b["forces"])
        # Energies are wave-reduced before the fixed-point add, and a wave of
        # the batch may straddle two replicas, so the truncation to 2^-32
        # kJ/mol can differ by a few units.
        assert abs(a["potential_energy"] - b["potential_energy"]) < 1e-7


def test_many_replicas_multi_workgroup_sort():
    # 33 boxes of 1,000 waters: 33 x 512 list cells, past what the sort's
    # one-workgroup scan handles, so the three-launch scan runs; replicas
    # still match lone contexts bit for bit.
    sys_ = [water(1000, 3.1, seed=300 + r) for r in range(33)]
    batch = ommhip.Batch(sys_, temperature=0.0, friction=0.0, seeds=list(range(300, 333)))
    for r in (0, 17, 32):
        single = ommhip.Context(sys_[r], temperature=0.0, friction=0.0, seed=300 + r)
        assert np.array_equal(batch.get_state(r)["forces"], single.get_state()["forces"])


def test_batch_replicas_match_oracle():
    sys_ = systems()
    batch = ommhip.Batch(sys_, temperature=0.0, friction=0.0)
    energies = batch.energies()
    for r, s in enumerate(sys_):
        F_ref, E_ref = orc.water_energy_forces(s)
        st = batch.get_state(r)
        frel = force_rms_rel(st["forces"], F_ref)
        erel = energy_rel(st["potential_energy"], E_ref)
        print(f"replica {r}: force_rms_rel={frel:.3e} energy_rel={erel:.3e}")
        assert frel <= 2e-6
        assert erel <= 1e-6
        assert energy_rel(energies[r], st["potential_energy"]) < 1e-12
        assert np.array_equal(st["positions"], s.positions)
        assert np.array_equal(batch.box(r), s.box)


def test_batch_trajectory_matches_single_contexts_bitwise():
    sys_ = systems()
    seeds = [21, 22, 23, 24]
    batch = ommhip.Batch(sys_, temperature=300.0, friction=1.0, seeds=seeds)
    batch.set_velocities_to_temperature(300.0, 40)
    batch.step(25)  # two graph replays and five eager steps
    for r, s in enumerate(sys_):
        single = ommhip.Context(s, temperature=300.0, friction=1.0, seed=seeds[r])
        single.set_velocities_to_temperature(300.0, 40 + r)
        single.step(25)
        a, b = batch.get_state(r), single.get_state()
        print(f"replica {r}: max |dx| = {np.abs(a['positions'] - b['positions']).max():.3e}")
        assert np.array_equal(a["positions"], b["positions"])
        assert np.array_equal(a["velocities"], b["velocities"])
        assert a["step"] == b["step"] == 25


def test_batch_rejects_mismatched_settings():
    a = water(216, 1.86, seed=1)
    b = water(216, 1.86, seed=2, cutoff=0.8)
    with pytest.raises(ommhip.OmmError):
        ommhip.Batch([a, b])
