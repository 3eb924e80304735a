"""Launch shapes: the FFT decomposition (plane passes or five line passes),
the spread tile edge, the FFT lines per workgroup and the pair grid change
no result bit, and the tuning cache replays a timed tuning."""
import numpy as np
import pytest

import ommhip
from helpers import water

pytestmark = pytest.mark.gpu

# (spread edge: 0 for the atom-owned spread, FFT lines per line-pass
# workgroup, plane passes, pair scale)
SHAPES = [(8, 16, 0, 1), (6, 4, 1, 2), (8, 2, 1, -2), (6, 8, 0, 1), (0, 16, 1, 1), (0, 4, 0, 2)]


def _run(system, shape, steps=12, seed=5, precision=ommhip.PRECISION_MIXED):
    c = ommhip.Context(system, temperature=300.0, friction=1.0, seed=seed, precision=precision)
    c.set_launch_tuning(*shape)
    t = c.launch_tuning()
    assert (t["spread_tile_edge"], t["fft_lines_per_workgroup"], t["fft_plane"], t["pair_scale"]) == shape
    first = c.get_state()
    c.set_velocities_to_temperature(300.0, 3)
    c.step(steps)
    return first, c.get_state()


@pytest.mark.parametrize("molecules,edge,cutoff", [(2000, 3.9, 0.9), (1558, 3.6, 1.0)])
def test_launch_shapes_change_no_bit(molecules, edge, cutoff):
    # A 35-point grid (radices 5 and 7) and the 29-point grid of the 1.0 nm
    # cutoff (a prime radix); the plane's y lines are padded in both.
    s = water(molecules, edge, seed=9, cutoff=cutoff)
    ref_first, ref_end = _run(s, SHAPES[0])
    for shape in SHAPES[1:]:
        first, end = _run(s, shape)
        print(shape, "grid", s.nonbonded.desc().pme_grid)
        assert np.array_equal(first["forces"], ref_first["forces"])
        assert first["potential_energy"] == ref_first["potential_energy"]
        assert np.array_equal(end["positions"], ref_end["positions"])
        assert np.array_equal(end["velocities"], ref_end["velocities"])


def test_plane_passes_in_plain_fp64():
    # The FP64 plane workgroup holds its half spectrum twice (16-byte
    # points).  Plain FP64 accumulates with device atomics, so the two
    # decompositions agree to rounding rather than bitwise.
    s = water(1000, 3.1, seed=4)
    a, _ = _run(s, (8, 16, 0, 1), steps=0, precision=ommhip.PRECISION_DOUBLE)
    b, _ = _run(s, (8, 16, 1, 1), steps=0, precision=ommhip.PRECISION_DOUBLE)
    assert np.abs(a["forces"] - b["forces"]).max() <= 1e-9 * np.abs(a["forces"]).max()
    assert abs(a["potential_energy"] - b["potential_energy"]) <= 1e-10 * abs(a["potential_energy"])


def _scaled_charges(system, factor):
    # Effective charges factor x q through a global parameter's offsets
    # (largest |q| above 2: the spreads' int64 path instead of int32 rounds).
    f = system.nonbonded
    f.add_global_parameter("qscale", factor - 1.0)
    w = ommhip.WATER
    for i in range(f.num_particles()):
        f.add_particle_parameter_offset("qscale", i, w["q_o"] if i % 3 == 0 else w["q_h"], 0.0, 0.0)
    return system


@pytest.mark.parametrize("case", ["small_grid", "large_charges", "flexible_box"])
def test_atom_owned_spread_matches_tile_spread(case):
    # The atom-owned spread (edge 0) adds the tile spread's fixed-point
    # contributions in another order: the energy evaluation (FP64 stencils),
    # the MD steps (FP32 stencils) and, with the flexible barostat, triclinic
    # boxes all give the same bits.
    s = water(500, 2.5, seed=31)
    steps = 20
    if case == "large_charges":
        # |q_O| = 2.0016: just past the int32 rounds' bound; a few steps (the
        # scaled charges pull hydrogens into neighbouring oxygens).
        s = _scaled_charges(s, 2.4)
        steps = 5
    runs = []
    for edge in (8, 0):
        c = ommhip.Context(s, temperature=300.0, friction=1.0, seed=7)
        c.set_launch_tuning(edge, 8, 0, 1)
        if case == "flexible_box":
            c.set_flexible_barostat(1.0, 300.0, 5, seed=3)
        first = c.get_state()
        c.set_velocities_to_temperature(300.0, 4)
        c.step(steps)
        runs.append((first, c.get_state(), c.box()))
    (fa, ea, ba), (fb, eb, bb) = runs
    print(case, "grid", s.nonbonded.desc().pme_grid, "box", ba.tolist())
    assert np.array_equal(fa["forces"], fb["forces"])
    assert fa["potential_energy"] == fb["potential_energy"]
    assert np.array_equal(ea["positions"], eb["positions"])
    assert np.array_equal(ea["velocities"], eb["velocities"])
    assert np.array_equal(ba, bb)


def test_atom_owned_spread_is_mixed_only():
    c = ommhip.Context(water(216, 1.86, seed=1), temperature=300.0, friction=1.0, seed=1,
                       precision=ommhip.PRECISION_DOUBLE)
    with pytest.raises(ommhip.OmmError):
        c.set_launch_tuning(0, 16, 0, 1)
    assert c.launch_tuning()["spread_tile_edge"] in (6, 8)


@pytest.mark.parametrize("edge", [6, 0])
def test_batch_shapes_match_lone_context(edge):
    sys_ = [water(216, 1.86, seed=70), water(800, 2.88, seed=71), water(500, 2.465, seed=72)]
    batch = ommhip.Batch(sys_, temperature=300.0, friction=1.0, seeds=[1, 2, 3])
    batch.set_launch_tuning(edge, 2, 1, 1)
    assert batch.launch_tuning()["fft_plane"] == 1
    batch.set_velocities_to_temperature(300.0, 8)
    batch.step(12)
    lone = ommhip.Context(sys_[1], temperature=300.0, friction=1.0, seed=2)
    lone.set_launch_tuning(8, 16, 0, 1)
    lone.set_velocities_to_temperature(300.0, 9)
    lone.step(12)
    assert np.array_equal(batch.get_state(1)["positions"], lone.get_state()["positions"])


def test_rejects_unknown_shapes():
    c = ommhip.Context(water(216, 1.86, seed=1), temperature=300.0, friction=1.0, seed=1)
    for bad in [(7, 16, 0, 1), (8, 3, 0, 1), (8, 16, 2, 1), (8, 16, 0, 3)]:
        with pytest.raises(ommhip.OmmError):
            c.set_launch_tuning(*bad)
    # The context keeps working with its own shapes.
    c.step(2)


def test_tuning_cache_replays_a_timed_tuning(tmp_path):
    path = tmp_path / "tuning.txt"
    ommhip.set_tuning_cache(path)
    try:
        s = water(700, 2.8, seed=12)
        a = ommhip.Context(s, temperature=300.0, friction=1.0, seed=3)
        ta = a.launch_tuning()
        assert ta["from_cache"] == 0
        lines = path.read_text().splitlines()
        assert len(lines) == 1 and "atoms=2100" in lines[0]
        b = ommhip.Context(s, temperature=300.0, friction=1.0, seed=3)
        tb = b.launch_tuning()
        assert tb["from_cache"] == 1
        for k in ("spread_tile_edge", "fft_lines_per_workgroup", "fft_plane", "pair_scale"):
            assert tb[k] == ta[k], k
        assert len(path.read_text().splitlines()) == 1  # a replay stores nothing
        for c in (a, b):
            c.set_velocities_to_temperature(300.0, 5)
            c.step(10)
        assert np.array_equal(a.get_state()["positions"], b.get_state()["positions"])
        # Another workload misses the cache and appends its own line.
        ommhip.Context(water(300, 2.1, seed=13), temperature=300.0, friction=1.0, seed=3).launch_tuning()
        assert len(path.read_text().splitlines()) == 2
    finally:
        ommhip.set_tuning_cache(None)
