"""Host-side checks of the C ABI (no device compute): the library loads,
exports every entry point include/ommhip.h declares, validates descriptors the
way the reference's builders assert, and generates the synthetic water boxes."""
import ctypes as C
import math
import os
import re

import numpy as np
import pytest

import ommhip

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    text = open(os.path.join(ROOT, "include", "ommhip.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(omm_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = ommhip.library()
    names = declared_functions()
    assert len(names) >= 20
    for name in names:
        assert hasattr(lib, name), name


def test_abi_version_and_pair_key():
    lib = ommhip.library()
    header = open(os.path.join(ROOT, "include", "ommhip.h")).read()
    declared = int(re.search(r"#define OMMHIP_ABI_VERSION (\d+)", header).group(1))
    assert lib.omm_abi_version() == declared
    # splitmix64 finaliser of (i << 32 | j)
    z = ((3 << 32) | 7) + 0x9E3779B97F4A7C15
    z &= 2**64 - 1
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & (2**64 - 1)
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & (2**64 - 1)
    assert ommhip.pair_key(3, 7) == z ^ (z >> 31)


def test_desc_defaults_match_reference_new():
    d = ommhip.NbDesc()
    ommhip.library().omm_nb_desc_defaults(C.byref(d))
    # nonbonded.rs:129-162
    assert d.method == ommhip.NO_CUTOFF
    assert d.cutoff == 1.0 and d.reaction_field_dielectric == 78.3 and d.ewald_error_tolerance == 5e-4
    assert d.use_dispersion_correction == 1 and d.exceptions_use_pbc == 0 and d.use_switching_function == 0
    assert d.switching_distance == -1.0 and d.reciprocal_force_group == -1 and d.include_direct_space == 1


def _create(force):
    lib = ommhip.library()
    d = force.desc()
    h = C.c_void_p()
    st = lib.omm_nb_create(C.byref(d), ommhip.PRECISION_MIXED, C.byref(h))
    if st == 0:
        lib.omm_nb_destroy(h)
    return st, lib.omm_last_error().decode()


def test_create_validates_like_the_reference():
    f = ommhip.NonbondedForce()
    f.add_particle(0.1, 0.3, 0.5)
    assert _create(f)[0] == 0
    f.particles.append((0.1, -0.3, 0.5))  # bypass the Python builder's own check
    st, msg = _create(f)
    assert st == 1 and "nonnegative" in msg
    g = ommhip.NonbondedForce()
    g.add_particle(0.1, 0.3, 0.5)
    g.ewald_error_tolerance = 0.7
    st, msg = _create(g)
    assert st == 1 and "tolerance" in msg


def test_python_builder_asserts():
    f = ommhip.NonbondedForce()
    with pytest.raises(ValueError):
        f.add_particle(float("nan"), 0.3, 0.1)
    with pytest.raises(ValueError):
        f.set_cutoff_distance(0.0)
    with pytest.raises(ValueError):
        f.set_pme_parameters(2.0, 4, 10, 10)
    with pytest.raises(ValueError):
        f.set_nonbonded_method("Bogus")


def test_water_box_geometry_and_determinism():
    p = ommhip.water_box(200, 2.0, 42)
    q = ommhip.water_box(200, 2.0, 42)
    assert np.array_equal(p, q)
    assert not np.array_equal(p, ommhip.water_box(200, 2.0, 43))
    o, h1, h2 = p[0::3], p[1::3], p[2::3]
    assert np.allclose(np.linalg.norm(h1 - o, axis=1), 0.09572, atol=1e-12)
    assert np.allclose(np.linalg.norm(h2 - o, axis=1), 0.09572, atol=1e-12)
    cosang = ((h1 - o) * (h2 - o)).sum(1) / 0.09572**2
    assert np.allclose(np.degrees(np.arccos(cosang)), 104.52, atol=1e-9)
    assert o.min() > -0.03 and o.max() < 2.03
    # distinct lattice sites
    assert len(np.unique(np.floor(o / (2.0 / 6)).astype(int), axis=0)) == 200


def test_exceptions_from_bonds_match_reference_rules():
    f = ommhip.NonbondedForce()
    for _ in range(5):
        f.add_particle(0.1, 0.3, 0.4)
    f.create_exceptions_from_bonds([(0, 1), (1, 2), (2, 3), (3, 4)], 0.5, 0.5)
    pairs = {(e[0], e[1]): e for e in f.exceptions}
    assert set(pairs) == {(0, 1), (1, 2), (2, 3), (3, 4), (0, 2), (1, 3), (2, 4), (0, 3), (1, 4)}
    assert pairs[(0, 1)][2] == 0.0 and pairs[(0, 1)][4] == 0.0
    assert pairs[(0, 3)][2] == pytest.approx(0.5 * 0.01) and pairs[(0, 3)][4] == pytest.approx(0.5 * 0.4)


def test_no_device_calls_fail_cleanly_without_gpu():
    if ommhip.device_count() > 0:
        pytest.skip("a device is present")
    s = ommhip.water_system(10, 1.9, 1)
    op = ommhip.GpuNonbondedForce(s.nonbonded)
    with pytest.raises(ommhip.OmmError) as e:
        op.initialize(s.box)
    assert e.value.status == 5


def test_tuning_cache_path_is_host_state(tmp_path):
    # Setting and clearing the cache path touches no device.
    ommhip.set_tuning_cache(tmp_path / "t.txt")
    ommhip.set_tuning_cache(None)
    assert ctypes_sizeof_tuning() == 3 * 4 + 4 + 3 * 8 + 3 * 4 + 4


def ctypes_sizeof_tuning():
    return C.sizeof(ommhip.LaunchTuning)
