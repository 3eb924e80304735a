"""The charge mesh (PmeCache::evaluate_charges / charge_derivatives,
pme.rs:63-138) on the device: the second consumer of the PME kernels, used by
ConstantPotentialForce with per-call charges."""
import numpy as np
import pytest

import ommhip
import oracle as orc

pytestmark = pytest.mark.gpu


def test_reference_known_answer():
    # pme.rs:432-484 restated: three charges in an orthorhombic and a
    # triclinic box, alpha 2.4, grid 12 x 14 x 16.  The mesh energy equals
    # half the sum of q * dE/dq, and dE/dq matches central differences of the
    # energy in the charge.
    cache = ommhip.PmeCache()
    positions = np.array([0.1, 0.2, 0.3, 0.7, 0.4, 1.2, 1.4, 0.3, 0.8]).reshape(3, 3)
    for tilt in (0.0, 0.35):
        box = np.array([[3.0, 0, 0], [tilt, 3.2, 0], [-tilt, 0.1, 3.4]])
        params = (2.4, (12, 14, 16))
        charges = np.array([0.0, 0.3, -0.4])
        deriv = np.zeros(3)
        cache.charge_derivatives(positions, charges, box, params, [0, 1, 2], deriv)
        forces = np.zeros((3, 3))
        energy = cache.evaluate_charges(positions, charges, box, params, forces)
        assert abs(energy - 0.5 * (charges * deriv).sum()) < 1e-12
        for i in range(3):
            q = charges.copy()
            q[i] += 1e-5
            plus = cache.evaluate_charges(positions, q, box, params)
            q[i] -= 2e-5
            minus = cache.evaluate_charges(positions, q, box, params)
            assert abs((plus - minus) / 2e-5 - deriv[i]) < 2e-8
        # The same values from the oracle's restatement.
        f_ref, e_ref = orc.pme_mesh(positions, charges, box, *params)
        assert abs(energy - e_ref) <= 1e-12 * max(1.0, abs(e_ref))
        assert np.abs(forces - f_ref).max() <= 1e-10 * max(1.0, np.abs(f_ref).max())
        d_ref = orc.pme_charge_derivatives(positions, charges, box, *params)
        assert np.abs(deriv - d_ref).max() <= 1e-11 * np.abs(d_ref).max()


def test_water_box_against_oracle():
    s = ommhip.water_system(1000, 3.1, 61)
    alpha, grid = orc.effective_pme_parameters(5e-4, 0.9, s.box)
    rng = np.random.default_rng(3)
    charges = rng.normal(scale=0.4, size=len(s.positions))
    charges[::7] = 0.0  # uncharged particles do not spread
    cache = ommhip.PmeCache()
    forces = np.full(s.positions.shape, 1.5)  # results are added to what the buffer holds
    energy = cache.evaluate_charges(s.positions, charges, s.box, (alpha, grid), forces)
    f_ref, e_ref = orc.pme_mesh(s.positions, charges, s.box, alpha, grid)
    assert abs(energy - e_ref) <= 1e-12 * abs(e_ref)
    err = np.sqrt(((forces - 1.5 - f_ref) ** 2).sum() / (f_ref ** 2).sum())
    assert err <= 1e-12
    idx = np.array([0, 5, 17, 2999, 1234, 5])  # any order, repeats allowed
    deriv = np.full(len(idx), 2.0)
    cache.charge_derivatives(s.positions, charges, s.box, (alpha, grid), idx, deriv)
    d_ref = orc.pme_charge_derivatives(s.positions, charges, s.box, alpha, grid)[idx]
    assert np.abs(deriv - 2.0 - d_ref).max() <= 1e-11 * np.abs(d_ref).max()
    # A second workspace (one unit charge at the origin, the solver's
    # preconditioner call) and back: both stay cached and correct.
    one = np.zeros(1)
    cache.charge_derivatives(np.zeros((1, 3)), [1.0], s.box, (alpha, grid), [0], one)
    one_ref = orc.pme_charge_derivatives(np.zeros((1, 3)), [1.0], s.box, alpha, grid)
    assert abs(one[0] - one_ref[0]) <= 1e-12 * abs(one_ref[0])
    again = cache.evaluate_charges(s.positions, charges, s.box, (alpha, grid))
    assert abs(again - e_ref) <= 1e-12 * abs(e_ref)
    # A new alpha with the same grid and a scaled box reuse the workspace.
    box2 = s.box * 1.01
    e2 = cache.evaluate_charges(s.positions * 1.01, charges, box2, (alpha * 0.9, grid))
    e2_ref = orc.pme_mesh(s.positions * 1.01, charges, box2, alpha * 0.9, grid, with_forces=False)
    assert abs(e2 - e2_ref) <= 1e-12 * abs(e2_ref)


def test_rejects_bad_input():
    cache = ommhip.PmeCache()
    pos = np.zeros((2, 3))
    box = np.eye(3) * 3.0
    with pytest.raises(ommhip.OmmError):
        cache.evaluate_charges(pos, [np.nan, 1.0], box, (2.4, (12, 12, 12)))
    with pytest.raises(ommhip.OmmError):
        cache.evaluate_charges(pos, [1.0, 1.0], box, (-1.0, (12, 12, 12)))
    with pytest.raises(ommhip.OmmError):
        cache.evaluate_charges(pos, [1.0, 1.0], box, (2.4, (4, 12, 12)))
    with pytest.raises(ommhip.OmmError):
        cache.charge_derivatives(pos, [1.0, 1.0], box, (2.4, (12, 12, 12)), [2], np.zeros(1))


@pytest.mark.parametrize("box", [
    [[3.0, 0, 0], [0, 3.2, 0], [0, 0.1, 3.4]],
    [[3.0, 0, 0], [0.35, 3.2, 0], [-0.35, 0.1, 3.4]],
    [[3.1, 0, 0], [-0.9, 2.9, 0], [1.2, -1.1, 3.3]],
])
@pytest.mark.parametrize("grid", [(12, 14, 16), (13, 15, 17)])
def test_triclinic_mesh_energy_is_exact(box, grid):
    # In a triclinic box a point at a Nyquist index and its mirror have
    # different |k|^2; the reference's full-grid sum averages their
    # influence, which the half-spectrum convolution must reproduce.  The
    # NonbondedForce operator's reciprocal energy in plain FP64 then matches
    # the oracle to rounding, and the mixed form to its fixed-point grid.
    box = np.array(box)
    rng = np.random.default_rng(7)
    pos = rng.uniform(0, 3, size=(40, 3))
    q = rng.normal(scale=0.5, size=40)
    f = ommhip.NonbondedForce()
    f.set_nonbonded_method("PME")
    f.set_cutoff_distance(1.0)
    f.set_use_dispersion_correction(False)
    for x in q:
        f.add_particle(float(x), 1.0, 0.0)
    f.set_pme_parameters(2.4, *grid)
    F_ref, e_ref = orc.pme_mesh(pos, q, box, 2.4, grid)
    for precision, tol in ((ommhip.PRECISION_DOUBLE, 1e-12), (ommhip.PRECISION_MIXED, 1e-8)):
        op = ommhip.GpuNonbondedForce(f, precision)
        op.initialize(box)
        e = op.energy_parts(pos, box)[2]
        assert abs(e - e_ref) <= tol * abs(e_ref), (precision, e, e_ref)
    cache = ommhip.PmeCache()
    forces = np.zeros((40, 3))
    e = cache.evaluate_charges(pos, q, box, (2.4, grid), forces)
    assert abs(e - e_ref) <= 1e-12 * abs(e_ref)
    assert np.sqrt(((forces - F_ref) ** 2).sum() / (F_ref ** 2).sum()) <= 1e-12
