"""omm_nb_calc_forces_and_energy_device, the resident form of the operator:
positions and forces in device memory, ordered on the caller's stream.  Its
results are the host form's, bit for bit."""
import ctypes as C

import numpy as np
import pytest

import ommhip
import oracle as orc
from helpers import H2D, DeviceArray, energy_rel, force_rms_rel, hip_runtime, water

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("precision", [ommhip.PRECISION_MIXED, ommhip.PRECISION_DOUBLE])
def test_device_form_equals_host_form(precision):
    s = water(1000, 3.1, seed=81)
    op = ommhip.GpuNonbondedForce(s.nonbonded, precision)
    op.initialize(s.box)
    F, E = op.calc_forces_and_energy(s.positions, s.box)
    hip = hip_runtime()
    dpos, dfor = DeviceArray(hip, s.positions), DeviceArray(hip, np.zeros_like(s.positions))
    try:
        # The mixed form's fixed-point sums are bitwise reproducible; plain FP64
        # adds with device atomics in a varying order, to about 1e-15.
        same = (lambda a, b: np.array_equal(a, b)) if precision == ommhip.PRECISION_MIXED else \
            (lambda a, b: np.abs(np.asarray(a) - b).max() <= 1e-12 * np.abs(b).max())
        e, evaluated = op.calc_forces_and_energy_device(dpos.p, s.box, dfor.p)
        assert evaluated and same(e, E)
        assert same(dfor.get(), F)
        # Forces only (the MD step's path) against the host form's forces only.
        Ff = op.calc_forces(s.positions, s.box)
        e2, _ = op.calc_forces_and_energy_device(dpos.p, s.box, dfor.p, energy=False)
        assert e2 is None and same(dfor.get(), Ff)
        # No group selected: the reference's None, outputs untouched.
        before = dfor.get()
        _, evaluated = op.calc_forces_and_energy_device(dpos.p, s.box, dfor.p, groups=0)
        assert not evaluated and np.array_equal(dfor.get(), before)
    finally:
        dpos.free()
        dfor.free()


def test_device_form_reads_after_the_callers_stream():
    # New positions copied on the caller's stream just before the call: the
    # evaluation must see them (it waits for that stream), and match the oracle.
    s = water(1000, 3.1, seed=82)
    op = ommhip.GpuNonbondedForce(s.nonbonded)
    op.initialize(s.box)
    hip = hip_runtime()
    moved = s.positions + 0.01 * np.random.default_rng(1).standard_normal(s.positions.shape)
    moved = np.ascontiguousarray(moved)
    stream = C.c_void_p()
    assert hip.hipStreamCreate(C.byref(stream)) == 0
    dpos, dfor = DeviceArray(hip, s.positions), DeviceArray(hip, np.zeros_like(s.positions))
    try:
        assert hip.hipMemcpyAsync(dpos.p, moved.ctypes.data, moved.nbytes, H2D, stream) == 0
        e, _ = op.calc_forces_and_energy_device(dpos.p, s.box, dfor.p, stream=stream)
        F_ref, E_ref, _ = orc.nb_evaluate(s.nonbonded.desc(), moved, s.box)
        assert force_rms_rel(dfor.get(), F_ref) <= 2e-6
        assert energy_rel(e, E_ref) <= 1e-6
    finally:
        assert hip.hipStreamSynchronize(stream) == 0
        dpos.free()
        dfor.free()
        hip.hipStreamDestroy(stream)


def test_device_form_nonperiodic():
    s = water(200, 2.1, seed=83)
    s.nonbonded.set_nonbonded_method("CutoffNonPeriodic")
    op = ommhip.GpuNonbondedForce(s.nonbonded)
    op.initialize(None)
    F, E = op.calc_forces_and_energy(s.positions, None)
    hip = hip_runtime()
    dpos, dfor = DeviceArray(hip, s.positions), DeviceArray(hip, np.zeros_like(s.positions))
    try:
        e, _ = op.calc_forces_and_energy_device(dpos.p, None, dfor.p)
        assert e == E and np.array_equal(dfor.get(), F)
    finally:
        dpos.free()
        dfor.free()
