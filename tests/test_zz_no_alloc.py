"""omm_pme_charge_derivatives_noalloc, the allocation-free variant of
omm_pme_charge_derivatives, against that entry point on two calls in a row
(the second with fewer particles and a smaller grid), through the C ABI with
device buffers of the HIP runtime the library itself uses."""
import ctypes as C

import numpy as np
import pytest

import ommhip
import oracle as orc
from helpers import DeviceArray as _Dev, hip_runtime as _hip

pytestmark = pytest.mark.gpu

def _noalloc(hip, pos, charges, box, alpha, grid, idx, start):
    lib = ommhip.library()
    bufs = [_Dev(hip, np.asarray(pos, np.float64)), _Dev(hip, np.asarray(charges, np.float64)),
            _Dev(hip, np.asarray(idx, np.int32)), _Dev(hip, np.asarray(start, np.float64))]
    try:
        b = np.ascontiguousarray(box, np.float64).reshape(9)
        g = (C.c_int32 * 3)(*grid)
        st = lib.omm_pme_charge_derivatives_noalloc(len(pos), bufs[0].p, bufs[1].p, b.ctypes.data_as(C.c_void_p),
                                                    C.c_double(alpha), g, bufs[2].p, len(idx), bufs[3].p, None)
        if st != 0:
            raise ommhip.OmmError(st, lib.omm_last_error().decode())
        return bufs[3].get()
    finally:
        for d in bufs:
            d.free()


def _case(molecules, edge, seed, count):
    s = ommhip.water_system(molecules, edge, seed)
    alpha, grid = orc.effective_pme_parameters(5e-4, 0.9, s.box)
    rng = np.random.default_rng(seed)
    q = rng.normal(scale=0.4, size=len(s.positions))
    q[::5] = 0.0
    idx = rng.choice(len(s.positions), count, replace=False).astype(np.int32)
    return s.positions, q, s.box, alpha, grid, idx


def test_noalloc_matches_charge_derivatives_twice():
    hip = _hip()
    cache = ommhip.PmeCache()
    for molecules, edge, seed, count in ((1000, 3.1, 71, 400), (300, 2.1, 72, 37)):
        pos, q, box, alpha, grid, idx = _case(molecules, edge, seed, count)
        start = np.linspace(-1.0, 1.0, count)
        ref = start.copy()
        cache.charge_derivatives(pos, q, box, (alpha, grid), idx, ref)
        got = _noalloc(hip, pos, q, box, alpha, grid, idx, start)
        scale = np.abs(ref - start).max()
        print(f"{len(pos)} particles, grid {grid}: max |diff| {np.abs(got - ref).max():.3e} of {scale:.3e}")
        assert np.abs(got - ref).max() <= 1e-11 * scale


def test_noalloc_rejects_grid_above_limit():
    hip = _hip()
    pos, q, box, alpha, _, idx = _case(300, 2.1, 73, 5)
    with pytest.raises(ommhip.OmmError) as e:
        _noalloc(hip, pos, q, box, alpha, (65, 20, 20), idx, np.zeros(5))
    assert e.value.status == ommhip.STATUS_CODES["INVALID"]
