"""Shared fixtures of the parity tests: water boxes and error metrics."""
import numpy as np

import ommhip
import oracle as orc


def water(num_molecules, edge, seed=1, cutoff=0.9, tol=5e-4, method="PME"):
    """A flexible-water system whose PME parameters are frozen explicitly to
    the reference rule for its box (what NonbondedForce::initialize does), so
    the oracle and the device see identical (alpha, grid)."""
    s = ommhip.water_system(num_molecules, edge, seed, cutoff=cutoff, tolerance=tol, method=method)
    if method == "PME":
        a, g = orc.effective_pme_parameters(tol, cutoff, s.box)
        s.nonbonded.set_pme_parameters(a, *g)
    return s


def force_rms_rel(f, ref):
    """sqrt(sum |F - F_ref|^2 / sum |F_ref|^2) over all atoms."""
    return float(np.sqrt(((f - ref) ** 2).sum() / (ref**2).sum()))


def energy_rel(e, ref):
    return abs(e - ref) / abs(ref)


def has_gpu():
    try:
        return ommhip.device_count() > 0
    except Exception:
        return False
