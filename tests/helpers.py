"""Shared fixtures of the parity tests: water boxes, error metrics and device
buffers of the HIP runtime (for the entry points that take device pointers)."""
import ctypes as C

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


def force_max_atom_rel(f, ref):
    """Largest per-atom |F_i - F_i_ref| / |F_i_ref| (reported beside the RMS bar)."""
    return float((np.linalg.norm(f - ref, axis=1) / np.maximum(np.linalg.norm(ref, axis=1), 1e-300)).max())


def energy_rel(e, ref):
    return abs(e - ref) / abs(ref)


def has_gpu():
    try:
        return ommhip.device_count() > 0
    except Exception:
        return False


def lj_pair_system(box):
    """Two Lennard-Jones atoms (no charge, no dispersion correction) within the
    cutoff of a periodic box: the energy is 4 eps ((s/r)^12 - (s/r)^6)."""
    f = ommhip.NonbondedForce()
    f.set_nonbonded_method("CutoffPeriodic")
    f.set_cutoff_distance(1.0)
    f.set_use_dispersion_correction(False)
    for _ in range(2):
        f.add_particle(0.0, 0.3, 0.5)
    pos = np.array([[0.2, 0.3, 0.4], [0.62, 0.47, 0.66]])
    return f, pos, np.asarray(box, float)


def lj_energy(d):
    x = 0.3 / np.linalg.norm(d)
    return 4 * 0.5 * (x**12 - x**6)


LJ_BOX = [[3.0, 0, 0], [0.4, 4.0, 0], [-0.2, 0.3, 5.0]]


def lj_pair_closed_forms(pos, vel, masses, box):
    """The four pressure queries of two one-atom molecules from the pair
    energy of the deformed separation (context/pressure.rs)."""
    bar = orc.BAR_NM3_TO_KJ_PER_MOL
    d = pos[1] - pos[0]
    v = box[0, 0] * box[1, 1] * box[2, 2]
    kin = sum(0.5 * m * np.array([vv[i] * vv[j] for i, j in orc.COMPONENTS]) for m, vv in zip(masses, vel))
    iso = ((2 / 3) * kin[:3].sum() / v - (lj_energy(1.001 * d) - lj_energy(0.999 * d))
           / (v * (1.001**3 - 0.999**3))) / bar
    axis = np.array([(2 * kin[a] / v - (lj_energy(d * np.where(np.arange(3) == a, 1.001, 1.0))
                                          - lj_energy(d * np.where(np.arange(3) == a, 0.999, 1.0))) / (0.002 * v)) / bar
                     for a in range(3)])
    flex = np.zeros(6)
    stress = np.zeros(6)
    for c, (i, j) in enumerate(orc.COMPONENTS):
        e = []
        for delta in (0.001, -0.001):
            t = np.eye(3)
            t[i, j] += delta
            e.append(lj_energy(t @ d))
        stress[c] = ((e[0] - e[1]) / 0.002 - 2 * kin[c]) / (v * bar)
        if c < 3:
            flex[c] = (2 * kin[c] - (e[0] - e[1]) / 0.002) / (v * bar)
        else:
            flex[c] = 2 * kin[c] / (v * bar)  # positions fixed: the pair keeps its separation
    return iso, axis, flex, stress


H2D, D2H = 1, 2


def hip_runtime():
    """The HIP runtime the library links against, for the tests' own device buffers."""
    ommhip.library()
    hip = C.CDLL("libamdhip64.so.7")
    hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    hip.hipMemcpyAsync.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int, C.c_void_p]
    hip.hipStreamCreate.argtypes = [C.POINTER(C.c_void_p)]
    hip.hipStreamSynchronize.argtypes = [C.c_void_p]
    hip.hipStreamDestroy.argtypes = [C.c_void_p]
    hip.hipFree.argtypes = [C.c_void_p]
    return hip


class DeviceArray:
    """A device copy of a host array (hipMalloc / hipMemcpy of the test, not of the library)."""

    def __init__(self, hip, host):
        self.hip, self.host = hip, np.ascontiguousarray(host)
        self.p = C.c_void_p()
        assert hip.hipMalloc(C.byref(self.p), self.host.nbytes) == 0
        assert hip.hipMemcpy(self.p, self.host.ctypes.data, self.host.nbytes, H2D) == 0

    def get(self):
        out = np.empty_like(self.host)
        assert self.hip.hipMemcpy(out.ctypes.data, self.p, out.nbytes, D2H) == 0
        return out

    def free(self):
        self.hip.hipFree(self.p)
