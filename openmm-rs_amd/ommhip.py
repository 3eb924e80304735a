"""Python host binding of the ommhip C ABI (include/ommhip.h).

Mirrors the reference's Python surface for this path — the pyo3 classes
``NonbondedForce`` (crates/openmm-py/src/lib.rs:222-530) and the Context step /
state calls (lib.rs:1999-2403) — on top of ``libommhip.so`` through ctypes.
Builder calls keep the reference's names, argument meaning and assertion
behaviour (nonbonded.rs:165-659); evaluation goes to the gfx950 library.  There
is no CPU fallback: evaluation raises if the HIP library or a device is missing.
"""
from __future__ import annotations

import ctypes as C
import math
import os
from dataclasses import dataclass, field

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# OMMHIP_LIBRARY selects another build of the same library (the sanitizer
# build of the CPU tests); there is no fallback to anything else.
LIB_PATH = os.environ.get("OMMHIP_LIBRARY") or os.path.join(_HERE, "libommhip.so")

NO_CUTOFF, CUTOFF_NON_PERIODIC, CUTOFF_PERIODIC, EWALD, PME, LJPME = range(6)
METHODS = {
    "NoCutoff": NO_CUTOFF,
    "CutoffNonPeriodic": CUTOFF_NON_PERIODIC,
    "CutoffPeriodic": CUTOFF_PERIODIC,
    "Ewald": EWALD,
    "PME": PME,
    "LJPME": LJPME,
}
PRECISION_MIXED, PRECISION_DOUBLE = 0, 1
STATUS = {
    0: "OK",
    1: "INVALID",
    2: "CUTOFF_BOX",
    3: "NONFINITE",
    4: "OVERFLOW",
    5: "DEVICE",
    6: "STATE",
}
STATUS_CODES = {name: code for code, name in STATUS.items()}

# Water force field of the workloads (charges/LJ of a 3-site model; flexible
# O-H bonds and H-O-H angle; masses in Da).
WATER = dict(
    q_o=-0.834, q_h=0.417, sigma_o=0.315061, eps_o=0.636386, sigma_h=1.0, eps_h=0.0,
    mass_o=15.999, mass_h=1.008, r_oh=0.09572, k_bond=462750.0,
    theta0=104.52 * math.pi / 180.0, k_angle=836.8,
)

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int32)


class LaunchTuning(C.Structure):
    """ctypes image of ``omm_launch_tuning``."""

    _fields_ = [("spread_tile_edge", C.c_int32), ("fft_lines_per_workgroup", C.c_int32), ("pair_grid", C.c_int32),
                ("spread_ms", C.c_double), ("fft_ms", C.c_double), ("pair_ms", C.c_double),
                ("fft_plane", C.c_int32), ("pair_scale", C.c_int32), ("from_cache", C.c_int32)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}

    @classmethod
    def of(cls, spread_tile_edge, fft_lines_per_workgroup, fft_plane, pair_scale=1):
        t = cls()
        t.spread_tile_edge = spread_tile_edge
        t.fft_lines_per_workgroup = fft_lines_per_workgroup
        t.fft_plane = fft_plane
        t.pair_scale = pair_scale
        return t


def set_tuning_cache(path):
    """Process-wide tuning cache file (None: off); see omm_set_tuning_cache."""
    _check(library().omm_set_tuning_cache(None if path is None else str(path).encode()))


class BarostatInfo(C.Structure):
    """ctypes image of ``omm_barostat_info``."""

    _fields_ = [
        ("volume_scale", C.c_double),
        ("attempted", C.c_int64), ("accepted", C.c_int64),
        ("total_attempted", C.c_int64), ("total_accepted", C.c_int64),
        ("volume_old", C.c_double), ("volume_new", C.c_double),
        ("energy_old", C.c_double), ("energy_new", C.c_double), ("work", C.c_double),
        ("accepted_last", C.c_int32), ("skipped", C.c_int32),
        ("axis", C.c_int32),
        ("scale", C.c_double * 3),
        ("area_change", C.c_double),
        ("log_shape", C.c_double),
        ("box_new", C.c_double * 9),
    ]

    def as_dict(self):
        d = {name: getattr(self, name) for name, _ in self._fields_}
        d["scale"] = tuple(self.scale)
        d["box_new"] = np.array(self.box_new).reshape(3, 3)
        return d


class NbDesc(C.Structure):
    """ctypes image of ``omm_nb_desc``."""

    _fields_ = [
        ("num_particles", C.c_int32),
        ("charge", _dp), ("sigma", _dp), ("epsilon", _dp),
        ("num_exceptions", C.c_int32),
        ("exception_particle1", _ip), ("exception_particle2", _ip),
        ("exception_charge_prod", _dp), ("exception_sigma", _dp), ("exception_epsilon", _dp),
        ("method", C.c_int32),
        ("cutoff", C.c_double),
        ("reaction_field_dielectric", C.c_double),
        ("ewald_error_tolerance", C.c_double),
        ("pme_alpha", C.c_double),
        ("pme_grid", C.c_int32 * 3),
        ("use_dispersion_correction", C.c_int32),
        ("exceptions_use_pbc", C.c_int32),
        ("use_switching_function", C.c_int32),
        ("switching_distance", C.c_double),
        ("force_group", C.c_int32),
        ("reciprocal_force_group", C.c_int32),
        ("include_direct_space", C.c_int32),
        ("num_global_parameters", C.c_int32),
        ("global_parameter_names", C.POINTER(C.c_char_p)),
        ("global_parameter_defaults", _dp),
        ("num_particle_offsets", C.c_int32),
        ("particle_offset_parameter", _ip), ("particle_offset_index", _ip), ("particle_offset_scales", _dp),
        ("num_exception_offsets", C.c_int32),
        ("exception_offset_parameter", _ip), ("exception_offset_index", _ip), ("exception_offset_scales", _dp),
        ("ljpme_alpha", C.c_double),
        ("ljpme_grid", C.c_int32 * 3),
    ]


class MdDesc(C.Structure):
    """ctypes image of ``omm_md_desc``."""

    _fields_ = [
        ("num_particles", C.c_int32),
        ("masses", _dp),
        ("nonbonded", C.POINTER(NbDesc)),
        ("num_bonds", C.c_int32),
        ("bond_particle1", _ip), ("bond_particle2", _ip),
        ("bond_length", _dp), ("bond_k", _dp),
        ("num_angles", C.c_int32),
        ("angle_particle1", _ip), ("angle_particle2", _ip), ("angle_particle3", _ip),
        ("angle_theta0", _dp), ("angle_k", _dp),
        ("box", C.c_double * 9),
        ("temperature", C.c_double),
        ("friction", C.c_double),
        ("step_size", C.c_double),
        ("seed", C.c_uint64),
        ("num_constraints", C.c_int32),
        ("constraint_particle1", _ip), ("constraint_particle2", _ip),
        ("constraint_distance", _dp),
        ("constraint_tolerance", C.c_double),
    ]


def _f64(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.float64))


def _i32(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.int32))


def _ptr(a, t):
    return a.ctypes.data_as(C.POINTER(t)) if a is not None and a.size else None


_lib = None


def library():
    """Load libommhip.so (built in-tree by __graft_entry__.build())."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with __graft_entry__.build()")
        lib = C.CDLL(LIB_PATH)
        lib.omm_last_error.restype = C.c_char_p
        lib.omm_abi_version.restype = C.c_int32
        lib.omm_device_count.restype = C.c_int32
        lib.omm_pair_key.restype = C.c_uint64
        lib.omm_pair_key.argtypes = [C.c_int32, C.c_int32]
        lib.omm_nb_desc_defaults.restype = None
        lib.omm_water_box.argtypes = [C.c_int32, C.c_double, C.c_uint64, _dp]
        lib.omm_ctx_set_velocities_to_temperature.argtypes = [C.c_void_p, C.c_double, C.c_uint64]
        lib.omm_ctx_apply_constraints.argtypes = [C.c_void_p, C.c_double]
        lib.omm_ctx_set_anisotropic_barostat.argtypes = [C.c_void_p, _dp, C.c_double, C.c_int32, C.c_int32,
                                                         C.c_uint64]
        lib.omm_batch_set_anisotropic_barostat.argtypes = [C.c_void_p, _dp, C.c_double, C.c_int32, C.c_int32,
                                                           C.c_uint64]
        for fn in ("omm_ctx_set_flexible_barostat", "omm_batch_set_flexible_barostat"):
            getattr(lib, fn).argtypes = [C.c_void_p, C.c_double, C.c_double, C.c_int32, C.c_uint64]
        for fn in ("omm_ctx_set_membrane_barostat", "omm_batch_set_membrane_barostat"):
            getattr(lib, fn).argtypes = [C.c_void_p, C.c_double, C.c_double, C.c_double, C.c_int32, C.c_int32,
                                         C.c_int32, C.c_uint64]
        lib.omm_ctx_apply_velocity_constraints.argtypes = [C.c_void_p, C.c_double]
        lib.omm_ctx_set_integrator.argtypes = [C.c_void_p, C.c_double, C.c_double, C.c_double]
        lib.omm_batch_set_velocities_to_temperature.argtypes = [C.c_void_p, C.c_double, C.c_uint64]
        lib.omm_batch_set_integrator.argtypes = [C.c_void_p, C.c_double, C.c_double, C.c_double]
        lib.omm_ctx_set_barostat.argtypes = [C.c_void_p, C.c_double, C.c_double, C.c_int32, C.c_uint64]
        lib.omm_nb_set_global_parameter.argtypes = [C.c_void_p, C.c_char_p, C.c_double, C.POINTER(C.c_int32)]
        lib.omm_ctx_set_parameter.argtypes = [C.c_void_p, C.c_char_p, C.c_double, C.POINTER(C.c_int32)]
        lib.omm_batch_set_barostat.argtypes = [C.c_void_p, C.c_double, C.c_double, C.c_int32, C.c_uint64]
        for kind in ("ctx", "batch"):
            for q in ("current_pressure", "current_anisotropic_pressure", "current_membrane_pressure",
                      "current_flexible_pressure"):
                getattr(lib, f"omm_{kind}_compute_{q}").argtypes = [C.c_void_p, _dp]
            getattr(lib, f"omm_{kind}_compute_stress_tensor").argtypes = [C.c_void_p, C.c_int32, _dp]
        _lib = lib
    return _lib


class OmmError(RuntimeError):
    def __init__(self, status, message):
        super().__init__(f"{STATUS.get(status, status)}: {message}")
        self.status = status


def _check(status):
    if status != 0:
        raise OmmError(status, library().omm_last_error().decode())


def device_count() -> int:
    return int(library().omm_device_count())


def pair_key(i: int, j: int) -> int:
    return int(library().omm_pair_key(i, j))


def water_box(num_molecules: int, box_edge: float, seed: int) -> np.ndarray:
    """Seeded flexible-water coordinates, shape (3*num_molecules, 3)."""
    out = np.zeros((3 * num_molecules, 3))
    _check(library().omm_water_box(num_molecules, box_edge, seed, out.ctypes.data_as(_dp)))
    return out


def cubic_box(edge: float) -> np.ndarray:
    return np.diag([edge, edge, edge]).astype(np.float64)


@dataclass
class NonbondedForce:
    """Builder with the reference's NonbondedForce surface (nonbonded.rs:110-659)."""

    particles: list = field(default_factory=list)
    exceptions: list = field(default_factory=list)
    method: int = NO_CUTOFF
    cutoff: float = 1.0
    reaction_field_dielectric: float = 78.3
    ewald_error_tolerance: float = 5e-4
    pme: tuple = (0.0, (0, 0, 0))
    ljpme: tuple = (0.0, (0, 0, 0))
    use_dispersion_correction: bool = True
    exceptions_use_pbc: bool = False
    use_switching_function: bool = False
    switching_distance: float = -1.0
    force_group: int = 0
    reciprocal_force_group: int = -1
    include_direct_space: bool = True
    global_parameters: list = field(default_factory=list)   # (name, default)
    particle_offsets: list = field(default_factory=list)    # (parameter, index, charge, sigma, epsilon)
    exception_offsets: list = field(default_factory=list)

    @staticmethod
    def _validate(charge, sigma, epsilon):
        # nonbonded.rs:1252-1261
        if not (math.isfinite(charge) and math.isfinite(sigma) and math.isfinite(epsilon)
                and sigma >= 0.0 and epsilon >= 0.0):
            raise ValueError("nonbonded parameters must be finite with nonnegative sigma and epsilon")

    def add_particle(self, charge, sigma, epsilon):
        self._validate(charge, sigma, epsilon)
        self.particles.append((float(charge), float(sigma), float(epsilon)))
        return len(self.particles) - 1

    def add_exception(self, p1, p2, charge_prod, sigma, epsilon):
        self._validate(charge_prod, sigma, epsilon)
        self.exceptions.append((int(p1), int(p2), float(charge_prod), float(sigma), float(epsilon)))
        return len(self.exceptions) - 1

    def num_particles(self):
        return len(self.particles)

    def add_global_parameter(self, name, default):
        """nonbonded.rs:247-256."""
        if not name or not math.isfinite(default):
            raise ValueError("global parameters require a name and a finite default value")
        self.global_parameters.append((str(name), float(default)))
        return len(self.global_parameters) - 1

    def _offset(self, parameter, index, scales, count, what):
        if parameter not in [n for n, _ in self.global_parameters]:
            raise ValueError(f"unknown offset parameter '{parameter}'")
        if not 0 <= index < count:
            raise ValueError(f"{what} offset index out of range")
        if not all(math.isfinite(x) for x in scales):
            raise ValueError("offset scales must be finite")
        return (str(parameter), int(index)) + tuple(float(x) for x in scales)

    def add_particle_parameter_offset(self, parameter, index, charge, sigma, epsilon):
        """nonbonded.rs:305-321."""
        self.particle_offsets.append(self._offset(parameter, index, (charge, sigma, epsilon),
                                                  len(self.particles), "particle"))
        return len(self.particle_offsets) - 1

    def add_exception_parameter_offset(self, parameter, index, charge, sigma, epsilon):
        """nonbonded.rs:323-339."""
        self.exception_offsets.append(self._offset(parameter, index, (charge, sigma, epsilon),
                                                   len(self.exceptions), "exception"))
        return len(self.exception_offsets) - 1

    def num_exceptions(self):
        return len(self.exceptions)

    def set_nonbonded_method(self, method):
        if isinstance(method, str):
            if method not in METHODS:
                raise ValueError("unknown method; expected NoCutoff, CutoffNonPeriodic, "
                                 "CutoffPeriodic, Ewald, PME, or LJPME")
            method = METHODS[method]
        self.method = int(method)

    def set_cutoff_distance(self, cutoff):
        if not (math.isfinite(cutoff) and cutoff > 0.0):
            raise ValueError("cutoff must be positive and finite")
        self.cutoff = float(cutoff)

    def set_reaction_field_dielectric(self, eps):
        if math.isnan(eps) or eps < 1.0:
            raise ValueError("reaction field dielectric must be at least one")
        self.reaction_field_dielectric = float(eps)

    def set_ewald_error_tolerance(self, tol):
        if not (math.isfinite(tol) and 0.0 < tol < 0.5):
            raise ValueError("Ewald tolerance must be between zero and one half")
        self.ewald_error_tolerance = float(tol)

    def set_pme_parameters(self, alpha, nx, ny, nz):
        if not (math.isfinite(alpha) and alpha >= 0.0):
            raise ValueError("PME alpha must be nonnegative and finite")
        if alpha != 0.0 and min(nx, ny, nz) < 5:
            raise ValueError("PME grid dimensions must be at least five")
        self.pme = (float(alpha), (int(nx), int(ny), int(nz))) if alpha != 0.0 else (0.0, (0, 0, 0))

    def set_ljpme_parameters(self, alpha, nx, ny, nz):
        """nonbonded.rs:615-630."""
        if not (math.isfinite(alpha) and alpha >= 0.0):
            raise ValueError("LJPME alpha must be nonnegative and finite")
        if alpha != 0.0 and min(nx, ny, nz) < 5:
            raise ValueError("LJPME grid dimensions must be at least five")
        self.ljpme = (float(alpha), (int(nx), int(ny), int(nz))) if alpha != 0.0 else (0.0, (0, 0, 0))

    def set_use_dispersion_correction(self, enabled):
        self.use_dispersion_correction = bool(enabled)

    def set_exceptions_use_periodic_boundary_conditions(self, periodic):
        self.exceptions_use_pbc = bool(periodic)

    def set_use_switching_function(self, enabled):
        self.use_switching_function = bool(enabled)

    def set_switching_distance(self, distance):
        if not (math.isfinite(distance) and distance >= 0.0):
            raise ValueError("switching distance must be nonnegative and finite")
        self.switching_distance = float(distance)

    def set_force_group(self, group):
        if not 0 <= group < 32:
            raise ValueError("force group must be between zero and 31")
        self.force_group = int(group)

    def set_reciprocal_space_force_group(self, group):
        if not -1 <= group < 32:
            raise ValueError("reciprocal force group must be between -1 and 31")
        self.reciprocal_force_group = int(group)

    def set_include_direct_space(self, enabled):
        self.include_direct_space = bool(enabled)

    def uses_periodic_boundary_conditions(self):
        return self.method in (CUTOFF_PERIODIC, EWALD, PME, LJPME)

    def create_exceptions_from_bonds(self, bonds, coulomb14scale, lj14scale):
        """nonbonded.rs:434-500 (1-2 and 1-3 exclusions, scaled 1-4 pairs)."""
        n = len(self.particles)
        nbr = [[] for _ in range(n)]
        for a, b in bonds:
            nbr[a].append(b)
            nbr[b].append(a)
        excl, p14 = set(), set()
        for i in range(n):
            for j in nbr[i]:
                if i < j:
                    excl.add((i, j))
            for j in nbr[i]:
                for k in nbr[j]:
                    if k != i:
                        excl.add((min(i, k), max(i, k)))
            for j in nbr[i]:
                for k in nbr[j]:
                    if k == i:
                        continue
                    for l in nbr[k]:
                        if l in (j, i):
                            continue
                        key = (min(i, l), max(i, l))
                        if key not in excl:
                            p14.add(key)
        for i, j in sorted(excl):
            s = 0.5 * (self.particles[i][1] + self.particles[j][1])
            self.add_exception(i, j, 0.0, s, 0.0)
        for i, j in sorted(p14 - excl):
            qi, si, ei = self.particles[i]
            qj, sj, ej = self.particles[j]
            self.add_exception(i, j, coulomb14scale * qi * qj, 0.5 * (si + sj), lj14scale * math.sqrt(ei * ej))

    def desc(self):
        """Build an ``NbDesc`` (keeps the numpy arrays alive on the returned object)."""
        p = np.array(self.particles, dtype=np.float64).reshape(-1, 3)
        e = self.exceptions
        d = NbDesc()
        keep = {}
        keep["q"] = _f64(p[:, 0])
        keep["s"] = _f64(p[:, 1])
        keep["e"] = _f64(p[:, 2])
        keep["p1"] = _i32([x[0] for x in e])
        keep["p2"] = _i32([x[1] for x in e])
        keep["eq"] = _f64([x[2] for x in e])
        keep["es"] = _f64([x[3] for x in e])
        keep["ee"] = _f64([x[4] for x in e])
        d.num_particles = len(self.particles)
        d.charge, d.sigma, d.epsilon = (_ptr(keep[k], C.c_double) for k in ("q", "s", "e"))
        d.num_exceptions = len(e)
        d.exception_particle1 = _ptr(keep["p1"], C.c_int32)
        d.exception_particle2 = _ptr(keep["p2"], C.c_int32)
        d.exception_charge_prod = _ptr(keep["eq"], C.c_double)
        d.exception_sigma = _ptr(keep["es"], C.c_double)
        d.exception_epsilon = _ptr(keep["ee"], C.c_double)
        d.method = self.method
        d.cutoff = self.cutoff
        d.reaction_field_dielectric = self.reaction_field_dielectric
        d.ewald_error_tolerance = self.ewald_error_tolerance
        d.pme_alpha = self.pme[0]
        for a in range(3):
            d.pme_grid[a] = self.pme[1][a]
        d.use_dispersion_correction = int(self.use_dispersion_correction)
        d.exceptions_use_pbc = int(self.exceptions_use_pbc)
        d.use_switching_function = int(self.use_switching_function)
        d.switching_distance = self.switching_distance
        d.force_group = self.force_group
        d.reciprocal_force_group = self.reciprocal_force_group
        d.include_direct_space = int(self.include_direct_space)
        names = [n for n, _ in self.global_parameters]
        keep["gn"] = (C.c_char_p * max(1, len(names)))(*[n.encode() for n in names])
        keep["gd"] = _f64([v for _, v in self.global_parameters])
        d.num_global_parameters = len(names)
        d.global_parameter_names = C.cast(keep["gn"], C.POINTER(C.c_char_p))
        d.global_parameter_defaults = _ptr(keep["gd"], C.c_double)
        for tag, offs in (("po", self.particle_offsets), ("eo", self.exception_offsets)):
            keep[tag + "p"] = _i32([names.index(o[0]) for o in offs])
            keep[tag + "i"] = _i32([o[1] for o in offs])
            keep[tag + "s"] = _f64([list(o[2:5]) for o in offs]).reshape(-1)
        d.num_particle_offsets = len(self.particle_offsets)
        d.particle_offset_parameter = _ptr(keep["pop"], C.c_int32)
        d.particle_offset_index = _ptr(keep["poi"], C.c_int32)
        d.particle_offset_scales = _ptr(keep["pos"], C.c_double)
        d.num_exception_offsets = len(self.exception_offsets)
        d.exception_offset_parameter = _ptr(keep["eop"], C.c_int32)
        d.exception_offset_index = _ptr(keep["eoi"], C.c_int32)
        d.exception_offset_scales = _ptr(keep["eos"], C.c_double)
        d.ljpme_alpha = self.ljpme[0]
        for a in range(3):
            d.ljpme_grid[a] = self.ljpme[1][a]
        d._keep = keep
        return d


def _box_ptr(box):
    if box is None:
        return None, None
    b = _f64(box).reshape(9)
    return b, b.ctypes.data_as(_dp)


class GpuNonbondedForce:
    """The device operator: ``impl Force`` for one NonbondedForce (nonbonded.rs:1053-1155)."""

    def __init__(self, force: NonbondedForce, precision: int = PRECISION_MIXED):
        lib = library()
        self._desc = force.desc()
        self._n = force.num_particles()
        h = C.c_void_p()
        _check(lib.omm_nb_create(C.byref(self._desc), precision, C.byref(h)))
        self._h = h

    def initialize(self, box=None):
        b, bp = _box_ptr(box)
        _check(library().omm_nb_initialize(self._h, self._n, bp))

    def pme_parameters(self):
        alpha = C.c_double()
        grid = (C.c_int32 * 3)()
        _check(library().omm_nb_get_pme_parameters(self._h, C.byref(alpha), grid))
        return alpha.value, tuple(grid)

    def ljpme_parameters(self):
        alpha = C.c_double()
        grid = (C.c_int32 * 3)()
        _check(library().omm_nb_get_ljpme_parameters(self._h, C.byref(alpha), grid))
        return alpha.value, tuple(grid)

    def calc_forces_and_energy(self, pos, box=None):
        p = _f64(pos).reshape(-1, 3)
        b, bp = _box_ptr(box)
        f = np.zeros_like(p)
        e = C.c_double()
        _check(library().omm_nb_calc_forces_and_energy(self._h, p.ctypes.data_as(_dp), bp,
                                                       f.ctypes.data_as(_dp), C.byref(e)))
        return f, e.value

    def calc_forces_and_energy_device(self, d_pos, box=None, d_forces=None, groups=0xFFFFFFFF, stream=None,
                                      energy=True):
        """The resident form (omm_nb_calc_forces_and_energy_device): d_pos and d_forces are device
        addresses (ints or ctypes pointers, e.g. a torch tensor's data_ptr()), stream a hipStream_t
        handle (None: the null stream).  Returns (energy or None, evaluated)."""
        def addr(x):
            return x.value if isinstance(x, C.c_void_p) else x

        b, bp = _box_ptr(box)
        e = C.c_double()
        ev = C.c_int32()
        _check(library().omm_nb_calc_forces_and_energy_device(
            self._h, C.c_void_p(addr(d_pos)), bp, C.c_uint32(groups), C.c_void_p(addr(d_forces)),
            C.byref(e) if energy else None, C.byref(ev), C.c_void_p(addr(stream))))
        return (e.value if energy else None), bool(ev.value)

    def calc_forces(self, pos, box=None):
        """Forces only (energy_out = NULL): the MD step's FP32 path."""
        p = _f64(pos).reshape(-1, 3)
        b, bp = _box_ptr(box)
        f = np.zeros_like(p)
        _check(library().omm_nb_calc_forces_and_energy(self._h, p.ctypes.data_as(_dp), bp,
                                                       f.ctypes.data_as(_dp), None))
        return f

    def calc_forces_and_energy_groups(self, pos, box, groups):
        p = _f64(pos).reshape(-1, 3)
        b, bp = _box_ptr(box)
        f = np.zeros_like(p)
        e = C.c_double()
        ev = C.c_int32()
        _check(library().omm_nb_calc_forces_and_energy_groups(
            self._h, p.ctypes.data_as(_dp), bp, C.c_uint32(groups), f.ctypes.data_as(_dp),
            C.byref(e), C.byref(ev)))
        return (f, e.value) if ev.value else None

    def energy_parts(self, pos, box=None):
        p = _f64(pos).reshape(-1, 3)
        b, bp = _box_ptr(box)
        parts = np.zeros(5)
        _check(library().omm_nb_energy_parts(self._h, p.ctypes.data_as(_dp), bp, parts.ctypes.data_as(_dp)))
        return parts

    def pair_stats(self, pos, box=None):
        p = _f64(pos).reshape(-1, 3)
        b, bp = _box_ptr(box)
        n = C.c_int64()
        h = C.c_uint64()
        _check(library().omm_nb_pair_stats(self._h, p.ctypes.data_as(_dp), bp, C.byref(n), C.byref(h)))
        return n.value, h.value

    def set_global_parameter(self, name, value):
        """Force::set_global_parameter: True when the force has the parameter."""
        found = C.c_int32()
        _check(library().omm_nb_set_global_parameter(self._h, name.encode(), float(value), C.byref(found)))
        return bool(found.value)

    def get_global_parameter(self, name):
        v, found = C.c_double(), C.c_int32()
        _check(library().omm_nb_get_global_parameter(self._h, name.encode(), C.byref(v), C.byref(found)))
        return v.value if found.value else None

    def update_parameters_from(self, force: NonbondedForce):
        """NonbondedForce::update_parameters_from (nonbonded.rs:721-767)."""
        d = force.desc()
        _check(library().omm_nb_update_parameters(self._h, C.byref(d)))

    def close(self):
        if getattr(self, "_h", None):
            library().omm_nb_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class PmeCache:
    """The charge mesh of ``PmeCache`` (pme.rs:8-138) on the device: the two calls
    ``ConstantPotentialForce`` makes, with per-call charges, alpha and grid."""

    def __init__(self):
        h = C.c_void_p()
        _check(library().omm_pme_create(C.byref(h)))
        self._h = h

    @staticmethod
    def _args(positions, charges, box, parameters):
        pos = _f64(positions).reshape(-1, 3)
        q = _f64(charges).reshape(-1)
        if len(q) != len(pos):
            raise ValueError("one charge per position is required")
        b = _f64(box).reshape(9)
        alpha, dims = parameters
        g = (C.c_int32 * 3)(*[int(x) for x in dims])
        return pos, q, b, C.c_double(float(alpha)), g

    def evaluate_charges(self, positions, charges, box, parameters, forces=None):
        """Mesh energy of the charges; the mesh forces are added into `forces` (N x 3) when given."""
        pos, q, b, a, g = self._args(positions, charges, box, parameters)
        e = C.c_double()
        fp = None
        if forces is not None:
            if not (isinstance(forces, np.ndarray) and forces.dtype == np.float64 and forces.flags.c_contiguous
                    and forces.size == pos.size):
                raise ValueError("forces must be a C-contiguous float64 array of N x 3")
            fp = forces.ctypes.data_as(_dp)
        _check(library().omm_pme_evaluate_charges(self._h, len(pos), pos.ctypes.data_as(_dp), q.ctypes.data_as(_dp),
                                                  b.ctypes.data_as(_dp), a, g, fp, C.byref(e)))
        return e.value

    def charge_derivatives(self, positions, charges, box, parameters, indices, derivatives):
        """Adds dE_mesh/dq of each listed particle into `derivatives` (float64, len(indices))."""
        pos, q, b, a, g = self._args(positions, charges, box, parameters)
        idx = _i32(indices).reshape(-1)
        if not (isinstance(derivatives, np.ndarray) and derivatives.dtype == np.float64
                and derivatives.flags.c_contiguous and derivatives.size == idx.size):
            raise ValueError("derivatives must be a C-contiguous float64 array with one value per index")
        _check(library().omm_pme_charge_derivatives(self._h, len(pos), pos.ctypes.data_as(_dp),
                                                    q.ctypes.data_as(_dp), b.ctypes.data_as(_dp), a, g,
                                                    idx.ctypes.data_as(C.POINTER(C.c_int32)), len(idx),
                                                    derivatives.ctypes.data_as(_dp)))

    def close(self):
        if getattr(self, "_h", None):
            library().omm_pme_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


@dataclass
class WaterSystem:
    """A flexible-water system: NonbondedForce (PME) + bonds + angles + masses."""

    nonbonded: NonbondedForce
    positions: np.ndarray
    box: np.ndarray
    masses: np.ndarray
    bonds: np.ndarray      # (nb, 2) int32
    bond_params: np.ndarray  # (nb, 2) length, k
    angles: np.ndarray     # (na, 3) int32
    angle_params: np.ndarray  # (na, 2) theta0, k
    constraints: list = field(default_factory=list)  # (i, j, distance nm), System::add_constraint
    constraint_tolerance: float = 0.0  # 0: the integrator default 1e-5


def water_system(num_molecules: int, box_edge: float, seed: int, cutoff: float = 0.9,
                 tolerance: float = 5e-4, method: str = "PME") -> WaterSystem:
    """Flexible 3-site water box with the workloads' force field."""
    w = WATER
    pos = water_box(num_molecules, box_edge, seed)
    nb = NonbondedForce()
    nb.set_nonbonded_method(method)
    nb.set_cutoff_distance(cutoff)
    nb.set_ewald_error_tolerance(tolerance)
    for _ in range(num_molecules):
        nb.add_particle(w["q_o"], w["sigma_o"], w["eps_o"])
        nb.add_particle(w["q_h"], w["sigma_h"], w["eps_h"])
        nb.add_particle(w["q_h"], w["sigma_h"], w["eps_h"])
    so, sh = w["sigma_o"], w["sigma_h"]
    for m in range(num_molecules):
        o = 3 * m
        # create_exceptions_from_bonds for O-H1, O-H2 bonds: two 1-2 and one 1-3 exclusion.
        nb.add_exception(o, o + 1, 0.0, 0.5 * (so + sh), 0.0)
        nb.add_exception(o, o + 2, 0.0, 0.5 * (so + sh), 0.0)
        nb.add_exception(o + 1, o + 2, 0.0, 0.5 * (sh + sh), 0.0)
    masses = np.tile([w["mass_o"], w["mass_h"], w["mass_h"]], num_molecules).astype(np.float64)
    o = 3 * np.arange(num_molecules, dtype=np.int32)
    bonds = np.stack([np.stack([o, o + 1], 1), np.stack([o, o + 2], 1)], 1).reshape(-1, 2).astype(np.int32)
    bond_params = np.tile([w["r_oh"], w["k_bond"]], (len(bonds), 1))
    angles = np.stack([o + 1, o, o + 2], 1).astype(np.int32)
    angle_params = np.tile([w["theta0"], w["k_angle"]], (num_molecules, 1))
    return WaterSystem(nb, pos, cubic_box(box_edge), masses, bonds, bond_params, angles, angle_params)


def _md_desc(system: WaterSystem, temperature, friction, step_size, seed):
    """An ``omm_md_desc`` for the system plus the arrays it points into."""
    s = system
    nbd = s.nonbonded.desc()
    keep = dict(
        nb=nbd, m=_f64(s.masses), b1=_i32(s.bonds[:, 0]), b2=_i32(s.bonds[:, 1]),
        bl=_f64(s.bond_params[:, 0]), bk=_f64(s.bond_params[:, 1]),
        a1=_i32(s.angles[:, 0]), a2=_i32(s.angles[:, 1]), a3=_i32(s.angles[:, 2]),
        at=_f64(s.angle_params[:, 0]), ak=_f64(s.angle_params[:, 1]),
    )
    d = MdDesc()
    d.num_particles = len(s.masses)
    d.masses = _ptr(keep["m"], C.c_double)
    d.nonbonded = C.pointer(nbd)
    d.num_bonds = len(s.bonds)
    d.bond_particle1, d.bond_particle2 = _ptr(keep["b1"], C.c_int32), _ptr(keep["b2"], C.c_int32)
    d.bond_length, d.bond_k = _ptr(keep["bl"], C.c_double), _ptr(keep["bk"], C.c_double)
    d.num_angles = len(s.angles)
    d.angle_particle1 = _ptr(keep["a1"], C.c_int32)
    d.angle_particle2 = _ptr(keep["a2"], C.c_int32)
    d.angle_particle3 = _ptr(keep["a3"], C.c_int32)
    d.angle_theta0, d.angle_k = _ptr(keep["at"], C.c_double), _ptr(keep["ak"], C.c_double)
    for i, v in enumerate(_f64(s.box).reshape(9)):
        d.box[i] = v
    d.temperature, d.friction, d.step_size, d.seed = temperature, friction, step_size, seed
    cons = list(getattr(s, "constraints", None) or [])
    keep["c1"] = _i32([c[0] for c in cons])
    keep["c2"] = _i32([c[1] for c in cons])
    keep["cd"] = _f64([c[2] for c in cons])
    d.num_constraints = len(cons)
    d.constraint_particle1 = _ptr(keep["c1"], C.c_int32)
    d.constraint_particle2 = _ptr(keep["c2"], C.c_int32)
    d.constraint_distance = _ptr(keep["cd"], C.c_double)
    d.constraint_tolerance = float(getattr(s, "constraint_tolerance", 0.0) or 0.0)
    return d, keep


def _state_arrays(state, n):
    """Contiguous FP64 (box, positions, velocities, list positions) of a state dict; absent keys are None."""
    out = []
    for key, shape in (("box", (9,)), ("positions", (n, 3)), ("velocities", (n, 3)), ("list_positions", (n, 3))):
        v = state.get(key)
        out.append(None if v is None else _f64(v).reshape(shape))
    return out


def _ptrs(arrays):
    return [None if a is None else a.ctypes.data_as(_dp) for a in arrays]


PHASES = ("nlist_build", "direct_pairs", "exceptions", "pme_spread", "pme_fft_conv", "pme_gather",
          "bonded", "integrator")


class Context:
    """Device-resident MD context: LangevinMiddle + bonds/angles + NonbondedForce.

    Stands in for ``Context`` with a ``LangevinMiddleIntegrator`` (context.rs:795-928).
    """

    def __init__(self, system: WaterSystem, temperature=300.0, friction=1.0, step_size=0.001,
                 seed=1, precision=PRECISION_MIXED):
        lib = library()
        self.n = len(system.masses)
        # NoCutoff and CutoffNonPeriodic systems have no periodic box to save or restore.
        self.periodic = system.nonbonded.method not in (NO_CUTOFF, CUTOFF_NON_PERIODIC)
        self._d, self._keep = _md_desc(system, temperature, friction, step_size, seed)
        h = C.c_void_p()
        _check(lib.omm_ctx_create(C.byref(self._d), precision, C.byref(h)))
        self._h = h
        self.set_positions(system.positions)

    def set_positions(self, pos):
        p = _f64(pos).reshape(-1, 3)
        _check(library().omm_ctx_set_positions(self._h, p.ctypes.data_as(_dp)))

    def set_velocities(self, vel):
        v = _f64(vel).reshape(-1, 3)
        _check(library().omm_ctx_set_velocities(self._h, v.ctypes.data_as(_dp)))

    def set_velocities_to_temperature(self, temperature, seed):
        _check(library().omm_ctx_set_velocities_to_temperature(self._h, temperature, seed))

    def apply_constraints(self, tolerance=1e-5):
        _check(library().omm_ctx_apply_constraints(self._h, float(tolerance)))

    def apply_velocity_constraints(self, tolerance=1e-5):
        _check(library().omm_ctx_apply_velocity_constraints(self._h, float(tolerance)))

    def step(self, steps):
        _check(library().omm_ctx_step(self._h, int(steps)))

    def set_integrator(self, temperature, friction, step_size):
        """LangevinMiddleIntegrator set_temperature / set_friction / set_step_size."""
        _check(library().omm_ctx_set_integrator(self._h, temperature, friction, step_size))

    def set_graph_steps(self, steps):
        """Steps per captured HIP graph (0: eager launches)."""
        _check(library().omm_ctx_set_graph_steps(self._h, int(steps)))

    def set_parameter(self, name, value):
        """Context::set_parameter for the NonbondedForce's global parameters."""
        found = C.c_int32()
        _check(library().omm_ctx_set_parameter(self._h, name.encode(), float(value), C.byref(found)))
        return bool(found.value)

    def set_barostat(self, pressure, temperature, frequency=25, seed=1):
        """MonteCarloBarostat(pressure bar, temperature K, frequency) with rigid molecules."""
        _check(library().omm_ctx_set_barostat(self._h, pressure, temperature, int(frequency), int(seed)))

    def set_flexible_barostat(self, pressure, temperature, frequency=25, seed=1):
        """MonteCarloFlexibleBarostat(pressure bar, temperature K, frequency)."""
        _check(library().omm_ctx_set_flexible_barostat(self._h, pressure, temperature, int(frequency), int(seed)))

    def set_membrane_barostat(self, pressure, surface_tension, temperature, xy_mode=0, z_mode=0, frequency=25,
                              seed=1):
        """MonteCarloMembraneBarostat(pressure bar, surface tension bar nm, temperature K, xy/z modes)."""
        _check(library().omm_ctx_set_membrane_barostat(self._h, pressure, surface_tension, temperature, int(xy_mode),
                                                       int(z_mode), int(frequency), int(seed)))

    def set_anisotropic_barostat(self, pressure, temperature, scale_axes=(True, True, True), frequency=25, seed=1):
        """MonteCarloAnisotropicBarostat(pressure[3] bar, temperature K, scale_axes, frequency)."""
        p = (C.c_double * 3)(*[float(x) for x in pressure])
        mask = sum(1 << a for a in range(3) if scale_axes[a])
        _check(library().omm_ctx_set_anisotropic_barostat(self._h, p, temperature, mask, int(frequency), int(seed)))

    def barostat_info(self):
        info = BarostatInfo()
        _check(library().omm_ctx_barostat_info(self._h, C.byref(info)))
        return info.as_dict()

    def barostat_attempt(self):
        _check(library().omm_ctx_barostat_attempt(self._h))

    def _query(self, fn, count, *args):
        out = np.zeros(count)
        _check(getattr(library(), fn)(self._h, *args, out.ctypes.data_as(_dp)))
        return out

    def compute_current_pressure(self):
        """Context::compute_current_pressure of the MonteCarloBarostat (bar)."""
        return float(self._query("omm_ctx_compute_current_pressure", 1)[0])

    def compute_current_anisotropic_pressure(self):
        """Pressure along x, y, z (bar) for the anisotropic barostat."""
        return self._query("omm_ctx_compute_current_anisotropic_pressure", 3)

    def compute_current_membrane_pressure(self):
        """Pressure along x, y, z (bar) for the membrane barostat."""
        return self._query("omm_ctx_compute_current_membrane_pressure", 3)

    def compute_current_flexible_pressure(self):
        """Pressure components xx yy zz xy xz yz (bar) for the flexible barostat."""
        return self._query("omm_ctx_compute_current_flexible_pressure", 6)

    def compute_stress_tensor(self, include_kinetic):
        """Stress components xx yy zz xy xz yz (bar) for the flexible barostat."""
        return self._query("omm_ctx_compute_stress_tensor", 6, int(bool(include_kinetic)))

    def box(self):
        b = np.zeros(9)
        _check(library().omm_ctx_get_box(self._h, b.ctypes.data_as(_dp)))
        return b.reshape(3, 3)

    def set_capacity(self, list_cap=0, bin_cap=0):
        """Test hook: list entries per cluster and spread-bin atoms per tile (0: unchanged)."""
        _check(library().omm_ctx_set_capacity(self._h, int(list_cap), int(bin_cap)))

    def capacity(self):
        """(list capacity, bin capacity, runs of steps replayed after an overflow)."""
        lc, bc, r = C.c_int32(), C.c_int32(), C.c_int64()
        _check(library().omm_ctx_capacity(self._h, C.byref(lc), C.byref(bc), C.byref(r)))
        return lc.value, bc.value, r.value

    def list_stats(self):
        """(tiles, j-list entries) of the current neighbour list."""
        t, e = C.c_int64(), C.c_int64()
        _check(library().omm_ctx_list_stats(self._h, C.byref(t), C.byref(e)))
        return t.value, e.value

    def launch_tuning(self):
        """The autotuner's launch shapes (dict)."""
        t = LaunchTuning()
        _check(library().omm_ctx_launch_tuning(self._h, C.byref(t)))
        return t.as_dict()

    def set_launch_tuning(self, spread_tile_edge, fft_lines_per_workgroup, fft_plane, pair_scale=1):
        """Launch shapes instead of the autotuner's (results do not change)."""
        t = LaunchTuning.of(spread_tile_edge, fft_lines_per_workgroup, fft_plane, pair_scale)
        _check(library().omm_ctx_set_launch_tuning(self._h, C.byref(t)))

    def get_state(self):
        pos = np.zeros((self.n, 3))
        vel = np.zeros((self.n, 3))
        frc = np.zeros((self.n, 3))
        e = C.c_double()
        k = C.c_int64()
        _check(library().omm_ctx_get_state(self._h, pos.ctypes.data_as(_dp), vel.ctypes.data_as(_dp),
                                           frc.ctypes.data_as(_dp), C.byref(e), C.byref(k)))
        lp = np.zeros((self.n, 3))
        _check(library().omm_ctx_get_list_positions(self._h, lp.ctypes.data_as(_dp)))
        return dict(positions=pos, velocities=vel, forces=frc, potential_energy=e.value, step=k.value,
                    box=self.box() if self.periodic else None, list_positions=lp)

    def set_state(self, state):
        """Context::set_state (context.rs:274-293) from a get_state() dict: step count, box,
        positions, velocities (each optional) and, when present, the positions the saved
        neighbour list was built from, so the trajectory continues bit for bit."""
        arrays = _state_arrays(state, self.n)
        _check(library().omm_ctx_set_state(self._h, C.c_int64(int(state.get("step", 0))), *_ptrs(arrays)))

    def kinetic_energy(self):
        e = C.c_double()
        _check(library().omm_ctx_kinetic_energy(self._h, C.byref(e)))
        return e.value

    PHASES = PHASES

    def profile(self, reps=20):
        """Per-phase average device ms (HIP events on the context stream) and the pair count."""
        ms = np.zeros(8)
        npairs = C.c_int64()
        _check(library().omm_ctx_profile(self._h, int(reps), ms.ctypes.data_as(_dp), C.byref(npairs)))
        return dict(zip(PHASES, ms.tolist())), npairs.value

    def close(self):
        if getattr(self, "_h", None):
            library().omm_ctx_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Batch:
    """Many independent MD systems (replicas) advanced by the same launches.

    The reference would hold one ``Context`` per system; replica r here follows
    the trajectory of ``Context(systems[r], seed=seeds[r])`` bit for bit.
    """

    def __init__(self, systems, temperature=300.0, friction=1.0, step_size=0.001, seeds=None,
                 precision=PRECISION_MIXED):
        lib = library()
        systems = list(systems)
        if seeds is None:
            seeds = list(range(1, len(systems) + 1))
        self.sizes = [len(s.masses) for s in systems]
        descs = (MdDesc * len(systems))()
        self._keep = []
        for r, s in enumerate(systems):
            d, keep = _md_desc(s, temperature, friction, step_size, int(seeds[r]))
            descs[r] = d
            self._keep.append(keep)
        self._descs = descs
        h = C.c_void_p()
        _check(lib.omm_batch_create(descs, len(systems), precision, C.byref(h)))
        self._h = h
        for r, s in enumerate(systems):
            self.set_positions(r, s.positions)

    def __len__(self):
        return len(self.sizes)

    def set_positions(self, replica, pos):
        p = _f64(pos).reshape(-1, 3)
        _check(library().omm_batch_set_positions(self._h, int(replica), p.ctypes.data_as(_dp)))

    def set_velocities(self, replica, vel):
        v = _f64(vel).reshape(-1, 3)
        _check(library().omm_batch_set_velocities(self._h, int(replica), v.ctypes.data_as(_dp)))

    def set_velocities_to_temperature(self, temperature, seed):
        _check(library().omm_batch_set_velocities_to_temperature(self._h, temperature, seed))

    def step(self, steps):
        _check(library().omm_batch_step(self._h, int(steps)))

    def set_integrator(self, temperature, friction, step_size):
        _check(library().omm_batch_set_integrator(self._h, temperature, friction, step_size))

    def set_graph_steps(self, steps):
        _check(library().omm_batch_set_graph_steps(self._h, int(steps)))

    def set_barostat(self, pressure, temperature, frequency=25, seed=1):
        _check(library().omm_batch_set_barostat(self._h, pressure, temperature, int(frequency), int(seed)))

    def set_flexible_barostat(self, pressure, temperature, frequency=25, seed=1):
        _check(library().omm_batch_set_flexible_barostat(self._h, pressure, temperature, int(frequency), int(seed)))

    def set_membrane_barostat(self, pressure, surface_tension, temperature, xy_mode=0, z_mode=0, frequency=25,
                              seed=1):
        _check(library().omm_batch_set_membrane_barostat(self._h, pressure, surface_tension, temperature,
                                                         int(xy_mode), int(z_mode), int(frequency), int(seed)))

    def set_anisotropic_barostat(self, pressure, temperature, scale_axes=(True, True, True), frequency=25, seed=1):
        p = (C.c_double * 3)(*[float(x) for x in pressure])
        mask = sum(1 << a for a in range(3) if scale_axes[a])
        _check(library().omm_batch_set_anisotropic_barostat(self._h, p, temperature, mask, int(frequency),
                                                            int(seed)))

    def barostat_info(self, replica):
        info = BarostatInfo()
        _check(library().omm_batch_barostat_info(self._h, int(replica), C.byref(info)))
        return info.as_dict()

    def barostat_attempt(self):
        _check(library().omm_batch_barostat_attempt(self._h))

    def _query(self, fn, per, *args):
        out = np.zeros(len(self) * per)
        _check(getattr(library(), fn)(self._h, *args, out.ctypes.data_as(_dp)))
        return out if per == 1 else out.reshape(len(self), per)

    def compute_current_pressure(self):
        """Pressure of every replica (bar), MonteCarloBarostat."""
        return self._query("omm_batch_compute_current_pressure", 1)

    def compute_current_anisotropic_pressure(self):
        return self._query("omm_batch_compute_current_anisotropic_pressure", 3)

    def compute_current_membrane_pressure(self):
        return self._query("omm_batch_compute_current_membrane_pressure", 3)

    def compute_current_flexible_pressure(self):
        return self._query("omm_batch_compute_current_flexible_pressure", 6)

    def compute_stress_tensor(self, include_kinetic):
        return self._query("omm_batch_compute_stress_tensor", 6, int(bool(include_kinetic)))

    def get_state(self, replica):
        n = self.sizes[replica]
        pos, vel, frc = np.zeros((n, 3)), np.zeros((n, 3)), np.zeros((n, 3))
        e = C.c_double()
        k = C.c_int64()
        _check(library().omm_batch_get_state(self._h, int(replica), pos.ctypes.data_as(_dp),
                                             vel.ctypes.data_as(_dp), frc.ctypes.data_as(_dp), C.byref(e),
                                             C.byref(k)))
        lp = np.zeros((n, 3))
        _check(library().omm_batch_get_list_positions(self._h, int(replica), lp.ctypes.data_as(_dp)))
        return dict(positions=pos, velocities=vel, forces=frc, potential_energy=e.value, step=k.value,
                    box=self.box(replica), list_positions=lp)

    def set_state(self, replica, state):
        """Context::set_state of one replica (the step count is shared by the batch)."""
        arrays = _state_arrays(state, self.sizes[replica])
        _check(library().omm_batch_set_state(self._h, int(replica), C.c_int64(int(state.get("step", 0))),
                                             *_ptrs(arrays)))

    def energies(self):
        out = np.zeros(len(self.sizes))
        _check(library().omm_batch_energies(self._h, out.ctypes.data_as(_dp)))
        return out

    def kinetic_energy(self, replica):
        e = C.c_double()
        _check(library().omm_batch_kinetic_energy(self._h, int(replica), C.byref(e)))
        return e.value

    def box(self, replica):
        b = np.zeros(9)
        _check(library().omm_batch_get_box(self._h, int(replica), b.ctypes.data_as(_dp)))
        return b.reshape(3, 3)

    def list_stats(self):
        t, e = C.c_int64(), C.c_int64()
        _check(library().omm_batch_list_stats(self._h, C.byref(t), C.byref(e)))
        return t.value, e.value

    def launch_tuning(self):
        t = LaunchTuning()
        _check(library().omm_batch_launch_tuning(self._h, C.byref(t)))
        return t.as_dict()

    def set_launch_tuning(self, spread_tile_edge, fft_lines_per_workgroup, fft_plane, pair_scale=1):
        t = LaunchTuning.of(spread_tile_edge, fft_lines_per_workgroup, fft_plane, pair_scale)
        _check(library().omm_batch_set_launch_tuning(self._h, C.byref(t)))

    def profile(self, reps=20):
        ms = np.zeros(8)
        npairs = C.c_int64()
        _check(library().omm_batch_profile(self._h, int(reps), ms.ctypes.data_as(_dp), C.byref(npairs)))
        return dict(zip(PHASES, ms.tolist())), npairs.value

    def close(self):
        if getattr(self, "_h", None):
            library().omm_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
