"""ctypes wrapper of the FP64 CPU oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
CPU-baseline leg of bench.py, always as the checker, never as the product path.
The functions restate crates/openmm-core/src/forces/{nonbonded,pme,neighbors,
ewald,harmonic_bond,harmonic_angle,barostat}.rs, integrators/langevin.rs and
context/pressure.rs; see oracle.h for the per-function citations and the
unpinned last-ulp notes.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# OMM_ORACLE_LIB names another build of oracle.cpp: the mutation tests load
# deliberately broken builds through it to show that the pins catch them.
LIB_PATH = os.environ.get("OMM_ORACLE_LIB") or os.path.join(_HERE, "liboracle.so")
_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int32)
_lib = None


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        L.orc_nb_evaluate.restype = C.c_int
        L.orc_nb_evaluate.argtypes = [C.c_void_p, C.c_double, _ip, _dp, _dp, C.c_int, C.c_int, C.c_int,
                                      _dp, _dp, _dp]
        L.orc_nb_pairs.restype = C.c_int64
        L.orc_nb_pairs.argtypes = [C.c_void_p, _dp, _dp, _ip, _ip, C.c_int64, C.POINTER(C.c_uint64)]
        L.orc_cell_candidates.restype = C.c_int64
        L.orc_cell_candidates.argtypes = [_dp, C.c_int32, C.c_double, _dp, _ip, _ip, C.c_int64]
        L.orc_pair_key.restype = C.c_uint64
        L.orc_pair_key.argtypes = [C.c_int32, C.c_int32]
        L.orc_effective_pme_parameters.argtypes = [C.c_double, C.c_double, _dp, _dp, _ip]
        L.orc_effective_ljpme_parameters.argtypes = [C.c_double, C.c_double, _dp, _dp, _ip]
        L.orc_ewald_parameters.argtypes = [_dp, C.c_double, C.c_double, _dp, _ip]
        L.orc_spline.argtypes = [C.c_double, _dp, _dp]
        L.orc_reciprocal_box.restype = C.c_double
        L.orc_reciprocal_box.argtypes = [_dp, _dp]
        L.orc_fft.argtypes = [_dp, C.c_int32, C.c_int32]
        L.orc_bspline_moduli.argtypes = [C.c_int32, _dp]
        L.orc_pme_mesh.restype = C.c_double
        L.orc_pme_mesh.argtypes = [_dp, _dp, C.c_int32, _dp, C.c_double, _ip, C.c_int32, _dp]
        L.orc_pme_charge_derivatives.argtypes = [_dp, _dp, C.c_int32, _dp, C.c_double, _ip, _dp]
        L.orc_ewald_reciprocal.restype = C.c_double
        L.orc_ewald_reciprocal.argtypes = [_dp, _dp, C.c_int32, _dp, C.c_double, _ip, _dp]
        L.orc_dispersion_correction.restype = C.c_double
        L.orc_dispersion_correction.argtypes = [C.c_void_p, C.c_double]
        L.orc_dispersion_reciprocal_pair.argtypes = [C.c_double, C.c_double, C.c_double, _dp, _dp]
        L.orc_minimum_image.argtypes = [_dp, _dp]
        L.orc_harmonic_bonds.restype = C.c_double
        L.orc_harmonic_bonds.argtypes = [C.c_int32, _ip, _ip, _dp, _dp, _dp, _dp, C.c_int32, _dp]
        L.orc_harmonic_angles.restype = C.c_double
        L.orc_harmonic_angles.argtypes = [C.c_int32, _ip, _ip, _ip, _dp, _dp, _dp, _dp, C.c_int32, _dp]
        L.orc_lfmiddle_step.argtypes = [C.c_int32, _dp, _dp, _dp, _dp, C.c_double, C.c_double,
                                        C.c_double, _dp]
        L.orc_constraints_solve.restype = C.c_int32
        L.orc_constraints_solve.argtypes = [C.c_int32, C.c_int32, _ip, _ip, _dp, _dp, _dp, _dp, C.c_int32,
                                            C.c_double]
        L.orc_lfmiddle_step_constrained.restype = C.c_int32
        L.orc_lfmiddle_step_constrained.argtypes = [C.c_int32, _dp, _dp, _dp, _dp, C.c_double, C.c_double,
                                                    C.c_double, _dp, C.c_int32, _ip, _ip, _dp, C.c_double]
        L.orc_boltzmann.restype = C.c_double
        L.orc_molecules.restype = C.c_int32
        L.orc_molecules.argtypes = [C.c_int32, C.c_int32, _ip, _ip, _ip]
        L.orc_scale_positions.argtypes = [C.c_int32, _dp, C.c_double, C.c_int32, _ip, C.c_int32]
        L.orc_scale_positions_axes.argtypes = [C.c_int32, _dp, _dp, _ip, C.c_int32]
        L.orc_reduced_box.restype = C.c_int32
        L.orc_reduced_box.argtypes = [_dp]
        L.orc_barostat_work.restype = C.c_double
        L.orc_barostat_work.argtypes = [C.c_double, C.c_double, C.c_double, C.c_double, C.c_int32,
                                        C.c_double]
        _lib = L
    return _lib


def _f(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.float64))


def _i(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.int32))


def _p(a, t=C.c_double):
    return None if a is None else a.ctypes.data_as(C.POINTER(t))


def nb_evaluate(desc, pos, box, include_direct=True, include_reciprocal=True, use_cells=True,
                ljpme=None):
    """NonbondedForce::evaluate -> (forces (N,3), energy, parts[5])."""
    pos = _f(pos).reshape(-1, 3)
    b = None if box is None else _f(box).reshape(9)
    forces = np.zeros_like(pos)
    energy = C.c_double()
    parts = np.zeros(5)
    la, lg = (0.0, None) if ljpme is None else (ljpme[0], _i(ljpme[1]))
    rc = lib().orc_nb_evaluate(C.byref(desc), la, _p(lg, C.c_int32), _p(pos), _p(b), int(include_direct),
                               int(include_reciprocal), int(use_cells), _p(forces), C.byref(energy), _p(parts))
    if rc != 0:
        raise AssertionError(f"oracle evaluate failed a reference assertion (code {rc})")
    return forces, energy.value, parts


def nb_pairs(desc, pos, box, with_list=False):
    pos = _f(pos).reshape(-1, 3)
    b = None if box is None else _f(box).reshape(9)
    h = C.c_uint64()
    count = lib().orc_nb_pairs(C.byref(desc), _p(pos), _p(b), None, None, 0, C.byref(h))
    if not with_list:
        return count, h.value
    oi = np.zeros(count, np.int32)
    oj = np.zeros(count, np.int32)
    lib().orc_nb_pairs(C.byref(desc), _p(pos), _p(b), _p(oi, C.c_int32), _p(oj, C.c_int32), count,
                       C.byref(h))
    return count, h.value, oi, oj


def cell_candidates(pos, cutoff, box):
    pos = _f(pos).reshape(-1, 3)
    b = None if box is None else _f(box).reshape(9)
    n = lib().orc_cell_candidates(_p(pos), len(pos), cutoff, _p(b), None, None, 0)
    oi = np.zeros(n, np.int32)
    oj = np.zeros(n, np.int32)
    lib().orc_cell_candidates(_p(pos), len(pos), cutoff, _p(b), _p(oi, C.c_int32), _p(oj, C.c_int32), n)
    return oi, oj


def pair_key(i, j):
    return int(lib().orc_pair_key(i, j))


def effective_pme_parameters(tol, cutoff, box):
    a = C.c_double()
    g = np.zeros(3, np.int32)
    lib().orc_effective_pme_parameters(tol, cutoff, _p(_f(box).reshape(9)), C.byref(a), _p(g, C.c_int32))
    return a.value, tuple(int(x) for x in g)


def effective_ljpme_parameters(tol, cutoff, box):
    a = C.c_double()
    g = np.zeros(3, np.int32)
    lib().orc_effective_ljpme_parameters(tol, cutoff, _p(_f(box).reshape(9)), C.byref(a), _p(g, C.c_int32))
    return a.value, tuple(int(x) for x in g)


def ewald_parameters(widths, cutoff, tol):
    a = C.c_double()
    g = np.zeros(3, np.int32)
    lib().orc_ewald_parameters(_p(_f(widths)), cutoff, tol, C.byref(a), _p(g, C.c_int32))
    return a.value, tuple(int(x) for x in g)


def spline(fraction):
    w = np.zeros(5)
    d = np.zeros(5)
    lib().orc_spline(fraction, _p(w), _p(d))
    return w, d


def reciprocal_box(box):
    r = np.zeros(9)
    v = lib().orc_reciprocal_box(_p(_f(box).reshape(9)), _p(r))
    return r.reshape(3, 3), v


def fft(x, inverse=False):
    a = np.ascontiguousarray(np.asarray(x, dtype=np.complex128).copy())
    lib().orc_fft(a.view(np.float64).ctypes.data_as(_dp), len(a), int(inverse))
    return a


def bspline_moduli(n):
    out = np.zeros(n)
    lib().orc_bspline_moduli(n, _p(out))
    return out


def pme_mesh(pos, amplitudes, box, alpha, grid, dispersion=False, with_forces=True):
    pos = _f(pos).reshape(-1, 3)
    f = np.zeros_like(pos) if with_forces else None
    g = _i(grid)
    e = lib().orc_pme_mesh(_p(pos), _p(_f(amplitudes)), len(pos), _p(_f(box).reshape(9)), alpha,
                           _p(g, C.c_int32), int(dispersion), _p(f))
    return (f, e) if with_forces else e


def pme_charge_derivatives(pos, charges, box, alpha, grid):
    pos = _f(pos).reshape(-1, 3)
    out = np.zeros(len(pos))
    lib().orc_pme_charge_derivatives(_p(pos), _p(_f(charges)), len(pos), _p(_f(box).reshape(9)), alpha,
                                     _p(_i(grid), C.c_int32), _p(out))
    return out


def ewald_reciprocal(pos, charges, widths, alpha, limits):
    pos = _f(pos).reshape(-1, 3)
    f = np.zeros_like(pos)
    e = lib().orc_ewald_reciprocal(_p(pos), _p(_f(charges)), len(pos), _p(_f(widths)), alpha,
                                   _p(_i(limits), C.c_int32), _p(f))
    return f, e


def dispersion_correction(desc, volume):
    return lib().orc_dispersion_correction(C.byref(desc), volume)


def dispersion_reciprocal_pair(c6, r2, alpha):
    """(energy, dE/dr / r) of the dispersion complement (nonbonded.rs:1158-1191)."""
    e, f = C.c_double(), C.c_double()
    lib().orc_dispersion_reciprocal_pair(float(c6), float(r2), float(alpha), C.byref(e), C.byref(f))
    return e.value, f.value


def minimum_image(delta, box):
    d = _f(delta).copy()
    lib().orc_minimum_image(_p(d), _p(_f(box).reshape(9)))
    return d


def harmonic_bonds(bonds, params, pos, box=None, use_pbc=False):
    pos = _f(pos).reshape(-1, 3)
    f = np.zeros_like(pos)
    b = _i(bonds)
    p = _f(params)
    bx = None if box is None else _f(box).reshape(9)
    e = lib().orc_harmonic_bonds(len(b), _p(_i(b[:, 0]), C.c_int32), _p(_i(b[:, 1]), C.c_int32),
                                 _p(_f(p[:, 0])), _p(_f(p[:, 1])), _p(pos), _p(bx), int(use_pbc), _p(f))
    return f, e


def harmonic_angles(angles, params, pos, box=None, use_pbc=False):
    pos = _f(pos).reshape(-1, 3)
    f = np.zeros_like(pos)
    a = _i(angles)
    p = _f(params)
    bx = None if box is None else _f(box).reshape(9)
    e = lib().orc_harmonic_angles(len(a), _p(_i(a[:, 0]), C.c_int32), _p(_i(a[:, 1]), C.c_int32),
                                  _p(_i(a[:, 2]), C.c_int32), _p(_f(p[:, 0])), _p(_f(p[:, 1])), _p(pos),
                                  _p(bx), int(use_pbc), _p(f))
    return f, e


def lfmiddle_step(pos, vel, forces, masses, dt, temperature, friction, normals=None):
    pos = _f(pos).copy()
    vel = _f(vel).copy()
    lib().orc_lfmiddle_step(len(masses), _p(pos), _p(vel), _p(_f(forces)), _p(_f(masses)), dt, temperature,
                            friction, _p(None if normals is None else _f(normals)))
    return pos, vel


def constraints_solve(constraints, masses, reference, values, velocities, tolerance):
    """ConstraintSolver::solve (constraints.rs:80-149); constraints = [(i, j, d)].
    Returns the projected values; raises on the reference's errors."""
    c = list(constraints)
    c1 = _i([x[0] for x in c])
    c2 = _i([x[1] for x in c])
    d = _f([x[2] for x in c])
    out = _f(values).copy()
    rc = lib().orc_constraints_solve(len(masses), len(c), _p(c1, C.c_int32), _p(c2, C.c_int32), _p(d),
                                     _p(_f(masses)), _p(_f(reference)), _p(out), int(velocities), tolerance)
    if rc:
        raise AssertionError("constraint solve failed a reference check")
    return out


def lfmiddle_step_constrained(pos, vel, forces, masses, dt, temperature, friction, constraints, tolerance,
                              normals=None):
    pos = _f(pos).copy()
    vel = _f(vel).copy()
    c = list(constraints)
    c1 = _i([x[0] for x in c])
    c2 = _i([x[1] for x in c])
    d = _f([x[2] for x in c])
    rc = lib().orc_lfmiddle_step_constrained(len(masses), _p(pos), _p(vel), _p(_f(forces)), _p(_f(masses)), dt,
                                             temperature, friction, _p(None if normals is None else _f(normals)),
                                             len(c), _p(c1, C.c_int32), _p(c2, C.c_int32), _p(d), tolerance)
    if rc:
        raise AssertionError("constrained step failed a reference check")
    return pos, vel


def boltzmann():
    return lib().orc_boltzmann()


def molecules(n, bonds):
    b = _i(bonds).reshape(-1, 2)
    out = np.zeros(n, np.int32)
    k = lib().orc_molecules(n, len(b), _p(_i(b[:, 0]), C.c_int32), _p(_i(b[:, 1]), C.c_int32),
                            _p(out, C.c_int32))
    return out, k


def scale_positions(pos, scale, rigid, molecule_of, nmol):
    p = _f(pos).copy()
    lib().orc_scale_positions(len(p), _p(p), scale, int(rigid), _p(_i(molecule_of), C.c_int32), nmol)
    return p


def reduced_box(box):
    """flexible_barostat.rs:15-41; None when the box has no reduced form."""
    v = _f(box).reshape(9).copy()
    return None if lib().orc_reduced_box(_p(v)) else v.reshape(3, 3)


def scale_positions_axes(pos, scale, molecule_of, nmol):
    """Rigid-molecule scaling by a diagonal transform (context.rs:976-1029)."""
    p = _f(pos).copy()
    lib().orc_scale_positions_axes(len(p) // 3 if p.ndim == 1 else len(p), _p(p), _p(_f(scale)),
                                   _p(_i(molecule_of), C.c_int32), int(nmol))
    return p

BAR_NM3_TO_KJ_PER_MOL = 0.0602214076
# (row, column) of the tensor components xx yy zz xy xz yz (pressure.rs:10).
COMPONENTS = ((0, 0), (1, 1), (2, 2), (0, 1), (0, 2), (1, 2))


def transform_positions(pos, transform, molecule_of, nmol):
    """Rigid-molecule move by a 3x3 transform T (context.rs:976-1029): every
    atom of a molecule is shifted by (T - I) c, c the molecule's geometric
    centre."""
    p = _f(pos).reshape(-1, 3)
    mol = np.asarray(molecule_of)
    sums = np.zeros((nmol, 3))
    np.add.at(sums, mol, p)
    centres = sums * (1.0 / np.bincount(mol, minlength=nmol))[:, None]
    return p + (centres @ (np.asarray(transform, float) - np.eye(3)).T)[mol]


def kinetic_tensor(vel, masses, molecule_of, nmol):
    """pressure.rs:107-131 with rigid molecules: 1/2 P_i P_j / M summed over
    the molecules of mass M > 0 (components xx yy zz xy xz yz)."""
    v = _f(vel).reshape(-1, 3)
    m = _f(masses)
    mol = np.asarray(molecule_of)
    P = np.zeros((nmol, 3))
    np.add.at(P, mol, m[:, None] * v)
    M = np.bincount(mol, weights=m, minlength=nmol)
    keep = M > 0
    return np.array([np.sum(0.5 * P[keep, i] * P[keep, j] / M[keep]) for i, j in COMPONENTS])


def current_pressure(energy, pos, vel, masses, box, molecule_of, nmol):
    """Context::compute_current_pressure (pressure.rs:14-42), bar; energy(pos,
    box) is the potential energy of the system."""
    box = _f(box).reshape(3, 3)
    e = [energy(transform_positions(pos, s * np.eye(3), molecule_of, nmol), box * s) for s in (1.001, 0.999)]
    kin = kinetic_tensor(vel, masses, molecule_of, nmol)[:3].sum()
    v = box[0, 0] * box[1, 1] * box[2, 2]
    dv = v * (1.001 ** 3 - 0.999 ** 3)
    return ((2.0 / 3.0) * kin / v - (e[0] - e[1]) / dv) / BAR_NM3_TO_KJ_PER_MOL


def axis_pressures(energy, pos, vel, masses, box, molecule_of, nmol):
    """Anisotropic / membrane pressure per axis (pressure.rs:78-105), bar."""
    box = _f(box).reshape(3, 3)
    kin = kinetic_tensor(vel, masses, molecule_of, nmol)
    v = box[0, 0] * box[1, 1] * box[2, 2]
    out = np.zeros(3)
    for axis in range(3):
        e = []
        for f in (1.001, 0.999):
            scale = np.ones(3)
            scale[axis] = f
            e.append(energy(transform_positions(pos, np.diag(scale), molecule_of, nmol), box * scale[None, :]))
        out[axis] = (2.0 * kin[axis] / v - (e[0] - e[1]) / (0.002 * v)) / BAR_NM3_TO_KJ_PER_MOL
    return out


def flexible_pressure(energy, pos, vel, masses, box, molecule_of, nmol):
    """Context::compute_current_flexible_pressure (pressure.rs:133-170), bar."""
    box = _f(box).reshape(3, 3)
    kin = kinetic_tensor(vel, masses, molecule_of, nmol)
    v = box[0, 0] * box[1, 1] * box[2, 2]
    out = np.zeros(6)
    for c, (column, row) in enumerate(COMPONENTS):
        e = [0.0, 0.0]
        if c < 3 or box[row, column] != 0.0:
            for k, delta in enumerate((0.001, -0.001)):
                trial = box.copy()
                trial[row, column] *= 1.0 + delta
                trial = reduced_box(trial)
                if c < 3:
                    t = np.eye(3)
                    t[c, c] += delta
                    e[k] = energy(transform_positions(pos, t, molecule_of, nmol), trial)
                else:
                    e[k] = energy(_f(pos).reshape(-1, 3), trial)
        out[c] = (2.0 * kin[c] - (e[0] - e[1]) / 0.002) / (v * BAR_NM3_TO_KJ_PER_MOL)
    return out


def stress_tensor(energy, pos, vel, masses, box, molecule_of, nmol, include_kinetic):
    """Context::compute_stress_tensor (pressure.rs:172-209), bar."""
    box = _f(box).reshape(3, 3)
    kin = kinetic_tensor(vel, masses, molecule_of, nmol) if include_kinetic else np.zeros(6)
    v = box[0, 0] * box[1, 1] * box[2, 2]
    out = np.zeros(6)
    for c, (i, j) in enumerate(COMPONENTS):
        e = []
        for delta in (0.001, -0.001):
            trial = box.copy()
            trial[:, i] += delta * trial[:, j]
            trial = reduced_box(trial)
            t = np.eye(3)
            t[i, j] += delta
            e.append(energy(transform_positions(pos, t, molecule_of, nmol), trial))
        out[c] = ((e[0] - e[1]) / 0.002 - 2.0 * kin[c]) / (v * BAR_NM3_TO_KJ_PER_MOL)
    return out


def barostat_work(de, pressure, v_old, v_new, scaled, temperature):
    return lib().orc_barostat_work(de, pressure, v_old, v_new, scaled, temperature)


def water_energy_forces(system, ljpme=None, include_bonded=True):
    """Total FP64 forces/energy of a WaterSystem (nonbonded with its frozen PME
    parameters + harmonic bonds + angles): the reference Context's force sum."""
    desc = system.nonbonded.desc()
    f, e, parts = nb_evaluate(desc, system.positions, system.box, ljpme=ljpme)
    if include_bonded:
        fb, eb = harmonic_bonds(system.bonds, system.bond_params, system.positions)
        fa, ea = harmonic_angles(system.angles, system.angle_params, system.positions)
        f = f + fb + fa
        e = e + eb + ea
    return f, e
