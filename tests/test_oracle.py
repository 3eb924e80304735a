"""Pins the FP64 CPU oracle against known answers.

The reference commits no golden value fixtures (its generated fixtures are
git-ignored and need an external OpenMM build), so the oracle is pinned by the
reference's own known-answer tests restated here (nonbonded.rs:1338-1534,
pme.rs:428-498, neighbors.rs:184-264, harmonic_bond.rs:187-223,
langevin.rs:300-332) and by checks the mathematics fixes: the NaCl Madelung
constant, PME -> Ewald convergence, brute-force DFTs, finite-difference
gradients, partition of unity and exact leapfrog trajectories.
"""
import math

import numpy as np
import pytest

import ommhip
import oracle as orc
from helpers import LJ_BOX, lj_energy, lj_pair_closed_forms, lj_pair_system
from ommhip import NonbondedForce

K = 138.93545764438198  # ONE_4PI_EPS0, nonbonded.rs:16


def make_force(particles, method="NoCutoff", cutoff=1.0, tol=5e-4, exceptions=(), **kw):
    f = NonbondedForce()
    f.set_nonbonded_method(method)
    f.set_cutoff_distance(cutoff)
    f.set_ewald_error_tolerance(tol)
    for p in particles:
        f.add_particle(*p)
    for e in exceptions:
        f.add_exception(*e)
    for k, v in kw.items():
        getattr(f, k)(*v) if isinstance(v, tuple) else getattr(f, k)(v)
    return f


# --- nonbonded.rs unit tests (known answers) ---------------------------------

def test_coulomb_closed_form():
    # nonbonded.rs:1459-1472
    f = make_force([(0.5, 1.0, 0.0), (-1.5, 1.0, 0.0)])
    F, E, _ = orc.nb_evaluate(f.desc(), [[0, 0, 0], [2, 0, 0]], None)
    expected_f = -K * (-0.75) / 4.0
    assert F[0, 0] == pytest.approx(expected_f, rel=1e-14)
    assert F[1, 0] == pytest.approx(-expected_f, rel=1e-14)
    assert E == pytest.approx(K * (-0.75) / 2.0, rel=1e-14)


def test_lj_closed_form():
    # nonbonded.rs:1476-1494: sigma = (1.2+1.4)/2, eps = sqrt(1*2), r = 2
    f = make_force([(0.0, 1.2, 1.0), (0.0, 1.4, 2.0)])
    F, E, _ = orc.nb_evaluate(f.desc(), [[0, 0, 0], [2, 0, 0]], None)
    sigma, eps, x = 1.3, math.sqrt(2.0), 1.3 / 2.0
    assert E == pytest.approx(4 * eps * (x**12 - x**6), rel=1e-13)
    ef = 4 * eps * (-12 * x**12 + 6 * x**6) / 2.0
    assert F[0, 0] == pytest.approx(ef, rel=1e-13)
    assert F[1, 0] == pytest.approx(-ef, rel=1e-13)


def test_exclusions_zero_interaction():
    # nonbonded.rs:1498-1515
    f = make_force([(1.0, 1.0, 1.0)] * 3, exceptions=[(0, 1, 0.0, 1.0, 0.0), (0, 2, 0.0, 1.0, 0.0)])
    F, _, _ = orc.nb_evaluate(f.desc(), [[0, 0, 0], [1, 0, 0], [2, 0, 0]], None)
    assert np.all(F[0] == 0.0)


def test_newton_third_law():
    # nonbonded.rs:1519-1533
    f = make_force([(0.5, 0.3, 0.5), (-0.3, 0.4, 0.8), (0.1, 0.35, 0.3)])
    F, _, _ = orc.nb_evaluate(f.desc(), [[0, 0, 0], [0.5, 0, 0], [0, 0.7, 0]], None)
    assert np.abs(F.sum(0)).max() < 1e-10


@pytest.mark.parametrize("method", ["Ewald", "PME", "LJPME"])
def test_reciprocal_methods_match_energy_gradients(method):
    # nonbonded.rs:1365-1413 (positions kept in FP64 here)
    coords = np.array([[0.1, 0.2, 0.3], [0.6, 0.4, 0.2], [2.8, 0.7, 0.4]])
    f = make_force([(0.4, 0.3, 0.5), (-0.3, 0.32, 0.6), (0.1, 0.34, 0.7)], method=method, cutoff=1.2,
                   tol=1e-5, exceptions=[(0, 2, -0.02, 0.32, 0.4)])
    f.set_exceptions_use_periodic_boundary_conditions(True)
    skew = 0.0 if method == "Ewald" else 0.3
    box = np.array([[3.0, 0, 0], [skew, 3.0, 0], [-skew, skew, 3.0]])
    d = f.desc()
    F, _, _ = orc.nb_evaluate(d, coords, box)
    h = 1e-5
    for p in range(3):
        for a in range(3):
            cp, cm = coords.copy(), coords.copy()
            cp[p, a] += h
            cm[p, a] -= h
            ep = orc.nb_evaluate(d, cp, box)[1]
            em = orc.nb_evaluate(d, cm, box)[1]
            fd = -(ep - em) / (2 * h)
            assert F[p, a] == pytest.approx(fd, rel=1e-5, abs=1e-5)


# --- Ewald / PME against mathematics -------------------------------------

MADELUNG_NACL = 1.747564594633182


def nacl_lattice(cells, a):
    basis = []
    for x in range(2):
        for y in range(2):
            for z in range(2):
                basis.append(((x, y, z), 1.0 if (x + y + z) % 2 == 0 else -1.0))
    pos, q = [], []
    for i in range(cells):
        for j in range(cells):
            for k in range(cells):
                for (x, y, z), s in basis:
                    pos.append([(2 * i + x) * a / 2, (2 * j + y) * a / 2, (2 * k + z) * a / 2])
                    q.append(s)
    return np.array(pos), np.array(q)


@pytest.mark.parametrize("method", ["Ewald", "PME"])
def test_nacl_madelung_constant(method):
    a = 0.564
    pos, q = nacl_lattice(3, a)
    L = 3 * a
    box = np.diag([L, L, L])
    f = make_force([(qi, 0.3, 0.0) for qi in q], method=method, cutoff=0.8, tol=1e-9)
    if method == "PME":
        alpha = math.sqrt(-math.log(2e-9)) / 0.8
        f.set_pme_parameters(alpha, 96, 96, 96)
    _, E, _ = orc.nb_evaluate(f.desc(), pos, box)
    expected = -(len(q) / 2) * MADELUNG_NACL * K / (a / 2)
    assert E == pytest.approx(expected, rel=2e-7)


def test_pme_converges_to_ewald():
    rng = np.random.default_rng(5)
    n = 40
    pos = rng.uniform(0, 2.2, (n, 3))
    q = rng.normal(size=n)
    q -= q.mean()
    L = 2.2
    box = np.diag([L, L, L])
    alpha = 3.2
    _, e_ewald = orc.ewald_reciprocal(pos, q, [L, L, L], alpha, [31, 31, 31])
    errs = []
    for g in (16, 24, 36, 54):
        _, e_pme = orc.pme_mesh(pos, q, box, alpha, [g, g, g])
        errs.append(abs(e_pme - e_ewald) / abs(e_ewald))
    assert errs[-1] < 2e-6
    assert errs[0] > errs[1] > errs[2] > errs[3]


def test_pme_half_sum_q_phi_and_charge_derivatives():
    # pme.rs:432-484
    positions = np.array([0.1, 0.2, 0.3, 0.7, 0.4, 1.2, 1.4, 0.3, 0.8]).reshape(3, 3)
    for tilt in (0.0, 0.35):
        box = np.array([[3.0, 0, 0], [tilt, 3.2, 0], [-tilt, 0.1, 3.4]])
        alpha, grid = 2.4, [12, 14, 16]
        q = np.array([0.0, 0.3, -0.4])
        der = orc.pme_charge_derivatives(positions, q, box, alpha, grid)
        _, e = orc.pme_mesh(positions, q, box, alpha, grid)
        assert abs(e - 0.5 * (q * der).sum()) < 1e-12
        for i in range(3):
            qp, qm = q.copy(), q.copy()
            qp[i] += 1e-5
            qm[i] -= 1e-5
            ep = orc.pme_mesh(positions, qp, box, alpha, grid, with_forces=False)
            em = orc.pme_mesh(positions, qm, box, alpha, grid, with_forces=False)
            assert abs((ep - em) / 2e-5 - der[i]) < 2e-8


def test_spline_partition_and_derivatives():
    # pme.rs:486-497
    for frac in (0.0, 0.13, 0.5, 0.999):
        w, dw = orc.spline(frac)
        assert abs(w.sum() - 1.0) < 1e-14
        assert abs(dw.sum()) < 1e-14
        wp, _ = orc.spline(frac + 1e-6)
        wm, _ = orc.spline(frac - 1e-6)
        assert np.abs((wp - wm) / 2e-6 - dw).max() < 1e-9


def test_spline_is_cardinal_bspline_of_order_five():
    # M5 sampled at u, u+1, ..., u+4 (closed-form piecewise quartic, de Boor).
    def m5(x):
        return sum((-1) ** k * math.comb(5, k) * max(x - k, 0.0) ** 4 for k in range(6)) / 24.0

    for frac in (0.0, 0.21, 0.77):
        w, _ = orc.spline(frac)
        expected = [m5(frac + 4 - i) for i in range(5)]
        assert np.allclose(w, expected, atol=1e-14)


@pytest.mark.parametrize("n", [5, 6, 11, 12, 22, 29, 40, 55, 116])
def test_fft_matches_dft(n):
    rng = np.random.default_rng(n)
    x = rng.normal(size=n) + 1j * rng.normal(size=n)
    assert np.allclose(orc.fft(x), np.fft.fft(x), rtol=1e-12, atol=1e-11)
    assert np.allclose(orc.fft(x, inverse=True), n * np.fft.ifft(x), rtol=1e-12, atol=1e-11)


def test_bspline_moduli_are_the_spline_structure_factor():
    for n in (10, 22, 55):
        w, _ = orc.spline(0.0)
        expected = np.abs(np.fft.ifft(np.concatenate([w, np.zeros(n - 5)])) * n) ** 2
        # pme.rs:253-257: the zero of an even grid's Nyquist mode takes the
        # mean of its neighbours.
        for k in range(n):
            if expected[k] < 1e-7:
                expected[k] = 0.5 * (expected[(k - 1) % n] + expected[(k + 1) % n])
        assert np.allclose(orc.bspline_moduli(n), expected, rtol=1e-12)


def test_effective_pme_parameters_table():
    # alpha = sqrt(-ln(2 tol))/rc; n = max(6, ceil(2 alpha L / (3 tol^0.2)))
    a, g = orc.effective_pme_parameters(5e-4, 0.9, np.diag([4.48] * 3))
    assert a == pytest.approx(2.920290, abs=1e-6) and g == (40, 40, 40)
    a, g = orc.effective_pme_parameters(5e-4, 0.9, np.diag([6.17] * 3))
    assert g == (55, 55, 55)
    a, g = orc.effective_pme_parameters(5e-4, 1.0, np.diag([14.42] * 3))
    assert a == pytest.approx(2.628261, abs=1e-6) and g == (116, 116, 116)


def test_direct_plus_reciprocal_equals_total():
    # tests/force_groups.rs:57-97
    rng = np.random.default_rng(2)
    pos = rng.uniform(0, 2.0, (30, 3))
    q = rng.normal(size=30)
    f = make_force([(qi, 0.3, 0.5) for qi in q], method="PME", cutoff=0.9, tol=1e-4)
    box = np.diag([2.0] * 3)
    d = f.desc()
    F, E, _ = orc.nb_evaluate(d, pos, box)
    Fd, Ed, _ = orc.nb_evaluate(d, pos, box, include_reciprocal=False)
    Fr, Er, _ = orc.nb_evaluate(d, pos, box, include_direct=False)
    assert E == pytest.approx(Ed + Er, abs=1e-10)
    assert np.allclose(F, Fd + Fr, atol=1e-10)


def test_dispersion_correction_matches_pair_sum():
    rng = np.random.default_rng(3)
    n = 7
    sig = rng.uniform(0.2, 0.4, n)
    eps = rng.uniform(0.0, 1.0, n)
    rc = 0.9
    f = make_force([(0.0, s, e) for s, e in zip(sig, eps)], method="PME", cutoff=rc)
    # Pair sum over all ordered pairs incl. i == j, plus the diagonal, of
    # eps_ij sigma_ij^p, with Lorentz-Berthelot combination.
    def pm(p):
        s = 0.0
        for i in range(n):
            for j in range(n):
                s += math.sqrt(eps[i] * eps[j]) * (0.5 * (sig[i] + sig[j])) ** p
        return s + (eps * sig**p).sum()

    expected = 8 * math.pi * n / (n + 1) * (pm(12) * rc**-9 / 9 - pm(6) * rc**-3 / 3) / 2.5
    assert orc.dispersion_correction(f.desc(), 2.5) == pytest.approx(expected, rel=1e-12)


# --- neighbour cells and the pair set --------------------------------------

@pytest.mark.parametrize("box", [None, [[2.0, 0, 0], [0.7, 2.0, 0], [-0.4, 0.5, 2.0]], np.eye(3)])
@pytest.mark.parametrize("cutoff", [0.2, 0.7, 1.1])
def test_cell_candidates_cover_cutoff_pairs(box, cutoff):
    # neighbors.rs:227-263
    pos = (np.array([((i * 137 + 37) % 997) for i in range(240)], dtype=float) / 121.0 - 4.0).reshape(80, 3)
    ci, cj = orc.cell_candidates(pos, cutoff, None if box is None else np.array(box))
    cand = set(zip(ci.tolist(), cj.tolist()))
    assert len(cand) == len(ci), "duplicate candidate"
    assert all(i < j for i, j in cand)
    for i in range(80):
        for j in range(i + 1, 80):
            d = pos[j] - pos[i]
            if box is not None:
                d = orc.minimum_image(d, np.array(box))
            if d @ d < cutoff * cutoff:
                assert (i, j) in cand


def test_pair_set_cells_equal_exhaustive():
    rng = np.random.default_rng(9)
    n = 600
    L = 2.5
    pos = rng.uniform(-1, L + 1, (n, 3))
    f = make_force([(0.1, 0.3, 0.5)] * n, method="CutoffPeriodic", cutoff=0.8,
                   exceptions=[(i, i + 1, 0.0, 0.3, 0.0) for i in range(0, n - 1, 3)])
    box = np.array([[L, 0, 0], [0.3, L, 0], [-0.2, 0.4, L]])
    cnt, h, pi, pj = orc.nb_pairs(f.desc(), pos, box, with_list=True)
    excl = {(i, i + 1) for i in range(0, n - 1, 3)}
    expected = set()
    for i in range(n):
        for j in range(i + 1, n):
            if (i, j) in excl:
                continue
            d = orc.minimum_image(pos[j] - pos[i], box)
            if d[0] * d[0] + d[1] * d[1] + d[2] * d[2] < 0.64:
                expected.add((i, j))
    assert set(zip(pi.tolist(), pj.tolist())) == expected
    assert cnt == len(expected)
    assert h == sum(orc.pair_key(i, j) for i, j in expected) % 2**64


# --- bonded, integrator, barostat ------------------------------------------

def test_harmonic_bonds():
    # harmonic_bond.rs:187-208
    pos = np.array([[0, 2, 0], [0, 0, 0], [1, 0, 0]], float)
    F, E = orc.harmonic_bonds([[0, 1], [1, 2]], [[1.5, 0.8], [1.2, 0.7]], pos)
    assert F[0, 1] == pytest.approx(-0.4)
    assert F[2, 0] == pytest.approx(0.14)
    assert F[1, 0] == pytest.approx(-F[0, 0] - F[2, 0])
    assert E == pytest.approx(0.5 * 0.8 * 0.25 + 0.5 * 0.7 * 0.04)


def test_harmonic_bond_periodic():
    # harmonic_bond.rs:211-223
    _, E = orc.harmonic_bonds([[0, 1]], [[1.2, 0.8]], np.array([[0, 2, 0], [0, 0, 0]], float),
                              box=np.diag([3.0] * 3), use_pbc=True)
    assert E == pytest.approx(0.5 * 0.8 * 0.04)


def test_harmonic_angle_forces_are_gradients():
    rng = np.random.default_rng(4)
    pos = rng.normal(size=(3, 3)) * 0.1
    ang, par = [[0, 1, 2]], [[1.9, 300.0]]
    F, E = orc.harmonic_angles(ang, par, pos)
    v1, v2 = pos[0] - pos[1], pos[2] - pos[1]
    th = math.acos(v1 @ v2 / np.linalg.norm(v1) / np.linalg.norm(v2))
    assert E == pytest.approx(0.5 * 300.0 * (th - 1.9) ** 2, rel=1e-12)
    h = 1e-6
    for p in range(3):
        for a in range(3):
            pp, pm = pos.copy(), pos.copy()
            pp[p, a] += h
            pm[p, a] -= h
            fd = -(orc.harmonic_angles(ang, par, pp)[1] - orc.harmonic_angles(ang, par, pm)[1]) / (2 * h)
            assert F[p, a] == pytest.approx(fd, rel=1e-6, abs=1e-6)


def test_lfmiddle_zero_temperature_is_leapfrog():
    rng = np.random.default_rng(1)
    x = rng.normal(size=(5, 3))
    v = rng.normal(size=(5, 3))
    m = rng.uniform(1, 16, 5)
    f = rng.normal(size=(5, 3))
    dt = 0.002
    x1, v1 = orc.lfmiddle_step(x, v, f, m, dt, 0.0, 0.0)
    ve = v + f / m[:, None] * dt
    assert np.allclose(v1, ve, rtol=1e-15, atol=1e-15)
    assert np.allclose(x1, x + dt * ve, rtol=1e-14, atol=1e-15)


def test_lfmiddle_temperature_statistics():
    # langevin.rs:300-332: <KE> = N_dof kT / 2 within 12 %
    rng = np.random.default_rng(7)
    n, T, m = 32, 300.0, 12.0
    x = np.zeros((n, 3))
    v = np.zeros((n, 3))
    f = np.zeros((n, 3))
    masses = np.full(n, m)
    kes = []
    for step in range(3000):
        x, v = orc.lfmiddle_step(x, v, f, masses, 0.01, T, 3.0, rng.normal(size=3 * n))
        if step >= 1000:
            kes.append(0.5 * m * (v * v).sum())
    expected = 0.5 * 3 * n * orc.boltzmann() * T
    assert np.mean(kes) == pytest.approx(expected, rel=0.12)


def test_molecules_and_rigid_scaling():
    mol, k = orc.molecules(7, [[0, 1], [0, 2], [3, 4], [4, 5]])
    assert k == 3 and mol.tolist() == [0, 0, 0, 1, 1, 1, 2]
    rng = np.random.default_rng(0)
    pos = rng.normal(size=(7, 3))
    out = orc.scale_positions(pos, 1.01, True, mol, k)
    for mm in range(3):
        sel = mol == mm
        # internal geometry kept, centroid scaled
        assert np.allclose(out[sel] - out[sel].mean(0), pos[sel] - pos[sel].mean(0), atol=1e-14)
        assert np.allclose(out[sel].mean(0), 1.01 * pos[sel].mean(0), atol=1e-14)
    assert np.allclose(orc.scale_positions(pos, 1.01, False, mol, k), 1.01 * pos, rtol=1e-15)


def test_barostat_work():
    kt = orc.boltzmann() * 300.0
    w = orc.barostat_work(1.5, 1.0, 100.0, 101.0, 10, 300.0)
    assert w == pytest.approx(1.5 + 0.0602214076 * 1.0 - 10 * kt * math.log(1.01), rel=1e-14)


def _two_particles(q0, s0, e0):
    f = ommhip.NonbondedForce()
    f.add_particle(q0, s0, e0)
    f.add_particle(-0.3, 0.35, 0.5)
    return f


def test_parameter_offsets_equal_explicit_parameters():
    # nonbonded.rs:1437-1455: an offset at value v equals the explicit parameters.
    pos = np.array([[0.1, 0.2, 0.3], [0.6, 0.4, 0.2]])
    for value in (0.25, 0.7, -0.2, 0.25):
        f = _two_particles(0.2, 0.3, 0.4)
        f.add_global_parameter("lambda", value)
        f.add_particle_parameter_offset("lambda", 0, 0.4, 0.02, 0.2)
        explicit = _two_particles(0.2 + 0.4 * value, 0.3 + 0.02 * value, 0.4 + 0.2 * value)
        F1, E1, _ = orc.nb_evaluate(f.desc(), pos, None)
        F2, E2, _ = orc.nb_evaluate(explicit.desc(), pos, None)
        assert E1 == E2
        assert np.array_equal(F1, F2)


def test_later_offset_of_same_target_replaces_earlier():
    # nonbonded.rs:401-405: offsets walked last-first, the first seen per (parameter, target) wins.
    pos = np.array([[0.1, 0.2, 0.3], [0.6, 0.4, 0.2]])
    f = _two_particles(0.2, 0.3, 0.4)
    f.add_global_parameter("lambda", 0.5)
    f.add_particle_parameter_offset("lambda", 0, 10.0, 10.0, 10.0)
    f.add_particle_parameter_offset("lambda", 0, 0.4, 0.0, 0.0)
    explicit = _two_particles(0.2 + 0.4 * 0.5, 0.3, 0.4)
    assert orc.nb_evaluate(f.desc(), pos, None)[1] == orc.nb_evaluate(explicit.desc(), pos, None)[1]


def test_exception_offsets_equal_explicit_parameters():
    pos = np.array([[0.1, 0.2, 0.3], [0.6, 0.4, 0.2]])
    f = _two_particles(0.2, 0.3, 0.4)
    f.add_exception(0, 1, -0.1, 0.3, 0.2)
    f.add_global_parameter("mu", 0.3)
    f.add_exception_parameter_offset("mu", 0, 0.5, 0.1, 0.4)
    explicit = _two_particles(0.2, 0.3, 0.4)
    explicit.add_exception(0, 1, -0.1 + 0.3 * 0.5, 0.3 + 0.3 * 0.1, 0.2 + 0.3 * 0.4)
    assert orc.nb_evaluate(f.desc(), pos, None)[1] == orc.nb_evaluate(explicit.desc(), pos, None)[1]


# --- constraints (SHAKE / RATTLE) ------------------------------------------

def test_constraint_projection_known_answer():
    # tests/constraints.rs:15-38: masses 1 and 2 joined at 1 nm, started 1.2 nm apart.
    masses = [1.0, 2.0]
    c = [(0, 1, 1.0)]
    pos = np.array([[0.0, 0, 0], [1.2, 0, 0]])
    vel = np.array([[1.0, 2.0, 0.0], [-1.0, -1.0, 0.0]])
    p = orc.constraints_solve(c, masses, pos, pos, False, 1e-6).reshape(-1, 3)
    v = orc.constraints_solve(c, masses, p, vel, True, 1e-6).reshape(-1, 3)
    assert p[1, 0] - p[0, 0] == pytest.approx(1.0, abs=2e-6)
    assert p[0, 0] + 2.0 * p[1, 0] == pytest.approx(2.4, abs=2e-6)
    assert v[1, 0] - v[0, 0] == pytest.approx(0.0, abs=2e-6)
    assert v[0, 0] + 2.0 * v[1, 0] == pytest.approx(-1.0, abs=2e-6)
    assert v[0, 1] == 2.0 and v[1, 1] == -1.0


def test_constraint_projection_errors():
    # tests/constraints.rs:77-110: coincident particles are singular; a
    # massive-massless constraint is rejected.
    with pytest.raises(AssertionError):
        orc.constraints_solve([(0, 1, 1.0)], [1.0, 1.0], np.zeros(6), np.zeros(6), False, 1e-5)
    with pytest.raises(AssertionError):
        orc.constraints_solve([(0, 1, 1.0)], [1.0, 0.0], np.arange(6.0), np.arange(6.0), False, 1e-5)


def test_constrained_lfmiddle_keeps_water_rigid():
    # tests/constraints.rs:41-74 with LangevinMiddle: the coupled triangle
    # keeps its lengths over 100 steps; the projections conserve momentum.
    masses = np.array([16.0, 1.0, 1.0])
    c = [(0, 1, 0.1), (0, 2, 0.1), (1, 2, 0.16)]
    pos = np.array([[0.0, 0, 0], [0.08, 0.06, 0.0], [-0.08, 0.06, 0.0]])
    vel = np.tile([0.2, 0.1, 0.0], (3, 1))
    rng = np.random.default_rng(3)
    for _ in range(100):
        normals = rng.normal(size=9)
        pos, vel = orc.lfmiddle_step_constrained(pos, vel, np.zeros(9), masses, 0.001, 300.0, 2.0, c, 1e-5,
                                                 normals)
        p = pos.reshape(-1, 3)
        for i, j, d in c:
            assert np.linalg.norm(p[i] - p[j]) == pytest.approx(d, rel=2e-5)
    # Zero temperature and friction: the projections move no net momentum.
    v0 = rng.normal(size=(3, 3))
    p1, v1 = orc.lfmiddle_step_constrained(p, v0, np.zeros(9), masses, 0.001, 0.0, 0.0, c, 1e-10)
    mom0 = (masses[:, None] * v0).sum(0)
    mom1 = (masses[:, None] * v1.reshape(-1, 3)).sum(0)
    assert np.abs(mom1 - mom0).max() < 1e-12


def test_rigid_scaling_by_axes():
    # The anisotropic proposal's diagonal transform: equal factors are the
    # isotropic scaling; one factor moves molecules along that axis only,
    # rigidly.
    rng = np.random.default_rng(8)
    pos = rng.uniform(0, 2.0, size=(30, 3))
    bonds = np.array([[3 * m, 3 * m + 1] for m in range(10)] + [[3 * m, 3 * m + 2] for m in range(10)])
    mol, nmol = orc.molecules(30, bonds)
    s = 1.013
    iso = orc.scale_positions(pos, s, True, mol, nmol)
    assert np.abs(orc.scale_positions_axes(pos, [s, s, s], mol, nmol) - iso).max() < 1e-15
    ax = orc.scale_positions_axes(pos, [1.0, s, 1.0], mol, nmol).reshape(-1, 3)
    assert np.array_equal(ax[:, 0], pos[:, 0]) and np.array_equal(ax[:, 2], pos[:, 2])
    d0 = pos[1::3] - pos[0::3]
    d1 = ax[1::3] - ax[0::3]
    assert np.abs(d1 - d0).max() < 1e-15


def test_reduced_box():
    # flexible_barostat.rs:15-41 on closed-form cases: a vector with a
    # negative diagonal flips; b_x = 0.7 a_x becomes -0.3 a_x; c is reduced by
    # b then a; an already reduced box is unchanged; a zero diagonal fails.
    a = 3.0
    box = np.array([[a, 0, 0], [0.7 * a, 2.5, 0], [-1.9 * a + 0.1, 1.4 * 2.5, 2.8]])
    r = orc.reduced_box(box)
    assert r[1, 0] == pytest.approx(-0.3 * a)
    assert r[2, 1] == pytest.approx(1.4 * 2.5 - 2.5 * round(1.4))
    assert abs(r[2, 0]) <= 0.5 * a + 1e-12 and abs(r[1, 0]) <= 0.5 * a
    flipped = orc.reduced_box(np.array([[a, 0, 0], [0.1, -2.0, 0], [0, 0, 2.0]]))
    assert flipped[1, 1] == 2.0 and flipped[1, 0] == -0.1
    ok = np.array([[2.0, 0, 0], [0.3, 2.0, 0], [-0.4, 0.5, 2.0]])
    assert np.array_equal(orc.reduced_box(ok), ok)
    assert orc.reduced_box(np.diag([2.0, 0.0, 2.0])) is None


# --- context/pressure.rs -------------------------------------------------------

def test_transform_positions_pins():
    """The rigid transform: a diagonal one is the axis scaling of the C oracle,
    one-atom molecules move by T x, and molecules keep their internal shape."""
    s = ommhip.water_system(20, 1.2, seed=4)
    mol, nmol = orc.molecules(len(s.masses), s.bonds)
    sc = np.array([1.01, 0.98, 1.003])
    assert np.allclose(orc.transform_positions(s.positions, np.diag(sc), mol, nmol),
                       orc.scale_positions_axes(s.positions, sc, mol, nmol).reshape(-1, 3), rtol=0, atol=1e-15)
    T = np.array([[1.01, 0.02, -0.01], [0.003, 0.99, 0.004], [0.0, -0.02, 1.002]])
    singles = np.arange(len(s.masses))
    assert np.allclose(orc.transform_positions(s.positions, T, singles, len(s.masses)), s.positions @ T.T,
                       rtol=0, atol=1e-14)
    moved = orc.transform_positions(s.positions, T, mol, nmol)
    d0 = s.positions[1::3] - s.positions[0::3]
    assert np.allclose(moved[1::3] - moved[0::3], d0, rtol=0, atol=1e-14)


def test_pressure_kinetic_known_answer():
    """pressure_uses_raw_molecular_momentum (tests/barostat.rs:112-149): with no
    energy, P = 2 K / (3 V) of the molecular momenta; molecule {0, 1} of mass 4
    carries P = (-2, 2, 6), the massless particle none; V = 60 nm^3."""
    box = np.array([[3.0, 0, 0], [0.4, 4.0, 0], [-0.2, 0.3, 5.0]])
    pos = np.array([[0.2, 0.3, 0.4], [0.9, 0.3, 0.4], [1.4, 1.1, 0.8]])
    vel = np.array([[1.0, 2.0, 3.0], [-1.0, 0.0, 1.0], [10.0, 20.0, 30.0]])
    masses = np.array([1.0, 3.0, 0.0])
    mol, nmol = orc.molecules(3, np.array([[0, 1]]))
    p = orc.current_pressure(lambda q, b: 0.0, pos, vel, masses, box, mol, nmol)
    expected = 2.0 * (0.5 * (4.0 + 4.0 + 36.0) / 4.0) / (3.0 * 60.0 * orc.BAR_NM3_TO_KJ_PER_MOL)
    assert p == pytest.approx(expected, rel=1e-14)
    kin = orc.kinetic_tensor(vel, masses, mol, nmol)
    assert np.allclose(kin, 0.5 * np.array([4, 4, 36, -4, -12, 12]) / 4.0, rtol=1e-15, atol=0)


def test_pressure_queries_of_an_lj_pair_closed_form():
    """The oracle's pressure queries, over its own nonbonded energy, against the
    closed forms of an isolated Lennard-Jones pair in a triclinic box."""
    f, pos, box = lj_pair_system(LJ_BOX)
    vel = np.array([[0.3, -0.2, 0.5], [-0.1, 0.4, 0.2]])
    masses = np.array([2.0, 3.0])
    mol, nmol = np.arange(2), 2

    def energy(q, b):
        return orc.nb_evaluate(f.desc(), q, b)[1]

    assert energy(pos, box) == pytest.approx(lj_energy(pos[1] - pos[0]), rel=1e-14)
    iso, axis, flex, stress = lj_pair_closed_forms(pos, vel, masses, box)
    assert orc.current_pressure(energy, pos, vel, masses, box, mol, nmol) == pytest.approx(iso, rel=1e-9)
    assert np.allclose(orc.axis_pressures(energy, pos, vel, masses, box, mol, nmol), axis, rtol=1e-9, atol=1e-12)
    assert np.allclose(orc.flexible_pressure(energy, pos, vel, masses, box, mol, nmol), flex, rtol=1e-9, atol=1e-12)
    assert np.allclose(orc.stress_tensor(energy, pos, vel, masses, box, mol, nmol, True), stress, rtol=1e-9,
                       atol=1e-12)
