"""Pins of the FP64 oracle's pair terms by identities and closed forms.

Each test here fixes one term of the reference's NonbondedForce arithmetic by
something the mathematics determines independently of how the oracle writes
it, so that a dropped term, a flipped sign or a wrong constant in the oracle
fails at least one test (tests/test_oracle_mutations.py builds such broken
oracles and checks that they do):

- the Ewald exclusion identity: erf + erfc = 1, so excluding a pair changes
  the PME/Ewald energy by exactly the bare Coulomb pair energy, whatever the
  mesh does (nonbonded.rs:1021-1026, 1275-1288);
- the reaction field's defining conditions: E(rc) = 0, and F(rc) = 0 for an
  infinite dielectric (nonbonded.rs:1233-1250);
- the quintic switch S(t) = 1 - 10t^3 + 15t^4 - 6t^5: S(0) = 1, S(1) = 0 and
  S(t) + S(1 - t) = 1 (nonbonded.rs:887-899);
- the restated gradient test of tests/nonbonded_gradients.rs:17-56 and the
  dispersion complement test of nonbonded.rs:1414-1434;
- LJPME against a brute-force lattice sum of C6/r^6 as the dispersion grid is
  refined (nonbonded.rs:875-886, 974-989; pme.rs:358-362);
- the switched dispersion correction against adaptive quadrature of the
  homogeneous-fluid integral (nonbonded.rs:660-719, 1290-1336).
"""
import math

import numpy as np
import pytest
from scipy import integrate, special

import oracle as orc
from ommhip import NonbondedForce

K = 138.93545764438198  # ONE_4PI_EPS0, nonbonded.rs:16


def force(particles, method, cutoff=0.9, exceptions=(), pme=None, dispersion=True, switching=None, rf=None):
    f = NonbondedForce()
    f.set_nonbonded_method(method)
    f.set_cutoff_distance(cutoff)
    f.set_use_dispersion_correction(dispersion)
    if pme is not None:
        f.set_pme_parameters(*pme)
    if switching is not None:
        f.set_use_switching_function(True)
        f.set_switching_distance(switching)
    if rf is not None:
        f.set_reaction_field_dielectric(rf)
    for p in particles:
        f.add_particle(*p)
    for e in exceptions:
        f.add_exception(*e)
    return f


def lj(sigma, eps, r):
    return 4.0 * eps * ((sigma / r) ** 12 - (sigma / r) ** 6)


def dlj(sigma, eps, r):
    return 4.0 * eps * (-12.0 * sigma**12 / r**13 + 6.0 * sigma**6 / r**7)


def complement(c6, alpha, r):
    """C6 (1 - e^{-x}(1 + x + x^2/2)) / r^6 with x = alpha^2 r^2: the part of
    C6/r^6 that the dispersion mesh carries (the Ewald split of r^-6)."""
    x = (alpha * r) ** 2
    return c6 / r**6 * (1.0 - math.exp(-x) * (1.0 + x + 0.5 * x * x))


# --- the Ewald exclusion identity -------------------------------------------

PARTICLES = [(0.52, 0.31, 0.6), (-0.41, 0.31, 0.4), (0.30, 0.31, 0.5), (-0.41, 0.31, 0.2)]
POSITIONS = np.array([[0.40, 0.50, 0.60], [0.62, 0.71, 0.55], [1.30, 1.10, 0.90], [1.60, 0.40, 1.50]])


@pytest.mark.parametrize("method", ["PME", "Ewald", "LJPME"])
def test_excluding_a_pair_removes_exactly_its_bare_interaction(method):
    """E(pair 0-1 interacting) - E(pair 0-1 excluded) = k q0 q1 / r + LJ terms:
    the direct space holds k q q erfc(ar)/r for the interacting pair and the
    exception -k q q erf(ar)/r for the excluded one, and the mesh is the same
    in both.  For LJPME the dispersion complement cancels the same way and
    the direct space's cutoff shift LJ(rc) + D(rc) is left."""
    L = 2.4
    box = np.diag([L, L, L]) if method == "Ewald" else np.array([[L, 0, 0], [0.3, L, 0], [-0.2, 0.25, L]])
    pme = (3.1, 9, 9, 9) if method == "Ewald" else (3.1, 24, 26, 28)
    rc = 0.9

    def evaluate(excluded):
        exc = [(0, 1, 0.0, 0.31, 0.0)] if excluded else []
        f = force(PARTICLES, method, rc, exc, pme)
        if method == "LJPME":
            f.set_ljpme_parameters(2.6, 20, 22, 24)
            ljpme = (2.6, [20, 22, 24])
        else:
            ljpme = None
        F, E, _ = orc.nb_evaluate(f.desc(), POSITIONS, box, ljpme=ljpme)
        return F, E

    F_in, E_in = evaluate(False)
    F_ex, E_ex = evaluate(True)
    delta = POSITIONS[1] - POSITIONS[0]
    r = float(np.linalg.norm(delta))
    assert r < rc
    q0, q1 = PARTICLES[0][0], PARTICLES[1][0]
    sigma, eps = 0.31, math.sqrt(PARTICLES[0][2] * PARTICLES[1][2])
    expected = K * q0 * q1 / r + lj(sigma, eps, r)
    if method == "LJPME":
        c6 = 4.0 * eps * sigma**6  # geometric and arithmetic sigma agree here
        expected -= lj(sigma, eps, rc) + complement(c6, 2.6, rc)
    assert E_in - E_ex == pytest.approx(expected, rel=1e-10, abs=1e-10)
    # Forces: only the pair's own force changes, F_0 = -dE/dr (x_0 - x_1)/r.
    dedr = -K * q0 * q1 / r**2 + dlj(sigma, eps, r)
    f0 = dedr * delta / r
    dF = F_in - F_ex
    assert np.allclose(dF[0], f0, rtol=1e-9, atol=1e-9)
    assert np.allclose(dF[1], -f0, rtol=1e-9, atol=1e-9)
    assert np.abs(dF[2:]).max() < 1e-9


def test_pme_direct_space_is_the_erfc_screened_pair():
    """Direct space alone of an isolated charged LJ pair:
    k q q erfc(a r)/r + LJ(r) with scipy's erfc (nonbonded.rs:1263-1273)."""
    L, alpha, rc = 3.0, 3.3, 1.0
    f = force([(0.7, 0.3, 0.4), (-0.45, 0.34, 0.9)], "PME", rc, pme=(alpha, 20, 20, 20), dispersion=False)
    for r in (0.21, 0.5, 0.93):
        pos = np.array([[1.0, 1.0, 1.0], [1.0 + r * 0.6, 1.0 + r * 0.8, 1.0]])
        F, E, _ = orc.nb_evaluate(f.desc(), pos, np.diag([L] * 3), include_reciprocal=False)
        sigma, eps = 0.32, math.sqrt(0.4 * 0.9)
        qq = 0.7 * -0.45
        expected = K * qq * special.erfc(alpha * r) / r + lj(sigma, eps, r)
        assert E == pytest.approx(expected, rel=1e-12)
        dedr = (-K * qq * (special.erfc(alpha * r) / r**2 + 2 * alpha / math.sqrt(math.pi)
                           * math.exp(-(alpha * r) ** 2) / r) + dlj(sigma, eps, r))
        assert F[1] == pytest.approx(-dedr * np.array([0.6, 0.8, 0.0]), rel=1e-10, abs=1e-10)


# --- reaction field -------------------------------------------------------

@pytest.mark.parametrize("method", ["CutoffPeriodic", "CutoffNonPeriodic"])
@pytest.mark.parametrize("eps_rf", [78.3, 2.0, 1.0, math.inf])
def test_reaction_field_vanishes_at_the_cutoff(method, eps_rf):
    """E = k q q (1/r + k_rf r^2 - c_rf) with k_rf = (e-1)/((2e+1) rc^3)
    (Tironi et al.) and c_rf fixed by E(rc) = 0; F(rc) = 0 when e = inf."""
    rc, qq = 1.1, 0.8 * -0.6
    f = force([(0.8, 0.3, 0.0), (-0.6, 0.3, 0.0)], method, rc, dispersion=False, rf=eps_rf)
    box = np.diag([3.0] * 3) if method == "CutoffPeriodic" else None
    k_rf = 0.5 / rc**3 if math.isinf(eps_rf) else (eps_rf - 1.0) / (2.0 * eps_rf + 1.0) / rc**3
    c_rf = 1.0 / rc + k_rf * rc**2
    for r in (0.25, 0.6, 1.0, rc * (1 - 1e-12), rc * (1 + 1e-9)):
        pos = np.array([[1.0, 1.0, 1.0], [1.0, 1.0 + r, 1.0]])
        F, E, _ = orc.nb_evaluate(f.desc(), pos, box)
        if r >= rc:
            assert E == 0.0 and np.all(F == 0.0)
            continue
        assert E == pytest.approx(K * qq * (1.0 / r + k_rf * r * r - c_rf), rel=1e-11, abs=1e-9)
        assert F[1, 1] == pytest.approx(-K * qq * (-1.0 / r**2 + 2.0 * k_rf * r), rel=1e-11, abs=1e-9)
        if r > rc * (1 - 1e-9):
            assert abs(E) < 1e-8
            if math.isinf(eps_rf):
                assert abs(F[1, 1]) < 1e-8


# --- switching function ---------------------------------------------------

@pytest.mark.parametrize("method", ["CutoffNonPeriodic", "CutoffPeriodic", "PME"])
def test_switched_lj_is_the_quintic_smoothstep(method):
    """LJ x S(t), t = (r - rs)/(rc - rs): E(rs) = LJ(rs), E(rc) = 0, and at
    t = 1/2 exactly LJ/2 because S(t) + S(1 - t) = 1."""
    rs, rc = 0.7, 1.0
    f = force([(0.0, 0.32, 0.7), (0.0, 0.36, 0.9)], method, rc, dispersion=False, switching=rs,
              pme=(3.0, 24, 24, 24) if method == "PME" else None)
    box = None if method == "CutoffNonPeriodic" else np.diag([3.0] * 3)
    sigma, eps = 0.34, math.sqrt(0.7 * 0.9)

    def energy(r):
        pos = np.array([[1.0, 1.0, 1.0], [1.0 + r, 1.0, 1.0]])
        return orc.nb_evaluate(f.desc(), pos, box)[1]

    assert energy(0.5) == pytest.approx(lj(sigma, eps, 0.5), rel=1e-13)
    assert energy(rs) == pytest.approx(lj(sigma, eps, rs), rel=1e-13)
    assert energy(0.5 * (rs + rc)) == pytest.approx(0.5 * lj(sigma, eps, 0.5 * (rs + rc)), rel=1e-12)
    for t in (0.2, 0.35):
        a, b = rs + t * (rc - rs), rs + (1 - t) * (rc - rs)
        assert energy(a) / lj(sigma, eps, a) + energy(b) / lj(sigma, eps, b) == pytest.approx(1.0, rel=1e-12)
    assert abs(energy(rc * (1 - 1e-9))) < 1e-12


@pytest.mark.parametrize("method", ["NoCutoff", "CutoffNonPeriodic", "CutoffPeriodic", "PME"])
def test_switched_forces_are_negative_energy_gradients(method):
    """tests/nonbonded_gradients.rs:17-56 restated in FP64 (the reference
    rounds positions through f32; here the step is exact)."""
    box = np.array([[3.0, 0.0, 0.0], [0.4, 3.0, 0.0], [-0.2, 0.3, 3.0]])
    f = force([(0.4, 0.3, 0.5), (-0.3, 0.32, 0.6), (0.1, 0.34, 0.7)], method, 1.2,
              [(0, 2, -0.02, 0.32, 0.4)], switching=0.8)
    f.set_ewald_error_tolerance(1e-5)
    d = f.desc()
    pos = np.array([0.1, 0.2, 0.3, 0.9, 0.4, 0.2, 2.8, 0.7, 0.4]).reshape(3, 3)
    F, _, _ = orc.nb_evaluate(d, pos, box)
    h = 1e-4
    for p in range(3):
        for a in range(3):
            plus, minus = pos.copy(), pos.copy()
            plus[p, a] += h
            minus[p, a] -= h
            numeric = -(orc.nb_evaluate(d, plus, box)[1] - orc.nb_evaluate(d, minus, box)[1]) / (2 * h)
            assert abs(numeric - F[p, a]) < 5e-5 * (1.0 + abs(numeric)), (method, p, a)


# --- dispersion complement and LJPME ---------------------------------------

def test_dispersion_complement_at_overlap_and_its_derivative():
    """nonbonded.rs:1414-1434, plus continuity across the series branch at
    x = alpha^2 r^2 = 0.01 and the closed form above it."""
    alpha, c6 = 2.3, 0.04
    e, f = orc.dispersion_reciprocal_pair(c6, 0.0, alpha)
    assert e == pytest.approx(c6 * alpha**6 / 6.0, abs=1e-14)
    assert f == pytest.approx(-c6 * alpha**8 / 4.0, abs=1e-14)
    for r in (1e-5, 0.02, 0.1, 0.7):
        h = r * 1e-4
        plus = orc.dispersion_reciprocal_pair(c6, (r + h) ** 2, alpha)[0]
        minus = orc.dispersion_reciprocal_pair(c6, (r - h) ** 2, alpha)[0]
        deriv = orc.dispersion_reciprocal_pair(c6, r * r, alpha)[1]
        assert (plus - minus) / (2 * h) == pytest.approx(deriv * r, rel=1e-4, abs=1e-6)
    r0 = 0.1 / alpha  # x = 0.01
    below = orc.dispersion_reciprocal_pair(c6, (r0 * (1 - 1e-9)) ** 2, alpha)
    above = orc.dispersion_reciprocal_pair(c6, (r0 * (1 + 1e-9)) ** 2, alpha)
    # Above the branch the closed forms cancel to ~x^3/6 (energy) and ~x^4/24
    # (derivative) of their unit terms, so they carry ~1e-9 and ~3e-7 relative
    # rounding there.
    assert below[0] == pytest.approx(above[0], rel=1e-8)
    assert below[1] == pytest.approx(above[1], rel=2e-6)
    for r in (0.05, 0.3, 0.9):
        assert orc.dispersion_reciprocal_pair(c6, r * r, alpha)[0] == pytest.approx(complement(c6, alpha, r),
                                                                                    rel=1e-12)


def _lattice_dispersion_pairs():
    rng = np.random.default_rng(11)
    L, n = 2.0, 6
    while True:
        pos = rng.uniform(0, L, (n, 3))
        d = pos[:, None] - pos[None]
        d -= L * np.rint(d / L)
        if (np.sqrt((d**2).sum(-1)) + 9 * np.eye(n)).min() > 0.3:
            break
    return L, pos, rng.uniform(0.2, 0.8, n)


def test_ljpme_converges_to_the_lattice_sum():
    """Zero charges, equal sigma (so the direct space's Lorentz-Berthelot C6
    equals the mesh's geometric C6).  The exact LJPME energy is
        -1/2 sum_ij sum_n' C6_ij / r^6            (all images, brute force)
        + 1/2 sum' C6 g(a r) / r^6                (real-space tail, every
                                                  image except in-cutoff pairs)
        + sum_{r < rc} [4 eps sigma^12/r^12 - LJ(rc) - D(rc)]
    with g(x) = e^{-x}(1 + x + x^2/2), x = a^2 r^2 and D = C6 (1 - g)/r^6; the
    mesh error falls with each refinement of the dispersion grid."""
    L, pos, eps = _lattice_dispersion_pairs()
    n, sigma, rc, alpha = len(pos), 0.3, 0.9, 3.0
    R = 40.0
    m = int(math.ceil(R / L))
    rng_ = np.arange(-m, m + 1)
    shifts = np.stack(np.meshgrid(rng_, rng_, rng_, indexing="ij"), -1).reshape(-1, 3) * L
    shifts = shifts[(shifts**2).sum(1) <= R * R]
    expected = 0.0
    for i in range(n):
        for j in range(n):
            c6 = 4.0 * math.sqrt(eps[i] * eps[j]) * sigma**6
            v = pos[j] - pos[i] + shifts
            r2 = (v**2).sum(1)
            r2 = r2[r2 > 0]
            tail_keep = np.ones_like(r2, dtype=bool)
            if i != j:
                dmin = pos[j] - pos[i]
                dmin -= L * np.rint(dmin / L)
                rmin2 = (dmin**2).sum()
                if rmin2 < rc * rc:
                    tail_keep = np.abs(r2 - rmin2) > 1e-12
            x = alpha**2 * r2[tail_keep]
            expected += -0.5 * c6 * (r2**-3).sum() - 0.5 * c6 * 4 * math.pi / (3 * L**3 * R**3)
            expected += 0.5 * c6 * (np.exp(-x) * (1 + x + 0.5 * x * x) * r2[tail_keep] ** -3).sum()
    for i in range(n):
        for j in range(i + 1, n):
            dmin = pos[j] - pos[i]
            dmin -= L * np.rint(dmin / L)
            r = math.sqrt((dmin**2).sum())
            if r < rc:
                e = math.sqrt(eps[i] * eps[j])
                expected += 4 * e * (sigma / r) ** 12 - lj(sigma, e, rc) - complement(4 * e * sigma**6, alpha, rc)
    f = force([(0.0, sigma, e) for e in eps], "LJPME", rc, pme=(3.0, 16, 16, 16), dispersion=False)
    errs = []
    for g in (16, 24, 32, 48, 64):
        _, E, _ = orc.nb_evaluate(f.desc(), pos, np.diag([L] * 3), ljpme=(alpha, [g, g, g]))
        errs.append(abs(E - expected) / abs(expected))
    assert all(a > b for a, b in zip(errs, errs[1:])), errs
    assert errs[-1] < 3e-7, errs


# --- dispersion correction with switching ----------------------------------

def test_switched_dispersion_correction_is_the_fluid_integral():
    """For a fluid of n particles the correction is (2 pi n^2 / V) times the
    mean over unordered pairs i <= j of int_0^inf r^2 (1 - S(r)) U_ij(r) dr
    outside the direct space (S = 1 below rs, the switch between rs and rc,
    0 beyond), computed here by scipy's adaptive quadrature per pair class."""
    rng = np.random.default_rng(7)
    n, rs, rc, V = 5, 0.72, 1.0, 3.1
    sig = rng.uniform(0.25, 0.4, n)
    eps = rng.uniform(0.1, 0.9, n)

    def S(r):
        t = (r - rs) / (rc - rs)
        return 1.0 - 10 * t**3 + 15 * t**4 - 6 * t**5

    total, count = 0.0, 0
    for i in range(n):
        for j in range(i, n):
            s, e = 0.5 * (sig[i] + sig[j]), math.sqrt(eps[i] * eps[j])
            inner = integrate.quad(lambda r: r * r * (1.0 - S(r)) * lj(s, e, r), rs, rc, epsabs=0, epsrel=1e-13)[0]
            outer = integrate.quad(lambda r: r * r * lj(s, e, r), rc, np.inf, epsabs=0, epsrel=1e-13)[0]
            total += inner + outer
            count += 1
    expected = 2 * math.pi * n * n / V * total / count
    f = force([(0.0, s, e) for s, e in zip(sig, eps)], "PME", rc, switching=rs)
    assert orc.dispersion_correction(f.desc(), V) == pytest.approx(expected, rel=1e-9)
    f_plain = force([(0.0, s, e) for s, e in zip(sig, eps)], "PME", rc)
    assert abs(orc.dispersion_correction(f_plain.desc(), V) - expected) > 1e-4 * abs(expected)
