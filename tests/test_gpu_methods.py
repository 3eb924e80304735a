"""The NonbondedForce methods beyond PME on the device, against the FP64 oracle.

NoCutoff and CutoffNonPeriodic run on a virtual rectangular box chosen from
the positions (large enough that the minimum image never alters a displacement
in reach), so their results must equal the reference's non-periodic sums.

The method matrix is the reference's acceptance grid: the 41 four-particle
cases of tests/reference/nonbonded.cpp:13-86 (inputs only; that file needs an
external OpenMM build for its values, so each case is checked against the
oracle here), every one on the device.
"""
import copy

import numpy as np
import pytest

import ommhip
import oracle as orc
from helpers import energy_rel, force_rms_rel, water

FORCE_TOL = 2e-6
ENERGY_TOL = 1e-6

CASES = [
    "nocutoff", "cutoff", "periodic", "periodic_triclinic", "periodic_exception", "periodic_exception_wrapped",
    "periodic_dispersion", "periodic_switch_dispersion", "pme", "pme_triclinic", "pme_exception",
    "pme_exception_wrapped", "pme_charged", "pme_switch_dispersion",
    "ewald", "ewald_exception", "ewald_exception_wrapped", "ewald_charged_switch_dispersion",
    "ljpme", "ljpme_triclinic", "ljpme_exception", "ljpme_exception_wrapped", "ljpme_charged",
    "ljpme_explicit_triclinic", "ljpme_neutral",
    "nocutoff_exception_offset", "nocutoff_exception_offset_runtime", "nocutoff_exception_offset_runtime_update",
    "periodic_exception_offset_dispersion", "periodic_exception_offset_runtime_dispersion",
    "periodic_exception_offset_runtime_update_dispersion",
    "pme_exception_offset", "pme_exception_offset_runtime", "pme_exception_offset_runtime_update_dispersion",
    "ljpme_exception_offset", "ljpme_exception_offset_runtime", "ljpme_exception_offset_runtime_update",
    "ewald_resize", "pme_triclinic_resize", "ljpme_triclinic_resize", "ljpme_switch",
]
assert len(CASES) == 41

POSITIONS = np.array([[0.1, 0.2, 0.3], [0.6, 0.4, 0.2], [2.8, 0.7, 0.4], [1.3, 1.1, 0.9]])


def case_force(name):
    """The force and box of one case, as tests/reference/nonbonded.cpp:13-47 builds them."""
    pme, ljpme, ewald = name.startswith("pme"), name.startswith("ljpme"), name.startswith("ewald")
    periodic = pme or ljpme or ewald or name.startswith("periodic")
    tric = "triclinic" in name
    f = ommhip.NonbondedForce()
    method = ("LJPME" if ljpme else "Ewald" if ewald else "PME" if pme else "CutoffPeriodic" if periodic
              else "CutoffNonPeriodic" if name == "cutoff" else "NoCutoff")
    f.set_nonbonded_method(method)
    f.set_cutoff_distance(1.2)
    f.set_ewald_error_tolerance(1e-5)
    f.set_use_dispersion_correction("dispersion" in name)
    f.set_use_switching_function("switch" in name)
    f.set_switching_distance(0.8)
    f.set_exceptions_use_periodic_boundary_conditions("wrapped" in name)
    if "explicit" in name:
        f.set_pme_parameters(2.7, 24, 25, 28)
        f.set_ljpme_parameters(2.5, 18, 20, 21)
    charges = [0.4, -0.3, 0.1, 0.2 if "charged" in name else -0.2]
    for i in range(4):
        f.add_particle(0.0 if "neutral" in name else charges[i], 0.3 + 0.02 * i, 0.5 + 0.1 * i)
    if "exception" in name:
        f.add_exception(0, 2, -0.02, 0.32, 0.4)
        f.add_exception(0, 1, 0.0, 1.0, 0.0)
    if "offset" in name:
        f.add_global_parameter("lambda", 0.25)
        f.add_global_parameter("mu", 0.1)
        f.add_particle_parameter_offset("lambda", 0, 0.2, 0.04, 0.3)
        f.add_particle_parameter_offset("mu", 0, -0.1, 0.01, -0.2)
        f.add_particle_parameter_offset("lambda", 2, -0.15, -0.02, 0.2)
        f.add_exception_parameter_offset("lambda", 0, 0.04, 0.05, 0.1)
        f.add_exception_parameter_offset("mu", 1, 0.03, -0.2, 0.15)
    box = np.array([[3.0, 0, 0], [0.4 if tric else 0, 3.0, 0], [-0.2 if tric else 0, 0.3 if tric else 0, 3.0]])
    return f, (box if periodic else None)


def run_case(name, precision):
    """Device forces and energy of one case and the oracle's on the same
    effective inputs (runtime parameter values, updated parameters, frozen
    PME parameters after a resize)."""
    f, box = case_force(name)
    op = ommhip.GpuNonbondedForce(f, precision)
    op.initialize(box)
    values = dict(f.global_parameters)
    if "runtime" in name:
        for k, v in (("lambda", 0.7), ("mu", 0.2)):
            assert op.set_global_parameter(k, v)
            values[k] = v
    if "update" in name:
        op.calc_forces_and_energy(POSITIONS, box)
        f.particle_offsets[0] = ("lambda", 0, 0.3, 0.06, -0.1)
        f.exception_offsets[0] = ("lambda", 0, -0.02, 0.01, 0.2)
        f.particles[3] = (0.1, 0.38, 0.7)
        f.exceptions[0] = (0, 2, 0.03, 0.36, 0.2)
        op.update_parameters_from(f)
    if "resize" in name:
        box = np.array([[3.4, 0, 0], [box[1, 0], 3.2, 0], [box[2, 0], box[2, 1], 3.3]])
    F, E = op.calc_forces_and_energy(POSITIONS, box)
    ref = copy.deepcopy(f)
    ref.global_parameters = [(k, values[k]) for k, _ in f.global_parameters]
    ljpme = None
    if f.method in (ommhip.METHODS["PME"], ommhip.METHODS["Ewald"], ommhip.METHODS["LJPME"]):
        # alpha and the grid (PME) or wave-vector limits (Ewald) are frozen at
        # initialize (nonbonded.rs:1055-1107); the oracle takes them explicitly.
        alpha, grid = op.pme_parameters()
        ref.set_pme_parameters(alpha, *grid)
    if f.method == ommhip.METHODS["LJPME"]:
        ljpme = op.ljpme_parameters()
        if "explicit" in name:
            assert ljpme == (2.5, (18, 20, 21))
    F_ref, E_ref, _ = orc.nb_evaluate(ref.desc(), POSITIONS, box, ljpme=ljpme)
    return F, E, F_ref, E_ref


DEVICE_CASES = CASES


@pytest.mark.gpu
@pytest.mark.parametrize("name", DEVICE_CASES)
def test_method_matrix_plain_fp64(name):
    F, E, F_ref, E_ref = run_case(name, ommhip.PRECISION_DOUBLE)
    print(f"{name}: E={E:.12f} E_ref={E_ref:.12f} dF={np.abs(F - F_ref).max():.3e}")
    assert np.abs(F - F_ref).max() <= 1e-9 * max(1.0, np.abs(F_ref).max())
    assert abs(E - E_ref) <= 1e-9 * max(1.0, abs(E_ref))


@pytest.mark.gpu
@pytest.mark.parametrize("name", DEVICE_CASES)
def test_method_matrix_mixed(name):
    F, E, F_ref, E_ref = run_case(name, ommhip.PRECISION_MIXED)
    frel, erel = force_rms_rel(F, F_ref), energy_rel(E, E_ref)
    print(f"{name}: force_rms_rel={frel:.3e} energy_rel={erel:.3e}")
    assert frel <= FORCE_TOL
    assert erel <= ENERGY_TOL


def droplet(method, molecules=300, seed=61):
    """Waters of a 2.1 nm lattice cell with no periodic box."""
    s = water(molecules, 2.1, seed=seed)
    s.nonbonded.set_nonbonded_method(method)
    return s


@pytest.mark.gpu
@pytest.mark.parametrize("method,switching", [("NoCutoff", False), ("CutoffNonPeriodic", False),
                                              ("CutoffNonPeriodic", True)])
def test_nonperiodic_droplet_parity(method, switching):
    s = droplet(method)
    if switching:
        s.nonbonded.set_use_switching_function(True)
        s.nonbonded.set_switching_distance(0.75)
    desc = s.nonbonded.desc()
    op = ommhip.GpuNonbondedForce(s.nonbonded)
    op.initialize(None)
    F, E = op.calc_forces_and_energy(s.positions, None)
    F_ref, E_ref, _ = orc.nb_evaluate(desc, s.positions, None)
    frel, erel = force_rms_rel(F, F_ref), energy_rel(E, E_ref)
    print(f"{method} switching={switching}: force_rms_rel={frel:.3e} energy_rel={erel:.3e}")
    assert frel <= FORCE_TOL
    assert erel <= ENERGY_TOL
    count, h = op.pair_stats(s.positions, None)
    count_ref, h_ref = orc.nb_pairs(desc, s.positions, None)
    assert (count, h) == (count_ref, h_ref)
    if method == "NoCutoff":
        n = len(s.positions)
        assert count == n * (n - 1) // 2 - len(s.nonbonded.exceptions)
    plain = ommhip.GpuNonbondedForce(s.nonbonded, ommhip.PRECISION_DOUBLE)
    plain.initialize(None)
    F64, E64 = plain.calc_forces_and_energy(s.positions, None)
    assert np.abs(F64 - F_ref).max() <= 1e-9 * np.abs(F_ref).max()
    assert abs(E64 - E_ref) <= 1e-9 * abs(E_ref)


@pytest.mark.gpu
@pytest.mark.parametrize("method", ["NoCutoff", "CutoffNonPeriodic"])
def test_nonperiodic_virtual_box_follows_the_atoms(method):
    # The droplet is expanded about a far-off centre between evaluations: the
    # virtual box must be re-chosen (and the list rebuilt) so that results
    # keep matching the oracle.
    s = droplet(method, molecules=200, seed=62)
    op = ommhip.GpuNonbondedForce(s.nonbonded)
    op.initialize(None)
    desc = s.nonbonded.desc()
    centre = s.positions.mean(0)
    for scale, shift in ((1.0, 0.0), (1.6, 0.0), (1.6, -37.5), (0.8, 12.25)):
        mol = s.positions.reshape(-1, 3, 3)
        com = mol.mean(1, keepdims=True)
        pos = ((com - centre) * scale + centre + shift + (mol - com)).reshape(-1, 3)
        F, E = op.calc_forces_and_energy(pos, None)
        F_ref, E_ref, _ = orc.nb_evaluate(desc, pos, None)
        frel, erel = force_rms_rel(F, F_ref), energy_rel(E, E_ref)
        print(f"scale={scale} shift={shift}: force_rms_rel={frel:.3e} energy_rel={erel:.3e}")
        assert frel <= FORCE_TOL
        assert erel <= ENERGY_TOL
        assert op.pair_stats(pos, None) == orc.nb_pairs(desc, pos, None)


def _droplet_energy_forces(s):
    """The oracle's total for a non-periodic water system: nonbonded without a
    box, harmonic bonds and angles."""
    f, e, _ = orc.nb_evaluate(s.nonbonded.desc(), s.positions, None)
    fb, eb = orc.harmonic_bonds(s.bonds, s.bond_params, s.positions)
    fa, ea = orc.harmonic_angles(s.angles, s.angle_params, s.positions)
    return f + fb + fa, e + eb + ea


@pytest.mark.gpu
@pytest.mark.parametrize("method", ["NoCutoff", "CutoffNonPeriodic"])
def test_nonperiodic_md_context(method):
    """Context::step (context.rs:878-928) for the non-periodic methods: the
    state against the oracle, graph replays against eager steps bit for bit,
    and the state again against the oracle after the steps."""
    s = droplet(method, molecules=200, seed=66)
    a = ommhip.Context(s, temperature=300.0, friction=1.0, seed=4)
    st = a.get_state()
    F_ref, E_ref = _droplet_energy_forces(s)
    assert force_rms_rel(st["forces"], F_ref) <= FORCE_TOL
    assert energy_rel(st["potential_energy"], E_ref) <= ENERGY_TOL
    assert st["box"] is None
    b = ommhip.Context(s, temperature=300.0, friction=1.0, seed=4)
    b.set_graph_steps(0)
    for c in (a, b):
        c.set_velocities_to_temperature(300.0, 2)
        c.step(60)
    sa, sb = a.get_state(), b.get_state()
    assert np.array_equal(sa["positions"], sb["positions"])
    assert np.array_equal(sa["velocities"], sb["velocities"])
    moved = copy.copy(s)
    moved.positions = sa["positions"]
    F_ref, E_ref = _droplet_energy_forces(moved)
    assert force_rms_rel(sa["forces"], F_ref) <= FORCE_TOL
    assert energy_rel(sa["potential_energy"], E_ref) <= ENERGY_TOL
    # context.rs:93-101: no barostat on a non-periodic system.
    with pytest.raises(ommhip.OmmError):
        a.set_barostat(1.0, 300.0, 25)


@pytest.mark.gpu
def test_nonperiodic_context_follows_escaping_molecules():
    """Two outer molecules leave the droplet at 15 nm/ps: they leave the region the
    virtual box serves, and the steps since the last snapshot are replayed in
    a new box.  Graph and eager runs replay at the same run boundaries and stay
    bitwise equal; the final state matches the oracle."""
    s = droplet("CutoffNonPeriodic", molecules=150, seed=67)
    vel = np.zeros_like(s.positions)
    com = s.positions.reshape(-1, 3, 3).mean(1)
    fast_x, fast_y = int(np.argmax(com[:, 0])), int(np.argmin(com[:, 1]))  # outermost molecules
    vel[3 * fast_x:3 * fast_x + 3] = [15.0, 0.0, 0.0]
    vel[3 * fast_y:3 * fast_y + 3] = [0.0, -15.0, 0.0]
    runs = []
    for graph_steps in (10, 0):
        c = ommhip.Context(s, temperature=0.0, friction=0.0, seed=4)
        c.set_graph_steps(graph_steps)
        c.set_velocities(vel)
        c.step(200)  # 3 nm of flight
        runs.append((c.get_state(), c.capacity()[2]))
    (sa, ra), (sb, rb) = runs
    assert ra > 0 and rb > 0
    assert np.array_equal(sa["positions"], sb["positions"])
    assert sa["positions"][3 * fast_x, 0] - s.positions[3 * fast_x, 0] > 2.0
    assert s.positions[3 * fast_y, 1] - sa["positions"][3 * fast_y, 1] > 2.0
    moved = copy.copy(s)
    moved.positions = sa["positions"]
    F_ref, E_ref = _droplet_energy_forces(moved)
    assert force_rms_rel(sa["forces"], F_ref) <= FORCE_TOL
    assert energy_rel(sa["potential_energy"], E_ref) <= ENERGY_TOL


@pytest.mark.gpu
@pytest.mark.parametrize("precision,ftol,etol", [(ommhip.PRECISION_MIXED, FORCE_TOL, ENERGY_TOL),
                                                 (ommhip.PRECISION_DOUBLE, 1e-9, 1e-9)])
def test_ewald_water_parity(precision, ftol, etol):
    """Plain Ewald (ewald.rs:22-122) on 300 waters: forces, energy and pair set."""
    s = water(300, 2.1, seed=64, method="Ewald")
    desc = s.nonbonded.desc()
    op = ommhip.GpuNonbondedForce(s.nonbonded, precision)
    op.initialize(s.box)
    alpha, limits = op.pme_parameters()
    ref_alpha, ref_limits = orc.ewald_parameters(np.diag(s.box), s.nonbonded.cutoff,
                                                 s.nonbonded.ewald_error_tolerance)
    assert alpha == ref_alpha and tuple(limits) == tuple(ref_limits)
    F, E = op.calc_forces_and_energy(s.positions, s.box)
    F_ref, E_ref, parts_ref = orc.nb_evaluate(desc, s.positions, s.box)
    parts = op.energy_parts(s.positions, s.box)
    frel, erel = force_rms_rel(F, F_ref), energy_rel(E, E_ref)
    print(f"limits={limits} force_rms_rel={frel:.3e} energy_rel={erel:.3e} parts={parts} ref={parts_ref}")
    assert frel <= ftol
    assert erel <= etol
    assert np.abs(parts - parts_ref).max() <= etol * np.abs(parts_ref).max()
    assert op.pair_stats(s.positions, s.box) == orc.nb_pairs(desc, s.positions, s.box)


@pytest.mark.gpu
def test_ewald_md_context():
    """The MD step with Ewald: state forces against the oracle, then steps."""
    s = water(300, 2.1, seed=65, method="Ewald")
    ctx = ommhip.Context(s, temperature=300.0, friction=1.0, seed=3)
    st = ctx.get_state()
    F_ref, E_ref = orc.water_energy_forces(s)
    assert force_rms_rel(st["forces"], F_ref) <= FORCE_TOL
    assert energy_rel(st["potential_energy"], E_ref) <= ENERGY_TOL
    ctx.set_velocities_to_temperature(300.0, 5)
    ctx.step(20)
    assert np.isfinite(ctx.get_state()["potential_energy"])


# Host-side validation: runs without a GPU.
def test_ewald_rejects_triclinic_box():
    s = water(300, 2.1, seed=66, method="Ewald")
    op = ommhip.GpuNonbondedForce(s.nonbonded)
    with pytest.raises(ommhip.OmmError) as e:
        op.initialize(np.array([[2.1, 0, 0], [0.3, 2.1, 0], [0, 0, 2.1]]))
    assert e.value.status == 1


@pytest.mark.gpu
@pytest.mark.parametrize("precision,ftol,etol", [(ommhip.PRECISION_MIXED, FORCE_TOL, ENERGY_TOL),
                                                 (ommhip.PRECISION_DOUBLE, 1e-9, 1e-9)])
def test_ljpme_water_parity(precision, ftol, etol):
    """LJPME (dispersion mesh, nonbonded.rs:875-886, 974-989) on 1,000 waters."""
    s = water(1000, 3.1, seed=67, method="LJPME")
    desc = s.nonbonded.desc()
    op = ommhip.GpuNonbondedForce(s.nonbonded, precision)
    op.initialize(s.box)
    tol, rc = s.nonbonded.ewald_error_tolerance, s.nonbonded.cutoff
    assert op.ljpme_parameters() == orc.effective_ljpme_parameters(tol, rc, s.box)
    assert op.pme_parameters() == orc.effective_pme_parameters(tol, rc, s.box)
    F, E = op.calc_forces_and_energy(s.positions, s.box)
    F_ref, E_ref, parts_ref = orc.nb_evaluate(desc, s.positions, s.box)
    parts = op.energy_parts(s.positions, s.box)
    frel, erel = force_rms_rel(F, F_ref), energy_rel(E, E_ref)
    print(f"force_rms_rel={frel:.3e} energy_rel={erel:.3e} parts={parts} ref={parts_ref}")
    assert frel <= ftol
    assert erel <= etol
    assert np.abs(parts - parts_ref).max() <= etol * np.abs(parts_ref).max()
    F_only = op.calc_forces(s.positions, s.box)
    assert force_rms_rel(F_only, F_ref) <= ftol
    assert op.pair_stats(s.positions, s.box) == orc.nb_pairs(desc, s.positions, s.box)


@pytest.mark.gpu
def test_ljpme_md_context_with_barostat():
    """The MD step under LJPME: state against the oracle, then NPT steps; after
    the barostat has moved the box the forces still match the oracle with the
    parameters frozen at initialisation (the dispersion mesh follows the box)."""
    s = water(500, 2.5, seed=68, method="LJPME")
    tol, rc = s.nonbonded.ewald_error_tolerance, s.nonbonded.cutoff
    frozen = orc.effective_pme_parameters(tol, rc, s.box)
    frozen_lj = orc.effective_ljpme_parameters(tol, rc, s.box)
    ctx = ommhip.Context(s, temperature=300.0, friction=1.0, seed=4)
    st = ctx.get_state()
    F_ref, E_ref = orc.water_energy_forces(s)
    assert force_rms_rel(st["forces"], F_ref) <= FORCE_TOL
    assert energy_rel(st["potential_energy"], E_ref) <= ENERGY_TOL
    ctx.set_velocities_to_temperature(300.0, 6)
    ctx.set_barostat(1.0, 300.0, frequency=2, seed=9)
    ctx.step(40)
    info = ctx.barostat_info()
    box = ctx.box()
    st = ctx.get_state()
    moved = copy.deepcopy(s)
    moved.positions = st["positions"]
    moved.box = box
    moved.nonbonded.set_pme_parameters(frozen[0], *frozen[1])
    F_ref, E_ref = orc.water_energy_forces(moved, ljpme=frozen_lj)
    print("box", np.diag(box), "info", info)
    assert info["total_accepted"] > 0
    assert force_rms_rel(st["forces"], F_ref) <= FORCE_TOL
    assert energy_rel(st["potential_energy"], E_ref) <= ENERGY_TOL


def _chain_system(nchains=60, length=24, edge=3.0, seed=5):
    """Branched chains (bonds i-(i+1) and i-(i+2)): with 1-2, 1-3 and scaled
    1-4 exceptions (nonbonded.rs:434-500) an inner atom has 12 exclusion
    partners.  Chains are random walks of 0.14 nm steps that keep 0.18 nm
    (minimum image) from every atom but the two before them, restarting a
    chain that runs into a dead end."""
    rng = np.random.default_rng(seed)
    pos = np.zeros((0, 3))

    def clear_of(cand, others):
        if not len(others):
            return True
        d = others - cand
        d -= edge * np.round(d / edge)
        return np.min(np.einsum("ij,ij->i", d, d)) >= 0.18 ** 2

    while len(pos) < nchains * length:
        chain = [rng.uniform(0, edge, 3)]
        if not clear_of(chain[0], pos):
            continue
        tries = 0
        while len(chain) < length and tries < 200:
            tries += 1
            step = rng.normal(size=3)
            cand = chain[-1] + 0.14 * step / np.linalg.norm(step)
            far = np.concatenate([pos, np.array(chain[:-1]).reshape(-1, 3)[:-1]])
            if clear_of(cand, far):
                chain.append(cand)
        if len(chain) == length:
            pos = np.concatenate([pos, np.array(chain)])
    pos = pos % edge
    n = len(pos)
    f = ommhip.NonbondedForce()
    f.set_nonbonded_method("PME")
    f.set_cutoff_distance(0.9)
    charges = rng.normal(scale=0.3, size=n)
    charges -= charges.mean()
    for i in range(n):
        f.add_particle(float(charges[i]), 0.22, 0.4)
    bonds = []
    for c in range(nchains):
        b = c * length
        bonds += [(b + k, b + k + 1) for k in range(length - 1)]
        bonds += [(b + k, b + k + 2) for k in range(length - 2)]
    f.create_exceptions_from_bonds(bonds, 0.8333, 0.5)
    box = np.eye(3) * edge
    alpha, grid = orc.effective_pme_parameters(5e-4, 0.9, box)
    f.set_pme_parameters(alpha, *grid)
    return f, pos, box


@pytest.mark.gpu
@pytest.mark.parametrize("precision", [ommhip.PRECISION_MIXED, ommhip.PRECISION_DOUBLE])
def test_many_exclusions_per_particle(precision):
    """More than eight exclusion partners per particle: the list marks every
    excluded pair per entry, so the count per particle is not limited."""
    f, pos, box = _chain_system()
    partners = [set() for _ in range(len(pos))]
    for a, b, *_ in f.exceptions:
        partners[a].add(b)
        partners[b].add(a)
    assert max(len(p) for p in partners) == 12
    desc = f.desc()
    op = ommhip.GpuNonbondedForce(f, precision)
    op.initialize(box)
    F, E = op.calc_forces_and_energy(pos, box)
    F_ref, E_ref, _ = orc.nb_evaluate(desc, pos, box)
    frel, erel = force_rms_rel(F, F_ref), energy_rel(E, E_ref)
    print(f"exceptions={len(f.exceptions)} force_rms_rel={frel:.3e} energy_rel={erel:.3e}")
    ftol, etol = (FORCE_TOL, ENERGY_TOL) if precision == ommhip.PRECISION_MIXED else (1e-9, 1e-9)
    assert frel <= ftol
    assert erel <= etol
    assert op.pair_stats(pos, box) == orc.nb_pairs(desc, pos, box)
