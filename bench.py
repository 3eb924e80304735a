"""Throughput of the MD step around the NonbondedForce/PME operator on one MI355X.

Workload (the configuration the headline ns/day is quoted on): 7,850 flexible
3-site waters (23,550 atoms) in a 6.17 nm cubic box, PME (cutoff 0.9 nm,
tolerance 5e-4 -> alpha 2.920290, grid 55^3), harmonic O-H bonds and H-O-H
angles, LangevinMiddle 1 fs at 300 K with friction 1/ps.  Synthetic seeded
coordinates and Maxwell-Boltzmann velocities.  Mixed precision: FP32 pair,
spline and FFT math, int64 fixed-point accumulation, FP64 integrator; the step
is replayed from a HIP graph.

Prints one JSON line.  Besides the contract fields it carries the roofline of
the dominant kernel (direct-space pairs, 75 FLOP per in-cutoff pair against
the FP32 peak), the per-phase device times, the plain FP64 form on the same
workload, the CPU baseline (the FP64 C++ restatement of the reference,
single-threaded as the reference is) and the batched workload: 64
independent water boxes of 500-6,000 molecules (log-uniform from a seed, box
edge for 33.4 molecules/nm^3, each with its own PME grid) advanced by one set
of launches per phase, in replica-steps/s, beside the same replicas run as 64
separate plain FP64 contexts (per-system launches); and 128 replicas of
300-4,000 molecules under NPT (Monte Carlo barostat at 1 bar every 25 steps);
and one 300,000-atom box with the HBM rate of every phase.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path[:0] = [os.path.join(ROOT, "openmm-rs_amd"), os.path.join(ROOT, "oracle")]

import numpy as np  # noqa: E402

import ommhip  # noqa: E402

METRIC = "ns/day (23.5k-atom PME water) + batched replica-steps/s, mixed vs plain FP64"
DT_PS = 0.001
# MI355X published peaks.  The FP32 vector peak counts packed FP32 (two
# operations per lane per cycle); this build compiles with packed FP32 off,
# so its attainable FP32 rate is half of it.
FP32_PEAK_TFLOPS = 157.3
FP32_NONPACKED_TFLOPS = 78.6
HBM_PEAK_GBPS = 8000.0
# Algorithmic work per unit (the fixed convention of the roofline): FLOP per
# in-cutoff pair, per exception, per atom for the spread and the gather;
# bytes per grid point for the FFT passes and convolution, per atom for the
# FP64 integrator with the bonded terms.
FLOP_PER_PAIR = 75
FLOP_PER_EXCEPTION = 40
FLOP_SPREAD_PER_ATOM = 590
FLOP_GATHER_PER_ATOM = 1140
MESH_BYTES_PER_POINT = 24
STEP_BYTES_PER_ATOM = 128
TILE_SLOTS = 8 * 64  # pair slots of one k_pairs tile (8-atom cluster x 64 list entries)
# HBM bytes per k_pairs launch on the 23,550-atom workload: rocprofv3 FETCH_SIZE
# and WRITE_SIZE passes, each scaled by the same counter's reading for a plain
# copy kernel of known size (profiles/k_pairs_md_step_water_23550_atoms.md).
K_PAIRS_TRAFFIC_BYTES = 49.33e6


def workload(seed=2024):
    s = ommhip.water_system(7850, 6.17, seed, cutoff=0.9, tolerance=5e-4)
    return s


def ns_per_day(steps, seconds):
    return steps / seconds * DT_PS * 86400.0 / 1000.0


def equilibrate_then(ctx, equilibrate, warmup):
    """The seeded lattice with random orientations starts ~3e5 kJ/mol above
    liquid water; relaxing it heats the box to ~1000 K.  `equilibrate` steps
    at friction 50/ps remove that heat before the `warmup` steps at the
    workload's friction 1/ps; timed steps all run at 300 K, 1/ps."""
    if equilibrate:
        ctx.set_integrator(300.0, 50.0, DT_PS)
        ctx.step(equilibrate)
        ctx.set_integrator(300.0, 1.0, DT_PS)
    ctx.step(warmup)


def timed_steps(ctx, steps):
    t0 = time.perf_counter()
    ctx.step(steps)  # returns after the device finished (stream synchronised)
    return time.perf_counter() - t0


GRAPH_STEPS = 50  # longest captured graph (results do not depend on it); 0 with --eager


def graph_steps_for(steps):
    """Steps per captured graph for a timed run of `steps`: the largest divisor
    of `steps` not above GRAPH_STEPS, so every timed step is a graph replay."""
    if GRAPH_STEPS <= 0 or steps <= 0:
        return 0
    return max(d for d in range(1, min(GRAPH_STEPS, steps) + 1) if steps % d == 0)


def prime_graph(ctx, steps):
    """Set the graph length for `steps` timed steps and run one untimed graph,
    so capture and instantiation happen before the timed region."""
    g = graph_steps_for(steps)
    ctx.set_graph_steps(g)
    if g:
        ctx.step(g)
    return g


def run_single(steps, warmup, equilibrate, plain_steps, seed=2024):
    """23,550-atom step: mixed form timed after equilibration; the plain FP64
    form is then timed from the same equilibrated state."""
    s = workload(seed)
    ctx = ommhip.Context(s, temperature=300.0, friction=1.0, step_size=DT_PS, seed=seed,
                         precision=ommhip.PRECISION_MIXED)
    ctx.set_graph_steps(GRAPH_STEPS)
    ctx.set_velocities_to_temperature(300.0, seed + 1)
    equilibrate_then(ctx, equilibrate, 0)
    g = prime_graph(ctx, steps)
    ctx.step(warmup)
    secs = timed_steps(ctx, steps)
    st = ctx.get_state()
    plain = ommhip.Context(s, temperature=300.0, friction=1.0, step_size=DT_PS, seed=seed,
                           precision=ommhip.PRECISION_DOUBLE)
    plain.set_positions(st["positions"])
    plain.set_velocities(st["velocities"])
    plain.step(20)
    psecs = timed_steps(plain, plain_steps)
    plain.close()
    return ctx, secs, psecs, g


def batch_workload(count=64, seed=4096, lo=500, hi=6000, density=33.4):
    """Replica r: M_r = round(exp(U(ln lo, ln hi))) molecules, edge (M_r/density)^(1/3) nm."""
    rng = np.random.default_rng(seed)
    mols = np.rint(np.exp(rng.uniform(np.log(lo), np.log(hi), size=count))).astype(int)
    return [ommhip.water_system(int(m), float((m / density) ** (1.0 / 3.0)), seed + 1 + r, cutoff=0.9,
                                tolerance=5e-4) for r, m in enumerate(mols)]


def run_batch(steps, warmup, equilibrate, plain_steps, count=64, seed=4096):
    """Mixed batched form vs the same replicas as separate plain FP64 contexts."""
    systems = batch_workload(count, seed)
    seeds = [seed + 100 + r for r in range(count)]
    batch = ommhip.Batch(systems, temperature=300.0, friction=1.0, step_size=DT_PS, seeds=seeds,
                         precision=ommhip.PRECISION_MIXED)
    batch.set_graph_steps(GRAPH_STEPS)
    batch.set_velocities_to_temperature(300.0, seed + 1)
    equilibrate_then(batch, equilibrate, warmup)
    g = prime_graph(batch, steps)
    secs = timed_steps(batch, steps)
    phases, npairs = batch.profile(10)
    plain_secs = 0.0
    for r, s in enumerate(systems):
        st = batch.get_state(r)
        c = ommhip.Context(s, temperature=300.0, friction=1.0, step_size=DT_PS, seed=seeds[r],
                           precision=ommhip.PRECISION_DOUBLE)
        c.set_positions(st["positions"])
        c.set_velocities(st["velocities"])
        c.step(2)
        plain_secs += timed_steps(c, plain_steps)
        c.close()
    atoms = int(sum(len(s.masses) for s in systems))
    value = count * steps / secs
    plain = count * plain_steps / plain_secs
    gpoints = sum(int(np.prod(pme_rule(0.9, 5e-4, float(s.box[0, 0]))[1])) for s in systems)
    return {
        "replicas": count,
        "atoms": atoms,
        "molecules_min_max": [int(min(len(s.masses) for s in systems) // 3),
                              int(max(len(s.masses) for s in systems) // 3)],
        "replica_steps_per_s": round(value, 2),
        "ms_per_batch_step": round(secs / steps * 1e3, 4),
        "steps": steps,
        "graph_steps": g,
        "pairs_per_step": npairs,
        "step_roofline": step_roofline(npairs, atoms, atoms, gpoints, secs / steps),
        "phase_roofline": phase_roofline(phases, npairs, atoms, atoms, gpoints),
        "phases_ms": {k: round(v, 5) for k, v in phases.items()},
        "launch_tuning": batch.launch_tuning(),
        "plain_fp64_per_system": {"replica_steps_per_s": round(plain, 2), "steps_per_replica": plain_steps},
        "speedup_mixed_batched_over_plain": round(value / plain, 2),
    }


def run_npt(steps, equilibrate, count=128, seed=8192):
    """128 uneven replicas (300-4,000 molecules) under NPT: Langevin 1 fs at
    300 K plus an isotropic Monte Carlo barostat at 1 bar every 25 steps."""
    systems = batch_workload(count, seed, lo=300, hi=4000)
    seeds = [seed + 100 + r for r in range(count)]
    batch = ommhip.Batch(systems, temperature=300.0, friction=1.0, step_size=DT_PS, seeds=seeds,
                         precision=ommhip.PRECISION_MIXED)
    batch.set_graph_steps(GRAPH_STEPS)
    batch.set_velocities_to_temperature(300.0, seed + 1)
    equilibrate_then(batch, equilibrate, 50)  # NVT relaxation of the lattice start
    batch.set_barostat(1.0, 300.0, 25, seed=seed + 7)
    batch.step(100)
    a0 = sum(batch.barostat_info(r)["total_attempted"] for r in range(count))
    secs = timed_steps(batch, steps)
    infos = [batch.barostat_info(r) for r in range(count)]
    attempted = sum(i["total_attempted"] for i in infos) - a0
    accepted = sum(i["total_accepted"] for i in infos)
    vol = [float(np.prod(np.diag(batch.box(r)))) for r in range(count)]
    vol0 = [float(np.prod(np.diag(s.box))) for s in systems]
    return {
        "replicas": count,
        "atoms": int(sum(len(s.masses) for s in systems)),
        "replica_steps_per_s": round(count * steps / secs, 2),
        "ms_per_batch_step": round(secs / steps * 1e3, 4),
        "steps": steps,
        "barostat": {"pressure_bar": 1.0, "frequency": 25, "attempts_timed": attempted,
                     "accepted_total": accepted, "attempted_total": sum(i["total_attempted"] for i in infos),
                     "mean_volume_ratio": round(float(np.mean(np.array(vol) / np.array(vol0))), 5)},
        "alpha_grid": "fixed at initialisation (reference behaviour)",
    }


def run_large(steps, equilibrate, seed=30000):
    """One 300,000-atom system: 100,000 flexible waters in a 14.42 nm box,
    cutoff 1.0 nm, PME tolerance 5e-4 (alpha 2.628261, grid 116^3), Langevin
    1 fs at 300 K.  Besides ns/day it reports, per phase, the device time and
    the HBM rate implied by the phase's algorithmic bytes (the least data the
    phase must move, listed in ALGORITHMIC_BYTES)."""
    s = ommhip.water_system(100000, 14.42, seed, cutoff=1.0, tolerance=5e-4)
    ctx = ommhip.Context(s, temperature=300.0, friction=1.0, step_size=DT_PS, seed=seed,
                         precision=ommhip.PRECISION_MIXED)
    ctx.set_graph_steps(GRAPH_STEPS)
    ctx.set_velocities_to_temperature(300.0, seed + 1)
    equilibrate_then(ctx, equilibrate, 100)
    secs = timed_steps(ctx, steps)
    phases, npairs = ctx.profile(10)
    tiles, entries = ctx.list_stats()
    alpha, grid = pme_rule(1.0, 5e-4, 14.42)
    n = len(s.masses)
    g = int(np.prod(grid))
    nbytes = algorithmic_bytes(n, g, entries)
    rates = {k: round(nbytes[k] / (phases[k] * 1e-3) / 1e9, 1) for k in nbytes if phases.get(k, 0) > 0}
    gpts = int(np.prod(grid))
    sroof = step_roofline(npairs, n, n, gpts, secs / steps)
    proof = phase_roofline(phases, npairs, n, n, gpts)
    ctx.close()
    return {
        "workload": "100000 flexible 3-site waters (300000 atoms), 14.42 nm cubic box, PME cutoff 1.0 nm "
                    f"tol 5e-4 (alpha {alpha:.6f}, grid {grid[0]}x{grid[1]}x{grid[2]}), LangevinMiddle 1 fs 300 K",
        "atoms": n,
        "value": round(ns_per_day(steps, secs), 3),
        "unit": "ns/day",
        "ms_per_step": round(secs / steps * 1e3, 4),
        "steps": steps,
        "pairs_per_step": npairs,
        "pair_tflops": round(FLOP_PER_PAIR * npairs / (phases["direct_pairs"] * 1e-3) / 1e12, 3),
        "phases_ms": {k: round(v, 5) for k, v in phases.items()},
        "algorithmic_bytes": nbytes,
        "effective_gbps": rates,
        "effective_gbps_note": "algorithmic bytes / serial HIP-event time of each phase, repeated on the same "
                               "buffers: the working sets fit the 256 MB Infinity Cache, so these are effective "
                               "rates, not HBM traffic (measured HBM bytes per kernel: rocprofv3 FETCH_SIZE / "
                               "WRITE_SIZE summaries)",
        "hbm_peak_gbps": HBM_PEAK_GBPS,
        "step_roofline": sroof,
        "phase_roofline": proof,
    }


def pme_rule(cutoff, tol, edge):
    """alpha and grid of a cubic box by the reference's rule (nonbonded.rs:647-658)."""
    alpha = np.sqrt(-np.log(2.0 * tol)) / cutoff
    n = max(6, int(np.ceil(2.0 * alpha * edge / (3.0 * tol ** 0.2))))
    return float(alpha), (n, n, n)


def algorithmic_bytes(n, g, entries):
    """Least HBM bytes each phase moves: FP64 positions/velocities (24 B per
    atom each), int64 fixed-point forces (24 B), FP32 particle parameters
    (12 B), j-list entries (4 B), the int64 charge grid (8 B per point) and
    the mesh's 24 B per grid point over the five FFT passes."""
    return {
        "nlist_build": n * (24 + 8) + entries * 4,
        "direct_pairs": n * (24 + 12 + 24) + entries * 4,
        "pme_spread": n * (24 + 4) + g * 8,
        "pme_fft_conv": g * 24,
        "pme_gather": n * (24 + 4 + 24) + g * 4,
        "bonded": n * (24 + 24),
        "integrator": n * 128,
    }


def step_roofline(npairs, natoms, nexc, gpoints, seconds_per_step):
    """Time the step would take at the peaks for its algorithmic work, and the
    fraction of the measured step time that is:
    t_roof = (75 P + 40 E + 1730 N) / FP32 peak + (24 G + 128 N) / HBM peak."""
    t_flop = (FLOP_PER_PAIR * npairs + FLOP_PER_EXCEPTION * nexc
              + (FLOP_SPREAD_PER_ATOM + FLOP_GATHER_PER_ATOM) * natoms) / (FP32_PEAK_TFLOPS * 1e12)
    t_mem = (MESH_BYTES_PER_POINT * gpoints + STEP_BYTES_PER_ATOM * natoms) / (HBM_PEAK_GBPS * 1e9)
    t_roof = t_flop + t_mem
    return {"t_roof_us": round(t_roof * 1e6, 3), "t_measured_us": round(seconds_per_step * 1e6, 2),
            "frac": round(t_roof / seconds_per_step, 4),
            "work": {"pairs": npairs, "exceptions": nexc, "atoms": natoms, "grid_points": gpoints}}


def phase_roofline(phases, npairs, natoms, nexc, gpoints):
    """Each phase's algorithmic rate (serial HIP-event time) against its
    roofline: FLOP phases against the FP32 vector peak, byte phases against HBM."""
    work = {
        "direct_pairs": (FLOP_PER_PAIR * npairs, "TFLOP/s"),
        "exceptions": (FLOP_PER_EXCEPTION * nexc, "TFLOP/s"),
        "pme_spread": (FLOP_SPREAD_PER_ATOM * natoms, "TFLOP/s"),
        "pme_gather": (FLOP_GATHER_PER_ATOM * natoms, "TFLOP/s"),
        "pme_fft_conv": (MESH_BYTES_PER_POINT * gpoints, "GB/s"),
    }
    out = {}
    for k, (w, unit) in work.items():
        t = phases.get(k, 0.0) * 1e-3
        if t <= 0:
            continue
        if unit == "TFLOP/s":
            a = w / t / 1e12
            out[k] = {"achieved": round(a, 3), "unit": unit, "peak": FP32_PEAK_TFLOPS,
                      "frac": round(a / FP32_PEAK_TFLOPS, 5), "frac_of_nonpacked": round(a / FP32_NONPACKED_TFLOPS, 5)}
        else:
            a = w / t / 1e9
            out[k] = {"achieved": round(a, 1), "unit": unit, "peak": HBM_PEAK_GBPS, "frac": round(a / HBM_PEAK_GBPS, 5)}
    t = (phases.get("integrator", 0.0) + phases.get("bonded", 0.0)) * 1e-3
    if t > 0:
        a = STEP_BYTES_PER_ATOM * natoms / t / 1e9
        out["integrator_and_bonded"] = {"achieved": round(a, 1), "unit": "GB/s", "peak": HBM_PEAK_GBPS,
                                        "frac": round(a / HBM_PEAK_GBPS, 5)}
    return out


def host_info():
    """Cores this process may use and the CPU model (lscpu's 'Model name')."""
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    usable = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    return {"nproc": os.cpu_count(), "usable_cores": usable, "cpu_model": model}


def _oracle_md_steps(s, steps, seed):
    """Seconds for `steps` full MD steps of a water system with the FP64 oracle
    (nonbonded + bonds + angles + LangevinMiddle), single-threaded."""
    import oracle as orc

    desc = s.nonbonded.desc()
    rng = np.random.default_rng(seed)
    pos = s.positions.copy()
    vel = rng.normal(size=pos.shape) * np.sqrt(orc.boltzmann() * 300.0 / s.masses)[:, None]
    t0 = time.perf_counter()
    for _ in range(steps):
        F, _, _ = orc.nb_evaluate(desc, pos, s.box)
        fb, _ = orc.harmonic_bonds(s.bonds, s.bond_params, pos)
        fa, _ = orc.harmonic_angles(s.angles, s.angle_params, pos)
        pos, vel = orc.lfmiddle_step(pos, vel, F + fb + fa, s.masses, DT_PS, 300.0, 1.0,
                                     rng.normal(size=pos.size))
    return time.perf_counter() - t0


def _replica_steps(args):
    r, steps = args
    s = batch_workload()[r]
    return r, len(s.masses), _oracle_md_steps(s, steps, r)


_C1_REFERENCE = {}


def compare_c1_with_device():
    """The device operator on the 9,000-atom input the CPU baseline evaluated,
    against that evaluation: force RMS and largest per-atom relative error,
    energy relative error, pair set."""
    r = _C1_REFERENCE
    s = r["system"]
    op = ommhip.GpuNonbondedForce(s.nonbonded)
    op.initialize(s.box)
    F, E = op.calc_forces_and_energy(s.positions, s.box)
    ref = r["forces"]
    per_atom = np.linalg.norm(F - ref, axis=1) / np.maximum(np.linalg.norm(ref, axis=1), 1e-300)
    return {
        "workload": "3000 waters (9000 atoms), 4.48 nm, PME 0.9 nm tol 5e-4, mixed precision",
        "force_rms_rel": float(np.sqrt(((F - ref) ** 2).sum() / (ref ** 2).sum())),
        "force_max_atom_rel": float(per_atom.max()),
        "energy_rel": float(abs(E - r["energy"]) / abs(r["energy"])),
        "pair_set_equal": op.pair_stats(s.positions, s.box) == r["pairs"],
        "pairs": int(r["pairs"][0]),
        "bars": {"force_rms_rel": 2e-6, "energy_rel": 1e-6},
    }


def cpu_baseline(sample_steps, c1_evals, c4_steps):
    """The FP64 C++ restatement of the reference path (the oracle), timed on
    this host: the 23,550-atom MD step (ns/day extrapolated from a bounded
    sample), the 9,000-atom energy+force evaluation (median of c1_evals after
    one warmup) and the 64-replica workload with one replica per usable core
    (at most 16, the host share of one GPU) run side by side, beside the
    one-core rate of the same replicas."""
    import multiprocessing as mp

    import oracle as orc

    host = host_info()
    dt = _oracle_md_steps(workload(), sample_steps, 0)
    s1 = ommhip.water_system(3000, 4.48, 11, cutoff=0.9, tolerance=5e-4)
    d1 = s1.nonbonded.desc()
    orc.nb_evaluate(d1, s1.positions, s1.box)  # warmup
    times = []
    for _ in range(c1_evals):
        t0 = time.perf_counter()
        orc.nb_evaluate(d1, s1.positions, s1.box)
        times.append(time.perf_counter() - t0)
    c1 = 1.0 / float(np.median(times))
    # The last evaluation is kept as the checker of the device operator on the
    # same input (compare_c1_with_device, after the GPU runs).
    F1, E1, _ = orc.nb_evaluate(d1, s1.positions, s1.box)
    _C1_REFERENCE.update(system=s1, forces=F1, energy=E1, pairs=orc.nb_pairs(d1, s1.positions, s1.box))
    cores = max(1, min(16, host["usable_cores"], 64))
    ctx = mp.get_context("fork")  # the children only run the CPU oracle
    t0 = time.perf_counter()
    with ctx.Pool(cores) as pool:
        res = pool.map(_replica_steps, [(r, c4_steps) for r in range(cores)])
    wall = time.perf_counter() - t0
    per_replica = [sec / c4_steps for _, _, sec in res]
    return {
        "value": ns_per_day(sample_steps, dt),
        "unit": "ns/day",
        "cores": 1,
        "kind": "port",
        "extrapolated": True,
        "sample": f"{sample_steps} full MD steps of the same 23,550-atom workload with the single-threaded "
                  f"FP64 C++ restatement (cell list, exclusion hash set, c2c FFT), {dt:.1f} s, extrapolated to ns/day",
        "host": host,
        "c1_evaluations_per_s": {"value": round(c1, 4), "workload": "3000 waters (9000 atoms), 4.48 nm, PME 0.9 nm",
                                 "protocol": f"median of {c1_evals} energy+force evaluations after one warmup",
                                 "cores": 1},
        "c4_replica_steps_per_s": {
            "value": round(cores * c4_steps / wall, 3), "cores": cores,
            "one_core": round(1.0 / float(np.mean(per_replica)), 3),
            "sample": f"replicas 0..{cores - 1} of the 64-replica workload, one per core, {c4_steps} steps each, "
                      f"{wall:.1f} s wall; one_core = 1 / mean seconds per replica-step",
        },
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=1000)
    ap.add_argument("--equilibrate", type=int, default=2000)
    ap.add_argument("--plain-steps", type=int, default=200)
    ap.add_argument("--cpu-steps", type=int, default=12)
    ap.add_argument("--cpu-c1-evals", type=int, default=5)
    ap.add_argument("--cpu-c4-steps", type=int, default=2)
    ap.add_argument("--profile-reps", type=int, default=50)
    ap.add_argument("--batch-steps", type=int, default=300)
    ap.add_argument("--batch-equilibrate", type=int, default=1000)
    ap.add_argument("--batch-plain-steps", type=int, default=20)
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--npt-steps", type=int, default=250)
    ap.add_argument("--npt-equilibrate", type=int, default=500)
    ap.add_argument("--no-npt", action="store_true")
    ap.add_argument("--large-steps", type=int, default=300)
    ap.add_argument("--large-equilibrate", type=int, default=600)
    ap.add_argument("--no-large", action="store_true")
    ap.add_argument("--large-only", action="store_true",
                    help="run only the 300,000-atom workload and print its record (for profiling)")
    ap.add_argument("--batch-only", action="store_true",
                    help="run only the 64-replica workload and print its record (for profiling)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--eager", action="store_true",
                    help="launch every step eagerly instead of replaying captured graphs (profiling)")
    args = ap.parse_args()
    global GRAPH_STEPS
    if args.eager:
        GRAPH_STEPS = 0
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (replicas would shard "
              "across GPUs with no communication; see DESIGN.md)", file=sys.stderr)
        sys.exit(2)
    # The CPU baseline runs first: its worker processes are forked before this
    # process touches the GPU.
    cpu = None
    if not (args.no_cpu or args.large_only or args.batch_only):
        cpu = cpu_baseline(args.cpu_steps, args.cpu_c1_evals, args.cpu_c4_steps)
    if ommhip.device_count() < 1:
        print("no gfx950 device visible", file=sys.stderr)
        sys.exit(1)

    if args.large_only:
        print(json.dumps(run_large(args.large_steps, args.large_equilibrate)), flush=True)
        return
    if args.batch_only:
        rec = run_batch(args.batch_steps, 100, args.batch_equilibrate, args.batch_plain_steps)
        print(json.dumps(rec), flush=True)
        return
    ctx, secs, psecs, g = run_single(args.steps, args.warmup, args.equilibrate, args.plain_steps)
    value = ns_per_day(args.steps, secs)
    phases, npairs = ctx.profile(args.profile_reps)
    tiles, entries = ctx.list_stats()
    t_pairs = phases["direct_pairs"] * 1e-3
    achieved = FLOP_PER_PAIR * npairs / t_pairs / 1e12

    plain = ns_per_day(args.plain_steps, psecs)
    natoms = 23550
    nexc = 23550  # three exclusions per water
    gpoints = 55 ** 3

    line = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "ns/day",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(secs / args.steps * 1e3, 5),
        "graph_steps": g,
        "graph_replays_timed": args.steps // g if g else 0,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 pair/spline/FFT, i64 fixed-point accumulation, f64 integrator",
        "data": "synthetic: seeded jittered-lattice flexible water, random orientations, Maxwell-Boltzmann velocities",
        "config": {
            "workload": "7850 flexible 3-site waters (23550 atoms), 6.17 nm cubic box, PME cutoff 0.9 nm tol 5e-4 "
                        "(alpha 2.920290, grid 55^3), LangevinMiddle 1 fs 300 K 1/ps",
            "atoms": 23550,
            "parallelism": "single GPU",
        },
        "roofline": {
            # A FLOP-bound kernel; the schema's FLOP option is "mfma", but the
            # work runs on the FP32 vector ALUs (no matrix cores): see "unit_note".
            "bound": "mfma",
            "achieved": round(achieved, 3),
            "peak": FP32_PEAK_TFLOPS,
            "unit": "TFLOP/s",
            "frac": round(achieved / FP32_PEAK_TFLOPS, 4),
            "traffic": K_PAIRS_TRAFFIC_BYTES,
            "unit_note": "FP32 vector (VALU) work, no MFMA; peak 157.3 TFLOP/s counts packed FP32, which this build "
                         "disables, so 78.6 TFLOP/s is the attainable peak (frac_of_nonpacked)",
            "frac_of_nonpacked": round(achieved / FP32_NONPACKED_TFLOPS, 4),
            "kernel": "k_pairs (direct-space LJ + erfc-Ewald), FP32 VALU work: 75 FLOP x in-cutoff pairs per launch",
            "pairs_per_launch": npairs,
            "ms_per_launch": round(phases["direct_pairs"], 5),
            "tiles_per_launch": tiles,
            "pair_slot_efficiency": round(npairs / max(1, tiles * TILE_SLOTS), 4),
        },
        "step_roofline": step_roofline(npairs, natoms, nexc, gpoints, secs / args.steps),
        "phase_roofline": phase_roofline(phases, npairs, natoms, nexc, gpoints),
        "phases_ms": {k: round(v, 5) for k, v in phases.items()},
        "launch_tuning": ctx.launch_tuning(),
        "plain_fp64": {"value": round(plain, 3), "unit": "ns/day", "steps": args.plain_steps,
                       "speedup_mixed_over_plain": round(value / plain, 2)},
    }
    if not args.no_batch:
        line["batched"] = run_batch(args.batch_steps, 100, args.batch_equilibrate, args.batch_plain_steps)
    if not args.no_npt:
        line["batched_npt"] = run_npt(args.npt_steps, args.npt_equilibrate)
    if not args.no_large:
        line["large_system"] = run_large(args.large_steps, args.large_equilibrate)
    if cpu is not None:
        cpu["parity_vs_device"] = compare_c1_with_device()
        line["cpu_baseline"] = cpu
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
