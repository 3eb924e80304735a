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
workload and the CPU baseline (the FP64 C++ restatement of the reference,
single-threaded as the reference is).
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
FP32_PEAK_TFLOPS = 157.3
FLOP_PER_PAIR = 75


def workload(seed=2024):
    s = ommhip.water_system(7850, 6.17, seed, cutoff=0.9, tolerance=5e-4)
    return s


def ns_per_day(steps, seconds):
    return steps / seconds * DT_PS * 86400.0 / 1000.0


def run_gpu(precision, steps, warmup, seed=2024):
    s = workload(seed)
    ctx = ommhip.Context(s, temperature=300.0, friction=1.0, step_size=DT_PS, seed=seed, precision=precision)
    ctx.set_velocities_to_temperature(300.0, seed + 1)
    ctx.step(warmup)
    t0 = time.perf_counter()
    ctx.step(steps)  # returns after the device finished (stream synchronised)
    t1 = time.perf_counter()
    return ctx, t1 - t0


def cpu_baseline(sample_steps):
    import oracle as orc

    s = workload()
    box = s.box
    desc = s.nonbonded.desc()
    rng = np.random.default_rng(0)
    pos = s.positions.copy()
    vel = rng.normal(size=pos.shape) * np.sqrt(orc.boltzmann() * 300.0 / s.masses)[:, None]
    t0 = time.perf_counter()
    for _ in range(sample_steps):
        F, _, _ = orc.nb_evaluate(desc, pos, box)
        fb, _ = orc.harmonic_bonds(s.bonds, s.bond_params, pos)
        fa, _ = orc.harmonic_angles(s.angles, s.angle_params, pos)
        pos, vel = orc.lfmiddle_step(pos, vel, F + fb + fa, s.masses, DT_PS, 300.0, 1.0,
                                     rng.normal(size=pos.size))
    dt = time.perf_counter() - t0
    return {
        "value": ns_per_day(sample_steps, dt),
        "unit": "ns/day",
        "cores": 1,
        "kind": "port",
        "sample": f"{sample_steps} full MD steps of the same 23,550-atom workload with the single-threaded "
                  f"FP64 C++ restatement (cell list, exclusion hash set, c2c FFT), {dt:.1f} s",
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=500)
    ap.add_argument("--plain-steps", type=int, default=200)
    ap.add_argument("--cpu-steps", type=int, default=5)
    ap.add_argument("--profile-reps", type=int, default=50)
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (replicas would shard "
              "across GPUs with no communication; see DESIGN.md)", file=sys.stderr)
        sys.exit(2)
    if ommhip.device_count() < 1:
        print("no gfx950 device visible", file=sys.stderr)
        sys.exit(1)

    ctx, secs = run_gpu(ommhip.PRECISION_MIXED, args.steps, args.warmup)
    value = ns_per_day(args.steps, secs)
    phases, npairs = ctx.profile(args.profile_reps)
    t_pairs = phases["direct_pairs"] * 1e-3
    achieved = FLOP_PER_PAIR * npairs / t_pairs / 1e12

    _, psecs = run_gpu(ommhip.PRECISION_DOUBLE, args.plain_steps, 20)
    plain = ns_per_day(args.plain_steps, psecs)

    line = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "ns/day",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(secs / args.steps * 1e3, 5),
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
            "bound": "mfma",
            "achieved": round(achieved, 3),
            "peak": FP32_PEAK_TFLOPS,
            "unit": "TFLOP/s",
            "frac": round(achieved / FP32_PEAK_TFLOPS, 4),
            "traffic": None,
            "kernel": "k_pairs (direct-space LJ + erfc-Ewald), FP32 VALU work: 75 FLOP x in-cutoff pairs per launch",
            "pairs_per_launch": npairs,
            "ms_per_launch": round(phases["direct_pairs"], 5),
        },
        "phases_ms": {k: round(v, 5) for k, v in phases.items()},
        "plain_fp64": {"value": round(plain, 3), "unit": "ns/day", "steps": args.plain_steps,
                       "speedup_mixed_over_plain": round(value / plain, 2)},
    }
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(args.cpu_steps)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
