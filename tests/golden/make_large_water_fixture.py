"""Writes large_water_300000.npz: the FP64 oracle's nonbonded result for the
300,000-atom water box (100,000 flexible waters, 14.42 nm, seed 30000, PME
cutoff 1.0 nm, tolerance 5e-4: alpha 2.628261, grid 116^3).

The oracle needs minutes for this system, so the parity test reads this
fixture instead: total energy, energy parts, pair count and pair-set hash,
the forces of 4,096 atoms drawn from a fixed seed and the sum of all forces
(PME forces do not sum to zero exactly), plus a SHA-256 digest of
the generated positions so the test can confirm it evaluates the same input.

Run from the repository root after building: python tests/golden/make_large_water_fixture.py
"""
import hashlib
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path[:0] = [os.path.join(ROOT, "openmm-rs_amd"), os.path.join(ROOT, "oracle"), os.path.join(ROOT, "tests")]

import numpy as np  # noqa: E402

import oracle as orc  # noqa: E402
from helpers import water  # noqa: E402

MOLECULES, EDGE, SEED, CUTOFF = 100000, 14.42, 30000, 1.0
SAMPLE = 4096


def main():
    s = water(MOLECULES, EDGE, seed=SEED, cutoff=CUTOFF)
    desc = s.nonbonded.desc()
    forces, energy, parts = orc.nb_evaluate(desc, s.positions, s.box)
    npairs, phash = orc.nb_pairs(desc, s.positions, s.box)
    idx = np.sort(np.random.default_rng(0).choice(len(s.positions), SAMPLE, replace=False)).astype(np.int64)
    alpha, grid = orc.effective_pme_parameters(5e-4, CUTOFF, s.box)
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "large_water_300000.npz")
    np.savez_compressed(
        out,
        positions_sha256=np.frombuffer(hashlib.sha256(np.ascontiguousarray(s.positions).tobytes()).digest(),
                                       dtype=np.uint8),
        alpha=alpha, grid=np.asarray(grid, np.int64),
        energy=energy, parts=np.asarray(parts, np.float64),
        num_pairs=np.int64(npairs), pair_hash=np.uint64(phash),
        sample=idx, sample_forces=forces[idx],
        force_norm2=np.float64((forces ** 2).sum()),
        force_sum=forces.sum(axis=0),
    )
    print(f"wrote {out}: E={energy:.9f} pairs={npairs}")


if __name__ == "__main__":
    main()
