"""Mutation tests of the oracle's pins (CPU only).

Each case copies oracle/oracle.cpp, plants one plausible mistake in the copy
(a dropped term, a flipped sign, a wrong function), builds it into a temporary
shared library with the oracle's own flags and runs the oracle's pin tests
against it in a child process (OMM_ORACLE_LIB points oracle.py at the copy).
The case passes when the pins FAIL on the broken oracle: a mistake the pins
cannot see would make the GPU parity tests, which compare against the oracle,
blind to the same mistake in the device code.  The real oracle is never
modified.
"""
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
import tempfile

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
PINS = ["tests/test_oracle_pins.py", "tests/test_oracle.py"]

# (name, exact text in oracle.cpp, replacement): each text occurs once.
MUTANTS = [
    ("exception erf correction dropped (energy and force)",
     "                e -= corr;\n                f_over_r -= deriv;\n", ""),
    ("exception erf correction dropped (energy only)",
     "                e -= corr;\n", ""),
    ("reaction-field c_rf sign flipped",
     "c_rf = (3.0 * eps_rf) / (2.0 * eps_rf + 1.0) / rc;", "c_rf = -(3.0 * eps_rf) / (2.0 * eps_rf + 1.0) / rc;"),
    ("reaction-field c_rf sign flipped (infinite dielectric)",
     "c_rf = 1.5 / rc;", "c_rf = -1.5 / rc;"),
    ("switch polynomial sign flipped",
     "double sw = 1.0 + std::pow(t, 3) * (-10.0 + t * (15.0 - 6.0 * t));",
     "double sw = 1.0 - std::pow(t, 3) * (-10.0 + t * (15.0 - 6.0 * t));"),
    ("LJPME self term dropped",
     "            energy += self;\n            part[3] += self;\n        }\n    }\n    if (include_direct && use_pbc",
     "        }\n    }\n    if (include_direct && use_pbc"),
    ("LJPME direct-space complement sign flipped",
     "lj_e += corr - shift;", "lj_e += -corr - shift;"),
    ("LJPME exception complement dropped",
     "                e += corr;\n", ""),
    ("direct-space screen erfc replaced by erf",
     "double erfc_v = std::erfc(alpha * r);", "double erfc_v = std::erf(alpha * r);"),
    ("switched dispersion integral dropped",
     "return tail + integrate(f, sw, rc);", "return tail;"),
]


def _build_and_run(mutant, workdir):
    name, old, new = mutant
    src = open(os.path.join(ORACLE_DIR, "oracle.cpp")).read()
    assert src.count(old) == 1, f"mutation site of '{name}' not found exactly once"
    d = tempfile.mkdtemp(dir=workdir)
    path = os.path.join(d, "oracle.cpp")
    with open(path, "w") as fh:
        fh.write(src.replace(old, new))
    lib = os.path.join(d, "liboracle_mutant.so")
    # oracle.h is found through -iquote, and it finds include/ommhip.h itself.
    subprocess.run(["g++", "-O1", "-ffp-contract=off", "-fPIC", "-std=c++17", "-shared", "-iquote", ORACLE_DIR,
                    "-o", lib, path], check=True)
    env = dict(os.environ, OMM_ORACLE_LIB=lib)
    r = subprocess.run([sys.executable, "-m", "pytest", "-x", "-q", "-p", "no:cacheprovider", *PINS],
                       cwd=ROOT, env=env, capture_output=True, text=True, timeout=600)
    return r.returncode, r.stdout[-2000:]


@pytest.fixture(scope="module")
def mutant_results():
    work = tempfile.mkdtemp(prefix="oracle_mutants_")
    try:
        with cf.ThreadPoolExecutor(max_workers=4) as ex:
            futures = {m[0]: ex.submit(_build_and_run, m, work) for m in MUTANTS}
            yield {k: f.result() for k, f in futures.items()}
    finally:
        shutil.rmtree(work, ignore_errors=True)


@pytest.mark.parametrize("name", [m[0] for m in MUTANTS])
def test_pins_catch_planted_mistake(name, mutant_results):
    rc, out = mutant_results[name]
    assert rc == 1, f"the oracle pins did not catch: {name}\n{out}"
