"""Host code under AddressSanitizer and UndefinedBehaviorSanitizer (CPU only).

The library's host code (descriptor validation, topology and exclusion
set-up, water-box generation, the C ABI) and the FP64 oracle are built with
both sanitizers (`make sanitize` in openmm-rs_amd/ and oracle/; device code
is built as usual) and exercised by the CPU suites in child processes with
the sanitizer runtime preloaded.  Any report fails the child: UBSan is built
without recovery and ASan stops at the first error.  Leak checking is off
(the interpreter keeps its allocations to the end)."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "openmm-rs_amd")
ORACLE = os.path.join(ROOT, "oracle")
CLANG = "/opt/rocm/llvm/bin/clang"


def _run(env_extra, tests):
    env = dict(os.environ)
    env.update(env_extra)
    env["ASAN_OPTIONS"] = "detect_leaks=0:abort_on_error=0:halt_on_error=1"
    env["UBSAN_OPTIONS"] = "halt_on_error=1:print_stacktrace=1"
    r = subprocess.run([sys.executable, "-m", "pytest", "-q", "-x", "-s", "-p", "no:cacheprovider", "-m", "not gpu"] + tests,
                       cwd=ROOT, env=env, capture_output=True, text=True, timeout=900)
    out = r.stdout + r.stderr
    assert r.returncode == 0, out[-4000:]
    assert "ERROR: AddressSanitizer" not in out and "runtime error:" not in out, out[-4000:]
    return out


def test_oracle_under_sanitizers():
    subprocess.run(["make", "-s", "-C", ORACLE, "sanitize"], check=True)
    runtime = subprocess.run(["gcc", "-print-file-name=libasan.so"], capture_output=True, text=True).stdout.strip()
    ubsan = subprocess.run(["gcc", "-print-file-name=libubsan.so"], capture_output=True, text=True).stdout.strip()
    assert os.path.isabs(runtime) and os.path.isabs(ubsan)
    out = _run({"LD_PRELOAD": f"{runtime}:{ubsan}", "OMM_ORACLE_LIB": os.path.join(ORACLE, "liboracle_san.so")},
               ["tests/test_oracle.py", "tests/test_oracle_pins.py"])
    assert "passed" in out


def test_host_library_under_sanitizers():
    if not os.path.exists(CLANG):
        pytest.skip("the ROCm clang that built the library is not present")
    subprocess.run(["make", "-s", "-j", str(min(8, os.cpu_count() or 1)), "-C", PKG, "sanitize"], check=True)
    runtime = subprocess.run([CLANG, "-print-file-name=libclang_rt.asan-x86_64.so"], capture_output=True,
                             text=True).stdout.strip()
    assert os.path.isabs(runtime), runtime
    out = _run({"LD_PRELOAD": runtime, "OMMHIP_LIBRARY": os.path.join(PKG, "libommhip_san.so")},
               ["tests/test_abi.py"])
    assert "passed" in out


def test_sanitized_run_reports_a_planted_overflow(tmp_path):
    # A copy of the oracle with one heap read past the end of a vector in
    # orc_spline, built as `make sanitize` builds it: the same child run must
    # fail with an AddressSanitizer report, so the checks above are live.
    # The children run pytest with -s: a report goes to the inherited
    # stderr, not into pytest's capture, before the process aborts.
    src = open(os.path.join(ORACLE, "oracle.cpp")).read()
    old = "    spline(fraction, weights, derivatives);\n}"
    assert src.count(old) == 1
    planted = ("    std::vector<double> probe(5, 0.0);\n"
               "    volatile double past = probe.data()[5];\n"
               "    spline(fraction, weights, derivatives);\n"
               "    weights[0] += 0.0 * past;\n}")
    (tmp_path / "oracle.cpp").write_text(src.replace(old, planted))
    for h in ("oracle.h",):
        (tmp_path / h).write_text(open(os.path.join(ORACLE, h)).read())
    lib = str(tmp_path / "liboracle_planted.so")
    subprocess.run(["g++", "-O2", "-ffp-contract=off", "-fPIC", "-std=c++17", "-g", "-fno-omit-frame-pointer",
                    "-fsanitize=address,undefined", "-fno-sanitize-recover=undefined", "-I", os.path.join(ROOT, "include"),
                    "-shared", "-o", lib, str(tmp_path / "oracle.cpp")], check=True, cwd=ORACLE)
    runtime = subprocess.run(["gcc", "-print-file-name=libasan.so"], capture_output=True, text=True).stdout.strip()
    ubsan = subprocess.run(["gcc", "-print-file-name=libubsan.so"], capture_output=True, text=True).stdout.strip()
    env = dict(os.environ)
    env.update({"LD_PRELOAD": f"{runtime}:{ubsan}", "OMM_ORACLE_LIB": lib, "ASAN_OPTIONS": "detect_leaks=0"})
    r = subprocess.run([sys.executable, "-m", "pytest", "-q", "-x", "-s", "-p", "no:cacheprovider",
                        "tests/test_oracle.py::test_spline_partition_and_derivatives"], cwd=ROOT, env=env,
                       capture_output=True, text=True, timeout=300)
    assert r.returncode != 0
    assert "heap-buffer-overflow" in r.stdout + r.stderr
