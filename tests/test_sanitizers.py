"""The CPU code under AddressSanitizer and UndefinedBehaviorSanitizer.

The host-side program lowering (isosurface_amd/csrc/program.cpp) runs under
a seeded fuzz driver (tests/sanitize/program_fuzz.cpp), and the oracle's
own test files run against a sanitizer build of the oracle library
(oracle/Makefile, target sanitize; selected through ISO_ORACLE_LIB, with the
sanitizer runtime preloaded into the interpreter).  Any out-of-bounds
access, use after free or undefined behaviour fails the run."""
import os
import subprocess
import sys
from pathlib import Path

ROOT = Path(__file__).resolve().parents[1]
SAN = ROOT / "tests" / "sanitize"

ORACLE_TESTS = ["tests/test_oracle.py", "tests/test_oracle_dc.py", "tests/test_oracle_lhmc.py",
                "tests/test_oracle_emc.py", "tests/test_oracle_directed.py",
                "tests/test_oracle_point_cloud.py", "tests/test_numbering.py",
                "tests/test_band_certificate.py", "tests/test_tables.py"]


def build():
    subprocess.run(["make", "-s", "-C", str(SAN), "all"], check=True)


def test_program_lowering_under_sanitizers():
    build()
    r = subprocess.run([str(SAN / "_build" / "program_fuzz"), "20000"], capture_output=True,
                       text=True, timeout=600)
    print(r.stdout, r.stderr)
    assert r.returncode == 0, r.stderr


def test_oracle_under_sanitizers():
    build()
    asan = subprocess.run(["g++", "-print-file-name=libasan.so"], capture_output=True, text=True,
                          check=True).stdout.strip()
    env = dict(os.environ)
    env["LD_PRELOAD"] = asan
    env["ASAN_OPTIONS"] = "detect_leaks=0"  # the interpreter's own allocations
    env["UBSAN_OPTIONS"] = "print_stacktrace=1"
    env["ISO_ORACLE_LIB"] = str(ROOT / "oracle" / "_san" / "liboracle.so")
    r = subprocess.run([sys.executable, "-m", "pytest", "-q", "-x", "-p", "no:cacheprovider",
                        "-m", "not gpu", *ORACLE_TESTS], cwd=str(ROOT), env=env,
                       capture_output=True, text=True, timeout=1200)
    print(r.stdout[-3000:], r.stderr[-3000:])
    assert r.returncode == 0
    assert "passed" in r.stdout
