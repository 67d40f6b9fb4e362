"""The CPU oracle under AddressSanitizer and UndefinedBehaviorSanitizer
(gcc, host code): oracle/sanitize_check.c calls every entry point the
parity tests use on small inputs (codec at every member kind, NaN/Inf and
out-of-range values, range stores, the MPM phases and substeps of each
dimension and material in f32 and f64, quantized substeps, the adjoint with
and without the SDF collider) and checks that a range store leaves other
particles' words untouched and that decode is a projection.  Any invalid
access or undefined operation ends the run with a sanitizer report."""
import os
import shutil
import subprocess

import pytest

ORACLE = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle")


def _env():
    env = dict(os.environ)
    env["ASAN_OPTIONS"] = "detect_leaks=1:abort_on_error=0:exitcode=23"
    env["UBSAN_OPTIONS"] = "print_stacktrace=1:halt_on_error=1:exitcode=24"
    env["OMP_NUM_THREADS"] = "2"
    return env


@pytest.fixture(scope="module")
def sanitized(tmp_path_factory):
    if shutil.which("gcc") is None or shutil.which("make") is None:
        pytest.skip("gcc or make not available")
    # built into this process's own directory: parallel test workers never share the file
    exe = str(tmp_path_factory.mktemp("san") / "sanitize_check")
    r = subprocess.run(["make", "-s", "-C", ORACLE, "sanitize", f"SAN_EXE={exe}"], capture_output=True, text=True,
                       timeout=600)
    if r.returncode != 0 and "cannot find" in (r.stderr + r.stdout) and "asan" in (r.stderr + r.stdout):
        pytest.skip("the sanitizer runtimes are not installed")
    assert r.returncode == 0, r.stderr[-2000:]
    return exe


def test_oracle_is_clean_under_sanitizers(sanitized):
    r = subprocess.run([sanitized], capture_output=True, text=True, timeout=600, env=_env())
    assert r.returncode == 0, (r.returncode, r.stderr[-4000:])
    assert "sanitize check ok" in r.stdout
    assert "runtime error" not in r.stderr and "AddressSanitizer" not in r.stderr


def test_sanitizers_are_live(sanitized):
    """A planted heap overflow (a two-word record packed into one word) is
    reported: the clean run above is not clean for want of a sanitizer."""
    r = subprocess.run([sanitized, "--plant-overflow"], capture_output=True, text=True, timeout=120, env=_env())
    assert r.returncode != 0
    assert "AddressSanitizer" in r.stderr and "heap-buffer-overflow" in r.stderr
