"""Sanitizer builds of the host side (CPU only): tests/cpp/host_checks.cpp built by tests/cpp/Makefile
with AddressSanitizer + UBSan (the host library, the oracle and the checks: RenderJob's frame,
callback, cancel, wait-timeout and move paths, sol's test scenes loaded from file, procedural scenes
with their BVH builds, image files, oracle renders and intersections) and with ThreadSanitizer
(the RenderJob checks), run to completion without a report."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BUILD = os.path.join(ROOT, "tests", "cpp", "_build")


@pytest.fixture(scope="module")
def built():
    jobs = str(min(8, os.cpu_count() or 1))
    r = subprocess.run(["make", "-C", os.path.join(ROOT, "tests", "cpp"), "-j", jobs], capture_output=True, text=True)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-4000:]


def run(cmd, env_extra):
    env = dict(os.environ, **env_extra)
    r = subprocess.run(cmd, capture_output=True, text=True, env=env, timeout=600)
    out = r.stdout + r.stderr
    print(out[-3000:])
    assert r.returncode == 0, out[-4000:]
    assert "all host checks passed" in r.stdout
    assert "ERROR: AddressSanitizer" not in out and "WARNING: ThreadSanitizer" not in out
    assert "runtime error" not in out          # UBSan


def test_host_checks_under_address_and_ub_sanitizers(built, tmp_path):
    run([os.path.join(BUILD, "host_checks_asan"), os.path.join(ROOT, "tests", "golden", "cornell"), str(tmp_path)],
        {"ASAN_OPTIONS": "detect_leaks=1:halt_on_error=1:verify_asan_link_order=0", "UBSAN_OPTIONS": "print_stacktrace=1",
         "OMP_NUM_THREADS": "4"})


def test_render_job_under_thread_sanitizer(built):
    run([os.path.join(BUILD, "render_job_tsan")], {"TSAN_OPTIONS": "halt_on_error=1"})
