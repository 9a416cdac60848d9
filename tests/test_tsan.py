"""ThreadSanitizer on the oracle's threaded path (CPU only). The oracle stepped on 4 threads -- the
pool's chunked loops, the CAS float adds of the particle scatter (AtomicOperations.h:31-49), the CAS
min of the fluid SDF, the reductions -- reports no data race; the same build with the scatter's CAS
add planted as a plain add reports one, so the check sees the kind of race it is there for. Built by
`make tsan` in oracle/ (oracle/tsan_check.cpp)."""
import os
import shutil
import subprocess

import pytest

ORACLE = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle")


@pytest.fixture(scope="module")
def tsan_build():
    if shutil.which("g++") is None or shutil.which("make") is None:
        pytest.skip("g++ / make not available")
    r = subprocess.run(["make", "-s", "-C", ORACLE, "tsan"], capture_output=True, text=True, timeout=600)
    if r.returncode != 0 and "tsan" in r.stderr.lower() and ("cannot find" in r.stderr or "not supported" in r.stderr):
        pytest.skip("no ThreadSanitizer runtime for g++: " + r.stderr[-300:])
    assert r.returncode == 0, r.stderr[-3000:]
    return ORACLE


def _run(path):
    return subprocess.run([path], capture_output=True, text=True, timeout=600)


def test_threaded_oracle_has_no_data_race(tsan_build):
    r = _run(os.path.join(tsan_build, "tsan_check"))
    print(r.stdout.strip())
    assert r.returncode == 0 and "ThreadSanitizer" not in r.stderr, r.stderr[-4000:]
    assert "success 1" in r.stdout


def test_planted_race_is_reported(tsan_build):
    r = _run(os.path.join(tsan_build, "tsan_check_race"))
    assert r.returncode != 0 and "ThreadSanitizer: data race" in r.stderr, (r.returncode, r.stderr[-2000:])
    assert "atomicAddR" in r.stderr  # the planted race, not another
