"""AddressSanitizer + UndefinedBehaviorSanitizer runs of the host code
(`make sanitize`): the CPU oracle and the CPU baselines (gcc) over small
seeded inputs, and the C++ API's host code (Datalog front end, LIE, relations,
RA operators; hipcc, sanitizers on the host side only) compiling the shipped
.dl programs and malformed ones.  CPU only."""
import os
import subprocess

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_host_code_is_clean_under_asan_and_ubsan():
    r = subprocess.run(["make", "-s", "sanitize"], cwd=REPO, stdout=subprocess.PIPE,
                       stderr=subprocess.STDOUT, text=True, timeout=600)
    assert r.returncode == 0, r.stdout[-4000:]
    assert "oracle sanitizer run: ok" in r.stdout
    assert "front-end sanitizer run:" in r.stdout and "ok" in r.stdout.splitlines()[-1]
