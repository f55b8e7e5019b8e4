"""The drift kernel's j / N (gf::quot_count in parallelpopgen_amd/csrc/
gf_math.hpp: three float operations with r = RN(1/N), for N <= 2^22) equals
the IEEE float quotient (float)j / N -- the reference's and the oracle's
division -- bit for bit: every 0 <= j <= N <= 32768, and N sampled up to
2^22 with j near 0, N and every N/64, plus random j (tests/cpp/quotient.hip,
host code, built by make -C tests/cpp)."""
import os
import subprocess

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(REPO, "tests", "cpp", "_build", "quotient")


def test_quot_count_is_the_ieee_quotient():
    if not os.path.exists(EXE):
        pytest.fail("tests/cpp/_build/quotient is not built (make -C tests/cpp)")
    q = subprocess.run(["make", "-q", "-C", os.path.join(REPO, "tests", "cpp"), EXE], capture_output=True, text=True)
    if q.returncode != 0:
        pytest.fail("tests/cpp/_build/quotient is older than its sources (make -C tests/cpp)")
    r = subprocess.run([EXE], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    assert "all equal to the IEEE quotient" in r.stdout
