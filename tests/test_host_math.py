"""Host checks of two device-math pieces of parallelpopgen_amd/csrc/
gf_math.hpp, run as host code (tests/cpp, built by make -C tests/cpp):

* the drift kernel's j / N (gf::quot_count: three float operations with
  r = RN(1/N), for N <= 2^22) equals the IEEE float quotient (float)j / N
  -- the reference's and the oracle's division -- bit for bit: every
  0 <= j <= N <= 32768, and N sampled up to 2^22 with j near 0, N and every
  N/64, plus random j (quotient.hip);
* the Poisson inversion's exp(-m) table (gf::expneg_tab) is within 1.5 ulp
  of exp for every float in [2^-12, 6] and every 4096th below, is 1 at 0,
  and never steps up by more than one ulp (tables.hip)."""
import os
import subprocess

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BUILD = os.path.join(REPO, "tests", "cpp", "_build")


def run_host_check(name):
    exe = os.path.join(BUILD, name)
    if not os.path.exists(exe):
        pytest.fail("tests/cpp/_build/%s is not built (make -C tests/cpp)" % name)
    q = subprocess.run(["make", "-q", "-C", os.path.join(REPO, "tests", "cpp"), exe], capture_output=True, text=True)
    if q.returncode != 0:
        pytest.fail("tests/cpp/_build/%s is older than its sources (make -C tests/cpp)" % name)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    return r.stdout


def test_quot_count_is_the_ieee_quotient():
    assert "all equal to the IEEE quotient" in run_host_check("quotient")


def test_expneg_table_error_bound():
    out = run_host_check("tables")
    assert "max error" in out, out
