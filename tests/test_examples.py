"""Example programs written against the C++ API (include/go_fish.cuh,
spectrum.h, go_fish_data_struct.h): they build with hipcc (and g++ for the
plain C++ driver), and on the GPU reproduce the reference's published
example outputs statistically."""
import json
import os
import subprocess

import numpy as np
import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BUILD = os.path.join(REPO, "examples", "_build")
GOLDEN = os.path.join(REPO, "tests", "golden")


# the reference's own example_compilation sources (compiled unchanged where
# the reference checkout is present; the binary travels with the build)
REF_EXAMPLE = "/root/reference/examples/example_compilation"


def test_examples_build():
    r = subprocess.run(["make", "-s", "-C", os.path.join(REPO, "examples")], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-3000:]
    for exe in ("speed", "dadi", "compilation", "custom"):
        assert os.access(os.path.join(BUILD, exe), os.X_OK)
    if os.path.exists(os.path.join(REF_EXAMPLE, "run.cu")):
        assert os.access(os.path.join(BUILD, "ref_compilation"), os.X_OK)


def run_example(name, *args, timeout=600):
    exe = os.path.join(BUILD, name)
    if not os.path.exists(exe):
        pytest.fail("example %s is not built (make -C examples)" % name)
    # the binary must be current: make's question mode exits non-zero if a
    # source or header is newer than what was built from it
    q = subprocess.run(["make", "-q", "-C", os.path.join(REPO, "examples"), os.path.join(BUILD, name)],
                       capture_output=True, text=True)
    if q.returncode != 0:
        pytest.fail("example %s is older than its sources (make -C examples)" % name)
    r = subprocess.run([exe] + [str(a) for a in args], capture_output=True, text=True, timeout=timeout, cwd=BUILD)
    assert r.returncode == 0, r.stderr[-2000:]
    return r.stdout


def parse_dadi(out):
    mean, sd, tail = [], [], {}
    for line in out.splitlines():
        t = line.split("\t")
        if len(t) == 4 and t[0].isdigit():
            mean.append(float(t[1]))
            sd.append(float(t[2]))
        elif ":" in line:
            k, v = line.split(":", 1)
            try:
                tail[k.strip()] = float(v)
            except ValueError:
                pass
    return np.array(mean), np.array(sd), tail


@pytest.mark.gpu
@pytest.mark.parametrize("L,key", [(2e9, "run_980"), (2e8, "run_980_2")])
def test_dadi_scenario_matches_reference_runs(L, key):
    """50 runs of the dadi scenario: every normalised SFS bin agrees with the
    reference's 50-run mean within the combined standard error (|z| < 4.5
    for every bin, and the mean squared z below 1.6); the mean
    number of mutations in the SFS agrees within 1 %."""
    ref = json.load(open(os.path.join(GOLDEN, "dadi_sfs.json")))[key]
    m, s, tail = parse_dadi(run_example("dadi", L, 50))
    rm, rs = np.array(ref["mean"]), np.array(ref["sd"])
    assert len(m) == len(rm) == 1000
    z = (m - rm) / np.sqrt(s ** 2 / 50 + rs ** 2 / 50)
    assert np.all(np.abs(z) < 4.5), (np.argmax(np.abs(z)) + 1, z[np.argmax(np.abs(z))])
    # Only an upper bound on the mean squared z: the 1,000 bins of one
    # 50-run mean are strongly positively correlated (one run's extra
    # mutations raise all its normalised bins together, and the same 50 runs
    # feed every bin), so mean(z^2) over bins does not concentrate at 1 --
    # it scatters like a chi-square with few effective degrees of freedom,
    # well below 1 as often as above (0.48 in one run of the L = 2e8 case).
    # Too large a value still flags a biased model; too small one flags
    # nothing.
    assert np.mean(z ** 2) < 1.6, np.mean(z ** 2)
    nm = tail["average number of mutations in SFS"]
    assert abs(nm / ref["average_number_of_mutations_in_SFS"] - 1) < 0.01, nm


@pytest.mark.gpu
def test_speed_example_output():
    ref = json.load(open(os.path.join(GOLDEN, "speed_snps.json")))
    vals = [r["snps"] for g in ("780M", "980") for r in ref[g] if r["xL"] == 20]
    out = run_example("speed", 20, 15)
    n = int([l for l in out.splitlines() if l.startswith("number of mutations")][0].split("\t")[1])
    assert abs(n - np.mean(vals)) < 4.5 * np.sqrt(np.mean(vals)) * 1.25, (n, vals)


def check_compilation_output(out, bfile):
    """the example_compilation scenario's printed summary and its bfile.dat"""
    lines = out.splitlines()
    ns = int([l for l in lines if l.startswith("number of time samples")][0].split(":")[1])
    m0 = int([l for l in lines if l.startswith("mutations in first time sample")][0].split(":")[1])
    m1 = int([l for l in lines if l.startswith("mutations in final time sample")][0].split(":")[1])
    assert ns == 2
    # reference example_compilation.out: 5288 and 16736 (a single run)
    assert 0.7 * 5288 < m0 < 1.3 * 5288 and 0.7 * 16736 < m1 < 1.3 * 16736, (m0, m1)
    text = open(bfile).read().splitlines()
    assert text[0].startswith("seed1\t-1091571699") and text[1].startswith("seed2\t-557122643")
    assert text[10].startswith("time sample:") and text[11].split("\t")[1:3] == ["20000", ""]
    rows = [l for l in text if l.startswith("(")]
    assert len(rows) == m1
    freqs = np.array([[float(x) for x in r.split("\t")[1:]] for r in rows[:500]])
    assert freqs.shape[1] == 4
    assert np.allclose(np.round(freqs * 1000), freqs * 1000, atol=1e-3)   # 1000 haploid chromosomes


@pytest.mark.gpu
def test_mixed_compiler_example(tmp_path):
    out = run_example("compilation", str(tmp_path / "bfile.dat"))
    check_compilation_output(out, tmp_path / "bfile.dat")


@pytest.mark.gpu
def test_reference_compilation_example_unchanged(tmp_path):
    """The reference's examples/example_compilation (run.cu through hipcc as
    HIP, main.cpp through g++, neither file changed) built against include/
    and libgofish.so: it runs and prints what the retyped copy prints.  It
    writes bfile.dat to its working directory."""
    exe = os.path.join(BUILD, "ref_compilation")
    if not os.path.exists(exe):
        pytest.skip("built only where the reference's example sources are present (make -C examples)")
    newest = max(os.path.getmtime(f) for f in
                 [os.path.join(REPO, "parallelpopgen_amd", "libgofish.so")] +
                 [os.path.join(REPO, "include", x) for x in os.listdir(os.path.join(REPO, "include"))])
    if os.path.getmtime(exe) < newest:
        pytest.fail("ref_compilation is older than the library or its headers (make -C examples)")
    r = subprocess.run([exe], capture_output=True, text=True, timeout=600, cwd=tmp_path)
    assert r.returncode == 0, r.stderr[-2000:]
    check_compilation_output(r.stdout, tmp_path / "bfile.dat")


@pytest.mark.gpu
def test_selection_functors_without_linear_form():
    """A selection functor with only the reference's operator() runs through
    gf_run_with_selection (the drift kernel compiled in the program with the
    functor): a functor with the values of selection_constant or
    selection_linear_frequency_dependent gives the same run bit for bit, on
    one and two populations; a non-linear functor (s = -1 above frequency
    1/2, which halves such a mutation every generation) keeps every final
    frequency near or below 1/2 while the neutral run has some far above.
    Also: the C++ joint two-population SFS of the two-population run sums
    to num_sites and its population-0 marginal equals the C++ marginal SFS
    (bins 1..n-1: 1e-3 of max(1, bin); each product is truncated to 2^-32,
    at most (n_b + 1) M 2^-32 per bin)."""
    out = run_example("custom")
    rows = {l.split("\t")[0]: l.split("\t")[1:] for l in out.strip().splitlines()}
    for key in ("constant", "linear", "two_pops"):
        a, b, n = rows[key]
        assert a == b, (key, rows[key])
        assert int(n) > 1000, (key, n)
    total, worst, npop = rows["joint"]
    assert abs(float(total) - 1.0) < 1e-4 and float(worst) < 1e-3 and int(npop) == 2, rows["joint"]
    mx_sel, mx_neutral, n = rows["halving"]
    assert float(mx_sel) < 0.6, rows["halving"]
    assert float(mx_neutral) > 0.8, rows["halving"]
