"""Edge cases of the GPU run_sim path against the oracle: a run in which no
mutation ever arises, the maximum of 8 populations, a time sample in every
generation, the reference's parameter checks (check_sim_parameters,
go_fish_impl.cuh:431-452) rejecting a run before anything is simulated, and
runs on device memory that earlier calls released."""
import numpy as np
import pytest

from parallelpopgen_amd import scenarios as S
from parallelpopgen_amd.capi import GF_INIT_BLANK, GoFishError
from test_gpu_parity import assert_same

pytestmark = pytest.mark.gpu


def test_no_mutations_at_all(gpu, oracle):
    """blank start and mu = 0: every sample is empty, every live count 0,
    and the SFS of an empty sample puts all sites in bin 0"""
    d = S.constant(N=1000, L=1e6, G=50, mu=0.0, init_mode=GF_INIT_BLANK, sample_every=10)
    g = gpu.run(d)
    o = oracle.run(d)
    assert_same(g, o, "empty")
    assert all(s["num_mutations"] == 0 for s in g["samples"])
    assert not g["live"].any() and len(g["ids"]) == 0
    empty = np.zeros(0, np.float32)
    s_g = gpu.site_frequency_spectrum(empty, 10, d.num_sites)
    s_o = oracle.site_frequency_spectrum(empty, 10, d.num_sites)
    assert np.array_equal(s_g.view(np.uint32), s_o.view(np.uint32))
    assert s_g[0] == np.float32(d.num_sites) and not s_g[1:].any()


def test_eight_populations(gpu, oracle):
    """P = 8 (the maximum): migration between all pairs, selection, one
    population inbred"""
    d = S.constant(N=1500, L=3e7, G=80, P=8, m=2e-3, s=S.selection_from_gamma(-5.0, 1500, 0.0), h=0.3,
                   sample_every=40)
    d.params["F"][:, 3] = 0.25
    g = gpu.run(d)
    o = oracle.run(d)
    assert_same(g, o, "P=8")
    assert g["samples"][-1]["num_mutations"] > 1000


def test_sample_every_generation(gpu, oracle):
    d = S.constant(N=800, L=2e7, G=40, s=S.selection_from_gamma(3.0, 800, 0.0), sample_every=1)
    g = gpu.run(d)
    o = oracle.run(d)
    assert len(g["samples"]) == 41
    assert_same(g, o, "every generation")


@pytest.mark.parametrize("field,value,text", [
    ("F", 1.5, "inbreeding coefficient > 1"),
    ("F", -0.1, "inbreeding coefficient < 0"),
    ("mu", -1e-9, "mu_rate < 0"),
])
def test_parameter_errors(gpu, oracle, field, value, text):
    """a bad parameter at generation 7 fails the whole run with the
    reference's message, on the GPU and in the oracle"""
    d = S.constant(N=1000, L=1e6, G=20)
    d.params[field][7, 0] = value
    with pytest.raises(GoFishError) as e:
        gpu.run(d)
    assert text in str(e.value) and "generation 7" in str(e.value)
    with pytest.raises(RuntimeError):
        oracle.run(d)


def test_migration_must_sum_to_one(gpu):
    d = S.constant(N=1000, L=1e6, G=10, P=2, m=0.01)
    d.migration[3, 0, 0] = 0.5      # column 0 of generation 3 now sums to 0.51
    with pytest.raises(GoFishError) as e:
        gpu.run(d)
    assert "migration rate does not sum to 1" in str(e.value)


def test_reused_device_memory(gpu, oracle):
    """device memory a call releases is kept by the library and handed to
    later calls with whatever it held: a larger run with other values
    between two identical runs must not change the second one"""
    d = S.plain_small(G=120)
    a = gpu.run(d)
    big = S.constant(N=3000, L=2e8, G=60, P=3, m=1e-3, s=S.selection_from_gamma(4.0, 3000, 0.0), h=0.2,
                     sample_every=20)
    gpu.run(big)
    b = gpu.run(d)
    assert_same(b, a, "after a larger run")
    assert_same(b, oracle.run(d), "oracle")


def test_reused_device_memory_batched(gpu):
    from test_gpu_batch import make_batch
    sims = make_batch(12, np.random.default_rng(5))
    a = gpu.batch_run(sims, compact_interval=16, sfs_n=30, keep_states=True)
    gpu.batch_run(make_batch(40, np.random.default_rng(6), L_scale=4.0), compact_interval=8, sfs_n=50)
    b = gpu.batch_run(sims, compact_interval=16, sfs_n=30, keep_states=True)
    assert np.array_equal(a["sfs"].view(np.uint32), b["sfs"].view(np.uint32))
    assert np.array_equal(a["num_mutations"], b["num_mutations"])
    for s, ((fa, ia), (fb, ib)) in enumerate(zip(a["states"], b["states"])):
        assert np.array_equal(fa.view(np.uint32), fb.view(np.uint32)), s
        assert np.array_equal(ia, ib), s


def test_release_device_cache(gpu, oracle):
    """gf_release_device_cache returns the kept blocks to the device; the
    next run allocates afresh and gives the same result.  A result's own
    device buffers (the final state kept for device-side spectra) outlive
    the run's stream and stay valid across the release."""
    d = S.plain_small(G=80)
    a = gpu.run(d)
    lib = gpu.load()
    h = gpu.C.c_void_p()
    gpu.check(lib.gf_run(gpu.C.byref(d.desc), gpu.C.byref(h)))
    try:
        assert lib.gf_release_device_cache(-1) == 0
        last = lib.gf_result_num_samples(h) - 1
        out = np.empty(50, np.float32)
        gpu.check(lib.gf_result_site_frequency_spectrum(h, last, 0, 50, float(d.num_sites),
                                                         out.ctypes.data_as(gpu.C.c_void_p)))
    finally:
        lib.gf_result_free(h)
    want = oracle.site_frequency_spectrum(a["samples"][-1]["freq"][0], 50, d.num_sites)
    assert np.array_equal(out.view(np.uint32), want.view(np.uint32))
    assert lib.gf_release_device_cache(-1) == 0
    assert_same(gpu.run(d), a, "after releasing the cache")


def test_device_cache_statistics(gpu):
    """gf_device_cache_stats: a run's released buffers stay kept (bytes held
    from the device >= bytes kept), a second run of the same shape takes
    them again without a new hipMalloc, and a release returns them all
    (kept 0, one hipFree per kept block)."""
    d = S.plain_small(G=60)
    assert gpu.load().gf_release_device_cache(-1) == 0
    gpu.run(d)
    a = gpu.device_cache_stats()
    assert a["kept_blocks"] > 0 and a["held_bytes"] >= a["kept_bytes"] > 0
    gpu.run(d)
    b = gpu.device_cache_stats()
    assert b["mallocs"] == a["mallocs"] and b["frees"] == a["frees"]
    assert gpu.load().gf_release_device_cache(-1) == 0
    c = gpu.device_cache_stats()
    assert c["kept_bytes"] == 0 and c["kept_blocks"] == 0
    assert c["frees"] == b["frees"] + b["kept_blocks"]
    assert c["held_bytes"] == b["held_bytes"] - b["kept_bytes"]
