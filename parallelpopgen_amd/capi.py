"""ctypes binding of include/gofish_capi.h (the library's C ABI).

The structures below mirror the C declarations field for field.  `load()`
opens the in-tree parallelpopgen_amd/libgofish.so and raises if it has not
been built: there is no CPU fallback behind this module.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libgofish.so")

GF_OK = 0
GF_MAX_POPULATIONS = 8
GF_INIT_MSE, GF_INIT_BLANK, GF_INIT_PREV = 0, 1, 2


class gf_mutid(C.Structure):
    _fields_ = [("origin_generation", C.c_int32), ("origin_population", C.c_int32),
                ("origin_threadID", C.c_int32), ("reserved", C.c_int32)]


class gf_pop_gen(C.Structure):
    _fields_ = [("N", C.c_int32), ("F", C.c_float), ("h", C.c_float), ("mu", C.c_float),
                ("sel_slope", C.c_float), ("sel_intercept", C.c_float)]


POP_GEN_DTYPE = np.dtype([("N", np.int32), ("F", np.float32), ("h", np.float32), ("mu", np.float32),
                          ("sel_slope", np.float32), ("sel_intercept", np.float32)])
MUTID_DTYPE = np.dtype([("origin_generation", np.int32), ("origin_population", np.int32),
                        ("origin_threadID", np.int32), ("reserved", np.int32)])


class gf_sim_desc(C.Structure):
    _fields_ = [("num_populations", C.c_int32), ("num_sites", C.c_float),
                ("seed1", C.c_int32), ("seed2", C.c_int32),
                ("num_generations", C.c_int32), ("compact_interval", C.c_int32),
                ("init_mode", C.c_int32), ("start_generation", C.c_int32),
                ("init_num_mutations", C.c_int32),
                ("init_freq", C.POINTER(C.c_float)), ("init_ids", C.POINTER(gf_mutid)),
                ("init_extinct", C.POINTER(C.c_uint8)),
                ("params", C.POINTER(gf_pop_gen)), ("migration", C.POINTER(C.c_float)),
                ("preserve", C.POINTER(C.c_uint8)), ("sample", C.POINTER(C.c_uint8)),
                ("device", C.c_int32)]


class gf_timing(C.Structure):
    _fields_ = [("advance_ms", C.c_double), ("advance_launches", C.c_int64),
                ("compact_ms", C.c_double), ("compact_launches", C.c_int64),
                ("init_ms", C.c_double), ("total_ms", C.c_double),
                ("advance_slot_generations", C.c_int64),
                ("advance_timed_launches", C.c_int64), ("compact_timed_launches", C.c_int64)]


class gf_batch_sim(C.Structure):
    _fields_ = [("N", C.c_int32), ("N1", C.c_int32), ("F", C.c_float), ("h", C.c_float), ("mu", C.c_float),
                ("s", C.c_float), ("num_sites", C.c_float), ("seed1", C.c_int32), ("seed2", C.c_int32),
                ("num_generations", C.c_int32)]


BATCH_SIM_DTYPE = np.dtype([("N", np.int32), ("N1", np.int32), ("F", np.float32), ("h", np.float32), ("mu", np.float32),
                            ("s", np.float32), ("num_sites", np.float32), ("seed1", np.int32),
                            ("seed2", np.int32), ("num_generations", np.int32)])

# name -> (restype, argtypes) for every symbol include/gofish_capi.h declares
SIGNATURES = {
    "gf_run": (C.c_int32, [C.POINTER(gf_sim_desc), C.POINTER(C.c_void_p)]),
    "gf_result_num_samples": (C.c_int32, [C.c_void_p]),
    "gf_result_sample_info": (C.c_int32, [C.c_void_p, C.c_int32, C.POINTER(C.c_int32), C.POINTER(C.c_int32),
                                          C.POINTER(C.c_int32), C.POINTER(C.c_uint8)]),
    "gf_result_sample_freq": (C.c_int32, [C.c_void_p, C.c_int32, C.c_void_p]),
    "gf_result_num_ids": (C.c_int32, [C.c_void_p]),
    "gf_result_ids": (C.c_int32, [C.c_void_p, C.c_void_p]),
    "gf_result_live_counts": (C.c_int32, [C.c_void_p, C.c_void_p]),
    "gf_result_timing": (C.c_int32, [C.c_void_p, C.POINTER(gf_timing)]),
    "gf_result_free": (None, [C.c_void_p]),
    "gf_check_sim": (C.c_int32, [C.POINTER(gf_sim_desc)]),
    "gf_release_device_cache": (C.c_int32, [C.c_int32]),
    "gf_device_cache_stats": (C.c_int32, [C.c_int32, C.c_void_p]),
    "gf_site_frequency_spectrum_noalloc": (C.c_int32, [C.c_void_p, C.c_int32, C.c_int32, C.c_float, C.c_void_p,
                                                      C.c_void_p]),
    "gf_population_histogram": (C.c_int32, [C.c_void_p, C.c_int32, C.c_int32, C.c_float, C.c_void_p, C.c_int32]),
    "gf_site_frequency_spectrum": (C.c_int32, [C.c_void_p, C.c_int32, C.c_int32, C.c_float, C.c_void_p, C.c_int32]),
    "gf_batch_run": (C.c_int32, [C.c_void_p, C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_int32,
                                 C.POINTER(C.c_void_p)]),
    "gf_batch_sfs": (C.c_int32, [C.c_void_p, C.c_void_p]),
    "gf_batch_num_mutations": (C.c_int32, [C.c_void_p, C.c_void_p]),
    "gf_batch_sim_state": (C.c_int32, [C.c_void_p, C.c_int32, C.c_void_p, C.c_void_p]),
    "gf_batch_live_mutation_generations": (C.c_int32, [C.c_void_p, C.POINTER(C.c_int64)]),
    "gf_batch_timing": (C.c_int32, [C.c_void_p, C.POINTER(gf_timing)]),
    "gf_batch_tuning": (C.c_int32, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    "gf_set_batch_region_scale": (C.c_int32, [C.c_float]),
    "gf_batch_free": (None, [C.c_void_p]),
    "gf_run_with_selection": (C.c_int32, [C.c_void_p, C.c_void_p, C.c_void_p]),
    "gf_result_site_frequency_spectrum": (C.c_int32, [C.c_void_p, C.c_int32, C.c_int32, C.c_int32, C.c_float, C.c_void_p]),
    "gf_joint_site_frequency_spectrum": (C.c_int32, [C.c_void_p, C.c_void_p, C.c_int32, C.c_int32, C.c_int32, C.c_float,
                                                     C.c_void_p, C.c_int32]),
    "gf_result_joint_site_frequency_spectrum": (C.c_int32, [C.c_void_p, C.c_int32, C.c_int32, C.c_int32, C.c_int32,
                                                            C.c_int32, C.c_float, C.c_void_p]),
    "gf_sg_create": (C.c_int32, [C.c_int32, C.c_uint64, C.c_int32, C.POINTER(C.c_void_p)]),
    "gf_sg_free": (None, [C.c_void_p]),
    "gf_sg_num_params": (C.c_int32, [C.c_void_p]),
    "gf_sg_get_params": (C.c_int32, [C.c_void_p, C.c_void_p]),
    "gf_sg_set_params": (C.c_int32, [C.c_void_p, C.c_void_p]),
    "gf_sg_gradients": (C.c_int32, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int32, C.c_float,
                                    C.c_void_p, C.c_void_p]),
    "gf_sg_train": (C.c_int32, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int32, C.c_int32, C.c_int32,
                                C.c_float, C.c_float, C.c_void_p, C.c_void_p]),
    "gf_sg_predict": (C.c_int32, [C.c_void_p, C.c_void_p, C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p]),
    "gf_sg_predict_device": (C.c_int32, [C.c_void_p, C.c_void_p, C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p]),
    "gf_sg_forward": (C.c_int32, [C.c_void_p, C.c_void_p, C.c_int32, C.c_void_p]),
    "gf_device_math": (C.c_int32, [C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                   C.c_int64, C.c_int32]),
    "gf_last_error": (C.c_char_p, []),
    "gf_set_kernel_timing": (C.c_int32, [C.c_int32]),
    "gf_set_generations_per_launch": (C.c_int32, [C.c_int32]),
    "gf_device_count": (C.c_int32, [C.POINTER(C.c_int32)]),
}

_lib = None


class GoFishError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("GO_Fish error %d: %s" % (code, msg))
        self.code = code


def load():
    """Open libgofish.so (in-tree build).  Raises if it is missing."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError("parallelpopgen_amd/libgofish.so is not built; run "
                           "`python -c 'import __graft_entry__ as g; g.build()'`")
    lib = C.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        f = getattr(lib, name, None)
        if f is None:
            continue
        f.restype = res
        f.argtypes = args
    _lib = lib
    return lib


def missing_symbols():
    """C ABI symbols declared in include/gofish_capi.h but absent from the library."""
    lib = load()
    return [n for n in SIGNATURES if getattr(lib, n, None) is None]


def device_cache_stats(device=-1):
    """the library's device-buffer cache: kept bytes and blocks, bytes held
    from the device, hipMalloc and hipFree calls"""
    out = np.zeros(5, np.int64)
    check(load().gf_device_cache_stats(device, out.ctypes.data_as(C.c_void_p)))
    return dict(kept_bytes=int(out[0]), kept_blocks=int(out[1]), held_bytes=int(out[2]),
                mallocs=int(out[3]), frees=int(out[4]))


def check(rc):
    if rc != GF_OK:
        raise GoFishError(rc, load().gf_last_error().decode())
    return rc


_hip = None


def hip_runtime():
    """The HIP runtime the library is linked against (opened by its soname
    after the library, so it is the instance the library uses), with the
    few calls DeviceArray needs."""
    global _hip
    if _hip is None:
        load()
        h = C.CDLL("libamdhip64.so.7")
        h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        h.hipFree.argtypes = [C.c_void_p]
        h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        for f in (h.hipMalloc, h.hipFree, h.hipMemcpy):
            f.restype = C.c_int
        _hip = h
    return _hip


class DeviceArray:
    """An array in device memory of the current device, allocated through
    the library's own HIP runtime, for the entry points that take device
    pointers (tests and benchmarks).  `ptr` is the device address."""

    def __init__(self, shape, dtype=np.float32):
        self.shape = tuple(int(v) for v in np.atleast_1d(shape))
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        p = C.c_void_p()
        rc = hip_runtime().hipMalloc(C.byref(p), max(self.nbytes, 1))
        if rc != 0:
            raise RuntimeError("hipMalloc of %d bytes failed (%d)" % (self.nbytes, rc))
        self.ptr = p.value

    @classmethod
    def from_host(cls, a):
        a = np.ascontiguousarray(a)
        d = cls(a.shape, a.dtype)
        if d.nbytes:
            rc = hip_runtime().hipMemcpy(d.ptr, a.ctypes.data, d.nbytes, 1)   # host to device
            if rc != 0:
                raise RuntimeError("hipMemcpy to the device failed (%d)" % rc)
        return d

    def to_host(self):
        out = np.empty(self.shape, self.dtype)
        if self.nbytes:
            rc = hip_runtime().hipMemcpy(out.ctypes.data, self.ptr, self.nbytes, 2)   # device to host
            if rc != 0:
                raise RuntimeError("hipMemcpy from the device failed (%d)" % rc)
        return out

    def free(self):
        if getattr(self, "ptr", None):
            hip_runtime().hipFree(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def _ptr(a, ctype):
    return a.ctypes.data_as(C.POINTER(ctype)) if a is not None else None


class SimDesc:
    """Owns the numpy tables behind a gf_sim_desc."""

    def __init__(self, P, num_sites, seed1, seed2, num_generations, params, migration,
                 preserve=None, sample=None, compact_interval=35, init_mode=GF_INIT_MSE,
                 start_generation=0, init_freq=None, init_ids=None, init_extinct=None, device=-1):
        G = int(num_generations)
        self.P = int(P)
        self.G = G
        self.params = np.ascontiguousarray(params, dtype=POP_GEN_DTYPE).reshape(G + 1, P)
        self.migration = np.ascontiguousarray(migration, dtype=np.float32).reshape(G + 1, P, P)
        self.preserve = np.zeros(G + 1, np.uint8) if preserve is None else np.ascontiguousarray(preserve, np.uint8)
        self.sample = np.zeros(G + 1, np.uint8) if sample is None else np.ascontiguousarray(sample, np.uint8)
        self.init_freq = None if init_freq is None else np.ascontiguousarray(init_freq, np.float32)
        self.init_ids = None if init_ids is None else np.ascontiguousarray(init_ids, MUTID_DTYPE)
        self.init_extinct = None if init_extinct is None else np.ascontiguousarray(init_extinct, np.uint8)
        d = gf_sim_desc()
        d.num_populations = self.P
        d.num_sites = float(num_sites)
        d.seed1 = _i32(seed1)
        d.seed2 = _i32(seed2)
        d.num_generations = G
        d.compact_interval = int(compact_interval)
        d.init_mode = int(init_mode)
        d.start_generation = int(start_generation)
        d.init_num_mutations = 0 if self.init_ids is None else len(self.init_ids)
        d.init_freq = _ptr(self.init_freq, C.c_float)
        d.init_ids = self.init_ids.ctypes.data_as(C.POINTER(gf_mutid)) if self.init_ids is not None else None
        d.init_extinct = _ptr(self.init_extinct, C.c_uint8)
        d.params = self.params.ctypes.data_as(C.POINTER(gf_pop_gen))
        d.migration = _ptr(self.migration, C.c_float)
        d.preserve = _ptr(self.preserve, C.c_uint8)
        d.sample = _ptr(self.sample, C.c_uint8)
        d.device = int(device)
        self.desc = d
        self.num_sites = float(num_sites)
        self.seed1, self.seed2 = _i32(seed1), _i32(seed2)
        self.compact_interval = int(compact_interval)


def _i32(v):
    v = int(v) & 0xFFFFFFFF
    return v - (1 << 32) if v >= (1 << 31) else v


def read_result(lib, prefix, handle, P, G):
    """Copies a gf_result (or the oracle's twin) into numpy arrays."""
    f = lambda n: getattr(lib, prefix + n)
    ns = f("result_num_samples")(handle)
    samples = []
    for s in range(ns):
        M = C.c_int32()
        gen = C.c_int32()
        Nc = (C.c_int32 * P)()
        ext = (C.c_uint8 * P)()
        check_any(f("result_sample_info")(handle, s, C.byref(M), C.byref(gen), Nc, ext), lib, prefix)
        freq = np.empty((P, M.value), np.float32)
        check_any(f("result_sample_freq")(handle, s, freq.ctypes.data_as(C.c_void_p)), lib, prefix)
        samples.append(dict(num_mutations=M.value, generation=gen.value, Nchrom_e=list(Nc),
                            extinct=[bool(x) for x in ext], freq=freq))
    nid = f("result_num_ids")(handle)
    ids = np.empty(nid, MUTID_DTYPE)
    if nid:
        f("result_ids")(handle, ids.ctypes.data_as(C.c_void_p))
    live = np.empty(G + 1, np.int64)
    f("result_live_counts")(handle, live.ctypes.data_as(C.c_void_p))
    return dict(samples=samples, ids=ids, live=live)


def check_any(rc, lib, prefix):
    if rc != 0:
        msg = lib.gf_last_error().decode() if prefix == "gf_" else "oracle error"
        raise GoFishError(rc, msg)


def check_sim(desc: SimDesc):
    """gf_check_sim: the run's parameter checks on the host (no device);
    raises GoFishError with the reference's message on a bad model."""
    check(load().gf_check_sim(C.byref(desc.desc)))


def run(desc: SimDesc, timing=False, final_sfs=(), final_joint_sfs=()):
    """Runs a simulation on the GPU; returns samples, IDs, live counts (and
    timing).  final_sfs: (population, sample_size) pairs whose sample SFS of
    the final generation is computed on the device before the state is
    copied to the host (out["final_sfs"][(population, n)]);
    final_joint_sfs: (pop_a, pop_b, n_a, n_b) joint spectra, likewise
    (out["final_sfs"][(pop_a, pop_b, n_a, n_b)])."""
    lib = load()
    h = C.c_void_p()
    check(lib.gf_run(C.byref(desc.desc), C.byref(h)))
    try:
        sfs = {}
        last = lib.gf_result_num_samples(h) - 1
        for pop, n in final_sfs:
            o = np.empty(int(n), np.float32)
            check(lib.gf_result_site_frequency_spectrum(h, last, int(pop), int(n), float(desc.num_sites),
                                                         o.ctypes.data_as(C.c_void_p)))
            sfs[(int(pop), int(n))] = o
        for pa, pb, na, nb in final_joint_sfs:
            o = np.empty((int(na) + 1, int(nb) + 1), np.float32)
            check(lib.gf_result_joint_site_frequency_spectrum(h, last, int(pa), int(pb), int(na), int(nb),
                                                               float(desc.num_sites), o.ctypes.data_as(C.c_void_p)))
            sfs[(int(pa), int(pb), int(na), int(nb))] = o
        out = read_result(lib, "gf_", h, desc.P, desc.G)
        out["final_sfs"] = sfs
        if timing:
            t = gf_timing()
            lib.gf_result_timing(h, C.byref(t))
            out["timing"] = {k: getattr(t, k) for k, _ in gf_timing._fields_}
    finally:
        lib.gf_result_free(h)
    return out


def population_frequency_histogram(freq_row, Nchrom_e, num_sites, device=-1):
    lib = load()
    f = np.ascontiguousarray(freq_row, np.float32)
    out = np.empty(int(Nchrom_e), np.float32)
    check(lib.gf_population_histogram(f.ctypes.data_as(C.c_void_p), len(f), int(Nchrom_e), float(num_sites),
                                      out.ctypes.data_as(C.c_void_p), device))
    return out


def joint_site_frequency_spectrum(freq_a, freq_b, n_a, n_b, num_sites, device=-1):
    """joint expected SFS [(n_a+1), (n_b+1)] of two populations' rows"""
    lib = load()
    a = np.ascontiguousarray(freq_a, np.float32)
    b = np.ascontiguousarray(freq_b, np.float32)
    if a.shape != b.shape:
        raise ValueError("the two rows must have the same length")
    out = np.empty((int(n_a) + 1, int(n_b) + 1), np.float32)
    check(lib.gf_joint_site_frequency_spectrum(a.ctypes.data_as(C.c_void_p), b.ctypes.data_as(C.c_void_p), len(a),
                                               int(n_a), int(n_b), float(num_sites), out.ctypes.data_as(C.c_void_p),
                                               device))
    return out


def site_frequency_spectrum(freq_row, sample_size, num_sites, device=-1):
    lib = load()
    f = np.ascontiguousarray(freq_row, np.float32)
    out = np.empty(int(sample_size), np.float32)
    check(lib.gf_site_frequency_spectrum(f.ctypes.data_as(C.c_void_p), len(f), int(sample_size), float(num_sites),
                                         out.ctypes.data_as(C.c_void_p), device))
    return out


def set_kernel_timing(on):
    """on: False / 0 off, True / 1 every drift launch timed, s > 1 every s-th"""
    check(load().gf_set_kernel_timing(int(on)))


def set_generations_per_launch(k):
    check(load().gf_set_generations_per_launch(int(k)))


def set_batch_region_scale(scale):
    check(load().gf_set_batch_region_scale(float(scale)))


MATH_OPS = dict(philox=0, u01=1, exp=2, log=3, normcdfinv=4, binom_draw=5, pois_draw=6, sqrt=7, div=8,
                round_away=9, copy=10, f2i=11)


def device_math(op, *inputs, device=-1):
    """Evaluates one device function on arrays (see gf_device_math)."""
    lib = load()
    code = MATH_OPS[op]
    if op == "philox":
        a0 = np.ascontiguousarray(inputs[0], np.uint32).reshape(-1, 6)
        n = a0.shape[0]
        out = np.empty((n, 4), np.uint32)
        ins = [a0, None, None, None]
    else:
        first = np.uint32 if op in ("u01", "binom_draw", "pois_draw") else np.float32
        a0 = np.ascontiguousarray(inputs[0], first).ravel()
        n = a0.size
        rest = [np.ascontiguousarray(np.broadcast_to(x, a0.shape), np.float32) for x in inputs[1:]]
        ins = [a0] + rest + [None] * (3 - len(rest))
        out = np.empty(n, np.int32 if op in ("binom_draw", "pois_draw", "f2i") else np.float32)
    ptrs = [None if x is None else x.ctypes.data_as(C.c_void_p) for x in ins]
    check(lib.gf_device_math(code, ptrs[0], ptrs[1], ptrs[2], ptrs[3], out.ctypes.data_as(C.c_void_p), n, device))
    return out


def batch_run(sims, compact_interval=35, sfs_n=50, keep_states=False, device=-1, timing=False):
    """Runs many single-population simulations in one batch.

    sims: numpy array of BATCH_SIM_DTYPE.  Returns dict with per-simulation
    SFS [S][sfs_n], final mutation counts, total live mutation-generations,
    optionally final states (freq, ids) and timing."""
    lib = load()
    arr = np.ascontiguousarray(sims, BATCH_SIM_DTYPE)
    S = len(arr)
    h = C.c_void_p()
    check(lib.gf_batch_run(arr.ctypes.data_as(C.c_void_p), S, int(compact_interval), int(sfs_n),
                           1 if keep_states else 0, device, C.byref(h)))
    try:
        sfs = np.empty((S, sfs_n), np.float32)
        check(lib.gf_batch_sfs(h, sfs.ctypes.data_as(C.c_void_p)))
        nm = np.empty(S, np.int32)
        check(lib.gf_batch_num_mutations(h, nm.ctypes.data_as(C.c_void_p)))
        live = C.c_int64()
        check(lib.gf_batch_live_mutation_generations(h, C.byref(live)))
        out = dict(sfs=sfs, num_mutations=nm, live_mutation_generations=live.value)
        if keep_states:
            states = []
            for s in range(S):
                f = np.empty(nm[s], np.float32)
                i = np.empty(nm[s], MUTID_DTYPE)
                check(lib.gf_batch_sim_state(h, s, f.ctypes.data_as(C.c_void_p), i.ctypes.data_as(C.c_void_p)))
                states.append((f, i))
            out["states"] = states
        if timing:
            t = gf_timing()
            lib.gf_batch_timing(h, C.byref(t))
            out["timing"] = {k: getattr(t, k) for k, _ in gf_timing._fields_}
        k, nb, nr = C.c_int32(), C.c_int32(), C.c_int32()
        check(lib.gf_batch_tuning(h, C.byref(k), C.byref(nb), C.byref(nr)))
        out["tuning"] = dict(generations_per_block=k.value, blocks=nb.value, relayouts=nr.value)
    finally:
        lib.gf_batch_free(h)
    return out
