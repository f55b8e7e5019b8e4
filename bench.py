"""Benchmark: GO_Fish plain method, large single simulation, on one MI355X.

Workload (one step): one full run_sim of the large single-population
scenario -- N = 100,000 diploids, F = 0, mu = 1e-9, L = 5e8 sites,
gamma = 2Ns = -10, h = 0.5, 5,000 generations from mutation-selection
equilibrium, compaction every 35 generations -- followed by the SFS of
samples of n = 100 and n = 1,000 chromosomes of the final generation.
Input is synthetic (the simulator's own equilibrium initialisation).

Metric: mutation-generations per second = sum over generations of the
live (segregating or preserved) mutations after each generation's update,
divided by the wall time of the step.

Prints ONE JSON line (rank 0, N = 1).  Also reports:
  roofline      the drift kernel (k_advance_fused) against MI355X HBM peak, with
                8 algorithmic bytes per mutation-generation (4 B read + 4 B
                written state), timed with HIP events on the library stream;
                traffic = measured HBM bytes per launch from rocprofv3
                FETCH_SIZE/WRITE_SIZE (profiles/advance_traffic.json), if any
  cpu_baseline  the CPU oracle (oracle/, a restatement of the same
                algorithm) on the same scenario for a bounded prefix of
                generations, on the host's cores (CPU model and thread
                counts stated)
  batched       the batched parameter sweep below (4,096 simulations)
                against the same simulations back to back (sims/s)
  surrogate     the surrogate below, trained on 262,144 draws: predictions/s
                with inputs and outputs in device memory (and with host
                arrays, copies included) against the batched simulator's
                SFS/s, and the held-out
                deviance ratio against two simulation replicates
                (--plain-only leaves both out)

Other workloads (--workload), each printing one JSON line of its own:
  three_population
             the three-population scenario (growth, bottleneck, inbreeding,
             migration, sine-wave selection, preserve every 100), 2,000
             generations + SFS n = 50 per population; metric as above
  batched    a parameter sweep of 4,096 single-population simulations
             (scenarios.parameter_sweep) in one gf_batch_run, against the
             same simulations run back to back through gf_run (a stratified
             sample, scaled to the whole sweep); metric simulations/s
  surrogate  trains the surrogate ensemble on batched simulations and
             reports its predictions/s, the batched simulator's SFS/s and
             the held-out deviance against two simulation replicates
  plain_small
             the small single-population scenario (N = 1e4 diploids,
             L = 2e7, neutral, 1,000 generations from equilibrium) + SFS
             n = 100: mean of 10 runs after 3 warm-up runs, and the whole
             run on the CPU oracle (one thread) beside it
  speed      the reference's own speed benchmark (examples/example_speed/
             run.cu:12-25: 2e5 haploid individuals, neutral, mu = 1e-9,
             1,000 generations from equilibrium, L = xL * 2e7 sites) at
             xL = 20 and 400, each with the compaction interval of the
             reference's GTX 980 run; mean run_sim time over 10 runs after
             10 warm-up runs, as the reference times it, against its
             published times (tests/golden/speed_snps.json, from
             documentation/780M_980_Intel_speed_results.xlsx)
"""
import argparse
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

HBM_PEAK_GBS = 8000.0
BYTES_PER_MUTGEN = 8.0
# every 8th drift launch is bracketed by HIP timing events: a pair of
# events adds a few microseconds to the stream, ~5 ms per step if every one
# of the 500 launches were timed
TIMING_STRIDE = 8


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--generations", type=int, default=5000)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-generations", type=int, default=0, help="generations of the CPU sample (0: auto)")
    ap.add_argument("--workload", choices=["plain_large", "plain_small", "three_population", "batched", "surrogate", "speed"],
                    default="plain_large")
    ap.add_argument("--sims", type=int, default=4096, help="batched: simulations in the sweep")
    ap.add_argument("--plain-sample", type=int, default=256, help="batched: simulations timed back to back")
    ap.add_argument("--train-draws", type=int, default=1000000, help="surrogate: training draws")
    ap.add_argument("--test-draws", type=int, default=10000, help="surrogate: held-out draws")
    ap.add_argument("--plain-only", action="store_true",
                    help="default workload: report the plain run only (no batched / surrogate sub-objects)")
    ap.add_argument("--default-train-draws", type=int, default=262144,
                    help="default workload's surrogate part: training draws")
    ap.add_argument("--default-test-draws", type=int, default=2000,
                    help="default workload's surrogate part: held-out draws")
    return ap.parse_args()


def gpu_step(capi, scenarios, G):
    d = scenarios.plain_large(G=G)
    t0 = time.perf_counter()
    # the final generation's SFS at n = 100 and 1000, binned on the device
    # before the state is copied out (gf_result_site_frequency_spectrum)
    r = capi.run(d, timing=True, final_sfs=[(0, 100), (0, 1000)])
    dt = time.perf_counter() - t0
    mutgens = int(r["live"][1:].sum())
    return dt, mutgens, r


def cpu_model():
    """the host CPU's model name and hardware thread count"""
    name = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    name = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        usable = len(os.sched_getaffinity(0))
    except AttributeError:
        usable = os.cpu_count() or 1
    return name, os.cpu_count() or 1, usable


def cpu_baseline(scenarios, G_gpu, want_gens):
    """Times the oracle on the same scenario, first `want_gens` generations,
    on at most 16 of the host's threads (the share of the host a one-GPU box
    gives this process)."""
    from oracle import oracle as O
    model, hw_threads, usable = cpu_model()
    threads = max(1, min(16, usable))
    O.set_threads(threads)
    G = want_gens
    if G <= 0:
        # calibrate: ~2e6 live mutations, aim for about 15 s of CPU work
        d = scenarios.plain_large(G=4)
        t0 = time.perf_counter()
        r = O.run(d)
        per_gen = max((time.perf_counter() - t0) / 4.0, 1e-3)
        G = int(max(4, min(G_gpu, 15.0 / per_gen)))
    d = scenarios.plain_large(G=G)
    t0 = time.perf_counter()
    r = O.run(d)
    dt = time.perf_counter() - t0
    mutgens = int(r["live"][1:].sum())
    return dict(value=mutgens / dt, unit="mutation-generations/s", cores=threads, kind="port",
                cpu_model=model, hardware_threads=hw_threads, usable_threads=usable,
                sample="plain_large scenario (N=1e5, L=5e8, gamma=-10), equilibrium start + first %d of %d "
                       "generations, %d threads, %.1f s" % (G, G_gpu, threads, dt))


def bench_batched(a, capi, scenarios):
    import numpy as np
    sims = scenarios.parameter_sweep(a.sims)
    interval = 32
    capi.batch_run(sims[:64], compact_interval=interval, sfs_n=50)     # warm-up
    times = []
    for _ in range(max(1, a.steps)):
        t0 = time.perf_counter()
        r = capi.batch_run(sims, compact_interval=interval, sfs_n=50)
        times.append(time.perf_counter() - t0)
    dt = min(times)
    # back-to-back plain runs: one simulation from each of `plain_sample`
    # equal strata of the expected work (expected live sites x generations)
    Nc = np.floor(2.0 * sims["N"] / (1.0 + sims["F"]))
    work = sims["num_sites"] * 2e-9 * Nc * scenarios.harmonic(Nc - 1) * sims["num_generations"]
    order = np.argsort(work)
    strata = np.array_split(order, a.plain_sample)
    rng = np.random.default_rng(7)
    est = 0.0
    d0 = scenarios.sweep_plain_desc(sims[strata[0][0]])
    capi.run(d0)
    for st in strata:
        k = int(rng.choice(st))
        d = scenarios.sweep_plain_desc(sims[k])
        t0 = time.perf_counter()
        res = capi.run(d)
        capi.site_frequency_spectrum(res["samples"][-1]["freq"][0], 50, d.num_sites)
        est += (time.perf_counter() - t0) * len(st)
    return {
        "metric": "simulations/s",
        "value": len(sims) / dt,
        "unit": "simulations/s",
        "n_gpus": 1,
        "steps": max(1, a.steps),
        "warmup": 1,
        "ms_per_step": 1e3 * dt,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (each simulation's own mutation-selection-equilibrium start)",
        "config": {
            "workload": "batched parameter sweep: %d simulations, N log-U[1e3,1e5], gamma U[-50,10], "
                        "h in {0,0.5,1}, F U[0,0.2], mu 1e-9, 4N/10 generations, E[live] log-U[1e3,1e6], "
                        "SFS n=50, compaction every %d generations" % (len(sims), interval),
            "mutation_generations": int(r["live_mutation_generations"]),
            "mutation_generations_per_s": r["live_mutation_generations"] / dt,
            "parallelism": "single GPU",
        },
        "tuning": r.get("tuning"),
        "plain_back_to_back": {
            "estimated_seconds": est,
            "simulations_per_s": len(sims) / est,
            "sample": "%d simulations (one per stratum of expected work), each gf_run + SFS n=50, "
                      "scaled by stratum size" % a.plain_sample,
            "batched_speedup": est / dt,
        },
    }


def bench_surrogate(a):
    from parallelpopgen_amd import surrogate as SG
    t0 = time.perf_counter()
    data = SG.make_dataset(a.train_draws + a.test_draws, seed=11)
    n_tr = a.train_draws
    x, y, t = data["x"][:n_tr], data["y"][:n_tr], data["t"][:n_tr]
    t1 = time.perf_counter()
    net, losses = SG.fit(x, y, t)
    t_fit = time.perf_counter() - t1
    thr = SG.threshold_from(net, data["x"][n_tr:])
    ev = SG.evaluate(SG.Surrogate(net, thr), a.test_draws, seed=12)
    rate = SG.prediction_rate_device(net)
    rate_host = SG.prediction_rate(net)
    sim_rate = len(data["x"]) / data["sim_seconds"]
    return {
        "metric": "SFS/s",
        "value": rate,
        "unit": "SFS predictions/s",
        "n_gpus": 1,
        "steps": 1,
        "warmup": 1,
        "ms_per_step": 1e3 * (1 << 20) / rate,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "bf16",
        "data": "synthetic (batched simulations of random parameter draws)",
        "config": {
            "workload": "surrogate: 4 MLPs 6->512x4 (GELU)->99 softmax + log-total, trained on %d draws "
                        "(gamma U[-50,10], h {0,0.5,1}, F U[0,0.2], growth logU[1,100] over logU[10,2000] "
                        "generations, n=100), %d held-out" % (n_tr, a.test_draws),
            "parallelism": "single GPU",
        },
        "simulator_sfs_per_s": sim_rate,
        "speedup_vs_batched_simulator": rate / sim_rate,
        "host_inclusive_sfs_per_s": rate_host,
        "training_seconds": t_fit,
        "data_seconds": data["sim_seconds"],
        "final_train_loss": float(losses[-1]),
        "disagreement_threshold": thr,
        "held_out": ev,
        "deviance_ratio": ev["deviance_surrogate"] / ev["deviance_replicates"],
        "total_seconds": time.perf_counter() - t0,
    }


def bench_speed(a, capi, scenarios):
    """The reference's speed benchmark at xL = 20 and 400 (value: xL = 400)."""
    with open(os.path.join(HERE, "tests", "golden", "speed_snps.json")) as f:
        pub = json.load(f)
    runs = []
    for xL in (20, 400):
        r980 = [r for r in pub["980"] if abs(r["xL"] - xL) < 1e-9][0]
        r780 = [r for r in pub["780M"] if abs(r["xL"] - xL) < 1e-9][0]
        d = scenarios.constant(N=200000, F=1.0, h=0.0, s=0.0, mu=1e-9, L=xL * 2e7, G=1000,
                               compact_interval=int(r980["compact_interval"]), seed1=0xbeeff00d, seed2=0xdecafbad)
        for _ in range(10):
            capi.run(d)
        t0 = time.perf_counter()
        for _ in range(10):
            r = capi.run(d)
        ms = 1e3 * (time.perf_counter() - t0) / 10
        mg = int(r["live"][1:].sum())
        runs.append({"xL": xL, "compact_interval": int(r980["compact_interval"]), "run_sim_ms": ms,
                     "snps": int(r["samples"][-1]["num_mutations"]), "mutation_generations": mg,
                     "mutation_generations_per_s": mg / (ms * 1e-3),
                     "gtx980_published_ms": r980["ms"], "gtx980_snps": r980["snps"],
                     "gtx780m_published_ms": r780["ms"], "speedup_vs_gtx980": r980["ms"] / ms})
    big = runs[-1]
    return {
        "metric": "run_sim ms (reference speed benchmark, xL=400)",
        "value": big["run_sim_ms"],
        "unit": "ms",
        "n_gpus": 1,
        "steps": 10,
        "warmup": 10,
        "ms_per_step": big["run_sim_ms"],
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": big["run_sim_ms"] / big["gtx980_published_ms"],
        "dtype": "f32",
        "data": "synthetic (simulator's own neutral-equilibrium start)",
        "config": {
            "workload": "speed: run_sim of 2e5 haploid individuals (F=1), neutral, mu=1e-9, 1000 generations "
                        "from equilibrium, L=xL*2e7, compaction interval of the reference's GTX 980 run; "
                        "mean of 10 runs after 10 warm-up runs",
            "parallelism": "single GPU",
        },
        "runs": runs,
    }


def bench_plain_small(a, capi, scenarios):
    """The small scenario: 8.4e3 live mutations, launch-bound on the GPU."""
    d = scenarios.plain_small()
    for _ in range(3):
        capi.run(d, final_sfs=[(0, 100)])
    t0 = time.perf_counter()
    for _ in range(10):
        r = capi.run(d, final_sfs=[(0, 100)])
    dt = (time.perf_counter() - t0) / 10
    mg = int(r["live"][1:].sum())
    out = {
        "metric": "mutation-generations/s",
        "value": mg / dt,
        "unit": "mutation-generations/s",
        "n_gpus": 1,
        "steps": 10,
        "warmup": 3,
        "ms_per_step": 1e3 * dt,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (simulator's own neutral-equilibrium start)",
        "config": {
            "workload": "plain_small: run_sim N=1e4 diploids, F=0, mu=1e-9, L=2e7, neutral, 1000 generations, "
                        "compact 35, + SFS n=100",
            "live_mutations_final": int(r["samples"][-1]["num_mutations"]),
            "mutation_generations_per_step": mg,
            "parallelism": "single GPU",
        },
    }
    if not a.no_cpu_baseline:
        from oracle import oracle as O
        O.load()
        O.set_threads(1)
        t0 = time.perf_counter()
        o = O.run(d)
        ct = time.perf_counter() - t0
        out["cpu_baseline"] = dict(value=int(o["live"][1:].sum()) / ct, unit="mutation-generations/s", cores=1,
                                   kind="port", sample="the whole plain_small run on the CPU oracle, 1 thread, "
                                                       "%.2f s" % ct)
    return out


def bench_three_population(a, capi, scenarios):
    """Three populations (exponential growth 2e4 -> 2e5, a bottleneck, F = 0.1,
    migration 1e-4, sine-wave selection, preserve every 100 generations),
    2,000 generations, final SFS n = 50 of each population on the device."""
    G = 2000
    capi.set_kernel_timing(TIMING_STRIDE)
    d = scenarios.three_population(G=G)
    pairs = [(0, 50), (1, 50), (2, 50)]
    for _ in range(a.warmup):
        capi.run(d, final_sfs=pairs)
    total_t, total_mg, adv_ms, adv_launch, adv_timed = 0.0, 0, 0.0, 0, 0
    for _ in range(a.steps):
        t0 = time.perf_counter()
        r = capi.run(d, timing=True, final_sfs=pairs)
        total_t += time.perf_counter() - t0
        total_mg += int(r["live"][1:].sum())
        adv_ms += r["timing"]["advance_ms"]
        adv_launch += r["timing"]["advance_launches"]
        adv_timed += r["timing"]["advance_timed_launches"]
    return {
        "metric": "mutation-generations/s",
        "value": total_mg / total_t,
        "unit": "mutation-generations/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": 1e3 * total_t / a.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (simulator's own mutation-selection-equilibrium start)",
        "config": {
            "workload": "three_population: P=3, N0=2e4 (population 0 grows exponentially to 2e5 over "
                        "generations 0-500, population 1 bottleneck 2e3 in [500,600), population 2 F=0.1), "
                        "migration 1e-4, mu=1e-9, L=2e8, h=0.3, sine-wave selection (gamma amplitude 5, "
                        "period 200), preserve every 100, 2000 generations, + SFS n=50 per population; "
                        "a mutation counts once per generation (not once per population)",
            "mutation_generations_per_step": total_mg // a.steps,
            "advance_launches_per_step": adv_launch // a.steps,
            "advance_avg_launch_ms": adv_ms / max(adv_timed, 1),
            "parallelism": "single GPU",
        },
    }


def main():
    a = parse()
    if a.gpus != 1:
        print("bench.py: this build runs on one GPU only (the batched and plain paths shard by whole "
              "simulations or mutation ranges; multi-GPU runs are not part of it); use --gpus 1",
              file=sys.stderr)
        sys.exit(2)
    from parallelpopgen_amd import capi, scenarios
    capi.load()
    if a.workload == "batched":
        print(json.dumps(bench_batched(a, capi, scenarios)))
        return
    if a.workload == "surrogate":
        print(json.dumps(bench_surrogate(a)))
        return
    if a.workload == "three_population":
        print(json.dumps(bench_three_population(a, capi, scenarios)))
        return
    if a.workload == "plain_small":
        print(json.dumps(bench_plain_small(a, capi, scenarios)))
        return
    if a.workload == "speed":
        print(json.dumps(bench_speed(a, capi, scenarios)))
        return
    capi.set_kernel_timing(TIMING_STRIDE)
    for _ in range(a.warmup):
        gpu_step(capi, scenarios, a.generations)
    total_t, total_mg = 0.0, 0
    adv_ms, adv_launch, adv_timed, last = 0.0, 0, 0, None
    for _ in range(a.steps):
        dt, mg, r = gpu_step(capi, scenarios, a.generations)
        total_t += dt
        total_mg += mg
        adv_ms += r["timing"]["advance_ms"]
        adv_launch += r["timing"]["advance_launches"]
        adv_timed += r["timing"]["advance_timed_launches"]
        last = r
    value = total_mg / total_t
    avg_launch_ms = adv_ms / max(adv_timed, 1)
    bytes_per_launch = BYTES_PER_MUTGEN * total_mg / max(adv_launch, 1)
    achieved = bytes_per_launch / (avg_launch_ms * 1e-3) / 1e9
    traffic = None
    tf = os.path.join(HERE, "profiles", "advance_traffic.json")
    if os.path.exists(tf):
        with open(tf) as f:
            traffic = json.load(f).get("hbm_bytes_per_launch")
    out = {
        "metric": "mutation-generations/s",
        "value": value,
        "unit": "mutation-generations/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": 1e3 * total_t / a.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (simulator's own mutation-selection-equilibrium start)",
        "config": {
            "workload": "plain_large: run_sim N=1e5 diploids, F=0, mu=1e-9, L=5e8, gamma=-10, h=0.5, "
                        "%d generations, compact 35, + SFS n=100 and n=1000" % a.generations,
            "live_mutations_final": int(last["samples"][-1]["num_mutations"]),
            "mutation_generations_per_step": total_mg // a.steps,
            "advance_launches_per_step": adv_launch // a.steps,
            "parallelism": "single GPU",
        },
        "roofline": {
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic,
            "traffic_source": "FETCH_SIZE + WRITE_SIZE per launch of this kernel on this workload, from "
                              "rocprofv3 --pmc runs (profiles/advance_traffic.json), not measured in this run",
            "kernel": "k_advance_fused (migration_selection_drift + add_new_mutations + tile-local compaction, "
                      "%.1f generations per launch)" % (a.generations * a.steps / max(adv_launch, 1)),
            "avg_launch_ms": avg_launch_ms,
            "timed_launches": adv_timed,
            "algorithmic_bytes_per_launch": bytes_per_launch,
        },
    }
    if not a.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(scenarios, a.generations, a.cpu_generations)
    if not a.plain_only:
        # the other two parts of the metric, on the same GPU in the same run;
        # two sweeps (the first one also tunes the block length), best kept
        steps = a.steps
        a.steps = min(steps, 2)
        b = bench_batched(a, capi, scenarios)
        a.steps = steps
        out["batched"] = {
            "metric": "simulations/s", "value": b["value"], "unit": "simulations/s",
            "seconds": b["ms_per_step"] / 1e3, "workload": b["config"]["workload"],
            "mutation_generations_per_s": b["config"]["mutation_generations_per_s"],
            "plain_back_to_back": b["plain_back_to_back"], "tuning": b["tuning"],
        }
        a.train_draws, a.test_draws = a.default_train_draws, a.default_test_draws
        sg = bench_surrogate(a)
        out["surrogate"] = {
            "metric": "SFS/s", "value": sg["value"], "unit": "SFS predictions/s",
            "workload": sg["config"]["workload"],
            "inputs_outputs": "device memory (host_inclusive_sfs_per_s: host arrays, PCIe copies included)",
            "host_inclusive_sfs_per_s": sg["host_inclusive_sfs_per_s"],
            "simulator_sfs_per_s": sg["simulator_sfs_per_s"],
            "speedup_vs_batched_simulator": sg["speedup_vs_batched_simulator"],
            "deviance_ratio": sg["deviance_ratio"], "held_out": sg["held_out"],
            "training_seconds": sg["training_seconds"],
        }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
