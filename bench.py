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
  roofline      the drift kernel (k_advance) against MI355X HBM peak, with
                8 algorithmic bytes per mutation-generation (4 B read + 4 B
                written state), timed with HIP events on the library stream;
                traffic = measured HBM bytes per launch from rocprofv3
                FETCH_SIZE/WRITE_SIZE (profiles/advance_traffic.json), if any
  cpu_baseline  the CPU oracle (oracle/, a restatement of the same
                algorithm) on the same scenario for a bounded prefix of
                generations, on the host's cores
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


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--generations", type=int, default=5000)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-generations", type=int, default=0, help="generations of the CPU sample (0: auto)")
    return ap.parse_args()


def gpu_step(capi, scenarios, G):
    d = scenarios.plain_large(G=G)
    t0 = time.perf_counter()
    r = capi.run(d, timing=True)
    last = r["samples"][-1]
    row = last["freq"][0]
    capi.site_frequency_spectrum(row, 100, d.num_sites)
    capi.site_frequency_spectrum(row, 1000, d.num_sites)
    dt = time.perf_counter() - t0
    mutgens = int(r["live"][1:].sum())
    return dt, mutgens, r


def cpu_baseline(scenarios, G_gpu, want_gens):
    """Times the oracle on the same scenario, first `want_gens` generations."""
    from oracle import oracle as O
    threads = max(1, min(16, os.cpu_count() or 1))
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
                sample="plain_large scenario (N=1e5, L=5e8, gamma=-10), equilibrium start + first %d of %d "
                       "generations, %d threads, %.1f s" % (G, G_gpu, threads, dt))


def main():
    a = parse()
    if a.gpus != 1:
        print("bench.py: this build runs on one GPU only (the batched and plain paths shard by whole "
              "simulations or mutation ranges; multi-GPU runs are not part of it); use --gpus 1",
              file=sys.stderr)
        sys.exit(2)
    from parallelpopgen_amd import capi, scenarios
    capi.load()
    capi.set_kernel_timing(True)
    for _ in range(a.warmup):
        gpu_step(capi, scenarios, a.generations)
    total_t, total_mg = 0.0, 0
    adv_ms, adv_launch, last = 0.0, 0, None
    for _ in range(a.steps):
        dt, mg, r = gpu_step(capi, scenarios, a.generations)
        total_t += dt
        total_mg += mg
        adv_ms += r["timing"]["advance_ms"]
        adv_launch += r["timing"]["advance_launches"]
        last = r
    value = total_mg / total_t
    achieved = BYTES_PER_MUTGEN * total_mg / (adv_ms * 1e-3) / 1e9
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
            "kernel": "k_advance (migration_selection_drift + add_new_mutations, temporal-blocked)",
            "avg_launch_ms": adv_ms / max(adv_launch, 1),
            "algorithmic_bytes_per_launch": BYTES_PER_MUTGEN * total_mg / max(adv_launch, 1),
        },
    }
    if not a.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(scenarios, a.generations, a.cpu_generations)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
