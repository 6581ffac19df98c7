"""Benchmark: photons/ms of the HIP transport engine on one MI355X.

Workload (default): the heterogeneous sphere case -- 100^3 uint8 volume,
1 mm voxels, background (0.005/mm, 1.0/mm, g 0.9, n 1.37) with a centred
R=15 sphere (0.05/mm, 5.0/mm, g 0.9, n 1.37), n=1 outside with Fresnel
boundaries, pencil at (50,50,0) along +z, 1e8 photons, 50 gates of 0.1 ns,
4 detectors (R=2) saving detector id + partial paths, seed 12345.

A step is one mcx_run_simulation call over that workload.  `value` uses the
reference's speed definition, photons / runtime, with runtime the
launch+sync time of the transport kernel (src/mcx_core.cu:3392, :3572,
:4009-4012); inputs are resident in HBM when it starts.  ms_per_step is the
whole call (uploads, tally merge, D2H, normalisation).

Prints one JSON line.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def _cpu_baseline(cfg_builder, budget_s):
    """Oracle (CPU restatement, OpenMP over virtual threads) on a bounded
    sample of the same workload."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from parity_util import oracle_run
    import oracle_py

    oracle_py.lib()
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    n0 = 20000
    r = oracle_run(cfg_builder(n0), 4096, want_field=True, ompthreads=threads)
    t0 = max(r["seconds"], 1e-3)
    n = int(min(5_000_000, max(n0, n0 * budget_s / t0)))
    r = oracle_run(cfg_builder(n), 16384, ompthreads=threads)
    return {
        "value": n / (r["seconds"] * 1e3),
        "unit": "photons/ms",
        "cores": threads,
        "kind": "port",
        "sample": "%d photons of the same workload (16384 virtual threads), %.1f s on %d host threads"
                  % (n, r["seconds"], threads),
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--nphoton", type=float, default=1e8)
    ap.add_argument("--workload", default="sphere100", choices=["sphere100", "cube60"])
    ap.add_argument("--tally", type=int, default=2, help="0 FP32+shadow, 1 FP32 hi/lo, 2 FP64, 3 FP64 per-XCD copies")
    ap.add_argument("--cpu-budget", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py runs on one GPU only: this build has no multi-GPU path (requested --gpus %d)" % args.gpus,
              file=sys.stderr)
        sys.exit(2)

    import mcx_amd
    from mcx_amd import workloads

    nph = int(args.nphoton)
    builder = workloads.sphere100 if args.workload == "sphere100" else workloads.cube60

    def mk(n):
        c = builder(n)
        c["quiet"] = True
        return c

    cfg = mk(nph)
    mcx_amd.set_tally_mode(args.tally)
    # warmup; the first one also counts work (steps, deposits) per photon
    stats_w = None
    for i in range(max(args.warmup, 1)):
        mcx_amd.set_counters(i == 0)
        res = mcx_amd.run(cfg)
        if i == 0:
            stats_w = mcx_amd.last_stats()
            detected_w = res["detp"].shape[1] if "detp" in res else 0
    mcx_amd.set_counters(False)
    kernel_ms, wall_ms, call_ms = [], [], []
    absorbed = None
    for _ in range(args.steps):
        t0 = time.perf_counter()
        res = mcx_amd.run(cfg)
        call_ms.append((time.perf_counter() - t0) * 1e3)
        s = mcx_amd.last_stats()
        kernel_ms.append(s["kernel_ms"] / max(s["nlaunches"], 1))
        wall_ms.append(s["wall_ms"])
        absorbed = res["stat"]["energyabs"] / res["stat"]["energytot"]
    total_ph = nph * args.steps
    value = total_ph / sum(wall_ms)

    steps_pp = stats_w["nstep"] / nph
    dep_pp = stats_w["ndeposit"] / nph
    reclen = 3 if args.workload == "sphere100" else 0
    # algorithmic bytes: 1 B label load per step; per deposit the tally
    # word(s) read and written (FP32: 4+4, FP32 hi/lo: 8+8, FP64: 8+8);
    # detected-photon records
    dep_bytes = {0: 8.0, 1: 16.0, 2: 16.0, 3: 16.0}[args.tally]
    bytes_pp = steps_pp * 1.0 + dep_pp * dep_bytes + reclen * 4.0 * detected_w / nph
    k_ms = sum(kernel_ms) / len(kernel_ms)
    achieved = bytes_pp * nph / (k_ms * 1e-3) / 1e9
    line = {
        "metric": "photons/ms at 1 GPU",
        "value": round(value, 1),
        "unit": "photons/ms",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(sum(call_ms) / len(call_ms), 2),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": "%s: %s, %d photons, %s tally" % (
                args.workload,
                "100^3 sphere R=15, 50 gates x 0.1 ns, 4 detectors DP, seed 12345" if args.workload == "sphere100"
                else "60^3 homogeneous, 1 gate 0-5 ns, seed 1648335518",
                nph, ["FP32+shadow", "FP32 hi/lo", "FP64", "FP64 per-XCD"][args.tally]),
            "kernel_ms": round(k_ms, 3),
            "launch_sync_ms": round(sum(wall_ms) / len(wall_ms), 3),
            "lanes": int(stats_w["nthread"]),
            "vgprs": int(stats_w["vgprs"]),
            "steps_per_photon": round(steps_pp, 2),
            "deposits_per_photon": round(dep_pp, 2),
            "absorbed_fraction": round(absorbed, 6),
        },
        "roofline": {
            "bound": "hbm",
            "achieved": round(achieved, 2),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 5),
            "traffic": None,
        },
    }
    if not args.no_cpu:
        line["cpu_baseline"] = _cpu_baseline(mk, args.cpu_budget)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
