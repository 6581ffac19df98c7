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

The second half of the metric, time to 5 % median voxel error on the 256^3
head volume for the plain kernel and for batched low-photon runs with the
learned denoiser (mcx_amd.time_to_error), is measured after the photons/ms
leg and reported in config.time_to_5pct (--no-tt5 skips it).

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
# HBM bytes per photon of the transport kernel on the sphere100 workload with
# the FP64 tally and the LDS hot-cell box: (FETCH_SIZE x f + WRITE_SIZE x w)
# per launch / photons, scale factors f, w from the FP64->FP32 merge kernel
# of the same run as a copy of known size; keyed by (workload, tally mode,
# blocked tally).  Linear tally: 1e7 photons, f = 2.000, w = 1.000
# (profiles/transport_sphere100_1e7_counters.md); blocked: the default
# 1e8-photon command, f = 0.859, w = 1.000
# (profiles/transport_sphere100_traffic_1e8.md)
TRAFFIC_B_PER_PHOTON = {("sphere100", 2, False): 8254.0, ("sphere100", 2, True): 5533.0}
ATOMIC_SIZES_MB = (4, 64, 400, 2048)


VALU_PEAK_TFLOPS = 157.3  # FP32 vector peak of MI355X
# VALU issue roof: a wave64 vector instruction holds a 16-lane SIMD for 4
# cycles; 256 CUs x 4 SIMDs at the 2.4 GHz peak engine clock
VALU_ISSUE_PEAK = 1024 * 2.4e9 / 4.0
# SQ_INSTS_VALU per wave loop iteration of the transport kernel, keyed as
# TRAFFIC_B_PER_PHOTON (profiles/transport_sphere100_valu_issue.md: 6.251e11
# per 1e8-photon launch over 8.80e8 wave loop iterations)
VALU_INSTS_PER_WAVE_ITER = {("sphere100", 2, True): 710.0}


def _valu(steps_pp, scat_pp, photons, seconds):
    """VALU fraction by the stated convention: 30 FP32 ops per propagation
    step and 60 per scattering event, per photon, against the FP32 vector
    peak.  The issued-instruction count is in the counter profile."""
    tflops = photons * (30.0 * steps_pp + 60.0 * scat_pp) / seconds / 1e12
    return {"achieved": round(tflops, 3), "peak": VALU_PEAK_TFLOPS, "unit": "TFLOP/s",
            "frac": round(tflops / VALU_PEAK_TFLOPS, 5), "convention": "30 flop/step + 60 flop/scatter"}


def _issue(per_iter, wave_iterations, seconds):
    """VALU wave-instructions issued per second against the issue roof: the
    profiled instruction count per wave loop iteration times this run's
    wave loop iterations (the kernel's own counters)."""
    if not per_iter or not wave_iterations:
        return None
    rate = per_iter * wave_iterations / seconds
    return {"achieved": round(rate / 1e9, 1), "peak": round(VALU_ISSUE_PEAK / 1e9, 1),
            "unit": "G VALU wave-instructions/s", "frac": round(rate / VALU_ISSUE_PEAK, 4),
            "source": "%.0f SQ_INSTS_VALU per wave loop iteration (profiles/transport_sphere100_valu_issue.md) "
                      "x this run's wave loop iterations" % per_iter}


def _atomic_roof(mode, nelem):
    """Scattered-atomic rate (atomics/s) over an array of nelem elements."""
    from mcx_amd import _lib
    r = _lib.atomic_bench(mode, int(nelem), 256, 4096)
    if abs(r["sum_ratio"] - 1.0) > 1e-9:
        raise RuntimeError("atomic microbenchmark lost updates: %r" % r)
    return r["atomics_per_s"]


def atomic_roof_table():
    """Microbenchmark of the device's scattered-atomic rate: every one of
    4096 x 256 lanes adds 1 at 256 uniformly random indices.  Modes: 0 FP64
    agent scope no return, 1 FP32 no return, 2 FP32 returning, 3 FP64
    workgroup scope into a private copy per XCD; coalesced forms: 4 FP64 with
    8 lanes per random 64-byte block, 5 FP64 with a wave on 64 consecutive
    elements, 6 FP32 with 16 lanes per 64-byte block."""
    from mcx_amd import _lib
    for mode in (0, 1, 2, 3, 4, 5, 6):
        for mb in ATOMIC_SIZES_MB:
            es = 4 if mode in (1, 2, 6) else 8
            r = _lib.atomic_bench(mode, (mb << 20) // es, 256, 4096)
            r.update(mode=mode, mbytes=mb)
            print(json.dumps(r), flush=True)


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


def bench_batch(args):
    """The 256-instance heterogeneous batch in one persistent launch (FP64
    tally by default: 34.1 k against 30.6 k photons/ms with the FP32 hi/lo
    tally, `--tally 1`), replayed from a captured hipGraph.  value = photons
    over the graph execution time (clear, transport, energy sums, merge and
    normalisation), inputs resident in HBM."""
    import mcx_amd
    from mcx_amd import workloads
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    cfgs = workloads.batch256()
    for c in cfgs:
        c["nphoton"] = max(1000, int(c["nphoton"] * args.batch_scale))
        c["quiet"] = True
    tally = args.tally if args.tally in (0, 1, 2) else 2
    mcx_amd.set_counters(True)
    res, info_w = mcx_amd.run_batch(cfgs, tally=tally, tune_fraction=1.0 / 32, download=False)
    mcx_amd.set_counters(False)
    times = []
    for _ in range(args.steps):
        t0 = time.perf_counter()
        _, info = mcx_amd.run_batch(cfgs, tally=tally, photons_per_lane=info_w["photons_per_lane"],
                                    block_size=info_w["block_size"], use_graph=True, repeat=1, download=False,
                                    tally_cache=info_w["tally_cache_cells"] or -1)
        times.append(((time.perf_counter() - t0) * 1e3, info["total_ms"]))
    nph = info_w["photons"]
    g_ms = sum(t[1] for t in times) / len(times)
    ctr = info_w["counters"]
    steps_pp, dep_pp = ctr["nstep"] / nph, ctr["ndeposit"] / nph
    bytes_pp = steps_pp * 1.0 + dep_pp * 16.0
    achieved = bytes_pp * nph / (g_ms * 1e-3) / 1e9
    total_cells = sum(int(c["vol"].size) * 10 for c in cfgs)
    aroof = _atomic_roof(2 if tally == 1 else 0, min(total_cells, 1 << 29))
    arate = ctr["ndeposit"] / (g_ms * 1e-3)
    line = {
        "metric": "photons/ms at 1 GPU (batched)", "value": round(nph / g_ms, 1), "unit": "photons/ms", "n_gpus": 1,
        "steps": args.steps, "warmup": 1, "ms_per_step": round(sum(t[0] for t in times) / len(times), 2),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "batch256: 256 heterogeneous instances (master seed 7, edges 40-160, 2-6 layers, "
                               "0-3 ellipsoids, 10 gates x 0.5 ns), %.4g photons, %s tally, one persistent launch "
                               "in a hipGraph" % (nph, ["FP32+shadow", "FP32 hi/lo", "FP64"][tally]),
                   "graph_ms": round(g_ms, 3), "lanes": info_w["lanes"], "photons_per_lane": info_w["photons_per_lane"],
                   "block_size": info_w["block_size"], "tally_cache_cells": info_w["tally_cache_cells"],
                   "vgprs": info_w["vgprs"],
                   "steps_per_photon": round(steps_pp, 2), "deposits_per_photon": round(dep_pp, 2),
                   "lane_occupancy": round(ctr["nstep"] / (64.0 * ctr["wave_iterations"]), 4)
                   if ctr.get("wave_iterations") else None},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": None,
                     "atomic": {"achieved": round(arate / 1e9, 3), "peak": round(aroof / 1e9, 3),
                                "unit": "G atomics/s", "frac": round(arate / aroof, 4)},
                     "valu": _valu(steps_pp, ctr["nscatter"] / nph, nph, g_ms * 1e-3)},
    }
    if not args.no_cpu:
        from parity_util import oracle_run
        import oracle_py
        oracle_py.lib()
        threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        n, secs = 0, 0.0
        for c in cfgs[:16]:
            cc = dict(c)
            cc["nphoton"] = min(c["nphoton"], 20000)
            r = oracle_run(cc, 4096, ompthreads=threads)
            n += cc["nphoton"]
            secs += r["seconds"]
            if secs > args.cpu_budget:
                break
        line["cpu_baseline"] = {"value": n / (secs * 1e3), "unit": "photons/ms", "cores": threads, "kind": "port",
                                "sample": "%d photons over the first batch instances (<= 2e4 each), %.1f s" % (n, secs)}
    print(json.dumps(line), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--nphoton", type=float, default=1e8)
    ap.add_argument("--workload", default="sphere100", choices=["sphere100", "cube60", "batch256"])
    ap.add_argument("--batch-scale", type=float, default=1.0, help="photon scale of the batch256 workload")
    ap.add_argument("--tally", type=int, default=2, help="0 FP32+shadow, 1 FP32 hi/lo, 2 FP64, 3 FP64 per-XCD copies")
    ap.add_argument("--tally-blocks", type=int, default=2,
                    help="FP64 tally in 2x2x2 voxel blocks summed per lane in LDS: 0 off, 1 on, 2 auto")
    ap.add_argument("--cpu-budget", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--atomic-roof", action="store_true", help="print the scattered-atomic microbenchmark and exit")
    ap.add_argument("--dn-timing", action="store_true", help="time the denoiser U-Net (train step, 256^3 inference)")
    ap.add_argument("--no-tt5", action="store_true", help="skip the time-to-5%% comparison on the head volume")
    ap.add_argument("--tt5-k", type=int, default=64, help="head volumes per batched launch in the time-to-5%% leg")
    args = ap.parse_args()
    if args.atomic_roof:
        atomic_roof_table()
        return
    if args.dn_timing:
        from mcx_amd import dn_experiment
        dn_experiment.timing()
        return
    if args.gpus != 1:
        print("bench.py runs on one GPU only: this build has no multi-GPU path (requested --gpus %d)" % args.gpus,
              file=sys.stderr)
        sys.exit(2)

    if args.workload == "batch256":
        bench_batch(args)
        return

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
    mcx_amd.set_tally_blocks(args.tally_blocks)
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
    hot_pp = stats_w["ndeposit_hot"] / nph
    # global tally atomics per photon: the deposits outside the LDS hot-cell
    # box, plus the box's flush (at most one atomic per cell per workgroup)
    flush_pp = stats_w["hot_cells"] * (stats_w["nthread"] / stats_w["nblock"]) / nph
    # requests the lanes send to the memory-side atomic unit: one per
    # global atomic, one per 2x2x2 block the blocked tally adds (its cells
    # share a 64-byte line), plus the box's flush
    req = stats_w.get("tally_requests", 0.0)
    glob_pp = (req / nph if req > 0 else dep_pp - hot_pp) + flush_pp
    reclen = 3 if args.workload == "sphere100" else 0
    # algorithmic bytes: 1 B label load per step; per tally deposit the
    # tally word(s) read and written (FP32: 4+4, FP32 hi/lo: 8+8, FP64:
    # 8+8), counted for every deposit whether the LDS box or a global
    # atomic takes it; detected-photon records
    dep_bytes = {0: 8.0, 1: 16.0, 2: 16.0, 3: 16.0}[args.tally]
    bytes_pp = steps_pp * 1.0 + dep_pp * dep_bytes + reclen * 4.0 * detected_w / nph
    k_ms = sum(kernel_ms) / len(kernel_ms)
    achieved = bytes_pp * nph / (k_ms * 1e-3) / 1e9
    key = (args.workload, args.tally, bool(stats_w.get("tally_blocks", 0)))
    tpp = TRAFFIC_B_PER_PHOTON.get(key)
    # the tally's atomic instruction: FP64 no-return (modes 2, 3), FP32
    # returning (1), FP32 no-return + rare returning shadow moves (0)
    nvox = int(cfg["vol"].size)
    ngates = int(round((cfg["tend"] - cfg["tstart"]) / cfg["tstep"]))
    amode = {0: 1, 1: 2, 2: 0, 3: 0}[args.tally]
    aroof = _atomic_roof(amode, nvox * ngates)
    arate = glob_pp * nph / (k_ms * 1e-3)
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
            "lanes": int(stats_w.get("lanes", stats_w["nthread"])),
            "virtual_threads": int(stats_w["nthread"]),
            "vgprs": int(stats_w["vgprs"]),
            "steps_per_photon": round(steps_pp, 2),
            "deposits_per_photon": round(dep_pp, 2),
            "lds_box_deposits_per_photon": round(hot_pp, 2),
            "lds_box_cells": int(stats_w["hot_cells"]),
            "global_atomics_per_photon": round(glob_pp, 2),
            "lane_occupancy": round(stats_w["nstep"] / (64.0 * stats_w["wave_iterations"]), 4)
            if stats_w.get("wave_iterations") else None,
            "tally_blocks": bool(stats_w.get("tally_blocks", 0)),
            "absorbed_fraction": round(absorbed, 6),
        },
        "roofline": {
            "bound": "hbm",
            # the roof the kernel runs at: the vector instruction issue rate
            # (`issue` below); the HBM and atomic fractions show the headroom
            "binding": "issue" if key in VALU_INSTS_PER_WAVE_ITER else None,
            "achieved": round(achieved, 2),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 5),
            "traffic": round(tpp * nph) if tpp else None,
            "traffic_source": "FETCH_SIZE and WRITE_SIZE per photon, scaled by the merge kernel as a copy of "
                              "known size, from the committed counter profile of this tally layout "
                              "(profiles/transport_sphere100_traffic_1e8.md for the blocked tally) x photons per "
                              "launch; a constant, not measured in this run" if tpp else None,
            "atomic": {"achieved": round(arate / 1e9, 3), "peak": round(aroof / 1e9, 3), "unit": "G atomics/s",
                       "frac": round(arate / aroof, 4),
                       "peak_source": "scattered-atomic microbenchmark over the tally's size, same run"},
            "valu": _valu(steps_pp, stats_w["nscatter"] / nph, nph, k_ms * 1e-3),
            "issue": _issue(VALU_INSTS_PER_WAVE_ITER.get(key), stats_w.get("wave_iterations"), k_ms * 1e-3),
        },
    }
    if not args.no_cpu:
        line["cpu_baseline"] = _cpu_baseline(mk, args.cpu_budget)
    if not args.no_tt5 and args.workload == "sphere100":
        line["config"]["time_to_5pct"] = _time_to_5pct(args.tt5_k)
    print(json.dumps(line), flush=True)


def _time_to_5pct(k):
    from mcx_amd import time_to_error
    r = time_to_error.measure(K=k, log=lambda m: print(m, file=sys.stderr, flush=True))
    if "skipped" in r:
        return r
    p, d = r["plain"], r["batched_denoised"]
    return {
        "workload": r["workload"], "reliable_voxels": r["reliable_voxels"],
        "plain_photons": round(p["photons_for_5pct"]), "plain_ms": round(p["ms_to_5pct"], 2),
        "plain_error": round(p["error_at"], 4),
        "batched_denoised_k": d["k"], "batched_denoised_photons": round(d["photons_for_5pct"]),
        "batched_denoised_ms_per_volume": round(d["ms_per_volume_to_5pct"], 2),
        "batched_denoised_error": round(d["error_at"], 4), "denoise_ms_per_volume": round(d["denoise_ms"], 2),
        "launch_ms": round(d["launch_ms"], 2),
        "ratio": round(r["ratio"], 4), "target_ratio": r["target_ratio"], "seconds": round(r["seconds"], 1),
        "points": {"plain": p["points"], "batched_denoised": d["points"]},
    }


if __name__ == "__main__":
    main()
