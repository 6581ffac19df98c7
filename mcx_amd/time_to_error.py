"""Time to 5 % median voxel error on the 256^3 head volume: the plain
kernel against batched low-photon runs denoised by the learned U-Net.

Workload: workloads.head256 (seed 99, 0.5 mm voxels, five labels, 20 gates
of 0.25 ns), Russian roulette below 1e-4 in every run of both arms and of
the reference (an unbiased estimator that shortens the long histories).

Reference: 10 independent runs of 1e8 photons (1e9 in all) in one batched
launch; the reference map is their mean time-integrated fluence, and a
voxel is reliable where that mean's coefficient of variation, estimated
from the 10 runs, is below 1 %.  The error of a map is the median over
reliable voxels of |x - ref| / ref.

Plain arm: the single-simulation kernel on one run (FP64 tally,
autopilot, lane refill at about one photon per virtual thread as in the
other arm) at increasing photon counts; the error is fitted as
c / sqrt(N), the count reaching 5 % is run three times, and the median of
their launch+sync wall times (the reference's runtime definition,
src/mcx_core.cu:4009-4012) is the plain time to 5 %.

Batched + denoised arm: K runs of the head volume in one launch of the
same kernel (mcx_amd_run_replicas: virtual thread idx belongs to run
idx % K, each run with its own streams, tally and launched energy; the
time-integrated maps are reduced on the device); the maps are denoised on
the GPU, 16 per U-Net call (feature kernels, U-Net, inverse kernels).
Time per volume = (the launch's device time: tally clear, transport,
reduction + the denoising of all K maps) / K.  The photon count reaching 5 % is found by a log-log
fit of the denoised error and then run to confirm it.  Maps and scores
stay on the device.
"""
import os
import time

import numpy as np

from . import denoise, dn_experiment as dx, workloads

WEIGHTS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "weights", "unet_head.npz")
MINENERGY = 1e-4
TARGET = 0.05


def _head(n=256, seed=99):
    hc = workloads.head256(nphoton=1_000_000, seed=seed, n=n, quiet=True, minenergy=MINENERGY)
    hc["autopilot"] = 1
    return hc


def _batched(hc, nph, seeds, shape=None):
    """K = len(seeds) runs in one launch: (time-integrated maps, device ms,
    launch info)."""
    import mcx_amd
    cfgs = [dict(hc, nphoton=int(nph), seed=int(s)) for s in seeds]
    kw = dict(shape) if shape else {}
    res, info = mcx_amd.run_batch(cfgs, tally=2, download="cw", **kw)
    return dx.cw_fluence(res, cfgs), float(info["total_ms"]), info


class _Scorer:
    """Median voxel error against the reference on the device: the
    reference's values at its reliable voxels and their indices in the
    x-fastest order the engine writes maps in."""

    def __init__(self, ref, mask):
        import torch
        idx = np.nonzero(np.asarray(mask).ravel(order="F"))[0]
        self.idx = torch.from_numpy(idx.astype(np.int64)).cuda()
        self.ref = torch.from_numpy(np.asarray(ref, dtype=np.float64).ravel(order="F")[idx]).cuda()

    def __call__(self, est):
        """est: device map (x fastest, any float type) -> median |est - ref| / ref,
        the mean of the two middle values for an even count as numpy's"""
        v = ((est.reshape(-1)[self.idx].double() - self.ref).abs() / self.ref).sort().values
        n = v.numel()
        return float((v[(n - 1) // 2] + v[n // 2]) / 2)


def _replicas(hc, nph, K, seed):
    """K runs of the head volume in one launch of the single-simulation
    kernel (mcx_amd_run_replicas: virtual thread idx belongs to run
    idx % K), the time-integrated maps left on the device: (device maps
    [K, nvox] in grid units, x fastest; device ms of the tally clear +
    transport + reduction; the run's statistics)."""
    import torch
    import mcx_amd
    nvox = int(np.prod(hc["vol"].shape))
    out = torch.empty((K, nvox), dtype=torch.float32, device="cuda")
    r = mcx_amd.run_replicas(dict(hc, nphoton=int(nph), seed=int(seed)), K, out=out)
    u = float(hc.get("unitinmm", 1.0))
    out.mul_(u * u)
    return out, r["clear_ms"] + r["kernel_ms"] + r["reduce_ms"], r


def reference(hc, runs=10, photons=1e8, log=print):
    t0 = time.perf_counter()
    maps, ms, info = _batched(hc, photons, [1000 + k for k in range(runs)])
    ref, mask = dx.reliable_mask(maps)
    log("time-to-5%%: reference %d x %.0e photons in %.0f ms (%.0f s), %d reliable voxels"
        % (runs, photons, ms, time.perf_counter() - t0, int(mask.sum())))
    return ref, mask


def plain_arm(hc, score, counts=(1e6, 2e6, 4e6, 8e6), log=print):
    """mcx_run_simulation's kernel on one run (a replica launch with K = 1:
    the same kernel and tally, its map reduced on the device); time = the
    kernel's launch+sync wall time, the reference's runtime."""

    def run(nph, seed):
        m, _, r = _replicas(hc, nph, 1, seed)
        return score(m[0]), float(r["wall_ms"])

    pts = []
    for k, n in enumerate(counts):
        e, ms = run(n, 7 + k)
        pts.append({"photons": float(n), "error": e, "ms": ms})
        log("time-to-5%%: plain %.3g photons -> error %.4f in %.1f ms" % (n, e, ms))
    ph = np.array([p["photons"] for p in pts])
    er = np.array([p["error"] for p in pts])
    c = float(np.exp(np.mean(np.log(er) + 0.5 * np.log(ph))))
    n5 = (c / TARGET) ** 2
    # the count reaching 5 %, three independent runs: the median time (a
    # launch's end is set by its longest histories, which vary run to run)
    for attempt in range(3):
        conf = [run(n5, 31 + 3 * attempt + i) for i in range(3)]
        for e, ms in conf:
            log("time-to-5%%: plain %.3g photons -> error %.4f in %.1f ms" % (n5, e, ms))
        e5 = float(np.median([e for e, _ in conf]))
        ms5 = float(np.median([ms for _, ms in conf]))
        if e5 <= TARGET:
            break
        n5 *= (e5 / TARGET) ** 2 * 1.01  # error ~ c / sqrt(N)
    return {"points": pts, "photons_for_5pct": n5, "ms_to_5pct": ms5, "error_at": e5,
            "runs_at_5pct": [{"error": e, "ms": ms} for e, ms in conf]}


def denoised_arm(hc, score, weights, K=64, scales=(1.25e5, 2.5e5, 5e5), log=print, roi_threshold=1e-5, nb=16):
    """K runs of the head volume per launch (replicas), each map denoised on
    the GPU, nb maps per U-Net call; time per volume = (tally clear +
    transport + reduction) / K + the denoising (boxes, feature kernels,
    U-Net, inverse kernels) / K."""
    import torch
    log("time-to-5%: denoised arm, loading the U-Net")
    net = denoise.UNet(seed=2024)
    net.load_state(np.load(weights)["p"])
    dn = dx.HeadDenoiser(net, hc, roi_threshold=roi_threshold)

    def point(nph, salt):
        t0 = time.perf_counter()
        maps, ms, r = _replicas(hc, nph, K, 5000 + 97 * salt)
        # the region the U-Net runs on: each map's box, all widened to one
        # size so the launch's volumes share one set of U-Net buffers
        # (allocated by an untimed first run)
        ev = [torch.cuda.Event(enable_timing=True) for _ in range(4)]
        ev[0].record()
        boxes = dn.common_size(dn.boxes(maps))
        ev[1].record()
        box = boxes[0]
        dn.run_many(maps[:nb], nph, boxes[:nb], nb=nb)
        torch.cuda.synchronize()
        ev[2].record()
        outs = dn.run_many(maps, nph, boxes, nb=nb)
        ev[3].record()
        torch.cuda.synchronize()
        dn_ms = (ev[0].elapsed_time(ev[1]) + ev[2].elapsed_time(ev[3])) / K
        full = dn.run(maps[0], nph, box=(0, hc["vol"].shape[0], 0, hc["vol"].shape[1], 0, hc["vol"].shape[2]))
        full_err = score(full)
        low = [score(maps[k]) for k in range(K)]
        den = [score(o) for o in outs]
        row = {"photons": float(nph), "launch_ms": ms, "clear_ms": r["clear_ms"], "kernel_ms": r["kernel_ms"],
               "reduce_ms": r["reduce_ms"], "ms_per_volume": ms / K + dn_ms, "denoise_ms": dn_ms,
               "low_error": float(np.median(low)), "error": float(np.median(den)), "box": [int(v) for v in box],
               "first_error_region": den[0], "first_error_whole_volume": full_err,
               "seconds": time.perf_counter() - t0}
        log("time-to-5%%: %d x %.3g photons in one launch: %.2f ms per volume (launch %.0f ms: clear %.1f, kernel %.1f, "
            "reduce %.1f; denoise %.2f ms on %s), error %.4f -> %.4f (first volume %.4f; on the whole volume %.4f)"
            % (K, nph, row["ms_per_volume"], ms, r["clear_ms"], r["kernel_ms"], r["reduce_ms"], dn_ms, box,
               row["low_error"], row["error"], den[0], full_err))
        return row

    pts = [point(n, k) for k, n in enumerate(scales)]
    ph = np.array([p["photons"] for p in pts])
    er = np.array([p["error"] for p in pts])
    slope, icpt = np.polyfit(np.log(ph), np.log(er), 1)
    n5 = float(np.exp((np.log(TARGET) - icpt) / slope)) if slope < 0 else float(ph[-1] * 4)
    n5 = float(min(max(n5, ph[0] / 4), ph[-1] * 16))
    # run the fitted count; while its error is above the target, scale the
    # count by the fitted power law and run again
    for attempt in range(3):
        conf = point(n5, 40 + attempt)
        if conf["error"] <= TARGET or slope >= 0:
            break
        n5 *= (conf["error"] / TARGET) ** (-1.0 / slope) * 1.01
    return {"k": K, "points": pts, "photons_for_5pct": n5, "ms_per_volume_to_5pct": conf["ms_per_volume"],
            "error_at": conf["error"], "denoise_ms": conf["denoise_ms"], "launch_ms": conf["launch_ms"]}


def measure(weights=WEIGHTS, K=64, log=print, seed=99):
    """The whole comparison on the head volume of this seed (99: the
    benchmark's; the U-Net was trained on heads of seeds 6000 on); returns
    the dict reported by bench.py."""
    if not os.path.exists(weights):
        return {"skipped": "no trained weights at %s" % os.path.relpath(weights)}
    os.environ.setdefault("MCX_AMD_VERBOSE", "1")  # phase lines of every batched call on stderr
    t0 = time.perf_counter()
    log("time-to-5%: building the head volume")
    hc = _head(seed=seed)
    ref, mask = reference(hc, log=log)
    score = _Scorer(ref, mask)
    # both arms run the transport kernel the fastest way for their small
    # photon counts: lane refill at about one photon per virtual thread
    # (the streams' biased first numbers shift the fluence by far less than
    # the 5 % being measured)
    import mcx_amd
    old = mcx_amd.set_refill("fast")
    try:
        plain = plain_arm(hc, score, log=log)
        den = denoised_arm(hc, score, weights, K=K, log=log)
    finally:
        mcx_amd.set_refill(old)
    out = {"workload": "head256: 256^3 five-label head volume (seed %d), 0.5 mm voxels, 20 gates x 0.25 ns, "
                       "Russian roulette below %g; reference 10 x 1e8 photons" % (seed, MINENERGY),
           "reliable_voxels": int(mask.sum()), "plain": plain, "batched_denoised": den,
           "ratio": den["ms_per_volume_to_5pct"] / plain["ms_to_5pct"], "target_ratio": 0.1,
           "seconds": time.perf_counter() - t0}
    log("time-to-5%%: plain %.1f ms, batched+denoised %.1f ms per volume: ratio %.3f"
        % (plain["ms_to_5pct"], den["ms_per_volume_to_5pct"], out["ratio"]))
    return out


def main(argv=None):
    """python -m mcx_amd.time_to_error --seeds 900,901 --out r.json: the
    comparison on other head volumes than the benchmark's (held out from
    training as well), one line per head."""
    import argparse
    import json
    import sys
    ap = argparse.ArgumentParser()
    ap.add_argument("--seeds", default="99")
    ap.add_argument("--out", required=True)
    ap.add_argument("--k", type=int, default=64)
    args = ap.parse_args(argv)
    rows = []
    for sd in [int(v) for v in args.seeds.split(",")]:
        r = measure(K=args.k, seed=sd, log=lambda m: print(m, file=sys.stderr, flush=True))
        row = {"seed": sd, "reliable_voxels": r["reliable_voxels"], "ratio": r["ratio"],
               "plain_photons": r["plain"]["photons_for_5pct"], "plain_ms": r["plain"]["ms_to_5pct"],
               "denoised_photons": r["batched_denoised"]["photons_for_5pct"],
               "denoised_ms_per_volume": r["batched_denoised"]["ms_per_volume_to_5pct"],
               "denoised_error": r["batched_denoised"]["error_at"]}
        rows.append(row)
        print(json.dumps(row), flush=True)
        with open(args.out, "w") as f:
            json.dump(rows, f, indent=1)
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
