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

Plain arm: mcx_run_simulation (FP64 tally, autopilot) at increasing photon
counts; the error is fitted as c / sqrt(N), the count reaching 5 % is run,
and its launch+sync wall time (the reference's runtime definition,
src/mcx_core.cu:4009-4012) is the plain time to 5 %.

Batched + denoised arm: K runs of the head volume with independent seeds
in one persistent launch (mcx_run_batch, FP64 tally, launch shape chosen
by its autotuner, time-integrated maps reduced on the device); each map is denoised on the GPU (feature kernel, U-Net, inverse
kernel).  Time per volume = the launch's device time
(clear, transport, energy sums, merge, reduction) / K + one volume's
denoising.  The host round trip of the maps is not counted (reported
separately as host_ms).  The photon count reaching 5 % is found by a
log-log fit of the denoised error and then run to confirm it.
"""
import os
import time

import numpy as np

from . import denoise, dn_experiment as dx, workloads

WEIGHTS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "weights", "unet_head.npz")
MINENERGY = 1e-4
TARGET = 0.05


def _head(n=256):
    hc = workloads.head256(nphoton=1_000_000, n=n, quiet=True, minenergy=MINENERGY)
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


def reference(hc, runs=10, photons=1e8, log=print):
    t0 = time.perf_counter()
    maps, ms, info = _batched(hc, photons, [1000 + k for k in range(runs)])
    ref, mask = dx.reliable_mask(maps)
    log("time-to-5%%: reference %d x %.0e photons in %.0f ms (%.0f s), %d reliable voxels"
        % (runs, photons, ms, time.perf_counter() - t0, int(mask.sum())))
    return ref, mask


def plain_arm(hc, ref, mask, counts=(1e6, 2e6, 4e6, 8e6), log=print):
    import mcx_amd

    def run(nph, seed):
        c = dict(hc, nphoton=int(nph), seed=int(seed))
        r = mcx_amd.run(c)
        st = mcx_amd.last_stats()
        return dx.median_error(dx.cw_fluence([r], [c])[0], ref, mask), float(st["wall_ms"])

    pts = []
    for k, n in enumerate(counts):
        e, ms = run(n, 7 + k)
        pts.append({"photons": float(n), "error": e, "ms": ms})
        log("time-to-5%%: plain %.3g photons -> error %.4f in %.1f ms" % (n, e, ms))
    ph = np.array([p["photons"] for p in pts])
    er = np.array([p["error"] for p in pts])
    c = float(np.exp(np.mean(np.log(er) + 0.5 * np.log(ph))))
    n5 = (c / TARGET) ** 2
    e5, ms5 = run(n5, 31)
    log("time-to-5%%: plain %.3g photons -> error %.4f in %.1f ms" % (n5, e5, ms5))
    return {"points": pts, "photons_for_5pct": n5, "ms_to_5pct": ms5, "error_at": e5}


def denoised_arm(hc, ref, mask, weights, K=8, scales=(1.25e5, 2.5e5, 5e5), log=print):
    import torch
    log("time-to-5%: denoised arm, loading the U-Net")
    net = denoise.UNet(seed=2024)
    net.load_state(np.load(weights)["p"])
    dn = dx.HeadDenoiser(net, hc)
    warm = [False]

    def point(nph, salt):
        log("time-to-5%%: batched launch of %d x %.3g photons" % (K, nph))
        t0 = time.perf_counter()
        maps, ms, info = _batched(hc, nph, [5000 + 97 * salt + i for i in range(K)])
        host_ms = (time.perf_counter() - t0) * 1e3 - ms
        log("time-to-5%%: launch done (%.0f ms on the device at %d photons per lane, %.1f s in all); denoising"
            % (ms, info["photons_per_lane"], time.perf_counter() - t0))
        dphis = [dn.upload(m) for m in maps]
        if not warm[0]:  # the U-Net's buffers are allocated on its first call
            dn.run(dphis[0], nph)
            warm[0] = True
        torch.cuda.synchronize()
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record()
        outs = [dn.run(d, nph) for d in dphis]
        b.record()
        torch.cuda.synchronize()
        dn_ms = a.elapsed_time(b) / K
        low = [dx.median_error(m, ref, mask) for m in maps]
        den = [dx.median_error(dn.download(o), ref, mask) for o in outs]
        row = {"photons": float(nph), "launch_ms": ms, "ms_per_volume": ms / K + dn_ms, "denoise_ms": dn_ms,
               "host_ms": host_ms, "low_error": float(np.median(low)), "error": float(np.median(den))}
        log("time-to-5%%: batched %d x %.3g photons: %.1f ms per volume (launch %.0f ms, denoise %.2f ms), "
            "error %.4f -> %.4f" % (K, nph, row["ms_per_volume"], ms, dn_ms, row["low_error"], row["error"]))
        return row

    pts = [point(n, k) for k, n in enumerate(scales)]
    ph = np.array([p["photons"] for p in pts])
    er = np.array([p["error"] for p in pts])
    slope, icpt = np.polyfit(np.log(ph), np.log(er), 1)
    n5 = float(np.exp((np.log(TARGET) - icpt) / slope)) if slope < 0 else float(ph[-1] * 4)
    n5 = float(min(max(n5, ph[0] / 4), ph[-1] * 16))
    conf = point(n5, 40)
    return {"k": K, "points": pts, "photons_for_5pct": n5, "ms_per_volume_to_5pct": conf["ms_per_volume"],
            "error_at": conf["error"], "denoise_ms": conf["denoise_ms"], "host_ms_per_volume": conf["host_ms"] / K}


def measure(weights=WEIGHTS, K=8, log=print):
    """The whole comparison; returns the dict reported by bench.py."""
    if not os.path.exists(weights):
        return {"skipped": "no trained weights at %s" % os.path.relpath(weights)}
    os.environ.setdefault("MCX_AMD_VERBOSE", "1")  # phase lines of every batched call on stderr
    t0 = time.perf_counter()
    log("time-to-5%: building the head volume")
    hc = _head()
    ref, mask = reference(hc, log=log)
    plain = plain_arm(hc, ref, mask, log=log)
    den = denoised_arm(hc, ref, mask, weights, K=K, log=log)
    out = {"workload": "head256: 256^3 five-label head volume (seed 99), 0.5 mm voxels, 20 gates x 0.25 ns, "
                       "Russian roulette below %g; reference 10 x 1e8 photons" % MINENERGY,
           "reliable_voxels": int(mask.sum()), "plain": plain, "batched_denoised": den,
           "ratio": den["ms_per_volume_to_5pct"] / plain["ms_to_5pct"], "target_ratio": 0.1,
           "seconds": time.perf_counter() - t0}
    log("time-to-5%%: plain %.1f ms, batched+denoised %.1f ms per volume: ratio %.3f"
        % (plain["ms_to_5pct"], den["ms_per_volume_to_5pct"], out["ratio"]))
    return out
