"""Denoised low-photon runs against plain high-photon runs.

Trains the U-Net of mcx_amd.denoise on pairs of low- and high-photon CW
fluence maps that the transport engine generates (batched launches), then
measures, against high-photon plain references:

 * held-out heterogeneous volumes like the training ones: median voxel
   error of the raw low-photon map, of the denoised map, and of plain runs
   with 10x and 100x the photons;
 * the large head-like volume (workloads.head256): median voxel error of
   plain runs of increasing photon count and of the denoised 1e6-photon
   run, and the wall time each needs to reach 5% median error.

Voxel error is |estimate - reference| / reference over the voxels where
the reference is reliable: its coefficient of variation, estimated from K
independent runs, is below 1%.

    python -m mcx_amd.dn_experiment --out results.json [--ntrain 256 ...]
    python -m mcx_amd.dn_experiment --timing     (U-Net step / inference times)
"""
import argparse
import json
import os
import sys
import time

import numpy as np

from . import denoise, workloads


def training_volume(master_seed, index, edge=64, photons=1e5):
    """A heterogeneous volume like the batch workload's (layers and
    ellipsoids, same property ranges), edge^3 voxels of 1 mm, one 0-5 ns
    gate (CW), with a pencil or a disk source (radius 1-6 voxels) entering
    the z=0 face along +z."""
    cfg = workloads.batch_instance(master_seed, index, edge_range=(edge, edge), photon_range=(photons, photons),
                                   ngates=1, tstep=5e-9)
    g = np.random.default_rng([int(master_seed), int(index), 17])
    if g.uniform() < 0.5:
        cfg["srctype"] = "disk"
        cfg["srcparam1"] = [float(g.uniform(1, 6)), 0, 0, 0]
    cfg["nphoton"] = int(photons)
    return cfg


def head_crop_volume(seed, n=128, crop=64):
    """A head-like volume (workloads.head256's generator, another seed, n^3
    voxels of 0.5 mm, one CW gate) for training crops: the crop is the
    crop^3 block under the source, mirrored in z so that the source enters
    at z = 0 along +z as in the cube volumes."""
    cfg = workloads.head256(nphoton=1e5, seed=int(seed), n=n, tstep=5e-9)
    lo = (n - crop) // 2
    cfg["_crop"] = (slice(lo, lo + crop), slice(lo, lo + crop), slice(n - crop, n))
    return cfg


def _training_head(seed, high, minenergy):
    """one full-size training head (a separate process builds it)"""
    hc = workloads.head256(nphoton=int(high), seed=int(seed), n=256, quiet=True, minenergy=minenergy)
    hc["autopilot"] = 1
    return hc


def head256_pairs(nheads, maps_per_head=4, crops_per_map=2, low=1e5, span=(3.0, 40.0), high=1e8, minenergy=1e-4,
                  edge=64, seed=6000, log=print):
    """Training pairs from full-size head volumes (workloads.head256's
    generator at 256^3, seeds seed .. seed+nheads-1, the test volume's 20
    gates summed over time): per head one target run of `high` photons and
    maps_per_head input runs with photon counts log-uniform in [low/span[0], low*span[1]], each map cut
    into crops_per_map edge^3 crops spread over the region the test volume's
    error is measured on (x, y origins 48-144, z origins 128-192, mirrored in
    z so the source side comes first).  Runs are single-run replica launches
    with the maps left on the device; only the crops come back.  Returns
    (features [n, edge, edge, edge, 8] float32, targets [n, edge^3])."""
    import multiprocessing
    from concurrent.futures import ProcessPoolExecutor
    from functools import partial
    import torch
    import mcx_amd
    g = np.random.default_rng(seed)
    feats, targs = [], []
    t0 = time.perf_counter()
    # the volumes are built in fresh host processes (numpy only; the
    # generator holds the GIL for much of its work) while the GPU runs the
    # previous ones
    pool = ProcessPoolExecutor(max(2, min(12, (os.cpu_count() or 4) - 2)),
                               mp_context=multiprocessing.get_context("spawn"))
    make = partial(_training_head, high=float(high), minenergy=float(minenergy))
    for h, hc in enumerate(pool.map(make, [int(seed + k) for k in range(nheads)])):
        X, Y, Z = hc["vol"].shape
        u = float(hc.get("unitinmm", 1.0))
        tgt = torch.empty((1, X * Y * Z), dtype=torch.float32, device="cuda")
        mcx_amd.run_replicas(dict(hc, seed=int(100000 + h)), 1, out=tgt)
        tgt3 = tgt.view(Z, Y, X)
        vol = np.asarray(hc["vol"])
        for m in range(maps_per_head):
            n = float(np.exp(g.uniform(np.log(low / span[0]), np.log(low * span[1]))))
            inp = torch.empty((1, X * Y * Z), dtype=torch.float32, device="cuda")
            mcx_amd.run_replicas(dict(hc, nphoton=int(n), seed=int(200000 + 16 * h + m)), 1, out=inp)
            inp3 = inp.view(Z, Y, X)
            for c in range(crops_per_map):
                x0, y0 = (int(v) for v in g.integers(48, 145, size=2))
                z0 = int(g.integers(128, 193))
                sl = (slice(x0, x0 + edge), slice(y0, y0 + edge), slice(z0, z0 + edge))
                # device maps are [z][y][x]; crops back in [x, y, z], mirrored in z
                def cut(t3):
                    return t3[z0:z0 + edge, y0:y0 + edge, x0:x0 + edge].permute(2, 1, 0).cpu().numpy()[:, :, ::-1] * (u * u)
                v = vol[sl][:, :, ::-1]
                feats.append(denoise.features(cut(inp3), v, hc["prop"], u, n))
                t = denoise.log_feature(cut(tgt3)).astype(np.float32)
                t[(v == 0) | (cut(tgt3) <= 0)] = np.nan
                targs.append(t.reshape(-1))
        if log and (h % 8 == 7 or h == nheads - 1):
            log("head256 training pairs: %d heads, %d crops (%.0f s)" % (h + 1, len(feats), time.perf_counter() - t0))
    pool.shutdown()
    return np.stack(feats), np.stack(targs)


def crop_flip(a, cfg):
    """The training crop of a [X, Y, Z] array (identity for uncropped)."""
    if "_crop" not in cfg:
        return a
    return a[cfg["_crop"]][:, :, ::-1]


def cw_fluence(results, cfgs):
    """CW fluence (sum over gates of flux x gate width) per result, in grid
    units (per voxel area)."""
    out = []
    for r, c in zip(results, cfgs):
        u = float(c.get("unitinmm", 1.0))
        if "cw" in r:  # reduced on the device: tstep x the sum over gates
            out.append(r["cw"] * (u * u))
        else:
            out.append(r["flux"].sum(axis=3) * float(c["tstep"]) * u * u)
    return out


def simulate(cfgs, nphoton, seed_salt, tally=2, minenergy=0.0):
    """Run every configuration with nphoton photons (a number, or one per
    configuration) and a fresh seed, with Russian roulette below minenergy;
    returns (CW fluence maps, kernel ms, photons)."""
    import mcx_amd
    runs = []
    counts = list(nphoton) if np.ndim(nphoton) else [nphoton] * len(cfgs)
    for i, c in enumerate(cfgs):
        d = {k: v for k, v in c.items() if not k.startswith("_")}
        d["nphoton"] = int(counts[i])
        d["minenergy"] = float(minenergy)
        d["seed"] = int((c["seed"] * 1000003 + seed_salt * 7919 + i) % (2 ** 31 - 2)) + 1
        d["quiet"] = True
        runs.append(d)
    res, info = mcx_amd.run_batch(runs, tally=tally, photons_per_lane=32, block_size=256, download="cw")
    return cw_fluence(res, runs), info["kernel_ms"], info["photons"]


def reliable_mask(runs, cv_max=0.01):
    """Mean of K independent maps and the voxels whose mean has a
    coefficient of variation below cv_max."""
    a = np.stack(runs).astype(np.float64)
    mean = a.mean(axis=0)
    sd = a.std(axis=0, ddof=1) / np.sqrt(a.shape[0])
    mask = (mean > 0) & (sd < cv_max * mean)
    return mean, mask


def median_error(est, ref, mask):
    return float(np.median(np.abs(est[mask] - ref[mask]) / ref[mask]))


def make_features(phis, cfgs, photons=None, flip_z=False):
    f = []
    counts = list(photons) if np.ndim(photons) else [photons] * len(cfgs)
    for phi, c, n in zip(phis, cfgs, counts):
        x = denoise.features(crop_flip(phi, c), crop_flip(c["vol"], c), c["prop"], float(c.get("unitinmm", 1.0)), n)
        f.append(x[:, :, ::-1] if flip_z else x)
    return np.stack(f)


def make_targets(phis, cfgs):
    """Scaled log fluence targets; NaN (left out of the loss) outside
    tissue and where the target run recorded no fluence."""
    out = []
    for phi, c in zip(phis, cfgs):
        p = crop_flip(phi, c)
        t = denoise.log_feature(p).astype(np.float32)
        t[(crop_flip(c["vol"], c) == 0) | (p <= 0)] = np.nan
        out.append(t.reshape(-1))
    return np.stack(out)


def to_device_bf16(x):
    import torch
    return torch.from_numpy(np.ascontiguousarray(x, dtype=np.float32)).to("cuda").to(torch.bfloat16)


def denoise_maps(net, feats, flip_z=False):
    """Denoised CW fluence maps for a stack of feature volumes."""
    import torch
    out = []
    for k in range(feats.shape[0]):
        pred = net.predict(to_device_bf16(feats[k:k + 1]))
        torch.cuda.synchronize()
        S = feats.shape[1]
        m = denoise.inv_log_feature(pred.cpu().numpy().reshape(S, S, S))
        out.append(m[:, :, ::-1] if flip_z else m)
    return out


class HeadDenoiser:
    """The denoised pipeline on one volume with everything resident on the
    GPU: feature kernel, U-Net, inverse kernel (the volume is mirrored in z
    because its source enters from +z).

    roi: the network runs on a box around the voxels whose input fluence
    exceeds roi_threshold x its maximum: per axis from the roi_quantile to
    the 1 - roi_quantile quantile of their coordinates (isolated voxels
    reached by a stray photon do not widen it), widened by roi_margin
    voxels and aligned to 8-voxel cells of the full grid (so the U-Net's
    pooling cells are the full-volume run's); voxels outside the box keep
    the input fluence."""

    def __init__(self, net, cfg, roi=True, roi_threshold=1e-5, roi_margin=16, roi_quantile=1e-3):
        import torch
        self.net = net
        self.shape = cfg["vol"].shape
        X, Y, Z = self.shape
        self.vol = torch.from_numpy(np.asfortranarray(cfg["vol"], dtype=np.uint8).ravel(order="F")).cuda()
        self.tab = torch.from_numpy(denoise._media_table(cfg["prop"], float(cfg.get("unitinmm", 1.0)))).cuda()
        self.roi, self.thr, self.margin, self.q = roi, float(roi_threshold), int(roi_margin), float(roi_quantile)
        self.last_box = (0, X, 0, Y, 0, Z)

    def upload(self, phi):
        import torch
        return torch.from_numpy(np.asfortranarray(phi, dtype=np.float32).ravel(order="F")).cuda()

    def box(self, dphi):
        """(x0, x1, y0, y1, z0, z1) of the region the network runs on"""
        X, Y, Z = self.shape
        if not self.roi:
            return (0, X, 0, Y, 0, Z)
        p = dphi.view(Z, Y, X)
        m = p > self.thr * p.max()
        out = []
        for ax, n in ((2, X), (1, Y), (0, Z)):
            other = tuple(a for a in (0, 1, 2) if a != ax)
            idx = m.any(dim=other[1]).any(dim=other[0]).nonzero()
            if idx.numel() == 0:
                return (0, X, 0, Y, 0, Z)
            lo, hi = int(idx.min()) - self.margin, int(idx.max()) + 1 + self.margin
            lo, hi = max(0, lo // 8 * 8), min(n, -(-hi // 8) * 8)
            out += [lo, hi]
        return tuple(out)

    def boxes(self, maps):
        """box() of every map of a [K, nvox] device tensor, with one host
        synchronisation"""
        import torch
        X, Y, Z = self.shape
        K = maps.shape[0]
        if not self.roi:
            return [(0, X, 0, Y, 0, Z)] * K
        # voxels above the threshold per x, y and z slice (two passes of the
        # signal-box kernels over the maps); the box spans the central 1 - 2q
        # of them along each axis
        L, st = self.net.L, self.net._stream()
        mx = torch.zeros(K, dtype=torch.float32, device="cuda")
        denoise._chk(L.mcx_dn_roi_max(denoise._ptr(maps), K, X * Y * Z, denoise._ptr(mx), st), "roi_max")
        thr = mx * self.thr
        xc = torch.zeros((K, X), dtype=torch.int32, device="cuda")
        rc = torch.empty((K, Z * Y), dtype=torch.int32, device="cuda")
        denoise._chk(L.mcx_dn_roi_counts(denoise._ptr(maps), K, X, Y, Z, denoise._ptr(thr), denoise._ptr(xc),
                                         denoise._ptr(rc), st), "roi_counts")
        r3 = rc.view(K, Z, Y).to(torch.float32)
        proj = [xc.to(torch.float32), r3.sum(dim=1), r3.sum(dim=2)]  # x, y, z
        lims = []
        for q in proj:
            n = q.shape[1]
            c = q.cumsum(dim=1)
            tot = c[:, -1:]
            lo = (c <= self.q * tot).sum(dim=1)
            hi = n - 1 - (c >= (1.0 - self.q) * tot).sum(dim=1) + 1
            hi = torch.where(tot[:, 0] > 0, hi, torch.full_like(hi, -1))
            lims.append(torch.stack([lo, hi], dim=1))
        L = torch.stack(lims, dim=1).cpu().numpy()  # [K, axis, lo/hi]
        out = []
        for k in range(K):
            r = []
            for a, n in enumerate((X, Y, Z)):
                lo, hi = int(L[k, a, 0]), int(L[k, a, 1])
                if hi < 0:
                    r += [0, n]
                    continue
                lo, hi = lo - self.margin, hi + 1 + self.margin
                r += [max(0, lo // 8 * 8), min(n, -(-hi // 8) * 8)]
            out.append(tuple(r))
        return out

    @staticmethod
    def union(boxes):
        b = list(boxes)
        return (min(x[0] for x in b), max(x[1] for x in b), min(x[2] for x in b), max(x[3] for x in b),
                min(x[4] for x in b), max(x[5] for x in b))

    def common_size(self, boxes):
        """every box widened about its centre to the largest extent of each
        axis among them (8-voxel cells, inside the grid): one U-Net buffer
        size for all"""
        b = list(boxes)
        X, Y, Z = self.shape
        ext = [max(x[2 * a + 1] - x[2 * a] for x in b) for a in range(3)]
        out = []
        for x in b:
            r = []
            for a, n in enumerate((X, Y, Z)):
                lo = (x[2 * a] + x[2 * a + 1] - ext[a]) // 2 // 8 * 8
                lo = min(max(lo, 0), n - ext[a])
                r += [lo, lo + ext[a]]
            out.append(tuple(r))
        return out

    def run(self, dphi, photons=None, box=None):
        """device fluence (x-fastest) made with `photons` photons -> device
        denoised fluence; box: the region to run on (default: self.box)"""
        import torch
        X, Y, Z = self.shape
        x0, x1, y0, y1, z0, z1 = self.last_box = box if box is not None else self.box(dphi)
        cx, cy, cz = x1 - x0, y1 - y0, z1 - z0
        full = (cx, cy, cz) == (X, Y, Z)
        if full:
            sub, vsub = dphi, self.vol
        else:
            sub = dphi.view(Z, Y, X)[z0:z1, y0:y1, x0:x1].contiguous().view(-1)
            vsub = self.vol.view(Z, Y, X)[z0:z1, y0:y1, x0:x1].contiguous().view(-1)
        f = torch.empty((1, cx, cy, cz, denoise.NFEAT), dtype=torch.bfloat16, device="cuda")
        L = self.net.L
        st = self.net._stream()
        denoise._chk(L.mcx_dn_features(denoise._ptr(sub), denoise._ptr(vsub), denoise._ptr(self.tab),
                                       int(self.tab.shape[0]), denoise._ptr(f), cx, cy, cz, 1, denoise.LOG_EPS,
                                       denoise.LOG_SHIFT, denoise.LOG_SCALE, st), "features")
        if photons:
            f[..., 6].fill_(denoise.photon_feature(photons))
        pred = self.net.predict(f)
        out = torch.empty(cx * cy * cz, dtype=torch.float32, device="cuda")
        denoise._chk(L.mcx_dn_inv_features(denoise._ptr(pred), denoise._ptr(out), cx, cy, cz, 1, denoise.LOG_EPS,
                                           denoise.LOG_SHIFT, denoise.LOG_SCALE, st), "inv_features")
        if full:
            return out
        res = dphi.clone()
        res.view(Z, Y, X)[z0:z1, y0:y1, x0:x1] = out.view(cz, cy, cx)
        return res

    def run_many(self, maps, photons, boxes, nb=16):
        """run() over the maps of a [K, nvox] device tensor, nb volumes per
        U-Net call: the boxes must share one size (common_size); each map's
        features are written into its slot of one batched input, the
        network runs once per nb maps, and each prediction is pasted into a
        copy of its map.  Returns K device maps."""
        import torch
        X, Y, Z = self.shape
        K = int(maps.shape[0])
        cx, cy, cz = (boxes[0][1] - boxes[0][0], boxes[0][3] - boxes[0][2], boxes[0][5] - boxes[0][4])
        if any((b[1] - b[0], b[3] - b[2], b[5] - b[4]) != (cx, cy, cz) for b in boxes):
            raise ValueError("run_many needs boxes of one size")
        L = self.net.L
        st = self.net._stream()
        n = cx * cy * cz
        outs = []
        for k0 in range(0, K, nb):
            ks = list(range(k0, min(K, k0 + nb)))
            f = torch.empty((len(ks), cx, cy, cz, denoise.NFEAT), dtype=torch.bfloat16, device="cuda")
            for j, k in enumerate(ks):
                x0, x1, y0, y1, z0, z1 = boxes[k]
                sub = maps[k].view(Z, Y, X)[z0:z1, y0:y1, x0:x1].contiguous().view(-1)
                vsub = self.vol.view(Z, Y, X)[z0:z1, y0:y1, x0:x1].contiguous().view(-1)
                denoise._chk(L.mcx_dn_features(denoise._ptr(sub), denoise._ptr(vsub), denoise._ptr(self.tab),
                                               int(self.tab.shape[0]), denoise._ptr(f[j]), cx, cy, cz, 1,
                                               denoise.LOG_EPS, denoise.LOG_SHIFT, denoise.LOG_SCALE, st), "features")
            if photons:
                f[..., 6].fill_(denoise.photon_feature(photons))
            pred = self.net.predict(f)
            for j, k in enumerate(ks):
                x0, x1, y0, y1, z0, z1 = boxes[k]
                out = torch.empty(n, dtype=torch.float32, device="cuda")
                denoise._chk(L.mcx_dn_inv_features(denoise._ptr(pred[j * n:(j + 1) * n]), denoise._ptr(out), cx, cy,
                                                   cz, 1, denoise.LOG_EPS, denoise.LOG_SHIFT, denoise.LOG_SCALE, st),
                             "inv_features")
                res = maps[k].clone()
                res.view(Z, Y, X)[z0:z1, y0:y1, x0:x1] = out.view(cz, cy, cx)
                outs.append(res)
        return outs

    def download(self, dout):
        return dout.cpu().numpy().reshape(self.shape, order="F")


def train(net, X, T, steps, batch=8, lr=1e-3, seed=2024, log=None, decay=True, augment=True):
    """X: bf16 device tensor [N, S, S, S, 8] ([x, y, z] voxel order); T: fp32
    device [N, S^3].  decay: cosine learning-rate decay to 5 % of lr;
    augment: each batch mirrored in x and/or y and/or with x and y swapped
    (symmetries of a source entering along +z; z is kept)."""
    import torch
    g = np.random.default_rng(seed)
    n = X.shape[0]
    S = X.shape[1]
    losses = []
    t0 = time.perf_counter()
    for s in range(steps):
        idx = torch.from_numpy(g.choice(n, size=batch, replace=False)).to(X.device)
        xb = X.index_select(0, idx)
        tb = T.index_select(0, idx).reshape(-1, S, S, S)
        if augment:
            k = int(g.integers(0, 8))
            dims = [d for d, bit in ((1, 1), (2, 2)) if k & bit]
            if dims:
                xb, tb = torch.flip(xb, dims), torch.flip(tb, dims)
            if k & 4:
                xb, tb = xb.transpose(1, 2).contiguous(), tb.transpose(1, 2).contiguous()
        tb = tb.reshape(-1)
        lr_s = lr * (0.05 + 0.95 * 0.5 * (1 + np.cos(np.pi * s / max(1, steps)))) if decay else lr
        loss = net.train_step(xb, tb, lr=lr_s)
        if s % 50 == 0 or s == steps - 1:
            losses.append((s, float(loss)))
            if log:
                log("step %d loss %.5f (%.1f s)" % (s, float(loss), time.perf_counter() - t0))
    torch.cuda.synchronize()
    return losses, time.perf_counter() - t0


def timing(nb=8, S=64, steps=20, infer_S=256):
    """Training-step and inference times of the U-Net on random inputs."""
    import torch
    g = np.random.default_rng(0)
    net = denoise.UNet(seed=1)
    X = to_device_bf16(g.standard_normal((nb, S, S, S, denoise.NFEAT)).astype(np.float32))
    T = torch.from_numpy(g.standard_normal(nb * S ** 3).astype(np.float32)).cuda()
    net.train_step(X, T)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        net.train_step(X, T)
    torch.cuda.synchronize()
    step_ms = (time.perf_counter() - t0) / steps * 1e3
    out = {"train_batch": nb, "train_edge": S, "train_step_ms": step_ms}
    if infer_S:
        Xi = to_device_bf16(g.standard_normal((1, infer_S, infer_S, infer_S, denoise.NFEAT)).astype(np.float32))
        net.predict(Xi)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(3):
            net.predict(Xi)
        torch.cuda.synchronize()
        out.update(infer_edge=infer_S, infer_ms=(time.perf_counter() - t0) / 3 * 1e3)
    # FLOPs of the convolutions (2 per multiply-add) per sample and edge
    fl = 0.0
    for name, cin, cout in denoise.CONVS:
        lvl = {"e1": 0, "e2": 1, "e3": 2, "b1": 3, "b2": 3, "d3": 2, "d2": 1, "d1": 0}[name.rstrip("a")]
        fl += 2.0 * 27 * cin * cout * (1.0 / 8 ** lvl)
    out["fwd_gflop_per_voxel_ratio"] = fl
    out["train_tflops"] = 3 * fl * nb * S ** 3 / (step_ms * 1e-3) / 1e12
    if infer_S:
        out["infer_tflops"] = fl * infer_S ** 3 / (out["infer_ms"] * 1e-3) / 1e12
    print(json.dumps(out), flush=True)
    return out


def head_eval(net, args, minenergy, log):
    """Plain runs of increasing photon count and the denoised low-photon run
    on the head volume, against K high-photon references."""
    import torch
    import mcx_amd
    hc = workloads.head256(nphoton=int(args.head_low), n=args.head, quiet=True, minenergy=minenergy)
    hc["autopilot"] = 1

    def plain(nph, seed):
        c = dict(hc)
        c["nphoton"] = int(nph)
        c["seed"] = seed
        r = mcx_amd.run(c)
        st = mcx_amd.last_stats()
        return cw_fluence([r], [c])[0], st["kernel_ms"]

    hrefs, href_ms = [], 0.0
    for k in range(args.refruns):
        m, ms = plain(args.head_ref, 1000 + k)
        hrefs.append(m)
        href_ms += ms
    log("head (minenergy %g) references: %d x %.0e photons, %.0f ms each" % (minenergy, args.refruns, args.head_ref,
                                                                          href_ms / args.refruns))
    ref, mask = reliable_mask(hrefs)
    ref_loo, mask_loo = reliable_mask(hrefs[1:])
    e_ref1 = median_error(hrefs[0], ref_loo, mask_loo & mask)
    plain_rows = []
    for k, v in enumerate(args.head_plain_scale.split(",")):
        nph = float(v) * args.head_low
        m, ms = plain(nph, 7 + k if k < 4 else 50 + k)
        plain_rows.append({"photons": nph, "kernel_ms": ms, "error": median_error(m, ref, mask)})
    plain_rows.append({"photons": args.head_ref, "kernel_ms": href_ms / args.refruns, "error": e_ref1})
    dn = HeadDenoiser(net, hc)
    den_rows = []
    dn_counts = [float(v) * args.head_low for v in args.head_dn_scale.split(",")]
    for k, nph in enumerate(dn_counts):
        low_map, low_ms_h = plain(nph, 11 + k)
        dphi = dn.upload(low_map)
        pf = nph if args.photon_feature else None
        dn.run(dphi, pf)  # warm-up (buffer allocation)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        dout = dn.run(dphi, pf)
        torch.cuda.synchronize()
        dn_ms = (time.perf_counter() - t0) * 1e3
        dmap = dn.download(dout)
        den_rows.append({"photons": nph, "kernel_ms": low_ms_h, "denoise_ms": dn_ms,
                         "low_error": median_error(low_map, ref, mask), "denoised_error": median_error(dmap, ref, mask)})
    # denoised arm: photons for 5 % by log-log interpolation/extrapolation of
    # its error, time by interpolation of (kernel + denoise) time
    dph = np.array([r["photons"] for r in den_rows])
    der = np.array([r["denoised_error"] for r in den_rows])
    dms = np.array([r["kernel_ms"] + r["denoise_ms"] for r in den_rows])
    slope, icpt = np.polyfit(np.log(dph), np.log(der), 1)
    dn5 = float(np.exp((np.log(0.05) - icpt) / slope)) if slope < 0 else float("inf")
    dms5 = float(np.exp(np.interp(np.log(dn5), np.log(dph), np.log(dms)))) if np.isfinite(dn5) else float("inf")
    if np.isfinite(dn5) and dn5 > dph[-1]:
        # beyond the measured counts: scale the last time linearly in photons
        dms5 = float(dms[-1] * dn5 / dph[-1])
    # plain: error ~ c / sqrt(N) fitted over the plain runs; kernel time at
    # that photon count by log-log interpolation of the measured runs (small
    # runs are bounded by the longest photon histories, not the count)
    ph = np.array([r["photons"] for r in plain_rows])
    er = np.array([r["error"] for r in plain_rows])
    kms = np.array([r["kernel_ms"] for r in plain_rows])
    cfit = float(np.exp(np.mean(np.log(er) + 0.5 * np.log(ph))))
    n5 = (cfit / 0.05) ** 2
    order = np.argsort(ph)
    ms5 = float(np.exp(np.interp(np.log(n5), np.log(ph[order]), np.log(kms[order]))))
    out = {"minenergy": minenergy, "edge": args.head, "reliable_voxels": int(mask.sum()),
           "ref_photons": args.head_ref * args.refruns, "plain": plain_rows,
           "plain_photons_for_5pct": n5, "plain_ms_to_5pct": ms5, "denoised": den_rows,
           "denoised_photons_for_5pct": dn5, "denoised_ms_to_5pct": dms5,
           "ratio_denoised_over_plain": dms5 / ms5}
    log("head: %s" % json.dumps(out))
    if args.batch_k > 0:
        out["batched"] = head_batch_eval(dn, args, hc, ref, mask, ms5, log)
    return out


def _time_to_5pct(photons, errors, ms, sqrt_law):
    """photons and time per volume at 5 % median error: errors fitted as
    c/sqrt(N) (plain) or by a log-log line (denoised), time interpolated in
    log-log between the measured counts (scaled linearly beyond them)"""
    ph, er, t = np.asarray(photons, float), np.asarray(errors, float), np.asarray(ms, float)
    if sqrt_law:
        c = float(np.exp(np.mean(np.log(er) + 0.5 * np.log(ph))))
        n5 = (c / 0.05) ** 2
    else:
        slope, icpt = np.polyfit(np.log(ph), np.log(er), 1)
        n5 = float(np.exp((np.log(0.05) - icpt) / slope)) if slope < 0 else float("inf")
    if not np.isfinite(n5):
        return n5, float("inf")
    o = np.argsort(ph)
    if n5 > ph[o][-1]:
        return n5, float(t[o][-1] * n5 / ph[o][-1])
    return n5, float(np.exp(np.interp(np.log(n5), np.log(ph[o]), np.log(t[o]))))


def head_batch_eval(dn, args, hc, ref, mask, single_plain_ms5, log):
    """The batched form on the head volume: K runs with independent seeds in
    one persistent launch of the batched kernel (mcx_run_batch), each then
    denoised, against K plain runs in one launch; time per volume = launch
    time / K (+ the volume's U-Net inference).  A lone low-photon launch
    lasts as long as its longest photon history; K runs in one launch share
    that tail."""
    import torch
    import mcx_amd
    K = args.batch_k

    def launch(nph, salt):
        cfgs = []
        for i in range(K):
            c = dict(hc)
            c["nphoton"] = int(nph)
            c["seed"] = 5000 + 97 * salt + i
            c["quiet"] = True
            cfgs.append(c)
        t0 = time.perf_counter()
        log("head batched: launch of %d x %.3g photons: autotuning" % (K, nph))
        _, info = mcx_amd.run_batch(cfgs, tally=1, download=False)  # autotunes the shape (cached per batch)
        log("head batched: tuned (%d photons/lane, block %d, cache %d) in %.1f s; timed run" %
            (info["photons_per_lane"], info["block_size"], info["tally_cache_cells"], time.perf_counter() - t0))
        res, info = mcx_amd.run_batch(cfgs, tally=1, photons_per_lane=info["photons_per_lane"],
                                      block_size=info["block_size"], tally_cache=info["tally_cache_cells"] or -1,
                                      download="cw")
        maps = cw_fluence(res, cfgs)
        log("head batched: %d x %.3g photons in %.1f ms (%.0f s with tuning and download)" %
            (K, nph, info["total_ms"], time.perf_counter() - t0))
        return maps, float(info["total_ms"])

    plain_rows = []
    for k, v in enumerate(args.batch_plain_scale.split(",")):
        nph = float(v) * args.head_low
        maps, ms = launch(nph, 1 + k)
        errs = [median_error(m, ref, mask) for m in maps]
        plain_rows.append({"photons": nph, "ms_per_volume": ms / K, "error": float(np.median(errs))})
    den_rows = []
    for k, v in enumerate(args.batch_dn_scale.split(",")):
        nph = float(v) * args.head_low
        maps, ms = launch(nph, 50 + k)
        errs_low, errs_dn, dn_ms = [], [], []
        log("head batched: denoising %d maps" % len(maps))
        for i, m in enumerate(maps):
            dphi = dn.upload(m)
            pf = nph if args.photon_feature else None
            if i == 0:
                dn.run(dphi, pf)  # warm-up (buffer allocation)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            dout = dn.run(dphi, pf)
            torch.cuda.synchronize()
            dn_ms.append((time.perf_counter() - t0) * 1e3)
            errs_low.append(median_error(m, ref, mask))
            errs_dn.append(median_error(dn.download(dout), ref, mask))
        den_rows.append({"photons": nph, "ms_per_volume": ms / K, "denoise_ms": float(np.mean(dn_ms)),
                         "low_error": float(np.median(errs_low)), "denoised_error": float(np.median(errs_dn))})
    pn5, pms5 = _time_to_5pct([r["photons"] for r in plain_rows], [r["error"] for r in plain_rows],
                              [r["ms_per_volume"] for r in plain_rows], True)
    dn5, dms5 = _time_to_5pct([r["photons"] for r in den_rows], [r["denoised_error"] for r in den_rows],
                              [r["ms_per_volume"] + r["denoise_ms"] for r in den_rows], False)
    out = {"k": K, "plain": plain_rows, "denoised": den_rows, "plain_photons_for_5pct": pn5,
           "plain_ms_per_volume_to_5pct": pms5, "denoised_photons_for_5pct": dn5,
           "denoised_ms_per_volume_to_5pct": dms5, "ratio_denoised_over_plain": dms5 / pms5,
           "ratio_over_single_plain": dms5 / single_plain_ms5}
    log("head batched: %s" % json.dumps(out))
    return out


def main(argv=None):
    if argv is None and len(sys.argv) > 1 and sys.argv[1] == "--timing":
        timing()
        return 0
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", required=True)
    ap.add_argument("--ntrain", type=int, default=384)
    ap.add_argument("--neval", type=int, default=16)
    ap.add_argument("--edge", type=int, default=64)
    ap.add_argument("--low", type=float, default=1e5)
    ap.add_argument("--high", type=float, default=2e7)
    ap.add_argument("--ref", type=float, default=1e7, help="photons per reference run (x --refruns)")
    ap.add_argument("--refruns", type=int, default=10)
    ap.add_argument("--steps", type=int, default=6000)
    ap.add_argument("--head-frac", type=float, default=0.6, help="fraction of training volumes that are head crops")
    ap.add_argument("--batch", type=int, default=8)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--head", type=int, default=256, help="edge of the head volume (0: skip)")
    ap.add_argument("--head-low", type=float, default=1e6)
    ap.add_argument("--head-ref", type=float, default=1e8, help="photons per head reference run (x --refruns)")
    ap.add_argument("--weights", default=None, help="save the trained parameters (npz)")
    ap.add_argument("--load-weights", default=None, help="skip training: load parameters saved by --weights")
    ap.add_argument("--head-minenergy", default="1e-4", help="comma list of roulette thresholds for the head runs "
                                                          "(applied to the plain and the denoised arm alike)")
    ap.add_argument("--minenergy", type=float, default=1e-4, help="roulette threshold of the training and held-out runs")
    ap.add_argument("--low-span", default="3,40", help="training inputs: photons log-uniform in [low/a, low*b]")
    ap.add_argument("--photon-feature", type=int, default=1, help="give the network the input photon count")
    ap.add_argument("--head-dn-scale", default="1,2,4", help="denoised head runs: photons as multiples of --head-low")
    ap.add_argument("--head-plain-scale", default="1,2,4,10",
                    help="plain head runs (besides the references): photons as multiples of --head-low")
    ap.add_argument("--batch-k", type=int, default=8, help="batched head evaluation: runs per launch (0: skip)")
    ap.add_argument("--batch-dn-scale", default="0.25,0.5,1",
                    help="batched denoised runs: photons per run as multiples of --head-low")
    ap.add_argument("--batch-plain-scale", default="2,4",
                    help="batched plain runs: photons per run as multiples of --head-low")
    ap.add_argument("--head256-train", type=int, default=0,
                    help="also train on crops of this many full-size head volumes (head256_pairs)")
    ap.add_argument("--head256-high", type=float, default=1e8, help="photons of each full-size head's target run")
    ap.add_argument("--head256-maps", type=int, default=4, help="input runs per full-size head")
    ap.add_argument("--head256-crops", type=int, default=2, help="crops per input run of a full-size head")
    args = ap.parse_args(argv)
    import torch

    def log(msg):
        print(msg, flush=True)

    R = {"args": vars(args)}
    t_all = time.perf_counter()
    if args.load_weights:
        net = denoise.UNet(seed=2024)
        net.load_state(np.load(args.load_weights)["p"])
        R["params"] = net.count
        R["train"] = {"loaded": args.load_weights}
        args.ntrain = 0
    # ---- training pairs --------------------------------------------------
    nhead = int(round(args.ntrain * args.head_frac))
    tr = [training_volume(2024, i, args.edge) for i in range(args.ntrain - nhead)]
    tr += [head_crop_volume(5000 + i, 2 * args.edge, args.edge) for i in range(nhead)]
    if not tr:
        return _evaluate(net, args, R, log, t_all)
    # input photon counts log-uniform in [low/3, low*10]: the network sees a
    # range of noise levels
    gl = np.random.default_rng(2024)
    sa, sb = (float(v) for v in args.low_span.split(","))
    lows = np.exp(gl.uniform(np.log(args.low / sa), np.log(args.low * sb), size=len(tr)))
    lo, lo_ms, _ = simulate(tr, lows, 1, minenergy=args.minenergy)
    log("train inputs: %d x %.0e-%.0e photons, %.0f ms" % (len(tr), args.low / sa, args.low * sb, lo_ms))
    hi, hi_ms, hi_ph = [], 0.0, 0.0
    for c0 in range(0, len(tr), 64):
        m, ms, ph = simulate(tr[c0:c0 + 64], args.high, 2 + c0, minenergy=args.minenergy)
        hi += m
        hi_ms += ms
        hi_ph += ph
        log("train targets: %d of %d volumes (%.0f ms)" % (min(c0 + 64, len(tr)), len(tr), hi_ms))
    log("train targets: %d x %.0e photons, %.0f ms (%.0f photons/ms)" % (len(tr), args.high, hi_ms, hi_ph / hi_ms))
    Xh = make_features(lo, tr, lows if args.photon_feature else None)
    Th = make_targets(hi, tr)
    del lo, hi
    if args.head256_train:
        fx, ft = head256_pairs(args.head256_train, maps_per_head=args.head256_maps, crops_per_map=args.head256_crops,
                               low=args.low, span=(sa, sb), high=args.head256_high, minenergy=args.minenergy,
                               edge=args.edge, log=log)
        if not args.photon_feature:
            fx[..., 6] = 0.0
        Xh, Th = np.concatenate([Xh, fx]), np.concatenate([Th, ft])
        del fx, ft
    X = to_device_bf16(Xh)
    T = torch.from_numpy(Th).to("cuda")
    del Xh, Th
    net = denoise.UNet(seed=2024)
    R["params"] = net.count
    losses, train_s = train(net, X, T, args.steps, args.batch, args.lr, log=log)
    R["train"] = {"losses": losses, "seconds": train_s, "steps": args.steps, "batch": args.batch,
                  "step_ms": train_s / max(1, args.steps) * 1e3}
    del X, T
    if args.weights:
        np.savez_compressed(args.weights, p=net.state())
    return _evaluate(net, args, R, log, t_all)


def _evaluate(net, args, R, log, t_all):
    if args.neval > 0:
        _heldout(net, args, R, log)
    # ---- head volume -------------------------------------------------------
    if args.head:
        R["head"] = []
        for me in [float(v) for v in args.head_minenergy.split(",")]:
            R["head"].append(head_eval(net, args, me, log))
    R["seconds"] = time.perf_counter() - t_all
    with open(args.out, "w") as f:
        json.dump(R, f, indent=1)
    log("wrote %s" % args.out)


def _heldout(net, args, R, log):
    ev = [training_volume(4048, i, args.edge) for i in range(args.neval)]
    me = args.minenergy
    refs = [simulate(ev, args.ref, 100 + k, minenergy=me)[0] for k in range(args.refruns)]
    low, low_ms, _ = simulate(ev, args.low, 3, minenergy=me)
    p10, p10_ms, _ = simulate(ev, args.low * 10, 4, minenergy=me)
    p100, p100_ms, _ = simulate(ev, args.low * 100, 5, minenergy=me)
    t0 = time.perf_counter()
    den = denoise_maps(net, make_features(low, ev, args.low if args.photon_feature else None))
    den_s = time.perf_counter() - t0
    rows = []
    for i in range(len(ev)):
        ref, mask = reliable_mask([r[i] for r in refs])
        rows.append({"reliable_voxels": int(mask.sum()), "low": median_error(low[i], ref, mask),
                     "denoised": median_error(den[i], ref, mask), "plain_x10": median_error(p10[i], ref, mask),
                     "plain_x100": median_error(p100[i], ref, mask)})
    R["heldout"] = {"volumes": rows, "denoise_seconds": den_s,
                    "median": {k: float(np.median([r[k] for r in rows])) for k in ("low", "denoised", "plain_x10",
                                                                                   "plain_x100")},
                    "low_photons": args.low, "ref_photons": args.ref * args.refruns}
    log("held-out: %s" % json.dumps(R["heldout"]["median"]))


if __name__ == "__main__":
    sys.exit(main())
