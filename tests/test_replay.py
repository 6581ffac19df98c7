"""Seed saving and photon replay (src/mcx_core.cu:1205-1215, :2019-2047,
:2208-2220; src/mcx_utils.c:3624-3699).

CPU tests pin the oracle: a detected photon relaunched from its saved RNG
state retraces its path and is detected again with the same record; the
replayed photons' absorbed fraction is the reference's known answer
(test/testmcx.sh:127-135, 'absorbed 3[0-8].x%').  GPU tests (marked gpu)
compare the engine with the oracle bit for bit."""
import numpy as np
import pytest

from mcx_amd import workloads
from parity_util import oracle_run, seq_sum, sort_rows

DETS = np.array([[29, 19, 0, 1], [29, 39, 0, 1], [19, 29, 0, 1], [39, 29, 0, 1]], dtype=np.float32)
R_C0 = np.float32(3.335640951981520e-12)


def baseline_cfg(n=200000, **kw):
    """cube60 with the reference bench's four detectors, no reflection"""
    return workloads.cube60(n, reflect=0, detpos=DETS, savedetflag="DP", issavedet=1, **kw)


def replay_arrays(detp, prop, tstart=0.0, tend=5e-9, replaydet=0):
    """detected weight / time of flight / detector id per photon from DP
    records, as mcx_validatecfg derives them (float32, per medium)"""
    prop = np.asarray(prop, dtype=np.float32)
    keep, w, tof, det = [], [], [], []
    for i, rec in enumerate(np.asarray(detp, dtype=np.float32)):
        d = int(rec[0])
        if replaydet > 0 and d != replaydet:
            continue
        wi, ti = np.float32(1.0), np.float32(0.0)
        for m in range(prop.shape[0] - 1):
            plen = np.float32(rec[1 + m])
            wi = np.float32(wi * np.exp(np.float32(-prop[m + 1, 0] * plen), dtype=np.float32))
            ti = np.float32(ti + np.float32(np.float32(plen * R_C0) * prop[m + 1, 3]))
        if ti < tstart or ti > tend:
            continue
        keep.append(i)
        w.append(wi)
        tof.append(ti)
        det.append(d)
    return (np.array(keep, dtype=np.int64), np.array(w, np.float32), np.array(tof, np.float32),
            np.array(det, np.int32))


def test_oracle_replay_retraces_detected_photons():
    cfg = baseline_cfg(100000)
    base = oracle_run(cfg, 1024, saveseed=1)
    n = base["detected"]
    assert n > 0 and base["seeds"].shape == (n, 16)
    keep, w, tof, det = replay_arrays(base["detp"], cfg["prop"])
    rcfg = dict(cfg, nphoton=len(keep), outputtype="jacobian")
    rep = oracle_run(rcfg, 256, saveseed=1,
                     replay={"seed": base["seeds"][keep], "weight": w, "tof": tof, "detid": det})
    assert rep["errors"] == 0
    # every replayed photon is detected again, with the same record and seed
    assert rep["detected"] == len(keep)
    assert np.array_equal(sort_rows(rep["detp"]), sort_rows(base["detp"][keep]))
    assert np.array_equal(sort_rows(rep["seeds"]), sort_rows(base["seeds"][keep]))
    tot, esc = seq_sum(rep["genergy"][1::2]), seq_sum(rep["genergy"][0::2])
    assert 0.30 <= (tot - esc) / tot < 0.39
    # the Jacobian tally: detected weight x path length per voxel, summing
    # to sum_i w_i L_i over the replayed photons (all paths in the medium)
    L = base["detp"][keep][:, 1].astype(np.float64)
    assert rep["field"].sum() == pytest.approx(float(np.sum(w.astype(np.float64) * L)), rel=1e-5)


def _gpu_run(cfg, nthread, tally=2):
    import mcx_amd
    c = dict(cfg)
    c.update(nthread=nthread, nblocksize=64, autopilot=0, quiet=True)
    old = mcx_amd.set_tally_mode(tally)
    try:
        return mcx_amd.run(c)
    finally:
        mcx_amd.set_tally_mode(old)


@pytest.mark.gpu
def test_gpu_seed_saving_matches_oracle():
    cfg = baseline_cfg(100000, issaveseed=1)
    res = _gpu_run(cfg, 1024)
    ora = oracle_run(cfg, 1024, saveseed=1)
    detp, seeds = res["detp"].T, res["seeds"].T
    assert detp.shape[0] == ora["detected"] == seeds.shape[0] > 0
    both = np.hstack([detp, seeds.view(np.uint32).astype(np.float64)])
    ref = np.hstack([ora["detp"], ora["seeds"].view(np.uint32).astype(np.float64)])
    assert np.array_equal(sort_rows(both), sort_rows(ref))


def _baseline_and_replay(outputtype, replaydet=0, nthread=256):
    cfg = baseline_cfg(100000, issaveseed=1)
    base = _gpu_run(cfg, 1024)
    detp, seeds = base["detp"], base["seeds"]
    rcfg = baseline_cfg(1, seed=seeds, detphotons=detp, outputtype=outputtype, replaydet=replaydet, issaveseed=0)
    rep = _gpu_run(rcfg, nthread)
    keep, w, tof, det = replay_arrays(detp.T, cfg["prop"], replaydet=replaydet)
    ora = oracle_run(dict(rcfg, nphoton=len(keep), seed=cfg["seed"]), nthread,
                     replay={"seed": seeds.T[keep], "weight": w, "tof": tof, "detid": det}, replaydet=replaydet)
    return base, rep, ora, keep, w


@pytest.mark.gpu
def test_gpu_replay_jacobian_matches_oracle():
    base, rep, ora, keep, w = _baseline_and_replay("jacobian")
    st = rep["stat"]
    assert st["nphoton"] == len(keep)
    assert st["energytot"] == seq_sum(ora["genergy"][1::2])
    assert st["energyabs"] == seq_sum(ora["genergy"][1::2]) - seq_sum(ora["genergy"][0::2])
    a = st["energyabs"] / st["energytot"]
    assert 0.30 <= a < 0.39, a  # test/testmcx.sh:133
    # replayed photons are detected again (test/testmcx.sh:127: simulated == detected)
    assert rep["detp"].shape[1] == len(keep)
    assert np.array_equal(sort_rows(rep["detp"].T), sort_rows(ora["detp"]))
    # Jacobian normalised by unitinmm / sum of detected weights
    ref = ora["field"] * (1.0 / float(np.sum(w, dtype=np.float32)))
    got = rep["flux"].astype(np.float64)
    sig = ref > 1e-4 * ref.max()
    assert np.abs(got[sig] - ref[sig]).max() / ref.max() < 1e-5
    assert np.all(got >= 0)


@pytest.mark.gpu
@pytest.mark.parametrize("otype", ["wp", "wm", "wptof", "wltof"])
def test_gpu_replay_scatter_tallies_match_oracle(otype):
    base, rep, ora, keep, w = _baseline_and_replay(otype)
    scale = 1.0 / float(np.sum(w, dtype=np.float32))
    ref = ora["field"] * scale
    got = rep["flux"].astype(np.float64)
    assert np.abs(got - ref).max() <= 1e-5 * np.abs(ref).max()


@pytest.mark.gpu
def test_gpu_replay_per_detector_blocks():
    base, rep, ora, keep, w = _baseline_and_replay("jacobian", replaydet=-1)
    got = rep["flux"]
    assert got.shape == (60, 60, 60, 1, 4)
    det = replay_arrays(base["detp"].T, baseline_cfg()["prop"], replaydet=-1)[3]
    for d in range(1, 5):
        wd = float(np.sum(w[det == d], dtype=np.float32))
        ref = ora["field"][..., d - 1] / wd if wd > 0 else ora["field"][..., d - 1]
        g = got[..., 0, d - 1].astype(np.float64)
        assert np.abs(g - ref).max() <= 1e-5 * max(np.abs(ref).max(), 1e-30)
