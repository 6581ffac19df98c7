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


def _baseline_and_replay(outputtype, replaydet=0, nthread=256, ncols=None, **rkw):
    ridx = int(rkw.pop("_replayindex", 0))
    cfg = baseline_cfg(100000, issaveseed=1)
    base = _gpu_run(cfg, 1024)
    detp, seeds = base["detp"], base["seeds"]
    if ncols is not None:
        detp, seeds = detp[:, :ncols], seeds[:, :ncols]
    rcfg = baseline_cfg(1, seed=seeds, detphotons=detp, outputtype=outputtype, replaydet=replaydet, issaveseed=0,
                        **rkw)
    rep = _gpu_run(rcfg, nthread)
    keep, w, tof, det = replay_arrays(detp.T, cfg["prop"], replaydet=replaydet)
    ora = oracle_run(dict(rcfg, nphoton=len(keep), seed=cfg["seed"]), nthread,
                     replay={"seed": seeds.T[keep], "weight": w, "tof": tof, "detid": det}, replaydet=replaydet,
                     replayindex=ridx)
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


# ------------------------------------------------------------- RF replay
# src/mcx_core.cu:1700-1703 (phasor of the saved time of flight at launch),
# :2208-2213 and :2254-2257 (otRF: -w L cos and -w L sin at voxel exits),
# :2033-2062 (otRFmus: w cos and w sin at scattering events); the
# imaginary plane follows the real one (:2909-2911).

OMEGA = np.float32(2 * np.pi * 5e8)  # 500 MHz: omega * tof spans several radians


def _rf_baseline(n=100000):
    cfg = workloads.cube60(n, reflect=0, detpos=DETS, savedetflag="DSP", issavedet=1)
    base = oracle_run(cfg, 1024, saveseed=1)
    detp = base["detp"]  # detid, nscat, ppath (one medium)
    w = np.exp(-np.float32(cfg["prop"][1][0]) * detp[:, 2], dtype=np.float32).astype(np.float32)
    tof = (detp[:, 2] * R_C0 * np.float32(cfg["prop"][1][3])).astype(np.float32)
    keep = (tof >= 0) & (tof <= 5e-9)
    rep = {"seed": base["seeds"][keep], "weight": w[keep], "tof": tof[keep], "detid": detp[keep, 0].astype(np.int32)}
    return cfg, detp[keep], rep


def _rf_replay(cfg, rep, otype, omega=OMEGA, **kw):
    rcfg = dict(cfg, nphoton=len(rep["weight"]), outputtype=otype, omega=float(omega), savedetflag="DP")
    return oracle_run(rcfg, 256, replay=rep, **kw)


def test_oracle_sincos_matches_libm():
    import oracle_py
    L = oracle_py.lib()
    x = np.concatenate([np.linspace(-60, 60, 20001), np.linspace(-1e-3, 1e-3, 101)]).astype(np.float32)
    s, c = ctypes_pair(L.mcx_oracle_sincosf, x)
    ref = x.astype(np.float64)
    assert np.abs(s - np.sin(ref)).max() < 3e-7
    assert np.abs(c - np.cos(ref)).max() < 3e-7
    with oracle_py.mutation(21):
        s2, c2 = ctypes_pair(L.mcx_oracle_sincosf, x)
    assert max(np.abs(s2 - np.sin(ref)).max(), np.abs(c2 - np.cos(ref)).max()) > 1e-6


def ctypes_pair(fn, x):
    import ctypes
    s, c = ctypes.c_float(), ctypes.c_float()
    out = np.zeros((2, len(x)))
    for i, v in enumerate(x):
        fn(float(v), ctypes.byref(s), ctypes.byref(c))
        out[0, i], out[1, i] = s.value, c.value
    return out[0], out[1]


def _phasor_sums(detp, rep, col):
    """sum_i w_i a_i exp(i omega tof_i) over the replayed photons, a_i = the
    record column (path length or scattering count)"""
    a = detp[:, col].astype(np.float64) * rep["weight"].astype(np.float64)
    ph = OMEGA.astype(np.float64) * rep["tof"].astype(np.float64)
    return float(np.sum(a * np.cos(ph))), float(np.sum(a * np.sin(ph)))


def test_oracle_rf_replay_phasor_sums():
    import oracle_py
    cfg, detp, rep = _rf_baseline()
    assert len(rep["weight"]) > 100
    # otRF: every photon deposits -w L exp(i omega tof) along its path
    r = _rf_replay(cfg, rep, "rf")
    assert r["errors"] == 0 and r["field"].shape == (60, 60, 60, 2)
    re, im = _phasor_sums(detp, rep, 2)
    assert r["field"][..., 0].sum() == pytest.approx(-re, rel=2e-5)
    assert r["field"][..., 1].sum() == pytest.approx(-im, rel=2e-5)
    # otRFmus: w exp(i omega tof) at each scattering event
    m = _rf_replay(cfg, rep, "rfmus")
    re, im = _phasor_sums(detp, rep, 1)
    assert m["field"][..., 0].sum() == pytest.approx(re, rel=2e-5)
    assert m["field"][..., 1].sum() == pytest.approx(im, rel=2e-5)
    # omega = 0: the real part is the Jacobian (negated) / the scattering
    # weight tally, the imaginary plane stays empty
    jac = _rf_replay(cfg, rep, "jacobian")["field"][..., 0]
    z = _rf_replay(cfg, rep, "rf", omega=0.0)["field"]
    assert np.array_equal(z[..., 0], -jac) and not z[..., 1].any()
    wp = _rf_replay(cfg, rep, "wp")["field"][..., 0]
    z = _rf_replay(cfg, rep, "rfmus", omega=0.0)["field"]
    assert np.array_equal(z[..., 0], wp) and not z[..., 1].any()
    # a planted mistake (sin and cos exchanged) leaves the sums
    with oracle_py.mutation(20):
        bad = _rf_replay(cfg, rep, "rf")
    re, im = _phasor_sums(detp, rep, 2)
    assert abs(bad["field"][..., 0].sum() + re) > 1e-2 * abs(re)


def test_oracle_reference_replay_numbering():
    """src/mcx_core.cu:1206, :1701, :2021: photon k of thread idx is saved
    photon idx*threadphoton + min(idx, oddphotons - 1) + k"""
    cfg, detp, rep = _rf_baseline()
    n = len(rep["weight"])
    nt = 7  # more photons than threads, with n % nt != 0
    assert n % nt
    rcfg = dict(cfg, nphoton=n, outputtype="jacobian", savedetflag="DP")
    # fewer photons than threads: both numberings replay photon idx on thread idx
    small = oracle_run(rcfg, n + 5, replay=rep)
    small_ref = oracle_run(rcfg, n + 5, replay=rep, replayindex=1)
    assert np.array_equal(small["field"], small_ref["field"])
    ref = oracle_run(rcfg, nt, replay=rep, replayindex=1)
    tp, odd = n // nt, n % nt
    idx = [i * tp + min(i, odd - 1) + k for i in range(nt) for k in range(tp + (i < odd))]
    assert len(set(idx)) == n - 1 and max(idx) == n - 2  # one photon twice, the last never
    wl = detp[:, 2].astype(np.float64) * rep["weight"].astype(np.float64)
    assert ref["field"].sum() == pytest.approx(float(wl[idx].sum()), rel=2e-5)
    cont = oracle_run(rcfg, nt, replay=rep)
    assert cont["field"].sum() == pytest.approx(float(wl.sum()), rel=2e-5)
    # every thread taking the same count: the reference reads photon -1
    m = n - odd
    with pytest.raises(RuntimeError):
        oracle_run(dict(rcfg, nphoton=m), nt, replay={k: v[:m] for k, v in rep.items()}, replayindex=1)


@pytest.mark.gpu
@pytest.mark.parametrize("otype,omega", [("rf", OMEGA), ("rfmus", OMEGA), ("rf", 0.0)])
def test_gpu_replay_rf_matches_oracle(otype, omega):
    base, rep, ora, keep, w = _baseline_and_replay(otype, omega=float(omega))
    got = rep["flux"]
    assert got.shape == (60, 60, 60, 1, 1, 2)
    assert rep["detp"].shape[1] == len(keep)
    assert np.array_equal(sort_rows(rep["detp"].T), sort_rows(ora["detp"]))
    scale = 1.0 / float(np.sum(w, dtype=np.float32))
    for plane in (0, 1):
        ref = ora["field"][..., plane] * scale
        g = got[..., 0, 0, plane].astype(np.float64)
        assert np.abs(g - ref).max() <= 1e-5 * max(np.abs(ora["field"][..., 0] * scale).max(), 1e-30)
    if omega == 0.0:
        assert not got[..., 1].any()
    else:
        assert np.abs(got[..., 1]).max() > 0.1 * np.abs(got[..., 0]).max()


@pytest.mark.gpu
def test_gpu_replay_rf_per_detector_blocks():
    base, rep, ora, keep, w = _baseline_and_replay("rf", replaydet=-1, omega=float(OMEGA))
    got = rep["flux"]
    assert got.shape == (60, 60, 60, 1, 4, 2)
    det = replay_arrays(base["detp"].T, baseline_cfg()["prop"], replaydet=-1)[3]
    for d in range(1, 5):
        wd = float(np.sum(w[det == d], dtype=np.float32))
        for plane in (0, 1):
            o = ora["field"][..., 4 * plane + d - 1]
            ref = o / wd if wd > 0 else o
            g = got[..., 0, d - 1, plane].astype(np.float64)
            assert np.abs(g - ref).max() <= 1e-5 * max(np.abs(ora["field"][..., d - 1] / max(wd, 1e-30)).max(), 1e-30)


@pytest.mark.gpu
def test_gpu_replay_reference_numbering():
    """The reference's replay numbering (src/mcx_core.cu:1206) on the GPU
    equals the oracle's with the same numbering and thread count."""
    import mcx_amd
    cfg = baseline_cfg(100000, issaveseed=1)
    base = _gpu_run(cfg, 1024)
    n = base["detp"].shape[1]
    nt = 64 if n % 64 else 128
    assert n > nt and n % nt
    old = mcx_amd.set_replay_index(True)
    try:
        base2, rep, ora, keep, w = _baseline_and_replay("jacobian", nthread=nt, _replayindex=1)
        assert len(keep) == n
        assert rep["stat"]["energytot"] == seq_sum(ora["genergy"][1::2])
        assert rep["stat"]["energyabs"] == seq_sum(ora["genergy"][1::2]) - seq_sum(ora["genergy"][0::2])
        ref = ora["field"] * (1.0 / float(np.sum(w, dtype=np.float32)))
        got = rep["flux"].astype(np.float64)
        assert np.abs(got - ref).max() <= 1e-5 * ref.max()
        # every thread with the same count: rejected (the reference reads photon -1)
        with pytest.raises(Exception):
            _baseline_and_replay("jacobian", nthread=nt, ncols=n - n % nt, _replayindex=1)
    finally:
        mcx_amd.set_replay_index(old)
