"""K independent runs of one volume in one launch (mcx_amd_run_replicas):
virtual thread idx tallies into run idx % K.  Checked against the oracle
on the same launch layout: each run's launched and escaped energy is the
sum of its lanes' oracle energies bit for bit, and the runs' tallies,
weighted back by their normalisation, add up to the oracle's whole
tally; one run (K = 1) equals mcx_run_simulation's time-integrated flux."""
import numpy as np
import pytest

import mcx_amd
from mcx_amd import workloads
from parity_util import oracle_run, seq_sum

pytestmark = pytest.mark.gpu


def _layout(cfg, nthread):
    return dict(cfg, nthread=nthread, nblocksize=64, autopilot=0, quiet=True)


@pytest.mark.parametrize("k,blocks", [(4, True), (8, False), (1, True)])
def test_replicas_match_oracle(k, blocks):
    nt = 2048
    cfg = {key: v for key, v in workloads.sphere100(8000, tstep=1e-9, tend=5e-9).items()
           if key not in ("detpos", "savedetflag", "issavedet")}
    old = mcx_amd.set_tally_blocks(blocks)
    try:
        rep = mcx_amd.run_replicas(_layout(cfg, nt), k)
    finally:
        mcx_amd.set_tally_blocks(old)
    full = dict(cfg, nphoton=cfg["nphoton"] * k)
    ora = oracle_run(full, nt)
    g = ora["genergy"]
    for m in range(k):
        lanes = np.arange(m, nt, k)
        assert rep["energytot"][m] == seq_sum(g[2 * lanes + 1])
        assert rep["energytot"][m] - rep["energyabs"][m] == seq_sum(g[2 * lanes])
    assert rep["cw"].shape == (k, 100, 100, 100)
    # weight each run's map back to its raw tally: cw = raw * unitinmm / (E * Vvox)
    raw = sum(rep["cw"][m].astype(np.float64) * rep["energytot"][m] for m in range(k))
    ref = ora["field"].sum(axis=3)
    sig = ref > 1e-4 * ref.max()
    assert np.abs(raw[sig] - ref[sig]).max() / ref.max() < 2e-6
    for m in range(k):
        assert rep["cw"][m].max() > 0
    if k == 1:
        old = mcx_amd.set_tally_mode(2)
        try:
            one = mcx_amd.run(_layout(cfg, nt))
        finally:
            mcx_amd.set_tally_mode(old)
        cw = one["flux"].astype(np.float64).sum(axis=3) * cfg["tstep"]
        s = cw > 1e-4 * cw.max()
        assert np.abs(rep["cw"][0][s] - cw[s]).max() / cw.max() < 2e-6


def test_replicas_head_volume_autopilot():
    hc = workloads.head256(250_000, quiet=True, minenergy=1e-4)
    rep = mcx_amd.run_replicas(hc, 8)
    print("head256 8 x 2.5e5: clear %.2f ms, kernel %.2f ms, reduce %.2f ms"
          % (rep["clear_ms"], rep["kernel_ms"], rep["reduce_ms"]), flush=True)
    assert rep["cw"].shape == (8, 256, 256, 256)
    # each run launches about 2.5e5 photons (unit weights; lanes split evenly)
    assert np.all(np.abs(rep["energytot"] - 250_000) < 2e3)
    a = rep["energyabs"] / rep["energytot"]
    assert np.all(a > 0) and np.all(a < 1) and a.std() < 0.02
    # the launch statistics describe this launch
    st = mcx_amd.last_stats()
    assert st["kernel_ms"] == pytest.approx(rep["kernel_ms"], rel=1e-6)
    assert st["nlaunches"] == 1 and st["errlanes"] == 0 and st["lanes"] > 0
