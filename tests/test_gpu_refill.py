"""Lane refill under autopilot: the launched lanes run R times as many
virtual threads, each lane taking the next unstarted one when its own has
launched all its photons (mode 1, the default: at least 64 photons per
virtual thread, at most 8 per lane; mode 2: about one photon per virtual
thread, at most 32 per lane).  Every virtual thread keeps its RNG stream,
photon count and energy sums, so a refilled launch must reproduce the
oracle with the same number of virtual threads bit for bit (energies,
detected records) and a static launch of one lane per virtual thread."""
import numpy as np
import pytest

import mcx_amd
from mcx_amd import workloads
from parity_util import normalize_like_engine, oracle_run, seq_sum, sort_rows

pytestmark = pytest.mark.gpu


def _run(cfg, refill=2, counters=False):
    old = mcx_amd.set_refill(refill)
    om = mcx_amd.set_tally_mode(2)
    mcx_amd.set_counters(counters)
    try:
        r = mcx_amd.run(dict(cfg, quiet=True))
        st = mcx_amd.last_stats()
    finally:
        mcx_amd.set_refill(old)
        mcx_amd.set_tally_mode(om)
        mcx_amd.set_counters(False)
    assert st["errlanes"] == 0
    return r, st


def test_refill_matches_oracle_cube60():
    cfg = workloads.cube60(2_000_000, detpos=np.array([[29, 19, 0, 2], [29, 39, 0, 2]], np.float32),
                           savedetflag="DP", issavedet=1)
    r, st = _run(cfg)
    nt = int(st["nthread"])
    assert nt > st["lanes"], (nt, st["lanes"])  # refilled: more virtual threads than lanes
    ora = oracle_run(cfg, nt)
    assert ora["errors"] == 0
    tot, esc = seq_sum(ora["genergy"][1::2]), seq_sum(ora["genergy"][0::2])
    assert r["stat"]["energytot"] == tot
    assert r["stat"]["energyabs"] == tot - esc
    ref, _ = normalize_like_engine(ora["field"], tot, cfg["tstep"])
    sig = ref > 1e-3 * ref.max()
    assert np.abs(r["flux"][sig] - ref[sig]).max() / ref.max() < 1e-6
    assert r["detp"].shape[1] == ora["detected"] > 0
    assert np.array_equal(sort_rows(r["detp"].T), sort_rows(ora["detp"]))


def test_refill_matches_static_launch_sphere100():
    cfg = workloads.sphere100(4_000_000)
    a, sa = _run(cfg)
    nt = int(sa["nthread"])
    lanes = int(sa["lanes"])
    assert nt == min(32, int(np.floor(4_000_000 / lanes + 0.5))) * lanes  # one photon per virtual thread, at most 32 per lane
    b, sb = _run(dict(cfg, nthread=nt, nblocksize=256, autopilot=0), refill=0)
    assert sb["nthread"] == nt and sb["lanes"] == nt
    assert a["stat"]["energytot"] == b["stat"]["energytot"]
    assert a["stat"]["energyabs"] == b["stat"]["energyabs"]
    fa, fb = a["flux"].astype(np.float64), b["flux"].astype(np.float64)
    sig = fb > 1e-4 * fb.max()
    assert np.abs(fa[sig] - fb[sig]).max() / fb.max() < 1e-6
    assert np.array_equal(sort_rows(a["detp"].T), sort_rows(b["detp"].T))


def test_refill_raises_lane_occupancy_head256():
    hc = workloads.head256(3_720_000, quiet=True, minenergy=1e-4)
    on, s_on = _run(hc, 2, counters=True)
    off, s_off = _run(hc, 0, counters=True)
    occ_on = s_on["nstep"] / (64.0 * s_on["wave_iterations"])
    occ_off = s_off["nstep"] / (64.0 * s_off["wave_iterations"])
    print("head256 3.72e6: lane occupancy %.3f -> %.3f, %.1f -> %.1f ms"
          % (occ_off, occ_on, s_off["wall_ms"], s_on["wall_ms"]), flush=True)
    assert occ_on > occ_off
    for r in (on, off):
        assert r["stat"]["energytot"] == 3_720_000


def test_default_refill_keeps_64_photons_per_thread():
    cfg = workloads.sphere100(100_000_000)
    r, st = _run(cfg, refill=1)
    lanes, nt = int(st["lanes"]), int(st["nthread"])
    assert nt == max(1, min(8, 100_000_000 // (64 * lanes))) * lanes and nt > lanes
    assert 100_000_000 / nt >= 64
    small, ss = _run(workloads.sphere100(4_000_000), refill=1)
    assert ss["nthread"] == ss["lanes"]  # 15 photons per lane: no refill
    assert r["stat"]["energytot"] == 100_000_000
