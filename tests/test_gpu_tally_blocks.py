"""The blocked FP64 tally (2 x 2 x 2 voxel blocks summed per lane in LDS,
added to the tally by the wave with 8 lanes to a block) against the
linear FP64 tally and the FP64 oracle: identical trajectories, fluence
within FP32 rounding of a block's few deposits (1e-6 relative), on even
and odd grid sizes, with the hot-cell box on and off."""
import numpy as np
import pytest

import mcx_amd
from mcx_amd import workloads
from parity_util import normalize_like_engine, oracle_run, seq_sum

pytestmark = pytest.mark.gpu


def _run(cfg, blocks, nthread=None, hot=True):
    c = dict(cfg, quiet=True)
    if nthread:
        c.update(nthread=nthread, nblocksize=64, autopilot=0)
    old = mcx_amd.set_tally_mode(2)
    ob = mcx_amd.set_tally_blocks(blocks)
    oh = mcx_amd.set_hot_box(hot)
    try:
        r = mcx_amd.run(c)
        st = mcx_amd.last_stats()
    finally:
        mcx_amd.set_tally_mode(old)
        mcx_amd.set_tally_blocks(ob)
        mcx_amd.set_hot_box(oh)
    assert st["tally_blocks"] == (1.0 if blocks else 0.0)
    assert st["errlanes"] == 0
    return r, st


def _close(a, b, rtol, floor):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    sig = np.abs(b) > floor * np.abs(b).max()
    assert np.abs(a[sig] - b[sig]).max() / np.abs(b[sig]).max() < rtol or \
        (np.abs(a[sig] - b[sig]) / np.abs(b[sig])).max() < rtol


def _odd_volume():
    rng = np.random.default_rng(11)
    vol = np.ones((61, 37, 23), dtype=np.uint8)
    vol[rng.uniform(size=vol.shape) < 0.3] = 2
    return vol


@pytest.mark.parametrize("case", ["sphere100", "odd", "odd_disk_nohot"])
def test_blocked_tally_matches_oracle(case):
    if case == "sphere100":
        cfg = workloads.sphere100(30000)
    else:
        cfg = workloads.cube60(30000, vol=_odd_volume(), srcpos=[30, 18, 0], tend=5e-9, tstep=1e-9,
                               prop=np.array([[0, 0, 1, 1], [0.01, 1.0, 0.9, 1.37], [0.05, 5.0, 0.8, 1.45]], np.float32))
        if case == "odd_disk_nohot":
            cfg.update(srctype="disk", srcparam1=[6, 0, 0, 0])
    hot = case != "odd_disk_nohot"
    nt = 2048
    b, _ = _run(cfg, True, nt, hot)
    u, _ = _run(cfg, False, nt, hot)
    ora = oracle_run(cfg, nt)
    tot, esc = seq_sum(ora["genergy"][1::2]), seq_sum(ora["genergy"][0::2])
    for r in (b, u):
        assert r["stat"]["energytot"] == tot
        assert r["stat"]["energyabs"] == tot - esc
    ref, _ = normalize_like_engine(ora["field"], tot, cfg["tstep"])
    assert b["flux"].shape == ref.shape
    _close(b["flux"], ref, 1e-6, 1e-4)
    _close(u["flux"], ref, 1e-6, 1e-4)
    # every cell of the volume, including the odd edges, lands where it should
    assert np.array_equal(b["flux"] != 0, u["flux"] != 0)


def test_blocked_tally_full_launch_sphere100():
    cfg = workloads.sphere100(10_000_000)
    b, sb = _run(cfg, True)
    u, su = _run(cfg, False)
    print("sphere100 1e7: blocked %.0f photons/ms, linear %.0f photons/ms"
          % (1e7 / sb["wall_ms"], 1e7 / su["wall_ms"]), flush=True)
    assert b["stat"]["energytot"] == u["stat"]["energytot"] == 1e7
    assert b["stat"]["energyabs"] == pytest.approx(u["stat"]["energyabs"], rel=1e-12)
    _close(b["flux"], u["flux"], 1e-6, 1e-4)
    assert b["detp"].shape == u["detp"].shape
