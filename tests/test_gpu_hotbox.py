"""GPU tests of the hot-cell box (mcx_amd_set_hot_box).

With the box on, the tally cells of the first gate next to the source are
summed in FP64 in each workgroup's LDS and added to the tally once per
workgroup; with it off every deposit is a global atomic.  The photon
histories are the same either way (same seeds, same lanes), so the two
runs' energies agree exactly and their fluence agrees to the tally's
summation-order error; a cell mapped to the wrong place in the box would
move whole deposits and fail by orders of magnitude more.  Against the
oracle's FP64 tally the boxed run agrees as the unboxed one does.
"""
import numpy as np
import pytest

import mcx_amd
from mcx_amd import workloads
from parity_util import normalize_like_engine, oracle_run, seq_sum

pytestmark = pytest.mark.gpu


def _run(cfg, tally, hot):
    old_t = mcx_amd.set_tally_mode(tally)
    old_h = mcx_amd.set_hot_box(hot)
    try:
        res = mcx_amd.run(cfg)
        return res, mcx_amd.last_stats()
    finally:
        mcx_amd.set_tally_mode(old_t)
        mcx_amd.set_hot_box(old_h)


def _layout(cfg, nthread=4096, block=64):
    c = dict(cfg)
    c.update(nthread=nthread, nblocksize=block, autopilot=0, quiet=True)
    return c


def _same_field(a, b, rtol):
    a = a.astype(np.float64)
    b = b.astype(np.float64)
    sig = np.abs(b) > 1e-6 * np.abs(b).max()
    rel = np.abs(a[sig] - b[sig]) / np.abs(b[sig])
    assert rel.max() < rtol, rel.max()
    if (~sig).any():
        assert np.abs(a[~sig] - b[~sig]).max() <= 1e-6 * np.abs(b).max()


@pytest.mark.parametrize("tally,rtol", [(1, 4e-7), (2, 4e-7), (3, 4e-7)])
def test_box_on_and_off_agree(tally, rtol):
    cfg = _layout(workloads.cube60(100000))
    off, s_off = _run(cfg, tally, False)
    on, s_on = _run(cfg, tally, True)
    assert s_off["hot_cells"] == 0
    assert s_on["hot_cells"] >= 512
    assert s_on["lds_bytes"] >= 8 * s_on["hot_cells"]
    assert on["stat"]["energytot"] == off["stat"]["energytot"]
    assert on["stat"]["energyabs"] == off["stat"]["energyabs"]
    _same_field(on["flux"], off["flux"], rtol)


def test_box_not_used_by_the_reference_tally():
    cfg = _layout(workloads.cube60(20000))
    _, st = _run(cfg, 0, True)
    assert st["hot_cells"] == 0


@pytest.mark.parametrize("block", [64, 256])
def test_box_at_a_corner_against_oracle(block):
    """Source in a corner voxel pointing into the cube: the box is clamped
    to the grid's corner; FP64 tally against the oracle's."""
    n = np.float32(1.0 / np.sqrt(3.0))
    cfg = workloads.cube60(40000, vol=np.ones((30, 30, 30), dtype=np.uint8), srcpos=[0.5, 0.5, 0.5],
                           srcdir=[float(n), float(n), float(n)])
    res, st = _run(_layout(cfg, nthread=2048, block=block), 2, True)
    assert st["hot_cells"] > 0
    ora = oracle_run(cfg, 2048)
    assert ora["errors"] == 0
    tot = seq_sum(ora["genergy"][1::2])
    esc = seq_sum(ora["genergy"][0::2])
    assert res["stat"]["energytot"] == tot
    assert res["stat"]["energyabs"] == tot - esc
    ref, _ = normalize_like_engine(ora["field"], tot, cfg["tstep"])
    _same_field(res["flux"], ref, 1e-6)


def test_box_on_a_thin_grid():
    """A grid thinner than the box on two axes: the box spans those axes
    whole and takes the rest of its cells along the third."""
    vol = np.ones((6, 80, 5), dtype=np.uint8)
    cfg = workloads.cube60(30000, vol=vol, srcpos=[3, 40, 0], srcdir=[0, 0, 1], reflect=1)
    lay = _layout(cfg, nthread=1024)
    off, _ = _run(lay, 2, False)
    on, st = _run(lay, 2, True)
    assert 0 < st["hot_cells"] <= 6 * 80 * 5
    assert st["hot_cells"] % 30 == 0
    assert on["stat"]["energyabs"] == off["stat"]["energyabs"]
    _same_field(on["flux"], off["flux"], 4e-7)


def test_box_with_gate_groups_and_reflectance():
    """Two-gate windows (the box covers the first gate of each window) and
    the negative diffuse-reflectance deposits in the void layer the
    source sits on."""
    vol = np.ones((40, 40, 40), dtype=np.uint8)
    vol[:, :, :2] = 0
    cfg = workloads.cube60(20000, vol=vol, srcpos=[19.5, 19.5, 0], reflect=0, issaveref=1, tend=5e-9, tstep=1e-9)
    lay = dict(_layout(cfg, nthread=1024), maxgate=2)
    off, _ = _run(lay, 2, False)
    on, st = _run(lay, 2, True)
    assert st["hot_cells"] > 0
    assert on["stat"]["energyabs"] == off["stat"]["energyabs"]
    _same_field(on["flux"], off["flux"], 4e-7)
    _same_field(on["dref"], off["dref"], 4e-7)


@pytest.mark.parametrize("shape,src,dirv", [((1, 60, 60), [0, 29, 0], [0, 0, 1]), ((60, 60, 1), [29, 0, 0], [0, 1, 0])])
def test_box_on_a_2d_domain_against_oracle(shape, src, dirv):
    """2-D domains (one axis of extent 1; the source direction has no
    component along it): the box spans the singular axis whole; FP64 tally
    against the oracle's."""
    cfg = workloads.cube60(40000, vol=np.ones(shape, dtype=np.uint8), srcpos=src, srcdir=dirv)
    res, st = _run(_layout(cfg, nthread=2048), 2, True)
    assert st["hot_cells"] > 0
    ora = oracle_run(cfg, 2048)
    assert ora["errors"] == 0
    tot = seq_sum(ora["genergy"][1::2])
    assert res["stat"]["energytot"] == tot
    assert res["stat"]["energyabs"] == tot - seq_sum(ora["genergy"][0::2])
    ref, _ = normalize_like_engine(ora["field"], tot, cfg["tstep"])
    _same_field(res["flux"], ref, 1e-6)


@pytest.mark.parametrize("hot", [False, True])
def test_hilo_exact_with_millions_of_deposits_per_cell(hot):
    """The single-simulation hi/lo tally (global cells, or the LDS box cells
    and their flush) against the FP64 tally of the same trajectories when
    the source voxel takes ~2e7 deposits."""
    cfg = _layout(workloads.cube60(20_000_000, tend=1e-9, tstep=1e-9), nthread=65536, block=256)
    hl, _ = _run(cfg, 1, hot)
    f6, _ = _run(cfg, 2, hot)
    assert hl["stat"]["energyabs"] == f6["stat"]["energyabs"]
    _same_field(hl["flux"], f6["flux"], 2.5e-7)
