"""The benchmark workloads at their full sizes on the GPU.

batch256: the 256 heterogeneous instances (master seed 7, edges 40-160,
  5.5e8 photons) in one persistent launch: exact launched energy per
  instance, the FP32 hi/lo batch against an FP64 batch of the same
  instances, graph replays against eager execution, and two instances
  reproduced lane for lane by single-simulation runs.
head256: the 256^3 five-label head volume: exact launched energy at 1e8
  photons, the hot-cell box on and off, bit-exact energies against the
  oracle at a small photon count with the same launch layout, and the
  absorbed fraction at 1e8 photons against an oracle sample.
"""
import numpy as np
import pytest

import mcx_amd
from mcx_amd import workloads
from parity_util import normalize_like_engine, oracle_run, seq_sum

pytestmark = pytest.mark.gpu


def _close(a, b, rtol, floor):
    sig = np.abs(b) > floor * np.abs(b).max()
    rel = np.abs(a[sig] - b[sig]) / np.abs(b[sig])
    assert rel.max() < rtol, rel.max()


@pytest.fixture(scope="module")
def batch256():
    cfgs = workloads.batch256()
    for c in cfgs:
        c["quiet"] = True
    return cfgs


def test_batch256_full(batch256):
    cfgs = batch256
    nph = sum(c["nphoton"] for c in cfgs)
    # FP32 hi/lo, autotuned, replayed twice from a captured graph
    hl, info = mcx_amd.run_batch(cfgs, tally=1, use_graph=True, repeat=2, download="cw")
    print("batch256: %.4g photons, graph %.0f ms (%.0f photons/ms), %d photons/lane, block %d, cache %d"
          % (info["photons"], info["total_ms"], info["photons"] / info["total_ms"], info["photons_per_lane"],
             info["block_size"], info["tally_cache_cells"]), flush=True)
    assert info["photons"] == nph
    shape = dict(photons_per_lane=info["photons_per_lane"], block_size=info["block_size"],
                 tally_cache=info["tally_cache_cells"] or -1)
    for c, r in zip(cfgs, hl):
        # unit launch weights: per-lane FP32 sums are exact, so the FP64 total is
        assert r["stat"]["energytot"] == c["nphoton"]
        assert 0.0 < r["stat"]["energyabs"] < r["stat"]["energytot"]
        assert np.all(np.isfinite(r["cw"])) and r["cw"].min() >= 0.0
    # the same instances and launch shape, eager, FP64 tally and FP32 hi/lo
    f64, _ = mcx_amd.run_batch(cfgs, tally=2, download="cw", **shape)
    eag, _ = mcx_amd.run_batch(cfgs, tally=1, download="cw", **shape)
    for a, b, e in zip(hl, f64, eag):
        # identical trajectories: identical launched energy, absorbed energy
        # to the FP64 summation order of the per-lane sums
        assert a["stat"]["energytot"] == b["stat"]["energytot"] == e["stat"]["energytot"]
        assert a["stat"]["energyabs"] == pytest.approx(b["stat"]["energyabs"], rel=1e-12)
        assert a["stat"]["energyabs"] == pytest.approx(e["stat"]["energyabs"], rel=1e-12)
        # hi/lo compensated tally against FP64 accumulation
        _close(a["cw"].astype(np.float64), b["cw"].astype(np.float64), 2e-6, 1e-4)
        _close(e["cw"].astype(np.float64), b["cw"].astype(np.float64), 2e-6, 1e-4)
    # two instances reproduced by single-simulation runs with the batched
    # instance's thread count (photons_per_lane rounded up to whole waves)
    ppl = shape["photons_per_lane"]
    for i in sorted(range(len(cfgs)), key=lambda k: cfgs[k]["nphoton"])[:2]:
        c = cfgs[i]
        nt = max(64, (-(-c["nphoton"] // ppl) + 63) // 64 * 64)
        old = mcx_amd.set_tally_mode(2)
        try:
            p = mcx_amd.run(dict(c, nthread=nt, nblocksize=64, autopilot=0))
        finally:
            mcx_amd.set_tally_mode(old)
        assert p["stat"]["energytot"] == f64[i]["stat"]["energytot"]
        assert p["stat"]["energyabs"] == pytest.approx(f64[i]["stat"]["energyabs"], rel=1e-12)
        cw = p["flux"].astype(np.float64).sum(axis=3) * c["tstep"]
        _close(f64[i]["cw"].astype(np.float64), cw, 2e-6, 1e-4)


@pytest.fixture(scope="module")
def head():
    return workloads.head256(100_000_000, quiet=True)


def test_head256_1e8_energy_and_hot_box(head):
    old = mcx_amd.set_tally_mode(2)
    try:
        on = mcx_amd.run(head)
        st_on = mcx_amd.last_stats()
        prev = mcx_amd.set_hot_box(False)
        try:
            off = mcx_amd.run(head)
        finally:
            mcx_amd.set_hot_box(prev)
    finally:
        mcx_amd.set_tally_mode(old)
    print("head256 1e8: absorbed %.6f, %.0f photons/ms (box %d cells), %.0f photons/ms without"
          % (on["stat"]["energyabs"] / on["stat"]["energytot"], 1e8 / max(on["stat"]["runtime"], 1),
             st_on["hot_cells"], 1e8 / max(off["stat"]["runtime"], 1)), flush=True)
    assert on["stat"]["energytot"] == 1e8
    assert st_on["errlanes"] == 0
    assert st_on["hot_cells"] > 0
    # same trajectories with and without the box (autopilot layouts equal)
    assert off["stat"]["energytot"] == on["stat"]["energytot"]
    assert off["stat"]["energyabs"] == pytest.approx(on["stat"]["energyabs"], rel=1e-12)
    assert on["flux"].shape == (256, 256, 256, 20)
    _close(on["flux"].astype(np.float64), off["flux"].astype(np.float64), 1e-6, 1e-4)


def test_head256_bitexact_and_absorbed_fraction(head):
    # a small run with an explicit layout: every lane's history is the oracle's
    nt = 2048
    small = dict(head, nphoton=20000, nthread=nt, nblocksize=64, autopilot=0)
    old = mcx_amd.set_tally_mode(2)
    try:
        res = mcx_amd.run(small)
    finally:
        mcx_amd.set_tally_mode(old)
    ora = oracle_run(small, nt)
    assert ora["errors"] == 0
    tot, esc = seq_sum(ora["genergy"][1::2]), seq_sum(ora["genergy"][0::2])
    assert res["stat"]["energytot"] == tot
    assert res["stat"]["energyabs"] == tot - esc
    ref, _ = normalize_like_engine(ora["field"], tot, head["tstep"], unitinmm=head["unitinmm"])
    _close(res["flux"].astype(np.float64), ref, 1e-6, 1e-3)
    # 1e8 photons under autopilot against an independent oracle sample
    big = mcx_amd.run(head)
    a_gpu = big["stat"]["energyabs"] / big["stat"]["energytot"]
    sample = oracle_run(dict(head, nphoton=200_000, seed=4711), 16384)
    g = sample["genergy"].astype(np.float64)
    a = (g[1::2] - g[0::2]) / np.maximum(g[1::2], 1e-30)
    se = a[g[1::2] > 0].std(ddof=1) / np.sqrt((g[1::2] > 0).sum())
    print("head256: GPU absorbed %.5f at 1e8, oracle %.5f +- %.5f at 2e5" % (a_gpu, sample["absorbed"], se), flush=True)
    assert abs(a_gpu - sample["absorbed"]) < 4 * se + 1e-4
