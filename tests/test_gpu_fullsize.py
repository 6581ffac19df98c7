"""GPU checks at the benchmark's full size (sphere100, 1e8 photons, the
default launch), through properties that do not depend on the photon
count:

* energy: every launched photon is counted once (energytot == nphoton
  exactly, pencil source, no specular loss);
* the absorbed fraction agrees with the oracle's estimate from a photon
  sample within that sample's statistical error;
* symmetry: the sphere is centred at (50, 50, 50) and the pencil enters at
  x = y = 50 along +z, so the fluence summed over each x-slab (y-slab)
  matches its mirror slab 99 - x (99 - y) within photon noise, except the
  pair at the source plane (the entry point lies in voxel 50);
* the LDS hot-cell box changes only where deposits are summed: with it off
  the energies are identical and the fluence agrees to FP64
  summation-order error.
"""
import numpy as np
import pytest

import mcx_amd
from mcx_amd import workloads
from parity_util import oracle_run, seq_sum

pytestmark = pytest.mark.gpu

NPH = 100_000_000


@pytest.fixture(scope="module")
def full_run():
    cfg = workloads.sphere100(NPH)
    cfg["quiet"] = True
    old = mcx_amd.set_tally_mode(2)
    try:
        res = mcx_amd.run(cfg)
        st = mcx_amd.last_stats()
    finally:
        mcx_amd.set_tally_mode(old)
    return cfg, res, st


def test_fullsize_energy_and_absorbed_fraction(full_run):
    cfg, res, st = full_run
    assert res["stat"]["energytot"] == NPH
    assert st["hot_cells"] > 0
    a_gpu = res["stat"]["energyabs"] / res["stat"]["energytot"]
    n = 200_000
    ora = oracle_run(workloads.sphere100(n), 16384, want_field=False)
    assert ora["errors"] == 0
    tot = seq_sum(ora["genergy"][1::2])
    a_ora = (tot - seq_sum(ora["genergy"][0::2])) / tot
    # weights in [0, 1]: the sample mean's standard error is below 0.5/sqrt(n)
    assert abs(a_gpu - a_ora) < 5 * 0.5 / np.sqrt(n), (a_gpu, a_ora)


def test_fullsize_mirror_symmetry(full_run):
    _, res, _ = full_run
    flux = res["flux"].astype(np.float64)  # [x, y, z, gate]
    for axis in (0, 1):
        other = tuple(a for a in range(4) if a != axis)
        slab = flux.sum(axis=other)
        mirror = slab[::-1]
        # slabs crossed by many photons, less the pair at the source plane:
        # the pencil enters at x = 50.0, which lies in voxel 50, so slab 50
        # (not 49) holds the unscattered photons' deposits
        core = np.r_[30:49, 51:70]
        rel = np.abs(slab[core] - mirror[core]) / mirror[core]
        assert rel.max() < 0.002, (axis, rel.max())
        assert slab[50] > slab[49]


def test_fullsize_hot_box_on_off(full_run):
    cfg, res, _ = full_run
    old_t = mcx_amd.set_tally_mode(2)
    old_h = mcx_amd.set_hot_box(False)
    try:
        off = mcx_amd.run(cfg)
        st = mcx_amd.last_stats()
    finally:
        mcx_amd.set_tally_mode(old_t)
        mcx_amd.set_hot_box(old_h)
    assert st["hot_cells"] == 0
    assert off["stat"]["energytot"] == res["stat"]["energytot"]
    assert off["stat"]["energyabs"] == res["stat"]["energyabs"]
    a, b = res["flux"].astype(np.float64), off["flux"].astype(np.float64)
    sig = b > 1e-6 * b.max()
    assert np.abs(a[sig] - b[sig]).max() <= 4e-7 * np.abs(b[sig]).max()
    assert (np.abs(a[sig] - b[sig]) / b[sig]).max() < 1e-6
