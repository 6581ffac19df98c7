"""GPU fluence against diffusion theory at 1e7 photons (the reference's
homogeneous validation, mcxlab/examples/demo_validation_homogeneous.m:
87-117): the cube60 medium (mua 0.005/mm, mus 1/mm, g 0.01, n 1.37) with
Reff = 0.493 for the n = 1.37 / 1 boundary, compared away from the source
(10-20 mm, about 10 transport lengths and more), where the diffusion
approximation holds.  The semi-infinite solutions are restated in
tests/diffusion.py and pinned by their own golden there.
"""
import numpy as np
import pytest

import mcx_amd
from mcx_amd import workloads
import diffusion

pytestmark = pytest.mark.gpu

MUA, MUSP, REFF = 0.005, 0.99, 0.493
SRC = [29, 29, 0]


def _run(cfg):
    old = mcx_amd.set_tally_mode(2)
    try:
        return mcx_amd.run(cfg)
    finally:
        mcx_amd.set_tally_mode(old)


def test_cw_fluence_matches_diffusion_theory():
    cfg = workloads.cube60(10_000_000)
    cfg["quiet"] = True
    res = _run(cfg)
    # normalised flux x gate width = CW fluence per photon (1/mm^2)
    f = res["flux"][..., 0].astype(np.float64) * cfg["tstep"]
    ks = np.arange(10, 21)
    det = np.stack([np.full(ks.shape, 29.5), np.full(ks.shape, 29.5), ks + 0.5], 1)
    cw = diffusion.cwdiffusion(MUA, MUSP, REFF, SRC, det)[:, 0]
    ratio = f[28:30, 28:30, ks].mean(axis=(0, 1)) / cw
    assert abs(ratio.mean() - 1) < 0.04 and np.all(np.abs(ratio - 1) < 0.10), ratio
    xs = np.arange(30, 42)
    det = np.stack([xs + 0.5, np.full(xs.shape, 29.5), np.full(xs.shape, 10.5)], 1)
    cw = diffusion.cwdiffusion(MUA, MUSP, REFF, SRC, det)[:, 0]
    ratio = 0.5 * (f[xs, 28, 10] + f[xs, 29, 10]) / cw
    assert abs(ratio.mean() - 1) < 0.04 and np.all(np.abs(ratio - 1) < 0.10), ratio


def test_time_resolved_fluence_matches_diffusion_theory():
    """Gates of 0.5 ns in a 120^3 cube of the same medium (its side faces
    60 mm from the source, so the semi-infinite solution holds to ~4 ns;
    in the 60^3 cube the diffusion front reaches the side faces by 2 ns):
    the gate-averaged fluence rate 15 mm below the source against the
    time-domain solution averaged over each gate, 1.5-4 ns.  Measured
    ratios 0.91-0.96, flat in time: the extrapolated-boundary model's
    amplitude at this depth, not a drift; the bound is that model's
    accuracy."""
    tstep = 5e-10
    cfg = workloads.cube60(10_000_000, vol=np.ones((120, 120, 120), dtype=np.uint8), srcpos=[59, 59, 0],
                           tend=4e-9, tstep=tstep)
    cfg["quiet"] = True
    res = _run(cfg)
    flux = res["flux"].astype(np.float64)  # fluence rate per photon, 1/(mm^2 s)
    v = 299792458000.0 / 1.37
    det = [[59.5, 59.5, 15.5]]
    got, want = [], []
    for gate in range(3, 8):
        t = np.linspace(gate * tstep, (gate + 1) * tstep, 101)
        td = diffusion.tddiffusion(MUA, MUSP, v, REFF, [59, 59, 0], det, t)[0]
        want.append(np.trapezoid(td, t) / tstep)
        got.append(flux[58:60, 58:60, 15, gate].mean())
    ratio = np.array(got) / np.array(want)
    print("time-resolved ratio", ratio)
    assert abs(ratio.mean() - 1) < 0.08 and np.all(np.abs(ratio - 1) < 0.12), ratio
    assert np.ptp(ratio) < 0.08, ratio  # no drift across the gates
