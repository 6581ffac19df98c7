"""The denoised arm's pipeline on the device (mcx_amd.dn_experiment.
HeadDenoiser): the batched U-Net call over several maps equals one call per
map, and the signal boxes from the two box kernels equal the same rule
computed with torch reductions."""
import numpy as np
import pytest

import mcx_amd
from mcx_amd import denoise, dn_experiment as dx, time_to_error, workloads

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


@pytest.fixture(scope="module")
def setup():
    hc = workloads.head256(20_000, n=64, quiet=True, minenergy=1e-4)
    hc["autopilot"] = 1
    K = 4
    maps, _, _ = time_to_error._replicas(hc, 20_000, K, 31)
    net = denoise.UNet(seed=2024)
    net.load_state(np.load(time_to_error.WEIGHTS)["p"])
    return hc, maps, net


def _torch_boxes(dn, maps):
    """the box rule with torch reductions (threshold x the map's maximum,
    central 1 - 2q of the voxels above it per axis, margin, 8-aligned)"""
    X, Y, Z = dn.shape
    K = maps.shape[0]
    p = maps.view(K, Z, Y, X)
    thr = p.reshape(K, -1).max(dim=1).values.clamp(min=0.0) * dn.thr
    m = (p > thr.view(K, 1, 1, 1)).to(torch.float32)
    proj = [m.sum(dim=2).sum(dim=1), m.sum(dim=3).sum(dim=1), m.sum(dim=3).sum(dim=2)]
    out = []
    for k in range(K):
        r = []
        for a, n in enumerate((X, Y, Z)):
            q = proj[a][k]
            c = q.cumsum(0)
            tot = c[-1:]
            lo = int((c <= dn.q * tot).sum())
            hi = n - int((c >= (1.0 - dn.q) * tot).sum()) + 1
            lo, hi = lo - dn.margin, hi + dn.margin
            r += [max(0, lo // 8 * 8), min(n, -(-hi // 8) * 8)]
        out.append(tuple(r))
    return out


def test_boxes_match_torch_rule(setup):
    hc, maps, net = setup
    for thr, margin in ((1e-5, 16), (1e-3, 4)):
        dn = dx.HeadDenoiser(net, hc, roi_threshold=thr, roi_margin=margin)
        assert dn.boxes(maps) == _torch_boxes(dn, maps)


def test_batched_unet_equals_one_call_per_map(setup):
    hc, maps, net = setup
    dn = dx.HeadDenoiser(net, hc, roi_threshold=1e-3, roi_margin=4)
    boxes = dn.common_size(dn.boxes(maps))
    many = dn.run_many(maps, 20_000, boxes, nb=3)  # a full and a partial batch
    for k in range(maps.shape[0]):
        one = dn.run(maps[k], 20_000, box=boxes[k])
        assert torch.equal(many[k], one), k
        # outside the box the map is the input
        x0, x1, y0, y1, z0, z1 = boxes[k]
        X, Y, Z = dn.shape
        inside = torch.zeros((Z, Y, X), dtype=torch.bool, device="cuda")
        inside[z0:z1, y0:y1, x0:x1] = True
        assert torch.equal(many[k].view(Z, Y, X)[~inside], maps[k].view(Z, Y, X)[~inside])
