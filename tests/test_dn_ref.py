"""CPU pins of the denoiser references in dn_ref.py: the numpy convolution,
its input and weight gradients, pooling and upsampling agree with torch's
CPU conv3d / max_pool3d / autograd in float64, and BF16 rounding is
nearest-even."""
import numpy as np
import pytest

import dn_ref

torch = pytest.importorskip("torch")
F = torch.nn.functional


def _t(a):
    return torch.tensor(np.asarray(a, dtype=np.float64)).permute(0, 4, 1, 2, 3)


def test_bf16_rounding_is_nearest_even():
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, 1.0 + 2 ** -9, -3.5, 0.0], dtype=np.float32)
    got = dn_ref.bf16_round(x)
    assert got.tolist() == [1.0, 1.0, 1.0 + 4 * 2 ** -8, 1.0, -3.5, 0.0]
    assert np.array_equal(dn_ref.from_bf16_bits(dn_ref.to_bf16_bits(x)), got)


def test_conv_and_gradients_match_torch():
    rng = np.random.default_rng(0)
    x = rng.standard_normal((2, 4, 5, 6, 8))
    w = rng.standard_normal((16, 27, 8))
    b = rng.standard_normal(16)
    mask = rng.standard_normal((2, 4, 5, 6, 16))
    xt = _t(x).requires_grad_(True)
    wt = torch.tensor(w.reshape(16, 3, 3, 3, 8)).permute(0, 4, 1, 2, 3).requires_grad_(True)
    bt = torch.tensor(b).requires_grad_(True)
    y = F.conv3d(xt, wt, bt, padding=1)
    ref = dn_ref.conv3_fwd(x, w, b, relu=False)
    assert np.allclose(y.detach().permute(0, 2, 3, 4, 1).numpy(), ref, atol=1e-10)
    assert np.allclose(np.maximum(ref, 0), dn_ref.conv3_fwd(x, w, b, relu=True))
    gy = rng.standard_normal(ref.shape)
    (y * _t(gy * (mask > 0))).sum().backward()
    gx = dn_ref.conv3_dgrad(gy, w, mask=mask)
    assert np.allclose(xt.grad.permute(0, 2, 3, 4, 1).numpy(), gx, atol=1e-9)
    dw, db = dn_ref.conv3_wgrad(x, gy, mask=mask)
    assert np.allclose(wt.grad.permute(0, 2, 3, 4, 1).reshape(16, 27, 8).numpy(), dw, atol=1e-9)
    assert np.allclose(bt.grad.numpy(), db, atol=1e-9)


def test_pool_and_upsample_match_torch():
    rng = np.random.default_rng(1)
    x = rng.standard_normal((2, 4, 6, 8, 3))
    xt = _t(x).requires_grad_(True)
    y = F.max_pool3d(xt, 2)
    ry, arg = dn_ref.maxpool(x)
    assert np.array_equal(y.detach().permute(0, 2, 3, 4, 1).numpy(), ry)
    gy = rng.standard_normal(ry.shape)
    (y * _t(gy)).sum().backward()
    assert np.allclose(xt.grad.permute(0, 2, 3, 4, 1).numpy(), dn_ref.maxpool_bwd(gy, arg, x.shape))
    u = dn_ref.upsample(ry)
    ut = F.interpolate(_t(ry), scale_factor=2, mode="nearest").permute(0, 2, 3, 4, 1).numpy()
    assert np.array_equal(u, ut)
    gu = rng.standard_normal(u.shape)
    assert np.allclose(dn_ref.upsample_bwd(gu), gu.reshape(2, 2, 2, 3, 2, 4, 2, 3).sum(axis=(2, 4, 6)))
