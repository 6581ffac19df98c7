"""GPU tests of the denoiser kernels against the float64 CPU references in
dn_ref.py, and of the whole U-Net's gradients against CPU autograd.

Inputs are BF16-representable, so products are exact and differences come
from FP32 accumulation order and the BF16 rounding of outputs: a BF16
output must be within one BF16 rounding (2^-8 relative) of the reference,
FP32 outputs within 1e-5 of the largest magnitude.
"""
import ctypes

import numpy as np
import pytest

import dn_ref
from mcx_amd import _lib

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
L = _lib.lib()


def _dev_bf16(a):
    h = dn_ref.to_bf16_bits(np.ascontiguousarray(a, dtype=np.float32))
    return torch.from_numpy(h.view(np.int16).copy()).to("cuda")


def _host_f32(t, shape):
    torch.cuda.synchronize()
    h = t.cpu().numpy()
    if h.dtype == np.int16:
        return dn_ref.from_bf16_bits(h.view(np.uint16)).reshape(shape).astype(np.float64)
    return h.reshape(shape).astype(np.float64)


def _p(t, off=0):
    return ctypes.c_void_p(t.data_ptr() + off * t.element_size())


def _bf_close(got, ref):
    tol = np.abs(ref) * 2.0 ** -8 + 1e-6 * (np.abs(ref).max() + 1e-30)
    bad = np.abs(got - ref) > tol
    assert not bad.any(), (np.abs(got - ref)[bad].max(), bad.sum())


def _weights(rng, cin, cout):
    kp = (27 * cin + 31) // 32 * 32
    w = dn_ref.bf16_round(rng.standard_normal((cout, 27, cin)) * 0.2)
    wpad = np.zeros((cout, kp), dtype=np.float32)
    wpad[:, :27 * cin] = w.reshape(cout, -1)
    return w.astype(np.float64), wpad, kp


SHAPE = (2, 6, 10, 12)


@pytest.mark.parametrize("cin,cout,shape", [(8, 16, SHAPE), (16, 32, SHAPE), (48, 16, SHAPE), (32, 64, SHAPE),
                                            (192, 64, SHAPE), (8, 16, (1, 9, 7, 37)), (16, 16, (1, 9, 7, 37)),
                                            (48, 32, (2, 5, 9, 33)), (96, 32, (1, 8, 8, 16)), (16, 48, SHAPE),
                                            (64, 64, (1, 9, 7, 37)), (128, 128, SHAPE)])
def test_conv3_forward(cin, cout, shape):
    """Both kernels (the LDS halo tile for cout <= 32, the per-wave tiles
    otherwise), on shapes that are and are not multiples of the tile."""
    rng = np.random.default_rng(cin * 100 + cout)
    x = dn_ref.bf16_round(rng.standard_normal(shape + (cin,)))
    w, wpad, kp = _weights(rng, cin, cout)
    b = rng.standard_normal(cout).astype(np.float32)
    dx, dw = _dev_bf16(x), _dev_bf16(wpad)
    db = torch.from_numpy(b).cuda()
    y = torch.zeros(int(np.prod(shape)) * cout, dtype=torch.int16, device="cuda")
    n, d, h, wd = shape
    assert L.mcx_dn_conv3_fwd(_p(dx), cin, None, 0, _p(dw), kp, _p(db), _p(y), cout, n, d, h, wd, cin, cout, 1, None) == 0
    ref = dn_ref.conv3_fwd(x, w, b, relu=True)
    _bf_close(_host_f32(y, ref.shape), ref)


@pytest.mark.parametrize("cu,cx,cout,shape", [(32, 16, 16, SHAPE), (64, 32, 32, SHAPE), (32, 16, 16, (1, 8, 8, 34))])
def test_conv3_concatenating_forward_equals_materialised(cu, cx, cout, shape):
    """conv(cat(up2(xu), x)) read on the fly equals the upsampling written
    into a concatenation buffer followed by the plain convolution, bit for
    bit (the same BF16 operands in the same K order)."""
    rng = np.random.default_rng(cu + cx)
    n, d, h, wd = shape
    cin = cu + cx
    xu = dn_ref.bf16_round(rng.standard_normal((n, d // 2, h // 2, wd // 2, cu)))
    x = dn_ref.bf16_round(rng.standard_normal(shape + (cx,)))
    w, wpad, kp = _weights(rng, cin, cout)
    b = rng.standard_normal(cout).astype(np.float32)
    dxu, dx, dw = _dev_bf16(xu), _dev_bf16(x), _dev_bf16(wpad)
    db = torch.from_numpy(b).cuda()
    nv = int(np.prod(shape))
    cat = torch.zeros(nv * cin, dtype=torch.int16, device="cuda")
    y1 = torch.zeros(nv * cout, dtype=torch.int16, device="cuda")
    y2 = torch.zeros(nv * cout, dtype=torch.int16, device="cuda")
    assert L.mcx_dn_upsample_fwd(_p(dxu), cu, _p(cat), cin, n, d // 2, h // 2, wd // 2, cu, None) == 0
    cat.view(nv, cin)[:, cu:] = dx.view(nv, cx)
    assert L.mcx_dn_conv3_fwd(_p(cat), cin, None, 0, _p(dw), kp, _p(db), _p(y1), cout, n, d, h, wd, cin, cout, 1, None) == 0
    assert L.mcx_dn_conv3_fwd_cat(_p(dxu), cu, cu, _p(dx), cx, _p(dw), kp, _p(db), _p(y2), cout, n, d, h, wd, cin, cout, 1,
                                  None) == 0
    torch.cuda.synchronize()
    assert torch.equal(y1, y2)
    ref = dn_ref.conv3_fwd(np.concatenate([xu.repeat(2, 1).repeat(2, 2).repeat(2, 3), x], axis=-1), w, b, relu=True)
    _bf_close(_host_f32(y2, ref.shape), ref)
    # odd dimensions and channel counts that are not multiples of 16 are refused
    assert L.mcx_dn_conv3_fwd_cat(_p(dxu), cu, cu, _p(dx), cx, _p(dw), kp, _p(db), _p(y2), cout, n, d + 1, h, wd, cin,
                                  cout, 1, None) == -22
    assert L.mcx_dn_conv3_fwd_cat(_p(dxu), cu, cu - 8, _p(dx), cx, _p(dw), kp, _p(db), _p(y2), cout, n, d, h, wd,
                                  cin - 8, cout, 1, None) == -22


def test_unet_inference_forward_equals_training_forward():
    """The inference forward (no concatenation buffers) predicts exactly
    what the training forward predicts."""
    from mcx_amd import denoise
    net = denoise.UNet(seed=11)
    rng = np.random.default_rng(12)
    nb, S = 2, 16
    x = dn_ref.bf16_round(rng.standard_normal((nb, S, S, S, denoise.NFEAT)) * 0.5)
    xd = _dev_bf16(x).view(torch.bfloat16).reshape(nb, S, S, S, denoise.NFEAT)
    td = torch.zeros(nb * S ** 3, dtype=torch.float32, device="cuda")
    p_train, _ = net.forward(xd, td)
    p_train = p_train.clone()
    p_infer, _ = net.forward(xd)
    torch.cuda.synchronize()
    assert torch.equal(p_train, p_infer)


def test_conv3_forward_channel_slices():
    """input and output as channel slices of wider buffers (skip concat)"""
    rng = np.random.default_rng(5)
    cin, cout, xcs, xoff, ycs, yoff = 16, 16, 48, 32, 48, 16
    xw = dn_ref.bf16_round(rng.standard_normal(SHAPE + (xcs,)))
    w, wpad, kp = _weights(rng, cin, cout)
    dx, dw = _dev_bf16(xw), _dev_bf16(wpad)
    y = torch.zeros(int(np.prod(SHAPE)) * ycs, dtype=torch.int16, device="cuda")
    n, d, h, wd = SHAPE
    assert L.mcx_dn_conv3_fwd(_p(dx, xoff), xcs, None, 0, _p(dw), kp, None, _p(y, yoff), ycs, n, d, h, wd, cin, cout, 0,
                              None) == 0
    ref = dn_ref.conv3_fwd(xw[..., xoff:xoff + cin], w, None, relu=False)
    got = _host_f32(y, SHAPE + (ycs,))
    _bf_close(got[..., yoff:yoff + cout], ref)
    assert np.all(got[..., :yoff] == 0) and np.all(got[..., yoff + cout:] == 0)


@pytest.mark.parametrize("cin,cout", [(16, 16), (96, 32), (64, 128)])
def test_conv3_input_gradient(cin, cout):
    rng = np.random.default_rng(7 + cin)
    gy = dn_ref.bf16_round(rng.standard_normal(SHAPE + (cout,)))
    mask = dn_ref.bf16_round(rng.standard_normal(SHAPE + (cout,)))
    w, wpad, kp = _weights(rng, cin, cout)
    kpt = (27 * cout + 31) // 32 * 32
    dgy, dmask, dw = _dev_bf16(gy), _dev_bf16(mask), _dev_bf16(wpad)
    wt = torch.zeros(cin * kpt, dtype=torch.int16, device="cuda")
    gx = torch.zeros(int(np.prod(SHAPE)) * cin, dtype=torch.int16, device="cuda")
    n, d, h, wd = SHAPE
    assert L.mcx_dn_weight_flip(_p(dw), kp, _p(wt), kpt, cin, cout, None) == 0
    assert L.mcx_dn_conv3_fwd(_p(dgy), cout, _p(dmask), cout, _p(wt), kpt, None, _p(gx), cin, n, d, h, wd, cout, cin, 0,
                              None) == 0
    ref = dn_ref.conv3_dgrad(gy, w, mask=mask)
    _bf_close(_host_f32(gx, ref.shape), ref)


ROWS = (1, 5, 6, 37)  # rows of 37 voxels: the all-taps row kernel, a full and a ragged segment


@pytest.mark.parametrize("cin,cout,nsplit,shape", [(8, 16, 3, SHAPE), (16, 16, 7, SHAPE), (48, 16, 1, SHAPE),
                                                   (16, 32, 2, SHAPE), (32, 32, 3, SHAPE), (32, 64, 5, SHAPE),
                                                   (192, 64, 2, SHAPE), (128, 128, 4, SHAPE),
                                                   (8, 16, 3, ROWS), (48, 16, 4, ROWS), (16, 32, 2, ROWS),
                                                   (96, 32, 5, ROWS), (48, 32, 3, ROWS), (32, 64, 2, ROWS),
                                                   (192, 64, 3, ROWS), (64, 128, 2, ROWS)])
def test_conv3_weight_and_bias_gradient(cin, cout, nsplit, shape):
    """SHAPE (rows of 12) runs the per-tap kernel; ROWS the all-taps row
    kernel for 16-64 output channels (input tiles split over workgroups:
    96 -> 32 in pairs, 48 -> 32 and 192 -> 64 one tile each)"""
    SHAPE_ = shape
    rng = np.random.default_rng(cin + cout + shape[3])
    x = dn_ref.bf16_round(rng.standard_normal(SHAPE_ + (cin,)))
    gy = dn_ref.bf16_round(rng.standard_normal(SHAPE_ + (cout,)))
    mask = dn_ref.bf16_round(rng.standard_normal(SHAPE_ + (cout,)))
    kp = (27 * cin + 31) // 32 * 32
    nv = int(np.prod(SHAPE_))
    dx, dgy, dmask = _dev_bf16(x), _dev_bf16(gy), _dev_bf16(mask)
    part = torch.zeros(nsplit * 27 * cout * cin, dtype=torch.float32, device="cuda")
    dw = torch.full((cout * kp,), 7.0, dtype=torch.float32, device="cuda")
    db = torch.zeros(cout, dtype=torch.float32, device="cuda")
    n, d, h, wd = SHAPE_
    assert L.mcx_dn_conv3_wgrad(_p(dx), cin, _p(dgy), cout, _p(dmask), cout, _p(part), nsplit, n, d, h, wd, cin, cout,
                                None) == 0
    assert L.mcx_dn_wgrad_reduce(_p(part), nsplit, _p(dw), kp, cin, cout, None) == 0
    assert L.mcx_dn_bias_grad(_p(dgy), cout, _p(dmask), cout, _p(part), nsplit, _p(db), nv, cout, None) == 0
    rw, rb = dn_ref.conv3_wgrad(x, gy, mask=mask)
    got = _host_f32(dw, (cout, kp))
    assert np.all(got[:, 27 * cin:] == 0)
    got = got[:, :27 * cin].reshape(cout, 27, cin)
    assert np.abs(got - rw).max() <= 1e-5 * np.abs(rw).max()
    assert np.abs(_host_f32(db, (cout,)) - rb).max() <= 1e-5 * np.abs(rb).max()


@pytest.mark.parametrize("c", [16, 4])
def test_pool_upsample_add(c):
    """c = 16 runs the 8-channel vector kernels, c = 4 the scalar ones"""
    rng = np.random.default_rng(3 + c)
    shape = (2, 4, 6, 8, c)
    x = dn_ref.bf16_round(rng.standard_normal(shape))
    dx = _dev_bf16(x)
    nlow = shape[0] * 2 * 3 * 4
    y = torch.zeros(nlow * c, dtype=torch.int16, device="cuda")
    arg = torch.zeros(nlow * c, dtype=torch.uint8, device="cuda")
    assert L.mcx_dn_maxpool_fwd(_p(dx), c, _p(y), c, _p(arg), 2, 4, 6, 8, c, None) == 0
    ry, rarg = dn_ref.maxpool(x)
    assert np.array_equal(_host_f32(y, ry.shape), ry)
    assert np.array_equal(arg.cpu().numpy().reshape(rarg.shape), rarg)
    gy = dn_ref.bf16_round(rng.standard_normal(ry.shape))
    g = torch.zeros(int(np.prod(shape)), dtype=torch.int16, device="cuda")
    assert L.mcx_dn_maxpool_bwd(_p(_dev_bf16(gy)), c, _p(arg), _p(g), c, 2, 4, 6, 8, c, None) == 0
    assert np.array_equal(_host_f32(g, shape), dn_ref.maxpool_bwd(gy, rarg, shape))
    # upsample from the pooled grid back up, into a 48-channel buffer slice
    up = torch.zeros(int(np.prod(shape[:4])) * 48, dtype=torch.int16, device="cuda")
    assert L.mcx_dn_upsample_fwd(_p(y), c, _p(up, 32), 48, 2, 2, 3, 4, c, None) == 0
    got = _host_f32(up, shape[:4] + (48,))
    assert np.array_equal(got[..., 32:32 + c], dn_ref.upsample(ry))
    assert np.all(got[..., :32] == 0) and np.all(got[..., 32 + c:] == 0)
    gu = dn_ref.bf16_round(rng.standard_normal(shape[:4] + (48,)))
    gl = torch.zeros(nlow * c, dtype=torch.int16, device="cuda")
    assert L.mcx_dn_upsample_bwd(_p(_dev_bf16(gu), 32), 48, _p(gl), c, 2, 2, 3, 4, c, None) == 0
    ref = dn_ref.upsample_bwd(gu[..., 32:32 + c].astype(np.float64))
    _bf_close(_host_f32(gl, ref.shape), ref)
    s = torch.zeros(int(np.prod(shape)), dtype=torch.int16, device="cuda")
    assert L.mcx_dn_add(_p(dx), c, _p(g), c, _p(s), c, int(np.prod(shape[:4])), c, None) == 0
    _bf_close(_host_f32(s, shape), x.astype(np.float64) + dn_ref.maxpool_bwd(gy, rarg, shape))


@pytest.mark.parametrize("c,gcs", [(16, 48), (4, 4)])
def test_relu_mask_in_place(c, gcs):
    """g[v][ch] = 0 where y is not > 0 (negative, zero, -0, NaN), in a
    channel slice of a wider buffer; c = 4 runs the scalar kernel"""
    rng = np.random.default_rng(21 + c)
    nv = 777
    y = dn_ref.bf16_round(rng.standard_normal((nv, c)))
    y[::7] = 0.0
    y[3::11] = -0.0
    y[5::13] = np.nan
    y[6::17] = 1e-40  # BF16 subnormal, > 0
    g = dn_ref.bf16_round(rng.standard_normal((nv, gcs)))
    dg = _dev_bf16(g)
    off = gcs - c
    assert L.mcx_dn_relu_mask(_p(dg, off), gcs, _p(_dev_bf16(y)), c, nv, c, None) == 0
    got = _host_f32(dg, (nv, gcs))
    yb = dn_ref.from_bf16_bits(dn_ref.to_bf16_bits(y.astype(np.float32))).astype(np.float64)
    want = g.astype(np.float64).copy()
    keep = np.greater(yb, 0.0, where=~np.isnan(yb), out=np.zeros_like(yb, dtype=bool))
    want[:, off:] = np.where(keep, want[:, off:], 0.0)
    assert np.array_equal(got, want)


def test_head_and_adam():
    rng = np.random.default_rng(9)
    nv, c = 3000, 16
    f = dn_ref.bf16_round(rng.standard_normal((nv, c)))
    res = dn_ref.bf16_round(rng.standard_normal((nv, 8)))
    w = rng.standard_normal(c).astype(np.float32)
    b = np.array([0.3], dtype=np.float32)
    tgt = rng.standard_normal(nv).astype(np.float32)
    df_, dres = _dev_bf16(f), _dev_bf16(res)
    dw, dbb, dt = (torch.from_numpy(a).cuda() for a in (w, b, tgt))
    pred = torch.zeros(nv, dtype=torch.float32, device="cuda")
    dpred = torch.zeros(nv, dtype=torch.float32, device="cuda")
    lpart = torch.zeros(64, dtype=torch.float32, device="cuda")
    assert L.mcx_dn_head_fwd(_p(df_), c, _p(dw), _p(dbb), _p(dres), 8, _p(dt), _p(pred), _p(dpred), _p(lpart), 64, nv, c,
                             None) == 0
    rp = f.astype(np.float64) @ w + b[0] + res[:, 0]
    assert np.abs(_host_f32(pred, (nv,)) - rp).max() < 1e-4
    rl = np.mean((rp - tgt) ** 2)
    assert abs(float(lpart.sum()) / nv - rl) < 1e-4 * rl
    gfeat = torch.zeros(nv * c, dtype=torch.int16, device="cuda")
    part = torch.zeros(64 * (c + 1), dtype=torch.float32, device="cuda")
    dwb = torch.zeros(c + 1, dtype=torch.float32, device="cuda")
    assert L.mcx_dn_head_bwd(_p(df_), c, _p(dw), _p(dpred), _p(gfeat), c, _p(part), 64, _p(dwb), nv, c, None) == 0
    rdp = 2 * (rp - tgt) / nv
    _bf_close(_host_f32(gfeat, (nv, c)), np.outer(rdp, w.astype(np.float64)))
    got = _host_f32(dwb, (c + 1,))
    assert np.abs(got[:c] - f.T.astype(np.float64) @ rdp).max() < 1e-5
    assert abs(got[c] - rdp.sum()) < 1e-5
    # Adam: two steps against the textbook update
    n = 1000
    p0 = rng.standard_normal(n).astype(np.float32)
    g0 = rng.standard_normal(n).astype(np.float32)
    P, M, V, G = (torch.from_numpy(a).cuda() for a in (p0.copy(), np.zeros(n, np.float32), np.zeros(n, np.float32), g0))
    PB = torch.zeros(n, dtype=torch.int16, device="cuda")
    rp_, rm, rv = p0.astype(np.float64), np.zeros(n), np.zeros(n)
    for step in (1, 2):
        assert L.mcx_dn_adam(_p(P), _p(M), _p(V), _p(G), _p(PB), n, 1e-3, 0.9, 0.999, 1e-8, step, None) == 0
        rm = 0.9 * rm + 0.1 * g0
        rv = 0.999 * rv + 0.001 * g0.astype(np.float64) ** 2
        rp_ = rp_ - 1e-3 * (rm / (1 - 0.9 ** step)) / (np.sqrt(rv / (1 - 0.999 ** step)) + 1e-8)
    assert np.abs(_host_f32(P, (n,)) - rp_).max() < 1e-6
    assert np.array_equal(_host_f32(PB, (n,)), dn_ref.bf16_round(P.cpu().numpy()).astype(np.float64))


def test_unet_gradients_match_cpu_autograd():
    """Whole-network forward and backward against float64 CPU autograd:
    tightly against a reference that rounds to BF16 where the GPU stores
    BF16, loosely against the exact (unrounded) network."""
    from mcx_amd import denoise
    net = denoise.UNet(seed=1)
    rng = np.random.default_rng(2)
    nb, S = 2, 16
    x = dn_ref.bf16_round(rng.standard_normal((nb, S, S, S, denoise.NFEAT)) * 0.5)
    tgt = (x[..., 0] + 0.1 * rng.standard_normal((nb, S, S, S))).astype(np.float32)
    # the network runs with BF16 conv weights: give the reference the same
    p = net.state()
    for name, cin, cout in denoise.CONVS:
        wo, bo, _, _, kp = net.layout[name]
        p[wo:wo + cout * kp] = dn_ref.bf16_round(p[wo:wo + cout * kp])
    net.load_state(p)
    xd = _dev_bf16(x).view(torch.bfloat16).reshape(nb, S, S, S, denoise.NFEAT)
    td = torch.from_numpy(tgt.reshape(-1)).cuda()
    pred, loss = net.forward(xd, td)
    net.backward()
    torch.cuda.synchronize()
    g = net.g.cpu().numpy().astype(np.float64)
    pred = pred.cpu().numpy()
    for bf16, tol_rel, tol_cos in ((True, 0.02, 0.9998), (False, 0.15, 0.99)):
        rloss, rpred, rgrad = dn_ref.unet_torch(p, net.layout, x, tgt, denoise.CONVS, net.head_off, bf16=bf16)
        assert abs(float(loss) - rloss) < 0.02 * rloss
        assert np.corrcoef(pred, rpred)[0, 1] > 0.999
        errs = []
        for name, cin, cout in denoise.CONVS + [("head", 16, 1)]:
            if name == "head":
                sl = slice(net.head_off, net.head_off + 17)
            else:
                wo, bo, _, _, kp = net.layout[name]
                sl = slice(wo, bo + cout)
            a, r = g[sl], rgrad[sl]
            rel = np.linalg.norm(a - r) / (np.linalg.norm(r) + 1e-30)
            cos = a @ r / (np.linalg.norm(a) * np.linalg.norm(r) + 1e-30)
            print("bf16-ref %d grad %-5s rel %.5f cos %.7f" % (bf16, name, rel, cos))
            errs.append((name, rel, cos))
        for name, rel, cos in errs:
            assert rel < tol_rel and cos > tol_cos, (bf16, name, rel, cos)


def test_unet_training_reduces_loss():
    from mcx_amd import denoise
    net = denoise.UNet(seed=3)
    rng = np.random.default_rng(4)
    nb, S = 2, 16
    clean = np.cumsum(rng.standard_normal((nb, S, S, S)), axis=3) * 0.1
    noisy = clean + 0.3 * rng.standard_normal(clean.shape)
    x = np.zeros((nb, S, S, S, denoise.NFEAT), dtype=np.float32)
    x[..., 0] = noisy
    x[..., 5] = 1.0
    xd = _dev_bf16(x).view(torch.bfloat16).reshape(x.shape)
    td = torch.from_numpy(clean.astype(np.float32).reshape(-1)).cuda()
    losses = [float(net.train_step(xd, td, lr=1e-3)) for _ in range(60)]
    assert np.all(np.isfinite(losses))
    assert min(losses[-5:]) < 0.5 * losses[0], (losses[0], losses[-5:])


def test_feature_kernels_match_host_features():
    from mcx_amd import denoise
    rng = np.random.default_rng(12)
    X, Y, Z = 6, 10, 12
    vol = rng.integers(0, 4, size=(X, Y, Z)).astype(np.uint8)
    prop = np.array([[0, 0, 1, 1], [0.01, 5.0, 0.9, 1.37], [0.05, 12.0, 0.8, 1.4], [0.002, 0.8, 0.95, 1.33]],
                    dtype=np.float32)
    phi = np.exp(rng.uniform(-20, 0, size=(X, Y, Z))) * (rng.uniform(size=(X, Y, Z)) > 0.1)
    neg = rng.uniform(size=(X, Y, Z)) < 0.05          # roulette can leave small negative voxels
    phi_in = np.where(neg, -1e-6, phi)
    phi = np.where(neg, 0.0, phi)
    for flip in (False, True):
        got = denoise.features_gpu(phi_in, vol, prop, unitinmm=0.5, flip_z=flip)
        torch.cuda.synchronize()
        g = got.float().cpu().numpy()[0]
        ref = denoise.features(phi_in, vol, prop, unitinmm=0.5)
        assert np.all(np.isfinite(g)) and np.all(np.isfinite(ref))
        if flip:
            ref = ref[:, :, ::-1]
        assert np.all(np.abs(g - ref) <= np.abs(ref) * 2.0 ** -8 + 1e-5)
        pred = torch.from_numpy(np.ascontiguousarray(ref[..., 0]).reshape(-1).astype(np.float32)).cuda()
        back = denoise.inverse_gpu(pred, (X, Y, Z), flip_z=flip)
        assert np.allclose(back, phi, rtol=1e-4, atol=1e-12)


def test_head_loss_skips_nan_targets():
    rng = np.random.default_rng(13)
    nv, c = 2000, 16
    f = dn_ref.bf16_round(rng.standard_normal((nv, c)))
    res = dn_ref.bf16_round(rng.standard_normal((nv, 8)))
    w = rng.standard_normal(c).astype(np.float32)
    b = np.zeros(1, dtype=np.float32)
    tgt = rng.standard_normal(nv).astype(np.float32)
    tgt[::3] = np.nan
    pred = torch.zeros(nv, dtype=torch.float32, device="cuda")
    dpred = torch.zeros(nv, dtype=torch.float32, device="cuda")
    lpart = torch.zeros(32, dtype=torch.float32, device="cuda")
    dw, dbb, dt = (torch.from_numpy(a).cuda() for a in (w, b, tgt))
    assert L.mcx_dn_head_fwd(_p(_dev_bf16(f)), c, _p(dw), _p(dbb), _p(_dev_bf16(res)), 8, _p(dt), _p(pred), _p(dpred),
                             _p(lpart), 32, nv, c, None) == 0
    rp = f.astype(np.float64) @ w + res[:, 0]
    ok = ~np.isnan(tgt)
    e = np.where(ok, rp - tgt, 0.0)
    assert abs(float(lpart.sum()) - float((e ** 2).sum())) < 1e-3 * float((e ** 2).sum())
    d = _host_f32(dpred, (nv,))
    assert np.all(d[~ok] == 0) and np.allclose(d[ok], 2 * e[ok] / nv, rtol=1e-4, atol=1e-9)


def test_trained_denoiser_beats_its_noisy_input():
    """Train the U-Net briefly on pairs of low/high-photon maps that the
    engine itself generates (32^3 volumes), then check on held-out volumes
    that the denoised map is closer to a high-photon reference than the
    noisy input it was given."""
    from mcx_amd import denoise, dn_experiment as dx
    tr = [dx.training_volume(77, i, edge=32) for i in range(96)]
    lo, _, _ = dx.simulate(tr, 2e4, 1)
    hi, _, _ = dx.simulate(tr, 2e6, 2)
    X = dx.to_device_bf16(dx.make_features(lo, tr))
    T = torch.from_numpy(dx.make_targets(hi, tr)).cuda()
    net = denoise.UNet(seed=5)
    losses, _ = dx.train(net, X, T, steps=1500, batch=8, lr=1e-3, seed=5)
    assert losses[-1][1] < losses[0][1]
    ev = [dx.training_volume(78, i, edge=32) for i in range(6)]
    refs = [dx.simulate(ev, 1e6, 100 + k)[0] for k in range(4)]
    low, _, _ = dx.simulate(ev, 2e4, 3)
    den = dx.denoise_maps(net, dx.make_features(low, ev))
    e_low, e_den = [], []
    for i in range(len(ev)):
        ref, mask = dx.reliable_mask([r[i] for r in refs], cv_max=0.02)
        e_low.append(dx.median_error(low[i], ref, mask))
        e_den.append(dx.median_error(den[i], ref, mask))
    print("median error raw %.4f denoised %.4f" % (np.median(e_low), np.median(e_den)))
    assert np.median(e_den) < 0.9 * np.median(e_low)


@pytest.mark.parametrize("nx,ny,nz", [(36, 5, 7), (256, 3, 2), (1024, 2, 3)])
def test_roi_kernels_match_torch(nx, ny, nz):
    """per-map maximum (clamped at 0) and the counts of voxels above a
    threshold per x and per x row, against torch; map 2 is all negative"""
    K = 3
    g = torch.Generator(device="cpu").manual_seed(nx)
    maps = torch.randn((K, nz, ny, nx), generator=g).cuda()
    maps[2] = -maps[2].abs() - 1.0
    mx = torch.zeros(K, dtype=torch.float32, device="cuda")
    assert L.mcx_dn_roi_max(_p(maps), K, nx * ny * nz, _p(mx), None) == 0
    want = maps.view(K, -1).max(dim=1).values.clamp(min=0.0)
    assert torch.equal(mx, want)
    thr = torch.tensor([0.5, -0.25, -1.5], dtype=torch.float32, device="cuda")
    xc = torch.zeros((K, nx), dtype=torch.int32, device="cuda")
    rc = torch.full((K, nz * ny), -1, dtype=torch.int32, device="cuda")
    assert L.mcx_dn_roi_counts(_p(maps), K, nx, ny, nz, _p(thr), _p(xc), _p(rc), None) == 0
    above = maps > thr.view(K, 1, 1, 1)
    assert torch.equal(xc, above.sum(dim=(1, 2)).to(torch.int32))
    assert torch.equal(rc, above.sum(dim=3).view(K, -1).to(torch.int32))
