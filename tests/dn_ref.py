"""CPU references for the denoiser kernels (test-only).

numpy float64 restatements of each kernel (3x3x3 convolution forward,
input gradient, weight gradient, pooling, upsampling, head, Adam) on
BF16-representable inputs, and a float64 torch-CPU autograd restatement of
the whole U-Net for gradient checks.  Layouts follow include/mcx_amd_denoise.h
(NDHWC, weights [cout][27*cin] tap-major).
"""
import numpy as np


def bf16_round(a):
    """float -> nearest-even BF16 value, returned as float32."""
    a = np.ascontiguousarray(a, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32)


def to_bf16_bits(a):
    a = bf16_round(a)
    return (a.view(np.uint32) >> 16).astype(np.uint16)


def from_bf16_bits(h):
    return (np.asarray(h, dtype=np.uint32) << 16).view(np.float32)


def _pad(x):
    return np.pad(x, ((0, 0), (1, 1), (1, 1), (1, 1), (0, 0)))


def conv3_fwd(x, w, b=None, relu=True, mask=None):
    """x [N,D,H,W,Cin], w [Cout, 27, Cin] -> y [N,D,H,W,Cout] (float64)."""
    x = np.asarray(x, dtype=np.float64)
    if mask is not None:
        x = x * (np.asarray(mask) > 0)
    N, D, H, W, C = x.shape
    xp = _pad(x)
    cout = w.shape[0]
    y = np.zeros((N, D, H, W, cout))
    for t in range(27):
        dz, dy, dx = t // 9, (t // 3) % 3, t % 3
        y += xp[:, dz:dz + D, dy:dy + H, dx:dx + W, :] @ np.asarray(w[:, t, :], dtype=np.float64).T
    if b is not None:
        y += b
    return np.maximum(y, 0) if relu else y


def conv3_dgrad(gy, w, mask=None):
    """Gradient wrt x of conv3_fwd (before any mask on x): gy [N,D,H,W,Cout]."""
    gy = np.asarray(gy, dtype=np.float64)
    if mask is not None:
        gy = gy * (np.asarray(mask) > 0)
    N, D, H, W, _ = gy.shape
    gp = _pad(gy)
    cin = w.shape[2]
    gx = np.zeros((N, D, H, W, cin))
    for t in range(27):
        # y[v] += x[v + o_t] w_t  =>  gx[u] += gy[u - o_t] w_t
        dz, dy, dx = 2 - t // 9, 2 - (t // 3) % 3, 2 - t % 3
        gx += gp[:, dz:dz + D, dy:dy + H, dx:dx + W, :] @ np.asarray(w[:, t, :], dtype=np.float64)
    return gx


def conv3_wgrad(x, gy, mask=None):
    """dW [Cout, 27, Cin] and db [Cout] of conv3_fwd."""
    x = np.asarray(x, dtype=np.float64)
    gy = np.asarray(gy, dtype=np.float64)
    if mask is not None:
        gy = gy * (np.asarray(mask) > 0)
    N, D, H, W, cin = x.shape
    cout = gy.shape[-1]
    xp = _pad(x)
    dw = np.zeros((cout, 27, cin))
    g2 = gy.reshape(-1, cout)
    for t in range(27):
        dz, dy, dx = t // 9, (t // 3) % 3, t % 3
        dw[:, t, :] = g2.T @ xp[:, dz:dz + D, dy:dy + H, dx:dx + W, :].reshape(-1, cin)
    return dw, g2.sum(axis=0)


def maxpool(x):
    N, D, H, W, C = x.shape
    r = x.reshape(N, D // 2, 2, H // 2, 2, W // 2, 2, C).transpose(0, 1, 3, 5, 2, 4, 6, 7).reshape(N, D // 2, H // 2,
                                                                                                  W // 2, 8, C)
    return r.max(axis=4), r.argmax(axis=4)


def maxpool_bwd(gy, arg, shape):
    N, D, H, W, C = shape
    g = np.zeros((N, D // 2, H // 2, W // 2, 8, C))
    np.put_along_axis(g, arg[:, :, :, :, None, :], gy[:, :, :, :, None, :], axis=4)
    return g.reshape(N, D // 2, H // 2, W // 2, 2, 2, 2, C).transpose(0, 1, 4, 2, 5, 3, 6, 7).reshape(N, D, H, W, C)


def upsample(x):
    return x.repeat(2, axis=1).repeat(2, axis=2).repeat(2, axis=3)


def upsample_bwd(gy):
    N, D, H, W, C = gy.shape
    return gy.reshape(N, D // 2, 2, H // 2, 2, W // 2, 2, C).sum(axis=(2, 4, 6))


def unet_torch(params, layout, x0, target, convs, head_off, bf16=False):
    """float64 CPU autograd restatement of mcx_amd.denoise.UNet; returns
    (loss, pred, grads as a flat array in the parameter layout).

    With bf16=True the reference rounds to BF16 wherever the GPU stores a
    BF16 tensor: every convolution output (forward) and the gradient of
    every stored activation (backward: input gradients of the convolutions,
    the head's feature gradient, upsampling and skip-sum gradients)."""
    import torch
    import torch.nn.functional as F
    p = torch.tensor(np.asarray(params, dtype=np.float64), requires_grad=True)
    x = torch.tensor(np.asarray(x0, dtype=np.float64)).permute(0, 4, 1, 2, 3)

    def rb(t):
        return t.to(torch.bfloat16).to(torch.float64)

    class Round(torch.autograd.Function):
        @staticmethod
        def forward(ctx, t):
            return rb(t)

        @staticmethod
        def backward(ctx, g):
            return g

    def stored(t):
        """a tensor the GPU keeps in BF16, whose gradient it keeps in BF16"""
        if bf16:
            t = Round.apply(t)
            t.register_hook(rb)
        return t

    def conv(name, t):
        wo, bo, cin, cout, kp = layout[name]
        w = p[wo:wo + cout * kp].reshape(cout, kp)[:, :27 * cin].reshape(cout, 3, 3, 3, cin).permute(0, 4, 1, 2, 3)
        return stored(F.relu(F.conv3d(t, w, p[bo:bo + cout], padding=1)))

    def cat(a, b):
        return stored(torch.cat([a, b], 1))

    def pool(t):
        return stored(F.max_pool3d(t, 2))

    e1 = conv("e1", conv("e1a", x))
    e2 = conv("e2", conv("e2a", pool(e1)))
    e3 = conv("e3", conv("e3a", pool(e2)))
    b2 = conv("b2", conv("b1", pool(e3)))
    up = lambda t: t.repeat_interleave(2, 2).repeat_interleave(2, 3).repeat_interleave(2, 4)
    d3 = conv("d3", conv("d3a", cat(up(b2), e3)))
    d2 = conv("d2", conv("d2a", cat(up(d3), e2)))
    d1 = conv("d1", conv("d1a", cat(up(d2), e1)))
    hw = p[head_off:head_off + 16]
    hb = p[head_off + 16]
    pred = (d1 * hw[None, :, None, None, None]).sum(1) + hb + x[:, 0]
    tgt = torch.tensor(np.asarray(target, dtype=np.float64)).reshape(pred.shape)
    loss = ((pred - tgt) ** 2).mean()
    loss.backward()
    return float(loss), pred.detach().numpy().reshape(-1), p.grad.numpy().copy()
