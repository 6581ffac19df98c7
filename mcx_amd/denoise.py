"""Learned fluence denoiser: a 4-level 3D U-Net trained and run on the GPU by
the project's own HIP kernels (include/mcx_amd_denoise.h).

Architecture (channels per level 16, 32, 64, 128; 3x3x3 convolutions with
ReLU; 2x2x2 max pooling down, nearest-neighbour x2 upsampling up, skip
connections by channel concatenation; a 1x1x1 head):

    x0 (8 feature channels) -> e1a(16) -> e1(16) -pool-> e2a(32) -> e2(32)
    -pool-> e3a(64) -> e3(64) -pool-> b1(128) -> b2(128)
    -up-> [up(b2) 128 | e3 64] -> d3a(64) -> d3(64)
    -up-> [up(d3) 64 | e2 32]  -> d2a(32) -> d2(32)
    -up-> [up(d2) 32 | e1 16]  -> d1a(16) -> d1(16) -> head(1) + x0[..., 0]

The network predicts a correction to the input log-fluence (residual).
Activations and gradients of activations are BF16, accumulation FP32,
master weights / weight gradients / Adam moments FP32.  PyTorch only
provides device memory and the stream; every operation is a kernel of
libmcx_amd.so launched through the C ABI.

Features (channel order): scaled log fluence, log mua, log mus, g, n,
tissue mask, two zero channels.
"""
import ctypes
import math

import numpy as np

from . import _lib

LOG_EPS = 1e-9
LOG_SHIFT = 10.0
LOG_SCALE = 5.0
NFEAT = 8


def kpad_of(cin):
    return (27 * cin + 31) // 32 * 32


def _torch():
    import torch
    return torch


def _ptr(t, off_elems=0):
    return ctypes.c_void_p(t.data_ptr() + off_elems * t.element_size())


def _chk(rc, what):
    if rc != 0:
        raise RuntimeError("%s failed (%d)" % (what, rc))


def log_feature(phi):
    """Scaled log fluence used as network input and target.  Negative
    voxels (Russian-roulette survivors deposit w0 - w with their boosted
    weight, as in the reference) count as zero."""
    return (np.log(np.maximum(np.asarray(phi, dtype=np.float64), 0.0) + LOG_EPS) + LOG_SHIFT) / LOG_SCALE


def inv_log_feature(f):
    return np.maximum(np.exp(np.asarray(f, dtype=np.float64) * LOG_SCALE - LOG_SHIFT) - LOG_EPS, 0.0)


def photon_feature(nphoton):
    """Input channel 6: the photon count the map was made with, (log10 N - 6) / 2
    (0 when unknown), so the network can weigh the map's noise level."""
    return 0.0 if not nphoton else (np.log10(float(nphoton)) - 6.0) / 2.0


def features(phi, vol, prop, unitinmm=1.0, photons=None):
    """Input features [X, Y, Z, 8] (float32) of a CW fluence map phi [X, Y, Z]
    made with `photons` photons on a label volume with media table prop (per
    mm; converted to per voxel)."""
    vol = np.asarray(vol)
    prop = np.asarray(prop, dtype=np.float64)
    f = np.zeros(vol.shape + (NFEAT,), dtype=np.float32)
    f[..., 0] = log_feature(phi)
    mua = prop[:, 0] * unitinmm
    mus = prop[:, 1] * unitinmm
    lab = vol.astype(np.int64)
    tissue = lab > 0
    f[..., 1] = np.where(tissue, np.log(np.maximum(mua[lab], 1e-6)) + 4.0, 0.0)
    f[..., 2] = np.where(tissue, np.log(np.maximum(mus[lab], 1e-6)), 0.0)
    f[..., 3] = np.where(tissue, (prop[lab, 2] - 0.8) * 10.0, 0.0)
    f[..., 4] = np.where(tissue, (prop[lab, 3] - 1.4) * 10.0, 0.0)
    f[..., 5] = tissue
    f[..., 6] = photon_feature(photons)
    return f


def _media_table(prop, unitinmm):
    prop = np.asarray(prop, dtype=np.float32)
    tab = prop.copy()
    tab[:, 0] *= unitinmm
    tab[:, 1] *= unitinmm
    return np.ascontiguousarray(tab)


def features_gpu(phi, vol, prop, unitinmm=1.0, flip_z=False, photons=None):
    """features() computed on the GPU by mcx_dn_features: a BF16 device
    tensor [1, X, Y, Z, 8].  phi and vol are [X, Y, Z] arrays."""
    torch = _torch()
    L = _lib.lib()
    X, Y, Z = vol.shape
    dphi = torch.from_numpy(np.asfortranarray(phi, dtype=np.float32).ravel(order="F")).to("cuda")
    dvol = torch.from_numpy(np.asfortranarray(vol, dtype=np.uint8).ravel(order="F")).to("cuda")
    dtab = torch.from_numpy(_media_table(prop, unitinmm)).to("cuda")
    f = torch.empty((1, X, Y, Z, NFEAT), dtype=torch.bfloat16, device="cuda")
    st = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    _chk(L.mcx_dn_features(_ptr(dphi), _ptr(dvol), _ptr(dtab), int(dtab.shape[0]), _ptr(f), X, Y, Z, int(flip_z), LOG_EPS,
                           LOG_SHIFT, LOG_SCALE, st), "features")
    if photons:
        f[..., 6].fill_(photon_feature(photons))
    return f


def inverse_gpu(pred, shape, flip_z=False):
    """Fluence [X, Y, Z] (float32, F-order) from a prediction tensor."""
    torch = _torch()
    L = _lib.lib()
    X, Y, Z = shape
    phi = torch.empty(X * Y * Z, dtype=torch.float32, device="cuda")
    st = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    _chk(L.mcx_dn_inv_features(_ptr(pred), _ptr(phi), X, Y, Z, int(flip_z), LOG_EPS, LOG_SHIFT, LOG_SCALE, st),
         "inv_features")
    return phi.cpu().numpy().reshape((X, Y, Z), order="F")


# (name, cin, cout) of the 3x3x3 convolutions in forward order
CONVS = [("e1a", NFEAT, 16), ("e1", 16, 16), ("e2a", 16, 32), ("e2", 32, 32), ("e3a", 32, 64), ("e3", 64, 64),
         ("b1", 64, 128), ("b2", 128, 128), ("d3a", 192, 64), ("d3", 64, 64), ("d2a", 96, 32), ("d2", 32, 32),
         ("d1a", 48, 16), ("d1", 16, 16)]


class UNet:
    """Parameters, optimiser state and the forward/backward launch sequence."""

    def __init__(self, seed=2024, device="cuda"):
        torch = _torch()
        self.L = _lib.lib()
        self.device = device
        self.layout = {}
        off = 0
        for name, cin, cout in CONVS:
            kp = kpad_of(cin)
            self.layout[name] = (off, off + cout * kp, cin, cout, kp)
            off += cout * kp + cout
        self.head_off = off
        off += 16 + 1
        self.nparam = off
        self.count = sum(cout * 27 * cin + cout for _, cin, cout in CONVS) + 17
        g = np.random.default_rng(seed)
        p = np.zeros(self.nparam, dtype=np.float32)
        for name, cin, cout in CONVS:
            wo, bo, _, _, kp = self.layout[name]
            w = np.zeros((cout, kp), dtype=np.float32)
            w[:, :27 * cin] = g.standard_normal((cout, 27 * cin)) * math.sqrt(2.0 / (27 * cin))
            p[wo:wo + cout * kp] = w.ravel()
        p[self.head_off:self.head_off + 16] = g.standard_normal(16) * math.sqrt(1.0 / 16) * 0.1
        self.p = torch.from_numpy(p).to(device)
        self.m = torch.zeros_like(self.p)
        self.v = torch.zeros_like(self.p)
        self.g = torch.zeros_like(self.p)
        self.pb = torch.zeros(self.nparam, dtype=torch.bfloat16, device=device)
        self.step_count = 0
        self._sync_bf16()
        self.flip = {}
        for name, cin, cout in CONVS:
            self.flip[name] = torch.zeros(cin * kpad_of(cout), dtype=torch.bfloat16, device=device)
        self._acts = None

    # ---- helpers ---------------------------------------------------------
    def _stream(self):
        return ctypes.c_void_p(_torch().cuda.current_stream().cuda_stream)

    def _sync_bf16(self):
        _chk(self.L.mcx_dn_f32_to_bf16(_ptr(self.p), _ptr(self.pb), self.nparam, self._stream()), "f32_to_bf16")

    def state(self):
        return self.p.detach().cpu().numpy().copy()

    def load_state(self, p):
        self.p.copy_(_torch().from_numpy(np.asarray(p, dtype=np.float32)).to(self.device))
        self._sync_bf16()

    def weights(self, name):
        """(w [cout, 27*cin] fp32, b [cout]) of a conv, for tests."""
        wo, bo, cin, cout, kp = self.layout[name]
        p = self.state()
        return p[wo:wo + cout * kp].reshape(cout, kp)[:, :27 * cin], p[bo:bo + cout]

    def _alloc(self, nb, S):
        """S: the full-resolution grid (X, Y, Z), each a multiple of 8"""
        torch = _torch()
        key = (nb, S)
        if self._acts is not None and self._acts["key"] == key:
            return self._acts
        bf = dict(dtype=torch.bfloat16, device=self.device)
        s1, s2, s3, s4 = (tuple(d // f for d in S) for f in (1, 2, 4, 8))
        n1, n2, n3, n4 = (nb * math.prod(s) for s in (s1, s2, s3, s4))
        A = {"key": key, "dims": (s1, s2, s3, s4), "n": (n1, n2, n3, n4)}
        A["x0"] = torch.zeros(n1 * NFEAT, **bf)
        for nm, n, c in [("e1a", n1, 16), ("cat1", n1, 48), ("p1", n2, 16), ("e2a", n2, 32), ("cat2", n2, 96),
                         ("p2", n3, 32), ("e3a", n3, 64), ("cat3", n3, 192), ("p3", n4, 64), ("b1", n4, 128),
                         ("b2", n4, 128), ("d3a", n3, 64), ("d3", n3, 64), ("d2a", n2, 32), ("d2", n2, 32),
                         ("d1a", n1, 16), ("d1", n1, 16)]:
            A[nm] = torch.zeros(n * c, **bf)
            A["g_" + nm] = torch.zeros(n * c, **bf)
        # inference: the encoder outputs in their own buffers (the decoder
        # reads them through the concatenating convolution)
        A["e1s"] = torch.zeros(n1 * 16, **bf)
        A["e2s"] = torch.zeros(n2 * 32, **bf)
        A["arg1"] = torch.zeros(n2 * 16, dtype=torch.uint8, device=self.device)
        A["arg2"] = torch.zeros(n3 * 32, dtype=torch.uint8, device=self.device)
        A["arg3"] = torch.zeros(n4 * 64, dtype=torch.uint8, device=self.device)
        A["pool1"] = torch.zeros(n1 * 16, **bf)
        A["pool2"] = torch.zeros(n2 * 32, **bf)
        A["pool3"] = torch.zeros(n3 * 64, **bf)
        A["pred"] = torch.zeros(n1, dtype=torch.float32, device=self.device)
        A["dpred"] = torch.zeros(n1, dtype=torch.float32, device=self.device)
        self.hblocks = 1024
        A["lpart"] = torch.zeros(self.hblocks, dtype=torch.float32, device=self.device)
        A["hpart"] = torch.zeros(self.hblocks * 17, dtype=torch.float32, device=self.device)
        maxpart = 1
        for name, cin, cout in CONVS:
            nv = self._nvox_of(name, A)
            ns = self._nsplit(nv, cin, cout)
            maxpart = max(maxpart, ns * 27 * cout * cin, ns * cout)
        A["wpart"] = torch.zeros(maxpart, dtype=torch.float32, device=self.device)
        self._acts = A
        return A

    @staticmethod
    def _nsplit(nvox, cin=None, cout=None):
        """Voxel-range splits of a weight gradient (partial sums reduced by
        a second kernel), measured per layer of the batch-8 64^3 step: one
        per 512 voxels (at most 2048) for the 16/32-channel layers, whose
        gradient kernel stages 32-voxel rows for all taps, one per 256
        voxels (at most 1024) for the wider ones, and no more than keep the
        partial sums (splits x 27 x cout x cin floats) within 256 MB."""
        per, cap = 256, 1024
        if cin is not None:
            if cout <= 32:
                per, cap = 512, 2048
            cap = min(cap, max(1, (64 << 20) // (27 * cout * cin)))
        return int(min(cap, max(1, nvox // per)))

    def _nvox_of(self, name, A):
        n1, n2, n3, n4 = A["n"]
        return {"e1a": n1, "e1": n1, "e2a": n2, "e2": n2, "e3a": n3, "e3": n3, "b1": n4, "b2": n4, "d3a": n3,
                "d3": n3, "d2a": n2, "d2": n2, "d1a": n1, "d1": n1}[name]

    def _dims_of(self, name, A):
        s1, s2, s3, s4 = A["dims"]
        return {"e1a": s1, "e1": s1, "e2a": s2, "e2": s2, "e3a": s3, "e3": s3, "b1": s4, "b2": s4, "d3a": s3,
                "d3": s3, "d2a": s2, "d2": s2, "d1a": s1, "d1": s1}[name]

    # ---- forward ---------------------------------------------------------
    def _conv(self, name, x, xoff, xcs, y, yoff, ycs, nb):
        wo, bo, cin, cout, kp = self.layout[name]
        S = self._dims_of(name, self._acts)
        _chk(self.L.mcx_dn_conv3_fwd(_ptr(x, xoff), xcs, None, 0, _ptr(self.pb, wo), kp, _ptr(self.p, bo), _ptr(y, yoff),
                                     ycs, nb, *S, cin, cout, 1, self._stream()), "conv3_fwd " + name)

    def _conv_cat(self, name, xu, xucs, cu, x, xcs, y, ycs, nb):
        """conv of cat(upsample2(xu), x) without writing the concatenation"""
        wo, bo, cin, cout, kp = self.layout[name]
        S = self._dims_of(name, self._acts)
        _chk(self.L.mcx_dn_conv3_fwd_cat(_ptr(xu), xucs, cu, _ptr(x), xcs, _ptr(self.pb, wo), kp, _ptr(self.p, bo),
                                         _ptr(y), ycs, nb, *S, cin, cout, 1, self._stream()), "conv3_fwd_cat " + name)

    def forward(self, x0, target=None):
        """x0: torch bf16 [nb, X, Y, Z, 8] on the device (X, Y, Z multiples
        of 8).  Returns the FP32 prediction [nb*X*Y*Z]; with a target (FP32
        [nb*X*Y*Z])
        also the MSE loss (a device scalar, summed on the host).  Without a
        target (inference) the two full-resolution skip concatenations are
        not written: the decoder reads the encoder outputs and the upsampled
        coarse maps through mcx_dn_conv3_fwd_cat (same result), and backward
        needs a forward with a target."""
        nb, S = int(x0.shape[0]), tuple(int(d) for d in x0.shape[1:4])
        if any(d % 8 for d in S):
            raise ValueError("the grid must be a multiple of 8 in every dimension")
        A = self._alloc(nb, S)
        A["x0"].copy_(x0.reshape(-1))
        L, st = self.L, self._stream()
        s1, s2, s3, s4 = A["dims"]
        n1, n2, n3, n4 = A["n"]
        infer = target is None
        e1, e1o, e1c = (A["e1s"], 0, 16) if infer else (A["cat1"], 32, 48)
        e2, e2o, e2c = (A["e2s"], 0, 32) if infer else (A["cat2"], 64, 96)
        self._conv("e1a", A["x0"], 0, NFEAT, A["e1a"], 0, 16, nb)
        self._conv("e1", A["e1a"], 0, 16, e1, e1o, e1c, nb)
        _chk(L.mcx_dn_maxpool_fwd(_ptr(e1, e1o), e1c, _ptr(A["p1"]), 16, _ptr(A["arg1"]), nb, *s1, 16, st), "pool1")
        self._conv("e2a", A["p1"], 0, 16, A["e2a"], 0, 32, nb)
        self._conv("e2", A["e2a"], 0, 32, e2, e2o, e2c, nb)
        _chk(L.mcx_dn_maxpool_fwd(_ptr(e2, e2o), e2c, _ptr(A["p2"]), 32, _ptr(A["arg2"]), nb, *s2, 32, st), "pool2")
        self._conv("e3a", A["p2"], 0, 32, A["e3a"], 0, 64, nb)
        self._conv("e3", A["e3a"], 0, 64, A["cat3"], 128, 192, nb)
        _chk(L.mcx_dn_maxpool_fwd(_ptr(A["cat3"], 128), 192, _ptr(A["p3"]), 64, _ptr(A["arg3"]), nb, *s3, 64, st), "pool3")
        self._conv("b1", A["p3"], 0, 64, A["b1"], 0, 128, nb)
        self._conv("b2", A["b1"], 0, 128, A["b2"], 0, 128, nb)
        _chk(L.mcx_dn_upsample_fwd(_ptr(A["b2"]), 128, _ptr(A["cat3"]), 192, nb, *s4, 128, st), "up3")
        self._conv("d3a", A["cat3"], 0, 192, A["d3a"], 0, 64, nb)
        self._conv("d3", A["d3a"], 0, 64, A["d3"], 0, 64, nb)
        if infer:
            self._conv_cat("d2a", A["d3"], 64, 64, e2, e2c, A["d2a"], 32, nb)
        else:
            _chk(L.mcx_dn_upsample_fwd(_ptr(A["d3"]), 64, _ptr(A["cat2"]), 96, nb, *s3, 64, st), "up2")
            self._conv("d2a", A["cat2"], 0, 96, A["d2a"], 0, 32, nb)
        self._conv("d2", A["d2a"], 0, 32, A["d2"], 0, 32, nb)
        if infer:
            self._conv_cat("d1a", A["d2"], 32, 32, e1, e1c, A["d1a"], 16, nb)
        else:
            _chk(L.mcx_dn_upsample_fwd(_ptr(A["d2"]), 32, _ptr(A["cat1"]), 48, nb, *s2, 32, st), "up1")
            self._conv("d1a", A["cat1"], 0, 48, A["d1a"], 0, 16, nb)
        self._conv("d1", A["d1a"], 0, 16, A["d1"], 0, 16, nb)
        ho = self.head_off
        tptr = _ptr(target) if target is not None else None
        _chk(L.mcx_dn_head_fwd(_ptr(A["d1"]), 16, _ptr(self.p, ho), _ptr(self.p, ho + 16), _ptr(A["x0"]), NFEAT, tptr,
                               _ptr(A["pred"]), _ptr(A["dpred"]) if target is not None else None,
                               _ptr(A["lpart"]) if target is not None else None, self.hblocks, n1, 16, st), "head_fwd")
        if target is None:
            return A["pred"], None
        return A["pred"], A["lpart"].sum() / n1

    # ---- backward --------------------------------------------------------
    def _conv_bwd(self, name, x, xoff, xcs, y, yoff, ycs, gy, gyoff, gycs, gx, gxoff, gxcs, nb, need_dx=True):
        """y = relu(conv(x)); gy: gradient wrt y.  Writes the weight and bias
        gradients into self.g and, when need_dx, gx = gradient wrt x."""
        A = self._acts
        L, st = self.L, self._stream()
        wo, bo, cin, cout, kp = self.layout[name]
        S = self._dims_of(name, A)
        nv = self._nvox_of(name, A)
        ns = self._nsplit(nv, cin, cout)
        # the ReLU derivative, applied to gy in place once (gy is final here
        # and read by nothing after this layer's three gradients)
        _chk(L.mcx_dn_relu_mask(_ptr(gy, gyoff), gycs, _ptr(y, yoff), ycs, nv, cout, st), "relu_mask " + name)
        _chk(L.mcx_dn_conv3_wgrad(_ptr(x, xoff), xcs, _ptr(gy, gyoff), gycs, None, 0, _ptr(A["wpart"]), ns, nb,
                                  *S, cin, cout, st), "wgrad " + name)
        _chk(L.mcx_dn_wgrad_reduce(_ptr(A["wpart"]), ns, _ptr(self.g, wo), kp, cin, cout, st), "wgrad_reduce " + name)
        _chk(L.mcx_dn_bias_grad(_ptr(gy, gyoff), gycs, None, 0, _ptr(A["wpart"]), ns, _ptr(self.g, bo), nv, cout,
                                st), "bias_grad " + name)
        if need_dx:
            kpt = kpad_of(cout)
            _chk(L.mcx_dn_weight_flip(_ptr(self.pb, wo), kp, _ptr(self.flip[name]), kpt, cin, cout, st), "flip " + name)
            _chk(L.mcx_dn_conv3_fwd(_ptr(gy, gyoff), gycs, None, 0, _ptr(self.flip[name]), kpt, None,
                                    _ptr(gx, gxoff), gxcs, nb, *S, cout, cin, 0, st), "dgrad " + name)

    def backward(self):
        """Gradients of the last forward's loss into self.g."""
        A = self._acts
        nb = A["key"][0]
        L, st = self.L, self._stream()
        s1, s2, s3, s4 = A["dims"]
        n1, n2, n3, n4 = A["n"]
        ho = self.head_off
        _chk(L.mcx_dn_head_bwd(_ptr(A["d1"]), 16, _ptr(self.p, ho), _ptr(A["dpred"]), _ptr(A["g_d1"]), 16,
                               _ptr(A["hpart"]), self.hblocks, _ptr(self.g, ho), n1, 16, st), "head_bwd")
        self._conv_bwd("d1", A["d1a"], 0, 16, A["d1"], 0, 16, A["g_d1"], 0, 16, A["g_d1a"], 0, 16, nb)
        self._conv_bwd("d1a", A["cat1"], 0, 48, A["d1a"], 0, 16, A["g_d1a"], 0, 16, A["g_cat1"], 0, 48, nb)
        _chk(L.mcx_dn_upsample_bwd(_ptr(A["g_cat1"]), 48, _ptr(A["g_d2"]), 32, nb, *s2, 32, st), "up1 bwd")
        self._conv_bwd("d2", A["d2a"], 0, 32, A["d2"], 0, 32, A["g_d2"], 0, 32, A["g_d2a"], 0, 32, nb)
        self._conv_bwd("d2a", A["cat2"], 0, 96, A["d2a"], 0, 32, A["g_d2a"], 0, 32, A["g_cat2"], 0, 96, nb)
        _chk(L.mcx_dn_upsample_bwd(_ptr(A["g_cat2"]), 96, _ptr(A["g_d3"]), 64, nb, *s3, 64, st), "up2 bwd")
        self._conv_bwd("d3", A["d3a"], 0, 64, A["d3"], 0, 64, A["g_d3"], 0, 64, A["g_d3a"], 0, 64, nb)
        self._conv_bwd("d3a", A["cat3"], 0, 192, A["d3a"], 0, 64, A["g_d3a"], 0, 64, A["g_cat3"], 0, 192, nb)
        _chk(L.mcx_dn_upsample_bwd(_ptr(A["g_cat3"]), 192, _ptr(A["g_b2"]), 128, nb, *s4, 128, st), "up3 bwd")
        self._conv_bwd("b2", A["b1"], 0, 128, A["b2"], 0, 128, A["g_b2"], 0, 128, A["g_b1"], 0, 128, nb)
        self._conv_bwd("b1", A["p3"], 0, 64, A["b1"], 0, 128, A["g_b1"], 0, 128, A["g_p3"], 0, 64, nb)
        # e3 feeds pool3 and the skip slice of cat3
        _chk(L.mcx_dn_maxpool_bwd(_ptr(A["g_p3"]), 64, _ptr(A["arg3"]), _ptr(A["pool3"]), 64, nb, *s3, 64, st), "pool3 bwd")
        _chk(L.mcx_dn_add(_ptr(A["pool3"]), 64, _ptr(A["g_cat3"], 128), 192, _ptr(A["g_cat3"], 128), 192, n3, 64, st), "add3")
        self._conv_bwd("e3", A["e3a"], 0, 64, A["cat3"], 128, 192, A["g_cat3"], 128, 192, A["g_e3a"], 0, 64, nb)
        self._conv_bwd("e3a", A["p2"], 0, 32, A["e3a"], 0, 64, A["g_e3a"], 0, 64, A["g_p2"], 0, 32, nb)
        _chk(L.mcx_dn_maxpool_bwd(_ptr(A["g_p2"]), 32, _ptr(A["arg2"]), _ptr(A["pool2"]), 32, nb, *s2, 32, st), "pool2 bwd")
        _chk(L.mcx_dn_add(_ptr(A["pool2"]), 32, _ptr(A["g_cat2"], 64), 96, _ptr(A["g_cat2"], 64), 96, n2, 32, st), "add2")
        self._conv_bwd("e2", A["e2a"], 0, 32, A["cat2"], 64, 96, A["g_cat2"], 64, 96, A["g_e2a"], 0, 32, nb)
        self._conv_bwd("e2a", A["p1"], 0, 16, A["e2a"], 0, 32, A["g_e2a"], 0, 32, A["g_p1"], 0, 16, nb)
        _chk(L.mcx_dn_maxpool_bwd(_ptr(A["g_p1"]), 16, _ptr(A["arg1"]), _ptr(A["pool1"]), 16, nb, *s1, 16, st), "pool1 bwd")
        _chk(L.mcx_dn_add(_ptr(A["pool1"]), 16, _ptr(A["g_cat1"], 32), 48, _ptr(A["g_cat1"], 32), 48, n1, 16, st), "add1")
        self._conv_bwd("e1", A["e1a"], 0, 16, A["cat1"], 32, 48, A["g_cat1"], 32, 48, A["g_e1a"], 0, 16, nb)
        self._conv_bwd("e1a", A["x0"], 0, NFEAT, A["e1a"], 0, 16, A["g_e1a"], 0, 16, None, 0, 0, nb, need_dx=False)

    def adam_step(self, lr=1e-3, beta1=0.9, beta2=0.999, eps=1e-8):
        self.step_count += 1
        _chk(self.L.mcx_dn_adam(_ptr(self.p), _ptr(self.m), _ptr(self.v), _ptr(self.g), _ptr(self.pb), self.nparam, lr,
                                beta1, beta2, eps, self.step_count, self._stream()), "adam")

    def train_step(self, x0, target, lr=1e-3):
        _, loss = self.forward(x0, target)
        self.backward()
        self.adam_step(lr)
        return loss

    def predict(self, x0):
        """Denoised scaled log fluence for input features x0 (bf16 tensor)."""
        pred, _ = self.forward(x0)
        return pred
