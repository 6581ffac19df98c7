"""Semi-infinite diffusion solutions used as analytic checks (test
infrastructure).  Restated from utils/cwdiffusion.m, utils/tddiffusion.m and
utils/getdistance.m (Boas 2002; Haskell 1994): an extrapolated-boundary image
source pair, source depth z0 = 1/(mua+musp), extrapolation distance
zb = 2D(1+Reff)/(1-Reff), D = 1/(3(mua+musp))."""
import numpy as np


def getdistance(srcpos, detpos):
    """separation[r, s] = |src_s - det_r| (utils/getdistance.m)"""
    s = np.atleast_2d(np.asarray(srcpos, dtype=np.float64))
    d = np.atleast_2d(np.asarray(detpos, dtype=np.float64))
    return np.linalg.norm(d[:, None, :] - s[None, :, :], axis=2)


def _images(mua, musp, Reff, srcpos):
    D = 1.0 / (3.0 * (mua + musp))
    zb = (1.0 + Reff) / (1.0 - Reff) * 2.0 * D
    z0 = 1.0 / (musp + mua)
    s = np.atleast_2d(np.asarray(srcpos, dtype=np.float64))
    real = np.concatenate([s[:, 0:2], s[:, 2:3] + z0], axis=1)
    image = np.concatenate([s[:, 0:2], s[:, 2:3] - z0 - 2.0 * zb], axis=1)
    return D, real, image


def cwdiffusion(mua, musp, Reff, srcpos, detpos):
    """CW fluence per unit source power, 1/mm^2 (utils/cwdiffusion.m)."""
    D, real, image = _images(mua, musp, Reff, srcpos)
    r = getdistance(real, detpos)
    r2 = getdistance(image, detpos)
    b = np.sqrt(3.0 * mua * musp)
    return 1.0 / (4.0 * np.pi * D) * (np.exp(-b * r) / r - np.exp(-b * r2) / r2)


def tddiffusion(mua, musp, v, Reff, srcpos, detpos, t):
    """Time-resolved fluence rate, 1/(mm^2 s) (utils/tddiffusion.m)."""
    D, real, image = _images(mua, musp, Reff, srcpos)
    r = getdistance(real, detpos)
    r2 = getdistance(image, detpos)
    t = np.asarray(t, dtype=np.float64)
    s = 4.0 * D * v * t
    return v / ((s * np.pi) ** 1.5) * np.exp(-mua * v * t) * (np.exp(-(r ** 2) / s) - np.exp(-(r2 ** 2) / s))
