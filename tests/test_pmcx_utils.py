"""The pmcx post-processing helpers (mcx_amd.utils) against the reference's
own known answers (pmcx/test/test_utils.py:44-154, copied here as data),
the diffusion oracle, and a .mch round trip through the product's writer."""
import ctypes

import numpy as np
import pytest

import diffusion
from mcx_amd import _lib, utils

PROP = np.array([[0, 0, 1, 1], [0.005, 1, 0, 1.37], [0.1, 10, 0.9, 1]])
PPATH = np.array([[4.0361433, 0], [40.828487, 0], [47.87069, 0]])
DETP = {"detid": np.array([1, 2, 1]), "ppath": PPATH, "prop": PROP,
        "data": np.array([[1, 2, 1], [4.0361433, 40.828487, 47.87069], [0, 0, 0]])}


def test_detweight_known_answer_and_errors():
    np.testing.assert_array_almost_equal(utils.detweight(DETP), [0.9800216, 0.81534624, 0.78713661])
    with pytest.raises(ValueError):
        utils.detweight({"ppath": PPATH})
    with pytest.raises(ValueError):
        utils.detweight({"prop": np.array([[1.0]]), "ppath": np.ones((2, 2))}, np.array([[1.0]]))


def test_meanpath_dettime_dettpsf_known_answers():
    np.testing.assert_array_almost_equal(utils.meanpath(DETP, PROP), [29.01278056, 0])
    with pytest.raises(ValueError):
        utils.meanpath({"ppath": PPATH})
    np.testing.assert_array_almost_equal(utils.dettime(DETP), [[1.84444811e-11, 1.86579167e-10, 2.18760825e-10]])
    np.testing.assert_array_almost_equal(utils.dettpsf(DETP, 1, PROP, np.array([0, 5e-9, 5e-9])), 1.767158)
    # a time grid of 3 bins puts each photon in its own bin
    tp = utils.dettpsf(DETP, 1, PROP, np.array([0, 3e-10, 1e-10]))
    np.testing.assert_array_almost_equal(tp[:, 0], [0.9800216, 0, 0.78713661])


def test_distance_and_diffusion_known_answers():
    np.testing.assert_array_almost_equal(utils.getdistance([[1, 5, 7], [43, 62, 4]], [[1, 3, 8], [40, 60, 2]]),
                                         [[2.23606798, 72.53275122], [67.60917098, 4.12310563]])
    c0 = 299792458000
    t = np.arange(5e-9 / 2, 5e-9, 5e-9)
    phi = utils.tddiffusion(0.005, 1.0, c0, 0, [29, 29, 0], np.array([[29, 13, 8], [0, 0, 0]]) + 1, t)
    np.testing.assert_array_almost_equal(phi, [[1888.13933534], [92.31054244]])
    det = np.array([[29, 29, 10], [29, 35, 0], [40, 29, 5]], dtype=float)
    cw, r = utils.cwdiffusion(0.005, 0.99, 0.493, [29, 29, 0], det)
    assert np.allclose(cw, diffusion.cwdiffusion(0.005, 0.99, 0.493, [29, 29, 0], det), rtol=1e-12)


def test_detphoton_splits_fields():
    out = utils.detphoton(DETP["data"], 1, "DP")
    assert set(out) == {"detid", "ppath"}
    assert np.array_equal(out["detid"], [1, 2, 1])
    assert np.array_equal(out["ppath"], [[4.0361433], [40.828487], [47.87069]])
    raw = np.arange(2 * 14, dtype=np.float32).reshape(14, 2)  # 1 + 3 x 2 media + 3 + 3 + 1 rows
    raw[0] = [1, 65536 * 3 + 2]
    full = utils.detphoton(raw, 2, "DSPMXVW")
    assert np.array_equal(full["detid"], [1, 2]) and np.array_equal(full["srcid"], [0, 3])
    assert full["nscat"].shape == (2, 2) and full["ppath"].shape == (2, 2) and full["mom"].shape == (2, 2)
    assert full["p"].shape == (2, 3) and full["v"].shape == (2, 3) and full["w0"].shape == (2, 1)
    assert np.array_equal(full["w0"][:, 0], raw[13]) and np.array_equal(full["v"], raw[10:13].T)


def test_loadmch_reads_the_writer_output(tmp_path):
    L = _lib.lib()
    c = _lib.Config()
    L.mcx_initcfg(ctypes.byref(c))
    c.parentid = _lib.mpPython
    c.flog = None
    c.session = str(tmp_path / "d").encode()
    c.outputformat = 0
    c.his.maxmedia, c.his.detnum, c.his.colcount = 2, 2, 6
    c.his.totalphoton, c.his.detected, c.his.savedphoton = 1000, 100, 4
    c.his.unitinmm, c.his.normalizer = 0.1, 1e6
    rows = np.arange(24, dtype=np.float32).reshape(4, 6)
    L.mcx_savedetphoton(rows.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), None, 4, 0, ctypes.byref(c))
    data, hdr, seeds = utils.loadmch(str(tmp_path / "d.mch"))
    assert np.array_equal(data, rows) and seeds is None
    assert hdr["magic"] == b"MCXH" and hdr["colcount"] == 6 and hdr["savedphoton"] == 4
    assert hdr["totalphoton"] == 1000 and hdr["unitinmm"] == pytest.approx(0.1)
    L.mcx_clearcfg(ctypes.byref(c))
