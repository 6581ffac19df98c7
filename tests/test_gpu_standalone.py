"""mcx_run_from_json end to end on the GPU.

The standalone path (the caller it replaces: src/mcx_utils.c:5578-5603 and
src/mcx_core.cu:3930-3960) reads the reference's JSON input, runs the
simulation and writes the .jnii volume and the _detp.jdat detected
photons with their seeds.  The written arrays must equal what the pmcx
dictionary path returns for the same problem: under autopilot both use
the same launch layout, so the detected records and seeds are the same
set and the fluence differs only by the order of FP64 additions.  The
standalone run happens in a child process (errors there exit the
process, as the reference's do)."""
import base64
import ctypes
import json
import os
import subprocess
import sys
import zlib

import numpy as np
import pytest

import mcx_amd
from mcx_amd import _lib
from parity_util import sort_rows

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _problem(tmp_path):
    return {
        "Session": {"ID": "e2e", "RootPath": str(tmp_path), "Photons": 1e6, "RNGSeed": 1648335518,
                    "DoMismatch": 1, "OutputFormat": "jnii", "DoSaveSeed": 1, "SaveDataMask": "DP"},
        "Domain": {"Dim": [60, 60, 60], "OriginType": 1,
                   "Media": [{"mua": 0.0, "mus": 0.0, "g": 1.0, "n": 1.0},
                             {"mua": 0.005, "mus": 1.0, "g": 0.01, "n": 1.37},
                             {"mua": 0.002, "mus": 5.0, "g": 0.9, "n": 1.0}]},
        "Shapes": [{"Grid": {"Tag": 1, "Size": [60, 60, 60]}}, {"Sphere": {"Tag": 2, "O": [30, 30, 30], "R": 10}}],
        "Forward": {"T0": 0.0, "T1": 5.0e-09, "Dt": 5.0e-09},
        "Optode": {"Source": {"Type": "pencil", "Pos": [29.0, 29.0, 0.0], "Dir": [0.0, 0.0, 1.0]},
                   "Detector": [{"Pos": [29.0, 19.0, 0.0], "R": 2.0}, {"Pos": [29.0, 39.0, 0.0], "R": 2.0}]},
    }


def _jdata(obj):
    raw = zlib.decompress(base64.b64decode(obj["_ArrayZipData_"]))
    dt = {"uint32": "<u4", "single": "<f4", "uint8": "u1"}[obj["_ArrayType_"]]
    order = "F" if obj.get("_ArrayOrder_") == "c" else "C"
    return np.frombuffer(raw, dtype=dt).reshape(obj["_ArraySize_"], order=order)


def test_run_from_json_writes_what_pmcx_returns(tmp_path):
    prob = _problem(tmp_path)
    path = tmp_path / "e2e.json"
    path.write_text(json.dumps(prob))
    code = ("import sys; sys.path.insert(0, %r); from mcx_amd import _lib; "
            "sys.exit(_lib.lib().mcx_run_from_json(%r.encode()))" % (ROOT, str(path)))
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=300)
    print(r.stdout[-600:], r.stderr[-600:], flush=True)
    assert r.returncode == 0, r.stderr
    assert (tmp_path / "e2e.jnii").exists() and (tmp_path / "e2e_detp.jdat").exists()

    # the same problem through the dictionary path: the volume from the
    # same JSON reader (shapes rasterised), everything else as keys
    L = _lib.lib()
    c = _lib.Config()
    L.mcx_initcfg(ctypes.byref(c))
    c.parentid = _lib.mpPython
    c.flog = None
    L.mcx_readconfig(json.dumps(prob).encode(), ctypes.byref(c))
    assert _lib.last_error()[0] == 0
    vol = (np.ctypeslib.as_array(c.vol, shape=(216000,)).copy() & 0x7FFFFFFF).reshape((60, 60, 60), order="F")
    L.mcx_clearcfg(ctypes.byref(c))
    cfg = {"nphoton": 1000000, "vol": vol.astype(np.uint8), "issrcfrom0": 1, "seed": 1648335518,
           "prop": np.array([[m["mua"], m["mus"], m["g"], m["n"]] for m in prob["Domain"]["Media"]], np.float32),
           "srcpos": [29, 29, 0], "srcdir": [0, 0, 1], "tstart": 0.0, "tend": 5e-9, "tstep": 5e-9, "isreflect": 1,
           "detpos": np.array([[29, 19, 0, 2], [29, 39, 0, 2]], np.float32), "savedetflag": "DP",
           "issaveseed": 1, "quiet": True}
    old = mcx_amd.set_tally_mode(2)
    try:
        res = mcx_amd.run(cfg)
    finally:
        mcx_amd.set_tally_mode(old)

    jn = json.load(open(tmp_path / "e2e.jnii"))
    assert jn["NIFTIHeader"]["Dim"] == [60, 60, 60, 1, 1]
    flux = _jdata(jn["NIFTIData"])[..., 0, 0].astype(np.float64)
    ref = res["flux"][..., 0].astype(np.float64)
    sig = ref > 1e-4 * ref.max()
    assert np.abs(flux[sig] - ref[sig]).max() / ref.max() < 1e-6

    jd = json.load(open(tmp_path / "e2e_detp.jdat"))["MCXData"]
    info, data = jd["Info"], jd["PhotonData"]
    n = res["detp"].shape[1]
    assert n > 100 and info["DetectedPhoton"] == n and info["DetNum"] == 2 and info["SeedByte"] == 16
    rows = np.hstack([_jdata(data["detid"]).astype(np.float64), _jdata(data["ppath"]).astype(np.float64),
                      _jdata(data["seed"]).view(np.uint32).astype(np.float64)])
    want = np.hstack([res["detp"].T.astype(np.float64), res["seeds"].T.copy().view(np.uint32).astype(np.float64)])
    assert np.array_equal(sort_rows(rows), sort_rows(want))
    # the log line carries the absorbed fraction of the same launched energy
    assert "absorbed:" in r.stdout
    a = float(r.stdout.split("absorbed:")[1].split("%")[0])
    assert a == pytest.approx(100 * res["stat"]["energyabs"] / res["stat"]["energytot"], abs=1e-4)
