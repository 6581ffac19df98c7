"""GPU parity of the property-word media formats (src/mcx_core.cu:582-687):
the HIP kernel's MediaWord instantiation against the oracle on
heterogeneous volumes of every format, with Fresnel reflection, a void
slab, a source launched from outside the grid (void skipping with specular
entry) and detectors.  Energies and detected-photon records are bit-exact;
the fluence agrees to FP64 tally order."""
import numpy as np
import pytest

import mcx_amd
from mcx_amd import media
from parity_util import normalize_like_engine, oracle_run, seq_sum, sort_rows

pytestmark = pytest.mark.gpu

N = 24
NT = 2048
PROP = np.array([[0, 0, 1, 1], [0.01, 1.0, 0.9, 1.37], [0.03, 2.5, 0.5, 1.5], [0.005, 0.6, 0.0, 1.2]],
                dtype=np.float32)


def _volume(fmt, rng):
    lab = rng.integers(1, 4, size=(N, N, N))
    lab[:2] = 0
    void = lab == 0
    if fmt == 101:
        m = rng.uniform(0.002, 0.05, size=(1, N, N, N)).astype(np.float32)
        m[0][void] = np.nan
        return m, {}
    if fmt == 100:
        m = np.stack([rng.uniform(0.002, 0.05, (N, N, N)), rng.uniform(0.3, 3.0, (N, N, N))]).astype(np.float32)
        m[0][void] = np.nan
        return m, {}
    if fmt == 102:
        w = media.half_words(rng.uniform(0.002, 0.05, (N, N, N)), rng.uniform(0.3, 3.0, (N, N, N)))
        w[void] = 0
        return w, {"mediabyte": 102}
    if fmt == 96:
        m = np.stack([rng.uniform(0.002, 0.05, (N, N, N)), rng.uniform(0.3, 3.0, (N, N, N)),
                      rng.uniform(0.0, 0.9, (N, N, N)), rng.uniform(1.0, 1.5, (N, N, N))]).astype(np.float32)
        m[0][void] = np.nan
        return m, {}
    if fmt == 98:
        s = np.stack([lab, rng.integers(1, 4, (N, N, N)), rng.integers(0, 32768, (N, N, N))]).astype(np.int16)
        s[2][lab % 2 == 0] = 0
        s[1][void] = 0
        return s, {}
    if fmt == 99:
        f = np.stack([rng.uniform(0.0, 1.0, (N, N, N)), rng.integers(0, 4, (N, N, N)), lab]).astype(np.float32)
        idx3 = f[1] == 3
        f[0][idx3] = rng.uniform(1.0, 1.5, idx3.sum())      # a refractive index
        f[0][f[1] == 0] *= 0.05                              # an absorption coefficient
        return f, {}
    if fmt == 103:
        b = rng.integers(0, 128, size=(4, N, N, N)).astype(np.int8)
        b[:, void] = 0
        return b, {}
    if fmt == 104:
        s = rng.integers(0, 65536, size=(2, N, N, N)).astype(np.uint16)
        s[:, void] = 0
        return s, {}
    raise AssertionError(fmt)


FORMATS = [96, 98, 99, 100, 101, 102, 103, 104]


@pytest.mark.parametrize("fmt", FORMATS)
def test_media_format_bitexact(fmt):
    rng = np.random.default_rng(fmt)
    vol, extra = _volume(fmt, rng)
    cfg = {
        "nphoton": 30000, "vol": vol, "prop": PROP, "srcpos": [12.5, 12.5, -1.0], "srcdir": [0, 0, 1],
        "issrcfrom0": 1, "tstart": 0.0, "tend": 5e-9, "tstep": 1e-9, "seed": 777, "isreflect": 1,
        "isspecular": 1, "detpos": [[12.0, 12.0, 0.0, 4.0], [2.0, 12.0, 12.0, 3.0]], "issavedet": 1,
        "savedetflag": "DPXVW", "nthread": NT, "nblocksize": 64, "autopilot": 0, "quiet": True,
    }
    cfg.update(extra)
    print("media format", fmt, flush=True)
    res = mcx_amd.run(cfg)
    ora = oracle_run(cfg, NT)
    st = res["stat"]
    tot, esc = seq_sum(ora["genergy"][1::2]), seq_sum(ora["genergy"][0::2])
    assert st["energytot"] == tot
    assert st["energyabs"] == tot - esc
    assert ora["errors"] == 0
    ref, _ = normalize_like_engine(ora["field"], st["energytot"], cfg["tstep"])
    got = res["flux"].astype(np.float64)
    sig = ref > 1e-3 * ref.max()
    assert np.abs(got[sig] - ref[sig]).max() <= 2e-5 * ref.max()
    # records: detector id + exit position/direction + initial weight
    n = ora["detected"]
    assert n > 0 and res["detp"].shape[1] == n
    assert res["detp"].shape[0] == ora["reclen"] == 1 + 3 + 3 + 1
    assert np.array_equal(sort_rows(res["detp"].T), sort_rows(ora["detp"][:n]))


def test_mua_float_reproduces_label_volume_on_gpu():
    lab = np.ones((N, N, N), dtype=np.uint8)
    lab[:2] = 0
    base = {"nphoton": 30000, "prop": PROP[:2], "srcpos": [12.5, 12.5, 0.0], "srcdir": [0, 0, 1], "issrcfrom0": 1,
            "tstart": 0.0, "tend": 5e-9, "tstep": 1e-9, "seed": 99, "isreflect": 1, "issavedet": 0,
            "nthread": NT, "nblocksize": 64, "autopilot": 0, "quiet": True}
    a = mcx_amd.run(dict(base, vol=lab))
    m = np.where(lab == 1, np.float32(0.01), np.float32(np.nan))[None].astype(np.float32)
    b = mcx_amd.run(dict(base, vol=m))
    assert a["stat"]["energyabs"] == b["stat"]["energyabs"]
    assert a["stat"]["energytot"] == b["stat"]["energytot"]
    assert np.allclose(a["flux"], b["flux"], rtol=1e-5, atol=1e-6 * a["flux"].max())


def test_split_voxels_bitexact():
    """Split voxels (97) with random interface planes and labels, some with
    the background as the lower side, partial paths per label recorded."""
    rng = np.random.default_rng(97)
    lower = rng.integers(1, 4, size=(N, N, N)).astype(np.uint8)
    upper = rng.integers(0, 4, size=(N, N, N)).astype(np.uint8)
    lower[:, :, 20:][rng.uniform(size=(N, N, 4)) < 0.3] = 0
    upper[lower == 0] = rng.integers(1, 4, size=int((lower == 0).sum()))
    lower[:2] = 0
    upper[:2] = 0
    raw = np.zeros((8, N, N, N), dtype=np.uint8)
    raw[0], raw[1] = lower, upper
    raw[2:8] = rng.integers(0, 256, size=(6, N, N, N))
    cfg = {
        "nphoton": 30000, "vol": raw.view(np.int8), "prop": PROP, "srcpos": [12.5, 12.5, -1.0], "srcdir": [0, 0, 1],
        "issrcfrom0": 1, "tstart": 0.0, "tend": 5e-9, "tstep": 1e-9, "seed": 4321, "isreflect": 1,
        "isspecular": 1, "detpos": [[12.0, 12.0, 0.0, 4.0], [2.0, 12.0, 12.0, 3.0]], "issavedet": 1,
        "savedetflag": "DSPMXVW", "nthread": NT, "nblocksize": 64, "autopilot": 0, "quiet": True,
    }
    print("split voxels", flush=True)
    res = mcx_amd.run(cfg)
    ora = oracle_run(cfg, NT)
    st = res["stat"]
    tot, esc = seq_sum(ora["genergy"][1::2]), seq_sum(ora["genergy"][0::2])
    assert st["energytot"] == tot
    assert st["energyabs"] == tot - esc
    assert ora["errors"] == 0
    ref, _ = normalize_like_engine(ora["field"], st["energytot"], cfg["tstep"])
    got = res["flux"].astype(np.float64)
    sig = ref > 1e-3 * ref.max()
    assert np.abs(got[sig] - ref[sig]).max() <= 2e-5 * ref.max()
    n = ora["detected"]
    assert n > 0 and res["detp"].shape == (ora["reclen"], n)
    assert np.array_equal(sort_rows(res["detp"].T), sort_rows(ora["detp"][:n]))
