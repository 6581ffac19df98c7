"""The allocation-free bias gradient (mcx_dn_bias_grad_noalloc) against the
workspace form (mcx_dn_bias_grad) through the C ABI: the same kernels in
the same order, so the results must be identical.  Two calls in a row, the
second on a smaller input, check that the fixed scratch array carries
nothing from one call to the next; sizes beyond the array are refused."""
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


def _p(t):
    return ctypes.c_void_p(t.data_ptr())


def _both(shape, cout, nsplit, seed, masked):
    rng = np.random.default_rng(seed)
    gy = _dev_bf16(dn_ref.bf16_round(rng.standard_normal(shape + (cout,))))
    mask = _dev_bf16(dn_ref.bf16_round(rng.standard_normal(shape + (cout,)))) if masked else None
    nv = int(np.prod(shape))
    part = torch.zeros(nsplit * cout, dtype=torch.float32, device="cuda")
    ref = torch.zeros(cout, dtype=torch.float32, device="cuda")
    got = torch.full((cout,), 9.0, dtype=torch.float32, device="cuda")
    mp = _p(mask) if masked else None
    assert L.mcx_dn_bias_grad(_p(gy), cout, mp, cout, _p(part), nsplit, _p(ref), nv, cout, None) == 0
    assert L.mcx_dn_bias_grad_noalloc(_p(gy), cout, mp, cout, nsplit, _p(got), nv, cout, None) == 0
    torch.cuda.synchronize()
    return ref.cpu().numpy(), got.cpu().numpy()


def test_noalloc_bias_grad_equals_workspace_form():
    # a large call filling most of the scratch array, then a smaller one
    ref, got = _both((2, 16, 16, 24), 64, 1000, 1, True)
    assert np.array_equal(ref, got)
    ref, got = _both((1, 6, 10, 12), 16, 7, 2, False)
    assert np.array_equal(ref, got)
    assert np.abs(ref).max() > 0


def test_noalloc_bias_grad_refuses_beyond_scratch():
    gy = torch.zeros(8 * 128, dtype=torch.int16, device="cuda")
    db = torch.zeros(128, dtype=torch.float32, device="cuda")
    assert L.mcx_dn_bias_grad_noalloc(_p(gy), 128, None, 128, 513, _p(db), 8, 128, None) == -7
    assert L.mcx_dn_bias_grad_noalloc(_p(gy), 128, None, 128, 512, _p(db), 8, 128, None) == 0
    torch.cuda.synchronize()
    assert np.all(db.cpu().numpy() == 0)
