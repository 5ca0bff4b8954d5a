"""jv_center_noalloc (fixed per-device scratch, no allocation) against
jv_center through the C ABI: two calls in a row, the second smaller, real and
complex data; results must be bit-identical (same kernels, same reduction
order).  Sizes beyond the fixed scratch are refused on the host."""
import ctypes

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

_lib = pytest.importorskip("vmc_jax_amd._lib")


def _center(name, O, w, cpx):
    # O holds n rows of p values (complex: 2p interleaved doubles)
    n, p = w.shape[0], O.shape[1] // (2 if cpx else 1)
    assert O.shape == (n, p * (2 if cpx else 1)) and O.is_contiguous()
    mean = torch.empty((p * (2 if cpx else 1),), dtype=torch.float64, device="cuda")
    data = torch.empty_like(O)
    _lib.call(name, _lib.ptr(O), _lib.ptr(w), n, p, int(cpx), _lib.ptr(mean), _lib.ptr(data), _lib.stream_ptr())
    return mean, data


def _inputs(rng, n, p, cpx):
    O = rng.normal(size=(n, p * (2 if cpx else 1)))
    w = rng.uniform(0.5, 1.5, size=n)
    w /= w.sum()
    return (torch.as_tensor(O, device="cuda").contiguous(), torch.as_tensor(w, device="cuda").contiguous())


@pytest.mark.parametrize("cpx", [False, True])
def test_center_noalloc_matches_center(cpx):
    rng = np.random.default_rng(7 + cpx)
    for (n, p) in ((8192, 1240), (1000, 333)):  # second call smaller
        O, w = _inputs(rng, n, p, cpx)
        m_ref, d_ref = _center("jv_center", O, w, cpx)
        m, d = _center("jv_center_noalloc", O, w, cpx)
        torch.cuda.synchronize()
        assert torch.equal(m, m_ref)
        assert torch.equal(d, d_ref)
        # the centred data have zero weighted mean (stats.py:197-205)
        dn = d.cpu().numpy()
        wn = w.cpu().numpy()
        assert np.abs((np.sqrt(wn)[:, None] * dn).sum(0)).max() < 1e-12


def test_center_noalloc_refuses_oversize():
    n, p = 256 * 65, 2 * 1024 * 1024 // 64  # 65 chunks x 32768 columns > 2^21 partial sums
    dummy = torch.empty((1,), dtype=torch.float64, device="cuda")
    rc = _lib.lib.jv_center_noalloc(_lib.ptr(dummy), _lib.ptr(dummy), n, p, 0, _lib.ptr(dummy), _lib.ptr(dummy),
                                    _lib.stream_ptr())
    assert rc != 0
    assert b"fixed scratch" in _lib.lib.jv_last_error()
