"""jv_stedc (tridiagonal divide and conquer on the device) against LAPACK's
tridiagonal eigensolver (scipy.linalg.eigh_tridiagonal) on matrices that
exercise every branch: leaves only (n <= 32), uneven splits, merges that
deflate everything (zero or tiny couplings), Givens deflation of close poles
(clusters, Wilkinson matrices, glued blocks), graded spectra spanning 18
decades (the spectrum of an NQS Gram matrix) and the largest size."""
import numpy as np
import pytest
import scipy.linalg
import torch

pytestmark = pytest.mark.gpu

jv = pytest.importorskip("vmc_jax_amd")


def _graded(n, rng):
    ev = np.logspace(-16, 2, n)
    Q, _ = np.linalg.qr(rng.normal(size=(n, n)))
    H = scipy.linalg.hessenberg((Q * ev) @ Q.T)
    return np.diag(H).copy(), np.diag(H, 1).copy()


def _wilkinson(m):
    return np.abs(np.arange(-m, m + 1)).astype(float), np.ones(2 * m)


def _glued(m, copies, glue):
    w, e = _wilkinson(m)
    d = np.tile(w, copies)
    ee = np.ones(len(d) - 1)
    ee[2 * m::2 * m + 1] = glue
    return d, ee


def _case(name, n):
    rng = np.random.default_rng(n + len(name))
    if name == "random":
        return rng.normal(size=n), rng.normal(size=max(n - 1, 0))
    if name == "graded":
        return _graded(n, rng)
    if name == "clusters":
        return np.repeat(rng.normal(size=4), n // 4) + 1e-12 * rng.normal(size=n // 4 * 4), \
            1e-3 * rng.normal(size=n // 4 * 4 - 1)
    if name == "near_diagonal":
        return np.ones(n), np.full(n - 1, 1e-9)
    if name == "zero_coupling":
        return rng.normal(size=n), np.zeros(n - 1)
    if name == "wilkinson":
        return _wilkinson(n // 2)
    if name == "glued":
        return _glued(50, n // 101, 1e-10)
    if name == "d0e1":
        return np.zeros(n), np.ones(n - 1)
    raise ValueError(name)


CASES = [("random", 1), ("random", 2), ("random", 31), ("random", 32), ("random", 33), ("random", 65),
         ("random", 100), ("random", 257), ("random", 1240), ("graded", 300), ("graded", 1240),
         ("clusters", 400), ("near_diagonal", 300), ("zero_coupling", 256), ("wilkinson", 201),
         ("glued", 404), ("d0e1", 500), ("random", 2496), ("random", 3276)]


def _run(d, e):
    n = len(d)
    dd = torch.as_tensor(d, device="cuda").contiguous()
    ed = torch.as_tensor(e if n > 1 else np.zeros(1), device="cuda").contiguous()
    z = torch.empty((n, n), dtype=torch.float64, device="cuda")
    jv._lib.call("jv_stedc", n, jv._lib.ptr(dd), jv._lib.ptr(ed), jv._lib.ptr(z), jv._lib.stream_ptr())
    torch.cuda.synchronize()
    return dd.cpu().numpy(), z.cpu().numpy(), ed.cpu().numpy()


@pytest.mark.parametrize("name,n", CASES)
def test_stedc_matches_lapack(name, n):
    d, e = _case(name, n)
    n = len(d)
    print("jv_stedc", name, n, flush=True)
    w, Z, e_after = _run(d, e)
    if n > 1:
        assert np.array_equal(e_after, e), "off-diagonal modified"
    wr = scipy.linalg.eigh_tridiagonal(d, e, eigvals_only=True) if n > 1 else d.copy()
    T = np.diag(d) + (np.diag(e, 1) + np.diag(e, -1) if n > 1 else 0)
    nrm = max(np.abs(wr).max(), 1e-300)
    V = Z.T  # rows of z are eigenvectors
    assert np.all(np.diff(w) >= 0)
    # backward-stable bounds: eigenvalues and residuals to a few eps ||T||, orthogonality to a few eps n^0.5
    assert np.abs(w - wr).max() <= 1e-13 * nrm, np.abs(w - wr).max() / nrm
    assert np.abs(T @ V - V * w).max() <= 1e-13 * nrm
    assert np.abs(V.T @ V - np.eye(n)).max() <= 2e-14 * max(np.sqrt(n), 1.0)


def test_stedc_sizes_in_sequence():
    """Consecutive calls with a larger, then a smaller size reuse the workspace."""
    for n in (700, 1240, 129, 1240):
        d, e = _case("random", n)
        w, Z, _ = _run(d, e)
        assert np.abs(w - scipy.linalg.eigh_tridiagonal(d, e, eigvals_only=True)).max() <= 1e-13 * np.abs(w).max()


def test_stedc_rejects_bad_sizes():
    d = torch.zeros(4000, dtype=torch.float64, device="cuda")
    z = torch.zeros((1, 1), dtype=torch.float64, device="cuda")
    with pytest.raises(jv._lib.HipError, match="3276"):
        jv._lib.call("jv_stedc", 4000, jv._lib.ptr(d), jv._lib.ptr(d), jv._lib.ptr(z), jv._lib.stream_ptr())
    with pytest.raises(jv._lib.HipError):
        jv._lib.call("jv_stedc", -1, jv._lib.ptr(d), jv._lib.ptr(d), jv._lib.ptr(z), jv._lib.stream_ptr())
