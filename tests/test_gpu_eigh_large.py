"""Complex Hermitian eigendecomposition past the LDS limit of the four-vector
complex Householder reduction (n > 2496, where v_{k-1} and w_{k-1} are rebuilt
from global memory in the column pass, up to n = 3276): eigenvalues against
LAPACK (numpy eigvalsh) at n = 2600; at the variant's largest size, and at
n = 4096 (the T of a 4096-sample complex MinSR step, rocSOLVER zheevd), through
properties that fix an eigendecomposition whatever its size: unitary
eigenvectors, the residual A V = V diag(ev), ascending order, the trace and the
Frobenius norm.  Then the complex MinSR T form (minsr.py:53-80) whose T is
2600 x 2600 with a cut spectrum, against the pseudo-inverse of numpy's eigh."""
import ctypes

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

jv = pytest.importorskip("vmc_jax_amd")


def _hermitian(n, seed):
    rng = np.random.default_rng(seed)
    A = rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))
    return (A + A.conj().T) / np.sqrt(2 * n)


def _eigh(A):
    n = A.shape[0]
    Ad = torch.as_tensor(A, device="cuda").contiguous()
    w = torch.empty(n, dtype=torch.float64, device="cuda")
    jv._lib.call("jv_eigh", n, 1, jv._lib.ptr(Ad), jv._lib.ptr(w), jv._lib.stream_ptr())
    return Ad, w


@pytest.mark.parametrize("n", [2600, 3276, 4096])
def test_complex_eigh_large(n):
    print("complex eigh n", n, flush=True)
    A = _hermitian(n, n)
    V, w = _eigh(A)
    Ag = torch.as_tensor(A, device="cuda")
    nrm = torch.linalg.norm(Ag).item()
    # residual and orthogonality to a few eps n^0.5 (backward stability)
    res = torch.linalg.norm(Ag @ V - V * w[None, :].to(torch.complex128)).item() / nrm
    VhV = V.conj().T.contiguous() @ V
    orth = torch.linalg.norm(VhV - torch.eye(n, dtype=torch.complex128, device="cuda")).item() / np.sqrt(n)
    ev = w.cpu().numpy()
    print("residual", res, "orthogonality", orth, flush=True)
    assert res < 1e-12 and orth < 1e-13
    assert np.all(np.diff(ev) >= 0)
    assert abs(ev.sum() - np.trace(A).real) <= 1e-11 * np.abs(ev).sum()
    assert abs(np.sum(ev ** 2) - nrm ** 2) <= 1e-11 * nrm ** 2
    if n <= 2600:
        ref = np.linalg.eigvalsh(A)
        assert np.abs(ev - ref).max() <= 1e-12 * np.abs(ref).max()


def test_complex_minsr_t_form_cut_large():
    """T = D D^H (N = 2600 samples, 1200 complex parameters) has rank 1200: the
    pseudo-inverse cuts the rest, so the eigendecomposition of T (the large-n
    reduction) runs; the update matches numpy's pseudo-inverse."""
    rng = np.random.default_rng(3)
    n, p = 2600, 1200
    D = (rng.normal(size=(n, p)) + 1j * rng.normal(size=(n, p))) / np.sqrt(p)
    e = rng.normal(size=n) + 1j * rng.normal(size=n)
    print("complex T form n", n, "p", p, flush=True)
    Dd = torch.as_tensor(D, device="cuda").contiguous()
    ed = torch.as_tensor(e, device="cuda").contiguous()
    upd = torch.empty(p, dtype=torch.complex128, device="cuda")
    info = jv._lib.JvSolveInfo()
    jv._lib.call("jv_minsr_solve", jv._lib.ptr(Dd), jv._lib.ptr(ed), n, p, 1, 1e-10, 0.0, 0,
                 jv._lib.ptr(upd), ctypes.byref(info), jv._lib.stream_ptr())
    u = upd.cpu().numpy()
    T = D @ D.conj().T
    wr, Vr = np.linalg.eigh(T)
    keep = np.abs(wr) > 1e-10 * np.abs(wr).max()
    inv = np.where(keep, 1.0 / np.where(keep, wr, 1.0), 0.0)
    ref = -D.conj().T @ ((Vr * inv) @ (Vr.conj().T @ e))
    assert info.rank == int(keep.sum()) == p
    assert np.linalg.norm(u - ref) / np.linalg.norm(ref) < 1e-8
