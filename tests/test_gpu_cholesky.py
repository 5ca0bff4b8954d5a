"""The Cholesky route of the real dual MinSR solve (jv_minsr_solve_dual) on
synthetic centred data, against the pseudo-inverse of numpy's eigh (the
algorithm of minsr.py:68-78 in its dual form): the route is taken exactly when
no eigenvalue of S = D^T D (+ shift) is at or below pinvTol * lambda_max, and
then gives the same update; otherwise the eigendecomposition runs.  Sizes
cover one partial panel, panel boundaries (63, 64, 65, 128, 129) and the 10x10
CNN's parameter count."""
import ctypes

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

jv = pytest.importorskip("vmc_jax_amd")


def _solve(D, e, rcond, shift=0.0):
    n, p = D.shape
    Dd = torch.as_tensor(D, device="cuda").contiguous()
    ed = torch.as_tensor(e.astype(np.complex128), device="cuda").contiguous()
    upd = torch.empty(p, dtype=torch.float64, device="cuda")
    info = jv._lib.JvSolveInfo()
    jv._lib.call("jv_minsr_solve_dual", jv._lib.ptr(Dd), jv._lib.ptr(ed), n, p, 0, float(rcond), float(shift), 1,
                 jv._lib.ptr(upd), ctypes.byref(info), jv._lib.stream_ptr())
    return upd.cpu().numpy(), info


def _reference(D, e, rcond, shift=0.0):
    S = D.T @ D + shift * np.eye(D.shape[1])
    w, V = np.linalg.eigh(S)
    keep = np.abs(w) > rcond * np.abs(w).max()
    inv = np.where(keep, 1.0 / np.where(keep, w, 1.0), 0.0)
    return -(V * inv) @ (V.T @ (D.T @ e.real)), int(keep.sum())


def _rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


@pytest.mark.parametrize("p", [1, 2, 63, 64, 65, 128, 129, 300, 1240, 2000])
def test_full_rank_takes_cholesky(p):
    rng = np.random.default_rng(p)
    n = p + 64 + p // 2
    D = rng.normal(size=(n, p)) / np.sqrt(n)
    e = rng.normal(size=n) + 1j * rng.normal(size=n)
    print("cholesky route p", p, flush=True)
    u, info = _solve(D, e, 1e-10)
    ref, rank = _reference(D, e, 1e-10)
    assert info.iterations == 1 and info.rank == p == rank
    assert _rel(u, ref) < 1e-11


@pytest.mark.parametrize("shift", [1e-3, 1.0])
def test_shift_takes_cholesky(shift):
    rng = np.random.default_rng(7)
    n, p = 300, 500  # N < P: singular D^T D, regular after the shift
    D = rng.normal(size=(n, p)) / np.sqrt(n)
    e = rng.normal(size=n) + 0j
    u, info = _solve(D, e, 1e-10, shift)
    ref, rank = _reference(D, e, 1e-10, shift)
    assert info.iterations == 1 and info.rank == p == rank
    assert _rel(u, ref) < 1e-10


@pytest.mark.parametrize("p,n", [(200, 150), (1240, 1000)])
def test_rank_deficient_uses_eigendecomposition(p, n):
    rng = np.random.default_rng(p + n)
    D = rng.normal(size=(n, p)) / np.sqrt(n)
    e = rng.normal(size=n) + 0j
    u, info = _solve(D, e, 1e-10)
    ref, rank = _reference(D, e, 1e-10)
    assert info.iterations == 0 and info.rank == rank == n
    assert _rel(u, ref) < 1e-8


def test_cut_spectrum_uses_eigendecomposition():
    """Singular values 1 .. 1e-7 (eigenvalues of S down to 1e-14): pinvTol 1e-10
    cuts the smallest, so the eigendecomposition runs.  Singular values 1 ..
    1e-5 (eigenvalues down to 1e-10) with pinvTol 1e-12: nothing is cut and the
    route's margin (8 p eps ||S||_F, 5e-13 here) admits the Cholesky solve."""
    rng = np.random.default_rng(11)
    n, p = 400, 256
    U, _ = np.linalg.qr(rng.normal(size=(n, p)))
    V, _ = np.linalg.qr(rng.normal(size=(p, p)))
    e = rng.normal(size=n) + 0j
    D = (U * np.logspace(0, -7, p)) @ V.T
    u, info = _solve(D, e, 1e-10)
    ref, rank = _reference(D, e, 1e-10)
    assert info.iterations == 0 and info.rank == rank < p
    assert _rel(u, ref) < 1e-6
    D = (U * np.logspace(0, -5, p)) @ V.T
    u, info = _solve(D, e, 1e-12)
    ref, rank = _reference(D, e, 1e-12)
    assert info.iterations == 1 and rank == p
    assert _rel(u, ref) < 1e-5  # condition number 1e10: both accurate to ~ cond * eps


def _solve_t(D, e, rcond, shift=0.0):
    n, p = D.shape
    cpx = np.iscomplexobj(D)
    Dd = torch.as_tensor(D, device="cuda").contiguous()
    ed = torch.as_tensor(e.astype(np.complex128), device="cuda").contiguous()
    upd = torch.empty(p, dtype=torch.complex128 if cpx else torch.float64, device="cuda")
    info = jv._lib.JvSolveInfo()
    jv._lib.call("jv_minsr_solve", jv._lib.ptr(Dd), jv._lib.ptr(ed), n, p, int(cpx), float(rcond), float(shift), 1,
                 jv._lib.ptr(upd), ctypes.byref(info), jv._lib.stream_ptr())
    return upd.cpu().numpy(), info


def _reference_t(D, e, rcond, shift=0.0):
    """minsr.py:53-80 in the T form: -D^dagger pinv(T) e, T = D D^dagger (+ shift
    for real data, where only Re e couples)."""
    cpx = np.iscomplexobj(D)
    T = D @ D.conj().T + (0 if cpx else shift) * np.eye(D.shape[0])
    w, V = np.linalg.eigh(T)
    keep = np.abs(w) > rcond * np.abs(w).max()
    inv = np.where(keep, 1.0 / np.where(keep, w, 1.0), 0.0)
    rhs = e if cpx else e.real
    return -D.conj().T @ ((V * inv) @ (V.conj().T @ rhs)), int(keep.sum())


@pytest.mark.parametrize("cpx,n,p", [(False, 300, 700), (False, 64, 65), (True, 200, 500), (True, 1100, 2000)])
def test_t_form_takes_cholesky(cpx, n, p):
    """The reference's own MinSR form with fewer samples than parameters: T is
    full rank, the route (complex T through its real 2n x 2n form) is taken."""
    rng = np.random.default_rng(n + p)
    D = rng.normal(size=(n, p)) + (1j * rng.normal(size=(n, p)) if cpx else 0)
    D /= np.sqrt(p)
    e = rng.normal(size=n) + 1j * rng.normal(size=n)
    print("T-form cholesky route", cpx, n, p, flush=True)
    u, info = _solve_t(D, e, 1e-10)
    ref, rank = _reference_t(D, e, 1e-10)
    assert info.iterations == 1 and info.rank == n == rank
    assert _rel(u, ref) < 1e-10


def test_t_form_singular_uses_eigendecomposition():
    rng = np.random.default_rng(5)
    n, p = 400, 150  # T = D D^T has rank 150
    D = rng.normal(size=(n, p)) / np.sqrt(p)
    e = rng.normal(size=n) + 0j
    u, info = _solve_t(D, e, 1e-10)
    ref, rank = _reference_t(D, e, 1e-10)
    assert info.iterations == 0 and info.rank == rank == p
    assert _rel(u, ref) < 1e-8


def _centred(O, w):
    """jv_center's layout: sqrt(w_n) (O_n - sum_m w_m O_m) (stats.py:197-205)."""
    return np.sqrt(w)[:, None] * (O - (w[:, None] * O).sum(0))


@pytest.mark.parametrize("cpx,n,p", [(False, 300, 700), (True, 200, 500), (True, 1100, 2000)])
def test_t_form_centred_deflates_null_vector(cpx, n, p):
    """Centred data with uniform weights: T = D D^dagger has the null vector
    1/sqrt(n), whose eigenvalue the pseudo-inverse cuts.  The route deflates it
    (rank n - 1) and gives the eigendecomposition's update."""
    rng = np.random.default_rng(n + p + 1)
    O = (rng.normal(size=(n, p)) + (1j * rng.normal(size=(n, p)) if cpx else 0)) / np.sqrt(p)
    D = _centred(O, np.full(n, 1.0 / n))
    e = rng.normal(size=n) + 1j * rng.normal(size=n)
    print("centred T form", cpx, n, p, flush=True)
    u, info = _solve_t(D, e, 1e-10)
    ref, rank = _reference_t(D, e, 1e-10)
    assert rank == n - 1
    assert info.iterations == 1 and info.rank == n - 1
    assert _rel(u, ref) < 1e-10


@pytest.mark.parametrize("case", ["weights", "rank"])
def test_t_form_centred_not_deflatable(case):
    """Non-uniform weights (the null vector is sqrt(w), not the constant: no
    deflation), or fewer parameters than samples (more than one eigenvalue
    cut): the eigendecomposition runs and gives the pseudo-inverse update."""
    rng = np.random.default_rng(9)
    n, p = (300, 700) if case == "weights" else (400, 150)
    O = (rng.normal(size=(n, p)) + 1j * rng.normal(size=(n, p))) / np.sqrt(p)
    w = rng.uniform(0.5, 1.5, n) if case == "weights" else np.ones(n)
    D = _centred(O, w / w.sum())
    e = rng.normal(size=n) + 1j * rng.normal(size=n)
    u, info = _solve_t(D, e, 1e-10)
    ref, rank = _reference_t(D, e, 1e-10)
    assert info.iterations == 0 and info.rank == rank == (n - 1 if case == "weights" else p)
    assert _rel(u, ref) < 1e-8


def test_cholesky_route_concurrent_streams():
    """Eight host threads, each on its own stream, run the route at once (the
    pattern of the batched field sweep): every workgroup of a panel launch must
    read the diagonal block's input values however late it starts, so each
    result is bit-identical to the same solve run alone."""
    import threading
    rng = np.random.default_rng(13)
    n, p = 1100, 2000
    cases = []
    for _ in range(8):
        O = (rng.normal(size=(n, p)) + 1j * rng.normal(size=(n, p))) / np.sqrt(p)
        cases.append((_centred(O, np.full(n, 1.0 / n)), rng.normal(size=n) + 1j * rng.normal(size=n)))
    alone = [_solve_t(D, e, 1e-10) for D, e in cases]
    results, errors = [None] * len(cases), []

    def work(k):
        try:
            s = torch.cuda.Stream()
            with torch.cuda.stream(s):
                results[k] = _solve_t(*cases[k], 1e-10)
            s.synchronize()
        except Exception as exc:  # reported below
            errors.append(repr(exc))

    threads = [threading.Thread(target=work, args=(k,)) for k in range(len(cases))]
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    assert not errors, errors
    for (u, info), (ua, infoa) in zip(results, alone):
        assert info.iterations == infoa.iterations == 1 and info.rank == n - 1
        assert np.array_equal(u, ua)


def _with_spectrum(rng, n, lam, centred):
    """D [n][p] (p = len(lam) + extra columns of zeros) whose T = D D^T has the
    eigenvalues lam on an orthonormal basis; centred: the basis is orthogonal
    to u0 = 1/sqrt(n), so u0 is an exact null vector."""
    k = len(lam)
    M = rng.normal(size=(n, k))
    if centred:
        M -= M.mean(0)
    U, _ = np.linalg.qr(M)
    V, _ = np.linalg.qr(rng.normal(size=(k + 40, k)))
    return (U * np.sqrt(lam)) @ V.T


@pytest.mark.parametrize("factor", [0.5, 1.01, 2.0, 10.0, 1e3])
def test_t_form_near_cutoff_matches_pinv(factor):
    """The smallest kept eigenvalue of the centred T placed at `factor` times the
    cutoff pinvTol * lambda_max: the solve either takes the route (and keeps
    exactly the reference's eigenvalues) or diagonalises; both give the
    pseudo-inverse update."""
    rng = np.random.default_rng(int(factor * 100))
    n, rcond = 200, 1e-8
    lam = np.logspace(0, -4, n - 1)
    lam[-1] = factor * rcond * lam[0]
    D = _with_spectrum(rng, n, lam, True)
    e = rng.normal(size=n) + 1j * rng.normal(size=n)
    u, info = _solve_t(D, e, rcond)
    ref, rank = _reference_t(D, e, rcond)
    print("near cutoff", factor, info.iterations, info.rank, rank, flush=True)
    assert info.rank == rank
    assert _rel(u, ref) < 1e-5
    if factor <= 1.01:
        assert info.iterations == 0  # a cut (or marginal) eigenvalue: no certificate


@pytest.mark.parametrize("eta", [1e-9, 1e-6])
def test_t_form_perturbed_null_vector_not_deflated(eta):
    """Data whose rows do not sum to zero exactly (||T u0|| far above rounding):
    u0 is no eigenvector to working precision, the deflation is refused and
    the eigendecomposition gives the pseudo-inverse update."""
    rng = np.random.default_rng(int(1 / eta) % 1000)
    n, rcond = 200, 1e-8
    lam = np.logspace(0, -3, n - 1)
    D = _with_spectrum(rng, n, lam, True)
    D = D + eta * rng.normal(size=(1, D.shape[1]))  # the same small offset on every row
    e = rng.normal(size=n) + 1j * rng.normal(size=n)
    u, info = _solve_t(D, e, rcond)
    ref, rank = _reference_t(D, e, rcond)
    assert info.iterations == 0 and info.rank == rank
    assert _rel(u, ref) < 1e-6


@pytest.mark.parametrize("factor", [0.5, 1.01, 2.0, 1e3])
def test_dual_near_cutoff_matches_pinv(factor):
    """The dual form with the smallest eigenvalue of S = D^T D at `factor` times
    the cutoff: the route is taken only with a certificate; the update is the
    pseudo-inverse's either way."""
    rng = np.random.default_rng(int(factor * 10) + 3)
    n, p, rcond = 600, 250, 1e-8
    U, _ = np.linalg.qr(rng.normal(size=(n, p)))
    V, _ = np.linalg.qr(rng.normal(size=(p, p)))
    sv = np.sqrt(np.logspace(0, -4, p))
    sv[-1] = np.sqrt(factor * rcond)
    D = (U * sv) @ V.T
    e = rng.normal(size=n) + 0j
    u, info = _solve(D, e, rcond)
    ref, rank = _reference(D, e, rcond)
    assert info.rank == rank
    assert _rel(u, ref) < 1e-5
    if factor <= 1.01:
        assert info.iterations == 0
