"""Lattice-symmetry orbits and the symmetrised network (CPU).

The product's map-based orbit construction is compared element by element with
the oracle's dense-matrix restatement of symmetries.py; the oracle SymNet is
pinned by the reference's own assertions (orbit sizes and dtypes,
symmetries_test.py:17-46; translation invariance, nets_test.py:48-64; symmetry
sectors, symmetries_test.py:49-124), by group closure, and by central finite
differences of its gradients.
"""
import itertools
import warnings

import numpy as np
import pytest
import torch

from oracle import nets as onets
from oracle import symmetries as osym

from vmc_jax_amd.util import symmetries as sym
from vmc_jax_amd.nets import sym_wrapper as sw

warnings.simplefilter("ignore", UserWarning)

NAMES1 = ("reflection", "translation", "spinflip")
NAMES2 = ("rotation", "reflection", "translation", "spinflip")


def subsets(names):
    for k in range(len(names) + 1):
        for c in itertools.combinations(names, k):
            yield c


# ---------------------------------------------------------------- orbits
def test_orbit_sizes_1d():
    """symmetries_test.py:30-44."""
    L = 3
    for symms in subsets(NAMES1):
        o = sym.get_orbit_1D(L, *symms)
        exp = (2 if "reflection" in symms else 1) * (L if "translation" in symms else 1) * \
              (2 if "spinflip" in symms else 1)
        assert o.orbit.shape[0] == exp
        assert np.issubdtype(o.orbit.dtype, np.integer)


def test_orbit_dtype_2d():
    """symmetries_test.py:17-28."""
    for symms in subsets(NAMES2):
        assert np.issubdtype(sym.get_orbit_2D_square(3, *symms).orbit.dtype, np.integer)


def test_full_2d_group_order():
    """The square lattice's space group with spin flip has 8 L^2 * 2 elements."""
    for L in (3, 4, 5):
        o = sym.get_orbit_2D_square(L, *NAMES2)
        assert o.orbit.shape[0] == 16 * L * L


@pytest.mark.parametrize("L", [3, 4, 5, 6])
def test_orbits_1d_match_dense_construction(L):
    facs = dict(translation_factor=np.exp(1j * 2 * np.pi / L), reflection_factor=-1., spinflip_factor=-1.)
    for symms in subsets(NAMES1):
        for kw in ({}, facs):
            o = sym.get_orbit_1D(L, *symms, **kw)
            d, f = osym.orbit_1d(L, *symms, **kw)
            np.testing.assert_array_equal(o.orbit, d)
            assert o.factor.dtype == f.dtype
            np.testing.assert_array_equal(o.factor, f)


@pytest.mark.parametrize("L", [2, 3, 4])
def test_orbits_2d_match_dense_construction(L):
    facs = dict(translation_factor=np.exp(1j * np.pi / L), reflection_factor=-1., rotation_factor=1j,
                spinflip_factor=-1.)
    for symms in subsets(NAMES2):
        for kw in ({}, facs):
            o = sym.get_orbit_2D_square(L, *symms, **kw)
            d, f = osym.orbit_2d_square(L, *symms, **kw)
            np.testing.assert_array_equal(o.orbit, d)
            assert o.factor.dtype == f.dtype
            np.testing.assert_array_equal(o.factor, f)


def test_group_closure_and_identity():
    """Every full orbit is a group: it holds the identity and the composition of
    any two elements (as signed permutation matrices)."""
    for o in (sym.get_orbit_1D(6, *NAMES1), sym.get_orbit_2D_square(4, *NAMES2)):
        d = o.orbit.astype(np.int64)
        keys = {m.tobytes() for m in d}
        assert np.eye(d.shape[1], dtype=np.int64).tobytes() in keys
        for a in d[::3]:
            for b in d[::5]:
                assert (a @ b).tobytes() in keys


def test_dense_orbit_roundtrip():
    o = sym.get_orbit_2D_square(3, "rotation", "spinflip")
    o2 = sym.LatticeSymmetry(o.orbit, o.factor)
    np.testing.assert_array_equal(o2.device_map(), o.device_map())
    with pytest.raises(NotImplementedError):
        sym.LatticeSymmetry(2 * o.orbit, o.factor)


# ---------------------------------------------------------------- oracle SymNet
def _symnet(onet, orbit, factor, avg):
    return osym.SymNet(onet, orbit, factor, avg)


def test_oracle_symnet_translation_invariance():
    """nets_test.py:48-64: log psi is equal on all translates of a configuration."""
    L = 5
    rng = np.random.default_rng(0)
    onet = onets.RBM((L,), numHidden=5, bias=True)
    d, f = osym.orbit_1d(L, "translation")
    net = _symnet(onet, d, f, "exp")
    P = 0.3 * rng.normal(size=onet.num_real_params())
    s0 = np.array([1, 0, 1, 1, 0])
    S = np.array([np.roll(s0, -i) for i in range(L)])
    psi = onets.NQS(net, P)(S)
    assert np.max(np.abs(psi - psi[0])) < 1e-12


def test_oracle_symmetry_sectors_1d():
    """symmetries_test.py:49-83: with factor -1 (translation: exp(i pi)) the
    coefficient changes sign under the symmetry.  A CpxRBM with visible bias
    stands in for CpxRBM_Nospinflip (both lack the spin-flip symmetry)."""
    L = 10
    rng = np.random.default_rng(1)
    onet = onets.CpxRBM((L,), numHidden=2, bias=False, visibleBias=True)
    P = 0.3 * rng.normal(size=onet.num_real_params())
    c1 = rng.permutation(np.array([0, 1] * (L // 2)))
    for name, c2 in (("reflection", c1[::-1]), ("translation", np.roll(c1, 1)), ("spinflip", 1 - c1)):
        d, f = osym.orbit_1d(L, name, reflection_factor=-1., translation_factor=np.exp(1j * np.pi),
                             spinflip_factor=-1.)
        psi = onets.NQS(_symnet(onet, d, f, "exp"), P)
        a, b = psi(c1[None])[0], psi(c2[None])[0]
        assert abs(1 + np.exp(a - b)) < 1e-10, name


def test_oracle_symmetry_sectors_2d():
    """symmetries_test.py:85-124."""
    L = 4
    rng = np.random.default_rng(2)
    onet = onets.CpxRBM((L, L), numHidden=2, bias=False, visibleBias=True)
    P = 0.3 * rng.normal(size=onet.num_real_params())
    c1 = rng.permutation(np.array([0, 1] * (L * L // 2))).reshape(L, L)
    cases = (("rotation", c1[:, ::-1].T), ("reflection", c1[::-1, :]), ("translation", np.roll(c1, 1, axis=0)),
             ("spinflip", 1 - c1))
    for name, c2 in cases:
        d, f = osym.orbit_2d_square(L, name, rotation_factor=-1., reflection_factor=-1.,
                                    translation_factor=np.exp(1j * np.pi), spinflip_factor=-1.)
        psi = onets.NQS(_symnet(onet, d, f, "exp"), P)
        a, b = psi(c1.reshape(1, -1))[0], psi(c2.reshape(1, -1))[0]
        assert abs(1 + np.exp(a - b)) < 1e-10, name


def test_oracle_identity_orbit_is_the_network():
    L = 7
    rng = np.random.default_rng(3)
    onet = onets.CpxRBM((L,), numHidden=3, bias=True)
    P = 0.4 * rng.normal(size=onet.num_real_params())
    d, f = osym.orbit_1d(L)
    s = rng.integers(0, 2, size=(20, L))
    a = onets.NQS(_symnet(onet, d, f, "exp"), P)(s)
    b = onets.NQS(onet, P)(s)
    assert np.max(np.abs(a.real - b.real)) < 1e-14
    assert np.max(np.abs(np.exp(1j * a.imag) - np.exp(1j * b.imag))) < 1e-14
    np.testing.assert_allclose(onets.NQS(_symnet(onet, d, f, "exp"), P).gradients(s),
                               onets.NQS(onet, P).gradients(s), rtol=1e-14)


def test_oracle_real_exp_negative_sum_is_nan():
    """jax logsumexp with real b: log|sum| and NaN where the sum is negative."""
    v = osym.avg_exp(np.array([[0.0, 1.0], [1.0, 0.0]]), np.array([1.0, -1.0]))
    assert np.isnan(v[0]) and abs(v[1] - np.log(np.e - 1)) < 1e-15


def _fd_case(kind):
    rng = np.random.default_rng(4)
    if kind == "cpxrbm_exp":
        L = 6
        onet = onets.CpxRBM((L,), numHidden=3, bias=True)
        d, f = osym.orbit_1d(L, "translation", "reflection", translation_factor=np.exp(2j * np.pi / L))
        avg, shape = "exp", (L,)
    elif kind == "rbm_exp_cpx_factor":
        L = 6
        onet = onets.RBM((L,), numHidden=4, bias=True)
        d, f = osym.orbit_1d(L, "translation", "spinflip", translation_factor=np.exp(2j * np.pi / L))
        avg, shape = "exp", (L,)
    elif kind == "cnn_log":
        onet = onets.CNN((3, 3), F=(2, 2), channels=(2, 2), strides=(1, 1))
        d, f = osym.orbit_2d_square(3, "rotation", "reflection")
        avg, shape = "log", (3, 3)
    else:  # rbm_sep
        L = 5
        onet = onets.RBM((L,), numHidden=3, bias=True)
        d, f = osym.orbit_1d(L, "translation", "spinflip")
        avg, shape = "sep", (L,)
    P = 0.3 * rng.normal(size=onet.num_real_params())
    s = rng.integers(0, 2, size=(12,) + shape)
    net = _symnet(onet, d, f, avg)
    # drop configurations whose symmetrised amplitude (nearly) cancels: there
    # log psi is ill-conditioned and differences are meaningless
    lp = onets.NQS(net, P)(s)
    keep = np.real(lp) > np.max(np.real(lp)) - 4
    return net, P, s[keep][:4]


@pytest.mark.parametrize("kind", ["cpxrbm_exp", "rbm_exp_cpx_factor", "cnn_log", "rbm_sep"])
def test_oracle_symnet_gradients_finite_differences(kind):
    """Central differences of log psi along each real parameter: holomorphic
    nets give O_k on the Re-slots (and i O_k on the Im-slots); real-parameter
    nets give d Re f + i d Im f (vqs.py:40-69)."""
    net, P, s = _fd_case(kind)
    psi = onets.NQS(net, P)
    G = psi.gradients(s)
    h = 1e-6
    fd = np.zeros_like(G)
    for k in range(len(P)):
        e = np.zeros_like(P)
        e[k] = h
        psi.set_parameters(P + e)
        a = psi(s)
        psi.set_parameters(P - e)
        b = psi(s)
        fd[:, k] = (a - b) / (2 * h)
    psi.set_parameters(P)
    if net.holomorphic:
        # the flat layout stores [g, i g]; derivatives along Re-slots are g
        npr = len(P)
        re_cols = np.concatenate([np.arange(o, o + l.size) for o, l in
                                  zip(np.cumsum([0] + [2 * l.size for l in net.leaves])[:-1], net.leaves)])
        assert np.max(np.abs(G[:, re_cols] - fd[:, re_cols])) < 1e-7
        assert npr == G.shape[1]
    else:
        assert np.max(np.abs(G - fd)) < 1e-7


# ---------------------------------------------------------------- product avgFun functions
def test_product_avg_functions_match_oracle():
    rng = np.random.default_rng(5)
    c = rng.normal(size=(6, 8)) + 1j * rng.normal(size=(6, 8))
    cr = rng.normal(size=(6, 8))
    b = np.exp(2j * np.pi * rng.uniform(size=8))
    br = np.where(rng.uniform(size=8) < 0.5, 1.0, 0.0)
    br[0] = 1.0
    for x, bb in ((c, b), (cr, b), (c, br), (cr, br)):
        np.testing.assert_allclose(sw.avgFun_Coefficients_Exp(torch.as_tensor(x), torch.as_tensor(bb)).numpy(),
                                   osym.avg_exp(x, bb), rtol=1e-13)
    np.testing.assert_allclose(sw.avgFun_Coefficients_Log(torch.as_tensor(c), None).numpy(), osym.avg_log(c),
                               rtol=1e-13)
    np.testing.assert_allclose(sw.avgFun_Coefficients_Sep(torch.as_tensor(c), None).numpy(), osym.avg_sep(c),
                               rtol=1e-13)
