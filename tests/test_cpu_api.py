"""CPU-side checks of the boundary: the C-ABI library loads, exports every
symbol declared in include/jvmc_hip.h, and its host-only layout logic agrees
with the oracle (no GPU compute calls here)."""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

from oracle import nets as onets
from oracle import operator as oop

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "jvmc_hip.h")
LIB = os.path.join(ROOT, "vmc_jax_amd", "libjvmc_hip.so")


def declared_symbols():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:int|const char\*)\s+(jv_\w+)\s*\(", text, flags=re.M)))


@pytest.fixture(scope="module")
def lib():
    if not os.path.exists(LIB):
        pytest.fail("libjvmc_hip.so is not built; run __graft_entry__.build()")
    from vmc_jax_amd import _lib
    return _lib


def test_library_exports_every_declared_symbol(lib):
    syms = declared_symbols()
    assert len(syms) >= 20
    out = subprocess.run(["nm", "-D", "--defined-only", LIB], capture_output=True, text=True, check=True).stdout
    exported = set(re.findall(r"\bT (jv_\w+)", out))
    missing = [s for s in syms if s not in exported]
    assert not missing, missing
    for s in syms:
        assert hasattr(lib.lib, s)
    # the Python binding declares exactly the header's entry points
    assert sorted(lib.exported_symbols()) == syms


def test_no_torch_types_in_header():
    text = open(HEADER).read()
    assert "torch" not in text.lower() and "at::" not in text


def _init(lib, d):
    rc = lib.lib.jv_net_init(ctypes.byref(d))
    assert rc == 0, lib.lib.jv_last_error()
    return d


def test_cnn_layout_matches_oracle(lib):
    from vmc_jax_amd import nets
    for shape, F, ch in [((10, 10), (3, 3), (8, 8, 8)), ((6,), (3,), (3, 2)),
                         ((4, 4), (2, 2), tuple([2] * 11 if False else [2] * 8))]:
        net = nets.CNN(F=F, channels=ch, strides=(1,) * len(shape))
        d = _init(lib, net.descriptor(shape))
        o = onets.CNN(shape, F=F, channels=ch, strides=(1,) * len(shape))
        assert d.num_real_params == o.num_real_params()
        off = 0
        for leaf in o.leaves:
            layer = int(leaf.name.split("/")[0].split("_")[1])
            if leaf.name.endswith("kernel"):
                assert d.off_conv_kernel[layer] == off, leaf.name
            else:
                assert d.off_conv_bias[layer] == off, leaf.name
            off += leaf.size


def test_rbm_layout_matches_oracle(lib):
    from vmc_jax_amd import nets
    for bias, vb in [(False, False), (True, False), (True, True)]:
        d = _init(lib, nets.CpxRBM(numHidden=20, bias=bias, visibleBias=vb).descriptor((10,)))
        o = onets.CpxRBM((10,), numHidden=20, bias=bias, visibleBias=vb)
        assert d.num_real_params == o.num_real_params()
        assert d.num_params == o.num_params()
    d = _init(lib, nets.CpxRBM(numHidden=20, bias=True, visibleBias=True).descriptor((10,)))
    assert (d.off_hidden_bias, d.off_kernel, d.off_visible_bias) == (0, 40, 440)
    assert d.num_real_params == 460  # C1: 230 complex parameters


def test_init_rejects_bad_descriptors(lib):
    from vmc_jax_amd import _lib as L
    d = L.JvNet()
    d.kind = 99
    d.ndim = 1
    d.shape[0] = 4
    assert lib.lib.jv_net_init(ctypes.byref(d)) == -1
    assert b"unknown" in lib.lib.jv_last_error()
    d = L.JvNet()
    d.kind = L.JV_NET_CNN
    d.ndim = 2
    d.shape[0] = d.shape[1] = 64
    d.num_layers = 3
    d.filter[0] = d.filter[1] = 3
    for l in range(3):
        d.channels[l] = 64
        d.layer_bias[l] = 1
    assert lib.lib.jv_net_init(ctypes.byref(d)) == -4  # exceeds one workgroup's LDS


def test_operator_tables_match_oracle():
    from vmc_jax_amd import operator as dop
    L = 5
    H, Ho = dop.BranchFreeOperator(), oop.BranchFreeOperator()
    for i in range(L):
        H.add(dop.scal_opstr(2., (dop.Sx(i),)))
        H.add(dop.scal_opstr(2., (dop.Sy(i), dop.Sz((i + 1) % L))))
        H.add((dop.Sz(i),))
        Ho.add(oop.scal_opstr(2., (oop.Sx(i),)))
        Ho.add(oop.scal_opstr(2., (oop.Sy(i), oop.Sz((i + 1) % L))))
        Ho.add((oop.Sz(i),))
    H.compile()
    Ho.compile()
    np.testing.assert_array_equal(H.idxH, Ho.idxC)
    np.testing.assert_array_equal(H.mapH, Ho.mapC)
    np.testing.assert_array_equal(H.matElsH, Ho.matElsC)
    np.testing.assert_array_equal(H.diagH, Ho.diag)
    f = lambda t: 3.0 * t
    H2 = dop.BranchFreeOperator()
    H2.add(dop.scal_opstr(f, (dop.Sp(0),)))
    H2.add(dop.scal_opstr(0.5, (dop.Sx(1),)))
    H2.compile()
    res = np.ones(2, np.complex128)
    for i, fn in H2.prefactor:
        res[i] = fn(2.0)
    np.testing.assert_allclose(res, [6.0, 0.5])


def test_distribute_sampling_semantics():
    from vmc_jax_amd import mpi_wrapper as mpi
    assert mpi.distribute_sampling(8192, 1, 512) == (16, 8192)
    assert mpi.globNumSamples == 8192
    assert mpi.distribute_sampling(1000, 1, 300) == (4, 1200)
    assert mpi.distribute_sampling(1024) == (1024, 1024)


def test_ctypes_signatures_match_header_arity(lib):
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    for name, params in re.findall(r"(?:int|const char\*)\s+(jv_\w+)\s*\(([^)]*)\)\s*;", text):
        params = params.strip()
        n = 0 if params in ("", "void") else params.count(",") + 1
        if name == "jv_last_error":
            continue
        assert len(lib._SIGS[name]) == n, "%s: header has %d parameters, binding %d" % (name, n, len(lib._SIGS[name]))


def test_sym_entry_points_reject_bad_orbits(lib):
    """Host-side validation of jv_sym_* (no device memory is touched)."""
    from vmc_jax_amd import _lib as L
    d = L.JvNet()
    d.kind = L.JV_NET_CPXRBM
    d.ndim = 1
    d.shape[0] = 6
    d.shape[1] = 1
    d.num_hidden = 4
    assert lib.lib.jv_net_init(ctypes.byref(d)) == 0
    dummy = ctypes.c_void_p(1)  # never dereferenced: validation fails first
    o = L.JvOrbit()
    o.num_elements = 3
    o.num_sites = 5  # the network has 6 sites
    o.avg = L.JV_AVG_EXP
    o.map = dummy
    o.factor = dummy
    assert lib.lib.jv_sym_forward(ctypes.byref(d), dummy, ctypes.byref(o), dummy, 4, dummy, None) == -1
    assert b"different number of sites" in lib.lib.jv_last_error()
    o.num_sites = 6
    o.avg = 7
    assert lib.lib.jv_sym_gradients(ctypes.byref(d), dummy, ctypes.byref(o), dummy, 4, 1, dummy, None, None) == -1
    o.avg = L.JV_AVG_SEP
    # Sep over a complex-parameter network is not holomorphic: no gradients
    assert lib.lib.jv_sym_gradients(ctypes.byref(d), dummy, ctypes.byref(o), dummy, 4, 1, dummy, None, None) == -1
    assert b"non-holomorphic" in lib.lib.jv_last_error()
    o.num_elements = 0
    assert lib.lib.jv_sym_forward(ctypes.byref(d), dummy, ctypes.byref(o), dummy, 4, dummy, None) == -1
    assert lib.lib.jv_sym_mc_sample(ctypes.byref(d), dummy, None, None, dummy, dummy, dummy, dummy, None) == -1


def test_holomorphic_tdvp_rejects_bad_column_maps(lib):
    dummy = ctypes.c_void_p(1)
    x = (ctypes.c_double * 2)(1.0, 0.0)
    info = lib.JvSolveInfo()
    args = (dummy, dummy, dummy, 10, 8)
    tail = (x, 0, 0.0, 0, 10.0, 2.0, 1e-8, 1e-8, dummy, dummy, dummy, dummy, dummy, ctypes.byref(info), None)
    # pr must be 2 * pc
    assert lib.lib.jv_tdvp_solve_holomorphic(*args, dummy, dummy, 3, *tail) == -1
    assert b"split pr = 2 pc" in lib.lib.jv_last_error()
    assert lib.lib.jv_tdvp_solve_holomorphic(*args, None, dummy, 4, *tail) == -1


def test_fermionic_table_host_side():
    """The compiled host tables carry the fermionic mask in factor order."""
    from vmc_jax_amd import operator as dop
    H = dop.BranchFreeOperator()
    H.add(dop.scal_opstr(1.0, (dop.creation(2), dop.annihilation(0))))
    H.add(dop.scal_opstr(0.5, (dop.number(1),)))
    H.compile()
    # factors are stored right to left: annihilation(0) first
    np.testing.assert_array_equal(H.fermionicH, [[1, 1], [0, 0]])
    np.testing.assert_array_equal(H.idxH, [[0, 2], [1, 0]])
    assert H.max_sites_changed == 2
    Ho = oop.BranchFreeOperator()
    Ho.add(oop.scal_opstr(1.0, (oop.creation(2), oop.annihilation(0))))
    Ho.add(oop.scal_opstr(0.5, (oop.number(1),)))
    Ho.compile()
    np.testing.assert_array_equal(Ho.fermiC, H.fermionicH)


def test_minsr_cholesky_retry_policy(lib):
    """Host-side route policy of the dual MinSR solve: a refused Cholesky
    attempt skips the next CHOLESKY_RETRY - 1 attempts; a skipped solve whose
    spectrum kept full rank re-enables the attempt; a taken route changes
    nothing."""
    from vmc_jax_amd.util.minsr import MinSR
    eq = MinSR(None, pinvTol=1e-8)

    def solved(tried, iterations, rank, full):
        eq.last_info = lib.JvSolveInfo()
        eq.last_info.iterations = iterations
        eq.last_info.rank = rank
        eq._note_route(tried, full, eq.last_info)

    assert eq._cholesky_hint() == 1
    solved(1, 0, 1100, 1240)          # refused: the spectrum was cut
    hints = [eq._cholesky_hint() for _ in range(MinSR.CHOLESKY_RETRY - 1)]
    assert hints == [0] * (MinSR.CHOLESKY_RETRY - 1) and eq._cholesky_hint() == 1
    solved(1, 0, 1100, 1240)
    assert eq._cholesky_hint() == 0
    solved(0, 0, 1240, 1240)          # skipped, nothing cut
    assert eq._cholesky_hint() == 1
    solved(1, 1, 1240, 1240)          # route taken
    assert eq._cholesky_hint() == 1
    solved(1, 0, 1237, 1240)          # refused again
    assert eq._cholesky_hint() == 0
    solved(0, 0, 1230, 1240)          # skipped and still cut: keep skipping
    assert eq._chol_skip == MinSR.CHOLESKY_RETRY - 2


def test_minsr_deferred_outcome_is_read_lazily(lib):
    """The overlapped solve's outcome (route, rank, E_loc mean and variance)
    is read only when needed: the route bookkeeping runs when the next route
    choice or last_info asks for it, the energy when ElocMean0 / ElocVar0 is
    read, each once; a failed factorisation raises when read."""
    from vmc_jax_amd.util.minsr import MinSR

    class Stub:  # stands for the pinned record of a non-blocking solve
        def __init__(self, hint, full, iterations, rank, e):
            self.hint, self.full, self.reads = hint, full, 0
            self.info = lib.JvSolveInfo()
            self.info.iterations, self.info.rank = iterations, rank
            self.info.e_mean[0], self.info.e_mean[1], self.info.e_var = e, 0.5, 2.0

        def resolve(self):
            self.reads += 1
            if self.info.iterations < 0:
                raise lib.HipError("factorisation failed")
            return self.info

    eq = MinSR(None, pinvTol=1e-8)
    st = Stub(1, 1240, 0, 1100, -3.25)     # a refused attempt, energy -3.25
    eq._pending = eq._energy_src = st
    assert st.reads == 0 and eq.last_outcome() is st
    assert eq._cholesky_hint() == 0         # the refusal is seen by the next choice
    assert eq._chol_skip == MinSR.CHOLESKY_RETRY - 2 and st.reads == 1
    assert eq.get_energy_mean() == -3.25 and eq.ElocMean0 == complex(-3.25, 0.5) and eq.ElocVar0 == 2.0
    assert eq.last_info.rank == 1100
    st2 = Stub(0, 1240, 0, 1240, 1.0)
    eq._pending = eq._energy_src = st2
    eq.ElocMean0 = 7.0                      # a later value replaces it after the read
    assert eq.ElocMean0 == 7.0 and eq._pending is None and eq._chol_skip == 0
    bad = Stub(1, 1240, -2, 0, 0.0)
    eq._pending = bad
    with pytest.raises(lib.HipError):
        eq.last_info
