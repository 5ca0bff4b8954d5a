"""ctypes binding of libjvmc_hip.so (the gfx950 HIP implementation).

The shared library is built in-tree by ``vmc_jax_amd/csrc/build.sh`` (or
``__graft_entry__.build()``).  There is no fallback: every compute entry point
of this package goes through the library, and importing it fails loudly when
the library is missing.  torch is imported first so that the HIP runtime the
library binds to is the one torch already loaded.
"""
import ctypes
import os

import torch  # noqa: F401  (loads the HIP runtime first)

_HERE = os.path.dirname(os.path.abspath(__file__))
# VMC_JAX_AMD_LIB overrides the in-tree library (e.g. to compare builds)
LIB_PATH = os.environ.get("VMC_JAX_AMD_LIB", os.path.join(_HERE, "libjvmc_hip.so"))

JV_NET_CPXRBM = 1
JV_NET_RBM = 2
JV_NET_CNN = 3
JV_MAX_LAYERS = 8
JV_PREC_FP64 = 0
JV_PREC_MIXED = 1

_i32 = ctypes.c_int32
_i64 = ctypes.c_int64
_dbl = ctypes.c_double
_vp = ctypes.c_void_p


class JvNet(ctypes.Structure):
    _fields_ = [
        ("kind", _i32), ("ndim", _i32), ("shape", _i32 * 2), ("num_hidden", _i32),
        ("hidden_bias", _i32), ("visible_bias", _i32), ("num_layers", _i32), ("filter", _i32 * 2),
        ("channels", _i32 * JV_MAX_LAYERS), ("layer_bias", _i32 * JV_MAX_LAYERS), ("precision", _i32),
        ("num_sites", _i32), ("is_complex", _i32), ("num_params", _i64), ("num_real_params", _i64),
        ("off_hidden_bias", _i64), ("off_kernel", _i64), ("off_visible_bias", _i64),
        ("off_conv_kernel", _i64 * JV_MAX_LAYERS), ("off_conv_bias", _i64 * JV_MAX_LAYERS),
    ]


class JvOpTable(ctypes.Structure):
    _fields_ = [("num_ops", _i32), ("max_len", _i32), ("l_dim", _i32), ("num_diag", _i32),
                ("max_sites_changed", _i32), ("reserved", _i32), ("idx", _vp), ("map", _vp), ("matel", _vp), ("diag", _vp),
                ("fermi", _vp)]


JV_AVG_EXP = 0
JV_AVG_LOG = 1
JV_AVG_SEP = 2


class JvOrbit(ctypes.Structure):
    _fields_ = [("num_elements", _i32), ("num_sites", _i32), ("avg", _i32), ("complex_factor", _i32),
                ("map", _vp), ("factor", _vp)]


class JvMcArgs(ctypes.Structure):
    _fields_ = [("num_chains", _i32), ("therm_steps", _i32), ("samples_per_chain", _i32),
                ("sweep_steps", _i32), ("mu", _dbl), ("seed", ctypes.c_uint64),
                ("call_index", ctypes.c_uint32), ("stream", ctypes.c_uint32)]


class JvSolveInfo(ctypes.Structure):
    _fields_ = [("residual", _dbl), ("pinv_cutoff", _dbl), ("iterations", _i32), ("rank", _i32),
                ("e_var", _dbl), ("tdvp_error", _dbl), ("e_mean", _dbl * 2)]


JV_BATCH_SAMPLE = 0
JV_BATCH_GRAD = 1
JV_BATCH_ELOC = 2


class JvBatchItem(ctypes.Structure):
    _fields_ = [("net", JvNet), ("params", _vp), ("op", JvOpTable), ("prefactor", _vp), ("mc", JvMcArgs),
                ("walkers", _vp), ("configs", _vp), ("logpsi", _vp), ("counters", _vp), ("eloc", _vp),
                ("grad", _vp), ("grad_layout", _i32), ("reserved", _i32)]


# name -> argtypes; every function returns int status
_SIGS = {
    "jv_version": [],
    "jv_set_device": [ctypes.c_int],
    "jv_init": [ctypes.c_int],
    "jv_shutdown": [],
    "jv_workspace_bytes": [_vp],
    "jv_timer_enable": [ctypes.c_int],
    "jv_timer_reset": [],
    "jv_timer_read": [ctypes.c_char_p, ctypes.POINTER(_dbl), ctypes.POINTER(_i64)],
    "jv_net_init": [ctypes.POINTER(JvNet)],
    "jv_net_forward": [ctypes.POINTER(JvNet), _vp, _vp, _i64, _vp, _vp],
    "jv_net_gradients": [ctypes.POINTER(JvNet), _vp, _vp, _i64, _i32, _vp, _vp, _vp],
    "jv_net_gradients_f32": [ctypes.POINTER(JvNet), _vp, _vp, _i64, _vp, _vp, _vp],
    "jv_sym_forward": [ctypes.POINTER(JvNet), _vp, ctypes.POINTER(JvOrbit), _vp, _i64, _vp, _vp],
    "jv_sym_gradients": [ctypes.POINTER(JvNet), _vp, ctypes.POINTER(JvOrbit), _vp, _i64, _i32, _vp, _vp, _vp],
    "jv_sym_eloc": [ctypes.POINTER(JvNet), _vp, ctypes.POINTER(JvOrbit), ctypes.POINTER(JvOpTable), _vp, _vp, _vp,
                    _i64, _vp, _vp],
    "jv_sym_mc_sample": [ctypes.POINTER(JvNet), _vp, ctypes.POINTER(JvOrbit), ctypes.POINTER(JvMcArgs), _vp, _vp,
                         _vp, _vp, _vp],
    "jv_get_s_primes": [ctypes.POINTER(JvOpTable), _vp, _vp, _i64, _i32, _vp, _vp, _vp, _vp],
    "jv_eloc": [ctypes.POINTER(JvNet), _vp, ctypes.POINTER(JvOpTable), _vp, _vp, _vp, _i64, _vp, _vp],
    "jv_mc_sample": [ctypes.POINTER(JvNet), _vp, ctypes.POINTER(JvMcArgs), _vp, _vp, _vp, _vp, _vp],
    "jv_exact_weights": [_vp, _i64, _dbl, _dbl, _vp, ctypes.POINTER(_dbl), _vp],
    "jv_mc_weights": [_vp, _i64, _dbl, _dbl, _vp, _vp],
    "jv_center": [_vp, _vp, _i64, _i64, _i32, _vp, _vp, _vp],
    "jv_center_f32": [_vp, _vp, _i64, _i64, _i32, _vp, _vp, _vp],
    "jv_center_noalloc": [_vp, _vp, _i64, _i64, _i32, _vp, _vp, _vp],
    "jv_col_norm2": [_vp, _i64, _i64, _i32, _vp, _vp],
    "jv_gram_nn": [_vp, _i64, _i64, _i32, _vp, _vp],
    "jv_gram_pp": [_vp, _i64, _i64, _i32, _vp, _vp],
    "jv_gram_pp_mixed": [_vp, _i32, _i64, _i64, _i32, _vp, _vp],
    "jv_gemm": [_i64, _i64, _i64, _vp, _i64, _i64, _i32, _vp, _i64, _i64, _i32, _vp, _i64, _i32, _dbl, _vp],
    "jv_gemv": [_i32, _vp, _vp, _i64, _i64, _i32, _i32, _dbl, _vp, _vp],
    "jv_eigh": [_i64, _i32, _vp, _vp, _vp],
    "jv_stedc": [_i64, _vp, _vp, _vp, _vp],
    "jv_minsr_solve": [_vp, _vp, _i64, _i64, _i32, _dbl, _dbl, _i32, _vp, ctypes.POINTER(JvSolveInfo), _vp],
    "jv_minsr_solve_dual": [_vp, _vp, _i64, _i64, _i32, _dbl, _dbl, _i32, _vp, ctypes.POINTER(JvSolveInfo), _vp],
    "jv_minsr_solve_dual_deferred": [_vp, _vp, _vp, _vp, _i64, _i64, _i32, _dbl, _dbl, _i32, _vp,
                                     ctypes.POINTER(JvSolveInfo),
                                     _vp],
    "jv_minsr_solve_dual_deferred_mixed": [_vp, _vp, _vp, _vp, _i64, _i64, _i32, _dbl, _dbl, _i32, _vp,
                                           ctypes.POINTER(JvSolveInfo), _vp],
    "jv_minsr_dual_gram": [_vp, _i32, _i64, _i64, _i32, _dbl, _vp, _vp],
    "jv_minsr_solve_dual_deferred_gram": [_vp, _i32, _vp, _vp, _vp, _vp, _i64, _i64, _i32, _dbl, _dbl, _i32, _vp,
                                          ctypes.POINTER(JvSolveInfo), _vp],
    "jv_minsr_solve_dual_deferred_gram_async": [_vp, _i32, _vp, _vp, _vp, _vp, _i64, _i64, _i32, _dbl, _dbl, _i32,
                                                _vp, _vp, _vp],
    "jv_tdvp_solve": [_vp, _vp, _vp, _i64, _i64, ctypes.POINTER(_dbl), _i32, _dbl, _i32, _dbl, _dbl, _dbl,
                      _dbl, _vp, _vp, _vp, _vp, _vp, ctypes.POINTER(JvSolveInfo), _vp],
    "jv_tdvp_solve_holomorphic": [_vp, _vp, _vp, _i64, _i64, _vp, _vp, _i64, ctypes.POINTER(_dbl), _i32, _dbl, _i32,
                                  _dbl, _dbl, _dbl, _dbl, _vp, _vp, _vp, _vp, _vp, ctypes.POINTER(JvSolveInfo), _vp],
    "jv_tdvp_solve_holomorphic_mixed": [_vp, _vp, _vp, _i64, _i64, _vp, _vp, _i64, ctypes.POINTER(_dbl), _i32, _dbl,
                                        _i32, _dbl, _dbl, _dbl, _dbl, _vp, _vp, _vp, _vp, _vp,
                                        ctypes.POINTER(JvSolveInfo), _vp],
    "jv_subset": [_vp, _vp, _vp, _i64, _i64, _i32, _i64, _i64, _i64, _vp, _vp, _vp, _vp],
    "jv_batch_create": [ctypes.POINTER(JvBatchItem), _i32, ctypes.POINTER(_vp)],
    "jv_batch_destroy": [_vp],
    "jv_batch_set_shape": [_vp, _i32, _i32, _i32],
    "jv_batch_phase": [_vp, _i32, _vp],
    "jv_batch_step": [_vp, _vp],
    "jv_graph_capture": [_vp, _vp, ctypes.POINTER(_vp)],
    "jv_graph_launch": [_vp, _vp],
    "jv_graph_destroy": [_vp],
    "jv_tdvp_cross_validate": [_vp, _vp, _i64, _i64, ctypes.POINTER(_dbl), _i32, _dbl, _vp, ctypes.POINTER(_dbl),
                               _vp],
}


class HipError(RuntimeError):
    pass


def _load():
    if not os.path.exists(LIB_PATH):
        raise ImportError("libjvmc_hip.so is not built (%s); run vmc_jax_amd/csrc/build.sh or "
                          "__graft_entry__.build()" % LIB_PATH)
    lib = ctypes.CDLL(LIB_PATH)
    for name, args in _SIGS.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = ctypes.c_int
    lib.jv_last_error.argtypes = []
    lib.jv_last_error.restype = ctypes.c_char_p
    return lib


lib = _load()


def check(rc, what=""):
    if rc != 0:
        msg = lib.jv_last_error().decode(errors="replace")
        raise HipError("%s failed (%d): %s" % (what, rc, msg))


def call(name, *args):
    check(getattr(lib, name)(*args), name)


def ptr(t):
    """Device pointer of a torch tensor (None -> NULL)."""
    if t is None:
        return None
    if not t.is_contiguous():
        raise ValueError("tensor must be contiguous")
    return ctypes.c_void_p(t.data_ptr())


def stream_ptr():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def exported_symbols():
    return sorted(list(_SIGS.keys()) + ["jv_last_error"])
