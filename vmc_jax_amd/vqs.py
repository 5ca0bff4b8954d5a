"""Variational state (vqs.py:85-491) with the network evaluated by HIP kernels.

Shapes follow jVMC: configurations carry a leading device axis of size 1,
``(1, B, *sampleShape)`` int32, and ``__call__`` returns ``(1, B)`` (complex128
for complex nets, float64 for real ones); ``gradients`` returns
``(1, B, numRealParameters)`` complex128 in jVMC's flat layout.  Parameters live
on the GPU as one flat float64 vector in jVMC's tree_flatten order.
"""
import numpy as np
import torch

from . import _lib
from . import global_defs
from .nets import sym_wrapper
from .nets.sym_wrapper import SymNet


def _as_configs(s):
    if not torch.is_tensor(s):
        s = torch.as_tensor(np.asarray(s))
    s = s.to(device=global_defs.device(), dtype=torch.int32)
    return s.contiguous()


class NQS:
    def __init__(self, net, logarithmic=True, batchSize=1000, seed=1234, orbit=None, avgFun=None, precision="fp64"):
        """``precision`` (extension, not in jVMC): "fp64" evaluates the network in
        FP64 as the reference does; "mixed" runs the CNN's Metropolis and
        local-energy window kernels in FP32 (log psi changes and O_loc summed in
        FP64) and returns the compact gradients (``layout=0``) in FP32, which
        the MinSR solve turns into the compensated-BF16 Gram; networks without
        an FP32 form keep FP64."""
        if precision not in ("fp64", "mixed"):
            raise ValueError("precision must be 'fp64' or 'mixed'")
        self.precision = precision
        if isinstance(net, (tuple, list)):
            raise NotImplementedError("two-network ansatz (TwoNets) is not implemented")
        if not logarithmic:
            raise NotImplementedError("only logarithmic networks are implemented")
        if orbit is not None:
            # vqs.py:165-166: the network is wrapped in a SymNet
            net = SymNet(orbit=orbit, net=net,
                         avgFun=avgFun if avgFun is not None else sym_wrapper.avgFun_Coefficients_Exp)
        self.net = net
        self.logarithmic = True
        self.batchSize = batchSize
        self.seed = seed
        self.initialized = False
        self._isGenerator = False
        self.desc = None
        self._params = None
        self._sym = None  # (JvOrbit, device map, device factors) of a SymNet

    # ------------------------------------------------------------ init
    def init_net(self, s):
        """vqs.py:187-218: lattice shape from the first configuration seen."""
        if self.initialized:
            return
        shape = tuple(int(x) for x in s.shape[2:])
        self.sampleShape = shape
        self.desc = self.net.descriptor(shape)
        self.desc.precision = _lib.JV_PREC_MIXED if self.precision == "mixed" else _lib.JV_PREC_FP64
        _lib.check(_lib.lib.jv_net_init(self.desc), "jv_net_init")
        self.holomorphic = bool(self.desc.is_complex)
        self.realParams = not self.holomorphic
        self._cpx_out = self.holomorphic
        self._cpx_weights = self.holomorphic
        if isinstance(self.net, SymNet):
            self._init_sym()
        rng = np.random.default_rng(self.seed)
        P = self.net.init_params(self.desc, rng)
        self._params = torch.as_tensor(P, dtype=torch.float64, device=global_defs.device())
        self.initialized = True

    def _init_sym(self):
        """Device form of the orbit; output and gradient types follow jax's
        promotion in SymNet.__call__ (sym_wrapper.py:66-77) and the
        Cauchy-Riemann test of vqs.py:199-210."""
        orb = self.net.orbit
        if orb.shape[1] != self.desc.num_sites:
            raise ValueError("the orbit acts on %d sites, the lattice has %d" % (orb.shape[1], self.desc.num_sites))
        dev = global_defs.device()
        fac = np.asarray(orb.factor)
        cfac = bool(np.iscomplexobj(fac))
        m = torch.as_tensor(orb.device_map(), device=dev).contiguous()
        f = torch.as_tensor(fac.astype(np.complex128), device=dev).contiguous()
        o = _lib.JvOrbit()
        o.num_elements = int(m.shape[0])
        o.num_sites = int(m.shape[1])
        o.avg = self.net.avg_code
        o.complex_factor = int(cfac)
        o.map = m.data_ptr()
        o.factor = f.data_ptr()
        self._sym = (o, m, f)
        avg = self.net.avg_code
        if avg == _lib.JV_AVG_SEP:
            # Sep is not holomorphic; its value is complex for every network
            self.holomorphic = False
            self._cpx_out = True
        elif avg == _lib.JV_AVG_EXP and cfac:
            self._cpx_out = True
            self._cpx_weights = True

    @property
    def numParameters(self):
        return int(self.desc.num_params)

    @property
    def numRealParameters(self):
        return int(self.desc.num_real_params)

    @property
    def is_generator(self):
        return False

    @property
    def params(self):
        return self._params

    @params.setter
    def params(self, P):
        self._params = P

    @property
    def parameters(self):
        return self._params

    # ------------------------------------------------------------ evaluation
    def _flat(self, s):
        s = _as_configs(s)
        self.init_net(s)
        return s, s.reshape(-1, self.desc.num_sites)

    def logpsi_complex(self, s):
        """log psi as complex128 (B,) for flat configurations (B, L)."""
        s = _as_configs(s)
        n = s.shape[0]
        out = torch.empty(n, dtype=torch.complex128, device=s.device)
        if self._sym is not None:
            _lib.call("jv_sym_forward", self.desc, _lib.ptr(self._params), self._sym[0], _lib.ptr(s), n,
                      _lib.ptr(out), _lib.stream_ptr())
        else:
            _lib.call("jv_net_forward", self.desc, _lib.ptr(self._params), _lib.ptr(s), n, _lib.ptr(out),
                      _lib.stream_ptr())
        return out

    def _out(self, lp):
        """Complex log psi -> the network's output dtype (real for real outputs)."""
        return lp if self._cpx_out else lp.real.contiguous()

    def __call__(self, s):
        """vqs.py:223-251."""
        s, flat = self._flat(s)
        return self._out(self.logpsi_complex(flat).reshape(s.shape[:2]))

    def gradients(self, s, layout=1):
        """vqs.py:271-287: O_k = d log psi / d theta_k, jVMC flat layout (layout=1)
        or the compact per-parameter layout (layout=0)."""
        s, flat = self._flat(s)
        if self._sym is not None and not self.holomorphic and not self.realParams:
            raise NotImplementedError("gradients of a non-holomorphic complex-parameter network "
                                      "(avgFun_Coefficients_Sep over a complex network) are not implemented")
        n = flat.shape[0]
        if layout == 0 and self.fp32_gradients:
            g = torch.empty((n, self.numParameters), dtype=torch.float32, device=flat.device)
            _lib.call("jv_net_gradients_f32", self.desc, _lib.ptr(self._params), _lib.ptr(flat), n, _lib.ptr(g), None,
                      _lib.stream_ptr())
            return g.reshape(s.shape[:2] + g.shape[1:])
        if layout == 1:
            g = torch.empty((n, self.numRealParameters), dtype=torch.complex128, device=flat.device)
        elif self._cpx_weights:
            g = torch.empty((n, self.numParameters), dtype=torch.complex128, device=flat.device)
        else:
            g = torch.empty((n, self.numParameters), dtype=torch.float64, device=flat.device)
        if self._sym is not None:
            _lib.call("jv_sym_gradients", self.desc, _lib.ptr(self._params), self._sym[0], _lib.ptr(flat), n, layout,
                      _lib.ptr(g), None, _lib.stream_ptr())
        else:
            _lib.call("jv_net_gradients", self.desc, _lib.ptr(self._params), _lib.ptr(flat), n, layout,
                      _lib.ptr(g), None, _lib.stream_ptr())
        return g.reshape(s.shape[:2] + g.shape[1:])

    @property
    def fp32_gradients(self):
        """The mixed form of a CNN returns compact gradients in FP32."""
        return (self.precision == "mixed" and self._sym is None and self.desc is not None
                and self.desc.kind == _lib.JV_NET_CNN)

    def get_sampler_net(self):
        """vqs.py:337-352: Re log psi and the current parameters."""
        return (lambda p, x: self.__call__(x).real), self._params

    # ------------------------------------------------------------ parameters
    def _to_dev(self, P):
        if torch.is_tensor(P):
            return P.to(global_defs.device())
        return torch.as_tensor(np.asarray(P), device=global_defs.device())

    def _param_unflatten(self, P):
        """vqs.py:430-444 for a flat vector that may be complex (the holomorphic
        MinSR update is complex, minsr.py:65): leaf = P[Re slot] + i P[Im slot]."""
        P = self._to_dev(P)
        if not torch.is_complex(P):
            return P.to(torch.float64).contiguous()
        if not self.holomorphic:
            return P.real.to(torch.float64).contiguous()
        out = torch.empty(self.numRealParameters, dtype=torch.float64, device=P.device)
        for off, size in self._complex_leaves():
            re = P[off:off + size]
            im = P[off + size:off + 2 * size]
            leaf = re + 1j * im
            out[off:off + size] = leaf.real
            out[off + size:off + 2 * size] = leaf.imag
        return out

    def _complex_leaves(self):
        d = self.desc
        leaves = []
        if d.off_hidden_bias >= 0:
            leaves.append((d.off_hidden_bias, d.num_hidden))
        leaves.append((d.off_kernel, d.num_sites * d.num_hidden))
        if d.off_visible_bias >= 0:
            leaves.append((d.off_visible_bias, d.num_sites))
        return leaves

    def holomorphic_columns(self):
        """Device int64 maps (col_g, col_ig) of a holomorphic network: flat
        gradient column col_ig[c] is i times column col_g[c] ([g, i g] per
        complex leaf, vqs.py:66-69)."""
        if getattr(self, "_holo_cols", None) is None:
            g, ig = [], []
            for off, size in self._complex_leaves():
                g.append(np.arange(off, off + size))
                ig.append(np.arange(off + size, off + 2 * size))
            dev = global_defs.device()
            self._holo_cols = (torch.as_tensor(np.concatenate(g), dtype=torch.int64, device=dev).contiguous(),
                               torch.as_tensor(np.concatenate(ig), dtype=torch.int64, device=dev).contiguous())
        return self._holo_cols

    def set_parameters(self, P):
        """vqs.py:409-427."""
        if not self.initialized:
            raise RuntimeError("Error in NQS.set_parameters(): Network not initialized. Evaluate net on example "
                               "input for initialization.")
        P = self._param_unflatten(P)
        if P.numel() != self.numRealParameters:
            raise ValueError("expected %d parameters, got %d" % (self.numRealParameters, P.numel()))
        self._params = P.clone()

    def update_parameters(self, deltaP):
        """vqs.py:384-406."""
        if not self.initialized:
            self.set_parameters(deltaP)
            return
        self._params = self._params + self._param_unflatten(deltaP)

    def get_parameters(self):
        """vqs.py:449-466: flat real vector (a device tensor)."""
        if not self.initialized:
            return None
        return self._params.clone()

    def expand_compact(self, v):
        """Compact complex update (one entry per complex parameter) -> jVMC flat
        layout [v_leaf, -i v_leaf] per leaf (the holomorphic MinSR update of
        minsr.py:65 in the [O, iO] gradient layout of vqs.py:66-69)."""
        out = torch.empty(self.numRealParameters, dtype=torch.complex128, device=v.device)
        c = 0
        for off, size in self._complex_leaves():
            out[off:off + size] = v[c:c + size]
            out[off + size:off + 2 * size] = -1j * v[c:c + size]
            c += size
        return out
