"""Networks: log-cosh, complex / real RBM and the real CNN, with analytic O_k.

Oracle (test infrastructure only).  Follows
* ``log_cosh``           activation_functions.py:19-22
* ``CpxRBM``             rbm.py:17-38 (plus an optional visible bias, an
                         extension: jVMC's CpxRBM has none)
* ``RBM``                rbm.py:67-88
* ``CNN``                cnn.py:18-81 (periodic boundary, strides 1)
* parameter flattening   vqs.py:430-466 (JAX ``tree_flatten`` order: dict keys
                         sorted as strings; complex leaves as [Re, Im])
* gradient layouts       vqs.py:46-69 (holomorphic: [g, i*g] per leaf;
                         real parameters: grad Re f + i grad Im f)
"""
import numpy as np

LOG2 = np.log(2.0)


def log_cosh(x):
    """Stable log(cosh(x)) with the sign taken from the sign bit of Re x
    (activation_functions.py:19-22)."""
    x = np.asarray(x)
    sgn = np.where(np.signbit(np.real(x)), -1.0, 1.0)
    y = x * sgn
    return y + np.log1p(np.exp(-2.0 * y)) - LOG2


def d_log_cosh(x):
    return np.tanh(x)


class Leaf:
    def __init__(self, name, shape, cpx):
        self.name = name
        self.shape = tuple(int(s) for s in shape)
        self.size = int(np.prod(self.shape)) if self.shape else 1
        self.cpx = cpx


class Net:
    """Common flat-vector handling.  ``leaves`` is in tree_flatten order."""
    holomorphic = False
    realParams = True

    def num_real_params(self):
        return sum(l.size * (2 if l.cpx else 1) for l in self.leaves)

    def num_params(self):
        return sum(l.size for l in self.leaves)

    def unflatten(self, P):
        """Flat vector -> list of leaf arrays (vqs.py:430-444).  Accepts complex
        P (the holomorphic MinSR update makes it complex, minsr.py:65)."""
        out, start = [], 0
        for l in self.leaves:
            if l.cpx:
                v = P[start:start + l.size] + 1.j * P[start + l.size:start + 2 * l.size]
                start += 2 * l.size
            else:
                v = P[start:start + l.size]
                start += l.size
            out.append(np.asarray(v).reshape(l.shape))
        return out

    def flatten(self, leaves):
        """Leaf arrays -> flat real vector (vqs.py:461-464)."""
        parts = []
        for l, v in zip(self.leaves, leaves):
            v = np.asarray(v).ravel()
            if l.cpx:
                parts += [v.real, v.imag]
            else:
                parts.append(np.real(v))
        return np.concatenate(parts).astype(np.float64)

    def grad_layout(self, leaf_grads):
        """Per-leaf derivative arrays (B, size) -> flat gradient (B, Pr) complex
        in jVMC's layout (vqs.py:46-69)."""
        parts = []
        for l, g in zip(self.leaves, leaf_grads):
            g = np.asarray(g).reshape(g.shape[0], -1)
            if self.holomorphic:
                parts += [g, 1.j * g]
            else:
                parts.append(g.astype(np.complex128))
        return np.concatenate(parts, axis=1)


class CpxRBM(Net):
    """rbm.py:17-38: log psi = sum_h log_cosh(W^T (2s-1) + b)."""
    holomorphic = True
    realParams = False

    def __init__(self, sampleShape, numHidden=2, bias=False, visibleBias=False):
        self.L = int(np.prod(sampleShape))
        self.H = int(numHidden)
        self.bias = bias
        self.visibleBias = visibleBias
        self.leaves = []
        if bias:
            self.leaves.append(Leaf("Dense_0/bias", (self.H,), True))
        self.leaves.append(Leaf("Dense_0/kernel", (self.L, self.H), True))
        if visibleBias:
            self.leaves.append(Leaf("visible_bias", (self.L,), True))

    def _split(self, leaves):
        i = 0
        b = None
        if self.bias:
            b = leaves[i]; i += 1
        W = leaves[i]; i += 1
        a = leaves[i] if self.visibleBias else None
        return b, W, a

    def theta(self, leaves, s):
        b, W, a = self._split(leaves)
        sig = (2 * np.asarray(s).reshape(len(s), -1) - 1).astype(np.float64)
        th = sig @ W
        if b is not None:
            th = th + b
        return sig, th

    def forward(self, leaves, s):
        b, W, a = self._split(leaves)
        sig, th = self.theta(leaves, s)
        out = np.sum(log_cosh(th), axis=1)
        if a is not None:
            out = out + sig @ a
        return out.astype(np.complex128)

    def leaf_grads(self, leaves, s):
        b, W, a = self._split(leaves)
        sig, th = self.theta(leaves, s)
        t = np.tanh(th)
        g = []
        if self.bias:
            g.append(t)
        g.append((sig[:, :, None] * t[:, None, :]).reshape(len(sig), -1))
        if self.visibleBias:
            g.append(sig.astype(np.complex128))
        return g


class RBM(Net):
    """rbm.py:67-88: real parameters, log psi = sum log(cosh(W^T(2s-1)+b))."""

    def __init__(self, sampleShape, numHidden=2, bias=False):
        if len(sampleShape) != 1:
            raise NotImplementedError("RBM acts on the last axis only; use 1D lattices")
        self.L = int(sampleShape[0])
        self.H = int(numHidden)
        self.bias = bias
        self.leaves = []
        if bias:
            self.leaves.append(Leaf("Dense_0/bias", (self.H,), False))
        self.leaves.append(Leaf("Dense_0/kernel", (self.L, self.H), False))

    def _th(self, leaves, s):
        W = leaves[-1]
        sig = (2 * np.asarray(s).reshape(len(s), -1) - 1).astype(np.float64)
        th = sig @ W
        if self.bias:
            th = th + leaves[0]
        return sig, th

    def forward(self, leaves, s):
        sig, th = self._th(leaves, s)
        return np.sum(np.log(np.cosh(th)), axis=1).astype(np.float64)

    def leaf_grads(self, leaves, s):
        sig, th = self._th(leaves, s)
        t = np.tanh(th)
        g = []
        if self.bias:
            g.append(t)
        g.append((sig[:, :, None] * t[:, None, :]).reshape(len(sig), -1))
        return g


class CNN(Net):
    """cnn.py:18-81 with periodicBoundary=True, strides 1, log-cosh activation.

    Layer l: circular pad (0, F-1) on every spatial axis, VALID convolution with
    an HWIO kernel, bias for l >= 1 when ``bias`` (layer 0 only when
    ``firstLayerBias``), then activation.  Output: sum / sqrt(#elements).
    """

    def __init__(self, sampleShape, F=(8,), channels=(10,), strides=(1,), bias=True,
                 firstLayerBias=False, periodicBoundary=True, actFun=("log_cosh",)):
        if not periodicBoundary:
            raise NotImplementedError("only periodic boundaries are supported")
        if any(int(s) != 1 for s in strides):
            raise NotImplementedError("only unit strides are supported")
        if any(a != "log_cosh" for a in actFun):
            raise NotImplementedError("only the log_cosh activation is supported")
        self.shape = tuple(int(x) for x in sampleShape)
        self.F = tuple(int(f) for f in F)
        if len(self.F) != len(self.shape):
            raise ValueError("filter rank must match the lattice rank")
        self.channels = tuple(int(c) for c in channels)
        self.nl = len(self.channels)
        self.layer_bias = [bias] * self.nl
        self.layer_bias[0] = firstLayerBias
        cin = [1] + list(self.channels[:-1])
        self.cin = cin
        entries = []
        for l in range(self.nl):
            name = "Conv_%d" % l
            if self.layer_bias[l]:
                entries.append((name, "bias", (self.channels[l],)))
            entries.append((name, "kernel", self.F + (cin[l], self.channels[l])))
        # tree_flatten sorts dict keys as strings ("Conv_10" < "Conv_2").
        entries.sort(key=lambda e: (e[0], e[1]))
        self.leaves = [Leaf(e[0] + "/" + e[1], e[2], False) for e in entries]
        self._where = {(e[0], e[1]): i for i, e in enumerate(entries)}
        nsites = int(np.prod(self.shape))
        self.nrm = np.sqrt(nsites * self.channels[-1])
        # Patch index table: out position p, kernel offset k -> input position.
        grids = np.meshgrid(*[np.arange(n) for n in self.shape], indexing="ij")
        offs = np.meshgrid(*[np.arange(f) for f in self.F], indexing="ij")
        pos = [g.reshape(-1, 1) for g in grids]
        off = [o.reshape(1, -1) for o in offs]
        idx = np.zeros((nsites, int(np.prod(self.F))), dtype=np.int64)
        stride = 1
        for d in reversed(range(len(self.shape))):
            idx += ((pos[d] + off[d]) % self.shape[d]) * stride
            stride *= self.shape[d]
        self.patch = idx  # (nsites, prod F)

    def kernel(self, leaves, l):
        return leaves[self._where[("Conv_%d" % l, "kernel")]]

    def biasv(self, leaves, l):
        k = ("Conv_%d" % l, "bias")
        return leaves[self._where[k]] if k in self._where else None

    def _forward_all(self, leaves, s):
        B = len(s)
        x = (2 * np.asarray(s).reshape(B, -1, 1) - 1).astype(np.float64)
        xs, zs = [x], []
        for l in range(self.nl):
            K = self.kernel(leaves, l).reshape(-1, self.channels[l])
            P = x[:, self.patch, :].reshape(B, x.shape[1], -1)  # (B, n, F*Cin)
            z = P @ K
            b = self.biasv(leaves, l)
            if b is not None:
                z = z + b
            zs.append(z)
            x = log_cosh(z)
            xs.append(x)
        return xs, zs

    def forward(self, leaves, s):
        xs, _ = self._forward_all(leaves, s)
        return np.sum(xs[-1], axis=(1, 2)) / self.nrm

    def leaf_grads(self, leaves, s):
        B = len(s)
        xs, zs = self._forward_all(leaves, s)
        n = xs[0].shape[1]
        grads = [None] * len(self.leaves)
        delta = np.tanh(zs[-1]) / self.nrm  # d out / d z_L
        for l in reversed(range(self.nl)):
            x = xs[l]
            P = x[:, self.patch, :].reshape(B, n, -1)
            gK = np.einsum("bpk,bpc->bkc", P, delta)
            grads[self._where[("Conv_%d" % l, "kernel")]] = gK.reshape(B, -1)
            if self.biasv(leaves, l) is not None:
                grads[self._where[("Conv_%d" % l, "bias")]] = delta.sum(axis=1)
            if l > 0:
                K = self.kernel(leaves, l).reshape(-1, self.cin[l], self.channels[l])
                # scatter back: input pos patch[p,k] receives delta[p,c] * K[k,ci,c]
                contrib = np.einsum("bpc,kic->bpki", delta, K)
                g = np.zeros((B, n, self.cin[l]))
                for k in range(self.patch.shape[1]):
                    np.add.at(g, (slice(None), self.patch[:, k]), contrib[:, :, k, :])
                delta = g * np.tanh(zs[l - 1])
        return grads


class NQS:
    """Minimal NQS (vqs.py:85-491) over a flat real parameter vector."""

    def __init__(self, net, params=None, batchSize=1000):
        self.net = net
        self.holomorphic = net.holomorphic
        self.realParams = net.realParams
        self.batchSize = batchSize
        self.logarithmic = True
        P = np.zeros(net.num_real_params()) if params is None else np.asarray(params)
        self.set_parameters(P)

    @property
    def numParameters(self):
        return self.net.num_params()

    def set_parameters(self, P):
        self._leaves = self.net.unflatten(np.asarray(P))

    def update_parameters(self, dP):
        d = self.net.unflatten(np.asarray(dP))
        self._leaves = [a + b for a, b in zip(self._leaves, d)]

    def get_parameters(self):
        return self.net.flatten(self._leaves)

    def __call__(self, s):
        s = np.asarray(s)
        out = [self.net.forward(self._leaves, s[i:i + self.batchSize])
               for i in range(0, len(s), self.batchSize)]
        return np.concatenate(out) if out else np.zeros(0)

    def gradients(self, s):
        s = np.asarray(s)
        out = []
        for i in range(0, len(s), self.batchSize):
            g = self.net.leaf_grads(self._leaves, s[i:i + self.batchSize])
            out.append(self.net.grad_layout(g))
        return np.concatenate(out)
