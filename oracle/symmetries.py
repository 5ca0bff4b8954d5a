"""Lattice-symmetry orbits and the symmetrised network (SymNet).

Oracle (test infrastructure only).  Follows
* orbit construction      symmetries.py:50-206, with dense integer matrices as
                          the reference builds them: o_t @ o_ref @ o_rot, the
                          negated copies for the spin flip, ``np.unique`` over
                          the flattened matrices with the factor of the first
                          occurrence
* ``SymNet``              sym_wrapper.py:29-83: x = 2s - 1, x_g = o_g x,
                          s_g = (x_g + 1) // 2, avgFun over the orbit
* ``avg_exp``             sym_wrapper.py:7-9 with jax.scipy.special.logsumexp's
                          conventions (entries with b = 0 dropped; real inputs:
                          log|sum|, NaN where the sum is negative)
* ``avg_log`` / ``avg_sep`` sym_wrapper.py:12-19
* gradients               the derivative of the average times the inner
                          network's O_k (vqs.py:40-69 through the chain rule);
                          pinned by central finite differences in the tests.
"""
import numpy as np

from . import nets as onets


def _onehot(maps, N):
    o = np.zeros((N, N), dtype=np.int32)
    o[np.arange(N), np.asarray(maps).ravel()] = 1
    return o


def orbit_1d(L, *args, translation_factor=1., reflection_factor=1., spinflip_factor=1.):
    """symmetries.py:156-206.  Returns (orbit int32 (|G|, L, L), factors)."""
    idx = np.arange(L)
    tm, tf = [idx], [1.]
    rm, rf = [idx], [1.]
    if "reflection" in args:
        rm.append(idx[::-1])
        rf.append(reflection_factor)
    if "translation" in args:
        for x in range(1, L):
            tm.append(np.roll(idx, x))
            tf.append(translation_factor ** x)
    orbits, factors = [], []
    for mt, ft in zip(tm, tf):
        for mr, fr in zip(rm, rf):
            orbits.append(_onehot(mt, L) @ _onehot(mr, L))
            factors.append(ft * fr)
    return _finish(np.array(orbits), np.array(factors), "spinflip" in args, spinflip_factor, L)


def orbit_2d_square(L, *args, translation_factor=1., reflection_factor=1., rotation_factor=1., spinflip_factor=1.):
    """symmetries.py:50-153.  Returns (orbit int32 (|G|, L^2, L^2), factors)."""
    idx = np.arange(L * L).reshape(L, L)
    rot, rotf = [idx], [1.]
    ref, reff = [idx], [1.]
    tr, trf = [idx], [1.]
    if "rotation" in args:
        rot += [idx[::-1, :].T, idx[::-1, ::-1], idx[:, ::-1].T]
        rotf += [rotation_factor, rotation_factor ** 2, rotation_factor ** 3]
    if "reflection" in args:
        ref += [idx[::-1, :], idx[:, ::-1], idx[::-1, ::-1]]
        reff += [reflection_factor, reflection_factor, reflection_factor ** 2]
    if "translation" in args:
        for x in range(L):
            for y in range(L):
                if x == 0 and y == 0:
                    continue
                tr.append(np.roll(idx, (y, x), axis=(0, 1)))
                trf.append(translation_factor ** (x + y))
    N = L * L
    orbits, factors = [], []
    for mt, ft in zip(tr, trf):
        for mr, fr in zip(ref, reff):
            for mo, fo in zip(rot, rotf):
                orbits.append(_onehot(mt, N) @ _onehot(mr, N) @ _onehot(mo, N))
                factors.append(ft * fr * fo)
    return _finish(np.array(orbits), np.array(factors), "spinflip" in args, spinflip_factor, N)


def _finish(orbits, factors, spinflip, spinflip_factor, N):
    if spinflip:
        orbits = np.concatenate([orbits, -orbits], axis=0)
        factors = np.concatenate([factors, spinflip_factor * factors])
    uniq, indices = np.unique(orbits.reshape((-1, N * N)), return_index=True, axis=0)
    return uniq.reshape(-1, N, N).astype(np.int32), factors[indices]


# ---------------------------------------------------------------- averaging
def avg_exp(c, b):
    """logsumexp(c, b=b) along the last axis with jax's conventions."""
    c = np.asarray(c)
    b = np.asarray(b)
    cpx = np.iscomplexobj(c) or np.iscomplexobj(b)
    a = c.astype(np.complex128) if cpx else c.astype(np.float64)
    a = np.where(b != 0, a, -np.inf)
    amax = np.max(np.real(a), axis=-1, keepdims=True)
    amax = np.where(np.isfinite(amax), amax, 0.0)
    with np.errstate(invalid="ignore", divide="ignore"):
        t = np.exp(a - amax) * b
        S = t.sum(axis=-1)
        if cpx:
            return np.log(S) + amax[..., 0]
        out = np.log(np.abs(S)) + amax[..., 0]
        return np.where(S < 0, np.nan, out)


def avg_exp_weights(c, b):
    """d avg_exp / d c_g (complex); zero for real samples whose sum is negative
    (the cotangent through jnp.where's NaN branch)."""
    c = np.asarray(c)
    b = np.asarray(b)
    cpx = np.iscomplexobj(c) or np.iscomplexobj(b)
    a = c.astype(np.complex128) if cpx else c.astype(np.float64)
    a = np.where(b != 0, a, -np.inf)
    amax = np.max(np.real(a), axis=-1, keepdims=True)
    amax = np.where(np.isfinite(amax), amax, 0.0)
    t = np.exp(a - amax) * b
    S = t.sum(axis=-1, keepdims=True)
    w = (t / S).astype(np.complex128)
    if not cpx:
        w = np.where(S < 0, 0.0, w)
    return w


def avg_log(c, b=None):
    return np.mean(c, axis=-1)


def avg_log_weights(c, b=None):
    c = np.asarray(c)
    return np.full(c.shape, 1.0 / c.shape[-1], dtype=np.complex128)


def avg_sep(c, b=None):
    c = np.asarray(c)
    re, im = np.real(c), np.imag(c)
    return 0.5 * np.log(np.mean(np.exp(2 * re), axis=-1)) + 1j * np.angle(np.mean(np.exp(1j * im), axis=-1))


def avg_sep_weights(c, b=None):
    """Real-parameter networks (Im c independent of the parameters)."""
    e = np.exp(2 * np.real(np.asarray(c)))
    return (e / e.sum(axis=-1, keepdims=True)).astype(np.complex128)


AVG = {"exp": (avg_exp, avg_exp_weights), "log": (avg_log, avg_log_weights), "sep": (avg_sep, avg_sep_weights)}


class SymNet(onets.Net):
    """sym_wrapper.py:29-83 around an oracle network of the same lattice."""

    def __init__(self, net, orbit, factor, avg="exp"):
        self.net = net
        self.orbit = np.asarray(orbit)
        self.factor = np.asarray(factor)
        self.avg = avg
        self.leaves = net.leaves
        self.holomorphic = net.holomorphic and avg != "sep"
        self.realParams = net.realParams
        self.sampleShape = getattr(net, "sampleShape", None)

    def transformed(self, s):
        """(B, *shape) -> (B, G, *shape): (o_g (2s - 1) + 1) // 2."""
        s = np.asarray(s)
        B = s.shape[0]
        x = 2 * s.reshape(B, -1).astype(np.int64) - 1
        xg = np.einsum("gij,bj->bgi", self.orbit.astype(np.int64), x)
        return ((xg + 1) // 2).reshape((B, self.orbit.shape[0]) + s.shape[1:])

    def coefficients(self, leaves, s):
        sg = self.transformed(s)
        B, G = sg.shape[:2]
        c = self.net.forward(leaves, sg.reshape((B * G,) + sg.shape[2:]))
        return c.reshape(B, G)

    def forward(self, leaves, s):
        c = self.coefficients(leaves, s)
        return AVG[self.avg][0](c, self.factor)

    def leaf_grads(self, leaves, s):
        sg = self.transformed(s)
        B, G = sg.shape[:2]
        flat = sg.reshape((B * G,) + sg.shape[2:])
        c = self.net.forward(leaves, flat).reshape(B, G)
        w = AVG[self.avg][1](c, self.factor)
        out = []
        for g in self.net.leaf_grads(leaves, flat):
            g = np.asarray(g).reshape(B, G, -1)
            out.append(np.einsum("bg,bgk->bk", w, g))
        return out
