"""Philox4x32-10 counter-based generator and the Metropolis proposal stream.

Oracle (test infrastructure only).  Philox4x32-10 is the published generator of
Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as easy as 1, 2, 3"
(SC'11).  jVMC draws its proposals from JAX threefry keys
(sampler.py:15-19 ``random.randint``, sampler.py:330-340 ``random.split`` /
``random.bernoulli``); those bit streams cannot be reproduced without JAX, so the
device sampler and this oracle both draw from Philox with the counter layout
below, which makes acceptance decisions comparable bit for bit.

Counter layout for Metropolis step ``j`` of chain ``c`` in sampler call ``k``:
``ctr = (j, c, k, stream)`` and ``key = (seed_lo, seed_hi)``.  From the four
output words:

* ``site = (r0 * numSites) >> 32``  (Lemire multiply-shift),
* ``u = ((r1 >> 5) * 2**26 + (r2 >> 6)) * 2**-53``  (53-bit uniform in [0, 1)),

and the proposal is accepted iff ``u < exp(mu * Re(log psi(s') - log psi(s)))``,
the rule of ``random.bernoulli(key, P)`` (sampler.py:336-340).
"""
import numpy as np

_M0 = np.uint64(0xD2511F53)
_M1 = np.uint64(0xCD9E8D57)
_W0 = np.uint32(0x9E3779B9)
_W1 = np.uint32(0xBB67AE85)
_MASK = np.uint64(0xFFFFFFFF)


def philox4x32_10(ctr, key):
    """Vectorised Philox4x32-10.

    ``ctr``: uint32 array of shape (..., 4); ``key``: uint32 array (..., 2)
    broadcastable to it.  Returns uint32 array (..., 4).
    """
    c = np.asarray(ctr, dtype=np.uint32)
    k = np.broadcast_to(np.asarray(key, dtype=np.uint32), c.shape[:-1] + (2,))
    c0, c1, c2, c3 = (c[..., i].astype(np.uint64) for i in range(4))
    k0 = k[..., 0].astype(np.uint32).copy()
    k1 = k[..., 1].astype(np.uint32).copy()
    with np.errstate(over="ignore"):
        for r in range(10):
            p0 = _M0 * c0
            p1 = _M1 * c2
            hi0, lo0 = p0 >> np.uint64(32), p0 & _MASK
            hi1, lo1 = p1 >> np.uint64(32), p1 & _MASK
            n0 = hi1 ^ c1 ^ k0.astype(np.uint64)
            n1 = lo1
            n2 = hi0 ^ c3 ^ k1.astype(np.uint64)
            n3 = lo0
            c0, c1, c2, c3 = n0, n1, n2, n3
            if r < 9:
                k0 = (k0 + _W0).astype(np.uint32)
                k1 = (k1 + _W1).astype(np.uint32)
    return np.stack([c0, c1, c2, c3], axis=-1).astype(np.uint32)


def split_seed(seed):
    """64-bit integer seed -> Philox key words (lo, hi)."""
    seed = int(seed) & 0xFFFFFFFFFFFFFFFF
    return np.array([seed & 0xFFFFFFFF, seed >> 32], dtype=np.uint32)


def proposal_stream(seed, call_index, step, num_chains, num_sites, stream=0):
    """Sites and uniforms of Metropolis step ``step`` for all chains.

    Returns ``(site int64 (C,), u float64 (C,))``.
    """
    C = int(num_chains)
    ctr = np.zeros((C, 4), dtype=np.uint32)
    ctr[:, 0] = np.uint32(step)
    ctr[:, 1] = np.arange(C, dtype=np.uint32)
    ctr[:, 2] = np.uint32(call_index)
    ctr[:, 3] = np.uint32(stream)
    r = philox4x32_10(ctr, split_seed(seed))
    site = (r[:, 0].astype(np.uint64) * np.uint64(num_sites)) >> np.uint64(32)
    a = (r[:, 1] >> np.uint32(5)).astype(np.float64)
    b = (r[:, 2] >> np.uint32(6)).astype(np.float64)
    u = (a * 67108864.0 + b) * (1.0 / 9007199254740992.0)
    return site.astype(np.int64), u
