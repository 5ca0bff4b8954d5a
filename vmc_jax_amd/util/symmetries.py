"""Lattice symmetry orbits (jVMC util/symmetries.py:1-206).

An orbit element is a signed permutation of the sites: (g s)_i = s_{m_i} or its
spin flip.  The reference stores each element as a dense integer matrix o with
(o x)_i = x_{m_i} on spins x = 2s - 1 (symmetries.py:54-72), multiplies the
translation, reflection and rotation matrices (symmetries.py:70), appends the
negated matrices for the spin flip (symmetries.py:145-148), and keeps the
unique matrices in ``np.unique`` order with the factor of the first occurrence
(symmetries.py:150-152).

Here every element is built as a site map and a sign; the dense matrices are
produced only when ``LatticeSymmetry.orbit`` is read.  ``np.unique`` runs on an
integer key per element whose lexicographic order is that of the flattened
dense matrices (see ``_unique_order``), so the element order and the factors are
those of the reference.
"""
import warnings

import numpy as np


class LatticeSymmetry:
    """symmetries.py:9-40: the orbit (``orbit``, int32 (|G|, N, N)) and the phase
    factor of each element (``factor``).

    Built either from a dense orbit array, as in the reference, or internally
    from (maps, signs).  ``maps[g, i] = j`` and ``signs[g, i] = +-1`` describe
    row i of element g: a single entry signs[g, i] in column j.
    """

    def __init__(self, orbit=None, factor=None, maps=None, signs=None):
        if orbit is not None:
            o = np.asarray(orbit)
            if o.ndim != 3 or o.shape[1] != o.shape[2]:
                raise ValueError("orbit must have shape (|G|, N, N)")
            maps, signs = _maps_from_dense(o)
            self._dense = o
        else:
            maps = np.asarray(maps, dtype=np.int64)
            signs = np.asarray(signs, dtype=np.int64)
            self._dense = None
        self._maps = maps
        self._signs = signs
        self._factor = np.asarray(factor) if factor is not None else np.ones(maps.shape[0])

    @property
    def factor(self):
        return self._factor

    @property
    def orbit(self):
        if self._dense is None:
            G, N = self._maps.shape
            d = np.zeros((G, N, N), dtype=np.int32)
            d[np.arange(G)[:, None], np.arange(N)[None, :], self._maps] = self._signs
            self._dense = d
        return self._dense

    @property
    def shape(self):
        return (self._maps.shape[0], self._maps.shape[1], self._maps.shape[1])

    @property
    def dtype(self):
        return self.orbit.dtype if self._dense is not None else np.dtype(np.int32)

    @property
    def maps(self):
        """(|G|, N) column of the nonzero entry of each row."""
        return self._maps

    @property
    def signs(self):
        """(|G|, N) sign of the nonzero entry of each row."""
        return self._signs

    def device_map(self):
        """int32 (|G|, N) site map of the device kernels: j for +1, -1 - j for -1."""
        return np.where(self._signs > 0, self._maps, -1 - self._maps).astype(np.int32)


def _maps_from_dense(o):
    """(maps, signs) of a dense orbit whose rows each hold one entry +1 or -1."""
    nz = o != 0
    if not np.all(nz.sum(axis=2) == 1):
        raise NotImplementedError("orbit matrices must be signed permutations (one nonzero per row)")
    maps = np.argmax(nz, axis=2)
    signs = np.take_along_axis(o, maps[..., None], axis=2)[..., 0].astype(np.int64)
    if not np.all(np.abs(signs) == 1):
        raise NotImplementedError("orbit matrix entries must be 0, +1 or -1")
    return maps.astype(np.int64), signs


def _unique_order(maps, signs):
    """Indices of the unique elements in the order ``np.unique(dense, axis=0)``
    gives the flattened dense matrices, with the first occurrence of each.

    Rows compare lexicographically from row 0.  Row sign*e_j versus sign'*e_j':
    every +e row is larger than every -e row; +e_j > +e_j' iff j < j'; -e_j <
    -e_j' iff j < j'.  The key N + (N-1-j) for +e_j and j for -e_j orders the
    rows the same way, so ``np.unique`` on the keys returns the same order."""
    N = maps.shape[1]
    keys = np.where(signs > 0, 2 * N - 1 - maps, maps)
    _, idx = np.unique(keys, axis=0, return_index=True)
    return idx


def get_2D_index(L):
    return np.arange(L ** 2).reshape((L, L))


def get_1D_index(L):
    return np.arange(L)


def _combine(site_maps, factors):
    """Products of the per-symmetry maps in the reference's loop order
    (symmetries.py:58-70 and 165-177): element (t, ref, rot) is the matrix
    o_t @ o_ref @ o_rot, whose row i reads site rot[ref[t[i]]]."""
    maps, facs = [], []

    def rec(k, m, f):
        if k == len(site_maps):
            maps.append(m)
            facs.append(f)
            return
        for mk, fk in zip(site_maps[k], factors[k]):
            rec(k + 1, mk.ravel()[m] if m is not None else mk.ravel(), fk if f is None else f * fk)

    rec(0, None, None)
    return np.array(maps), np.array(facs)


def get_maps_2D(L, **kwargs):
    """symmetries.py:75-110: site maps and factors of rotations, reflections and
    translations of the L x L square lattice."""
    idx = get_2D_index(L)
    rotation = {"maps": [idx], "factor": [1.]}
    reflection = {"maps": [idx], "factor": [1.]}
    translation = {"maps": [idx], "factor": [1.]}
    if kwargs["rotation"]["use"]:
        f = kwargs["rotation"]["factor"]
        rotation["maps"] += [idx[::-1, :].T, idx[::-1, ::-1], idx[:, ::-1].T]
        rotation["factor"] += [f, f ** 2, f ** 3]
    if kwargs["reflection"]["use"]:
        f = kwargs["reflection"]["factor"]
        reflection["maps"] += [idx[::-1, :], idx[:, ::-1], idx[::-1, ::-1]]
        reflection["factor"] += [f, f, f ** 2]
    if kwargs["translation"]["use"]:
        f = kwargs["translation"]["factor"]
        for x in range(L):
            for y in range(L):
                if x == 0 and y == 0:
                    continue
                translation["maps"].append(np.roll(idx, (y, x), axis=(0, 1)))
                translation["factor"].append(f ** (x + y))
    return translation, reflection, rotation


def _finish(maps, factors, N, use_spinflip, spinflip_factor):
    signs = np.ones_like(maps)
    if use_spinflip:
        warnings.warn("Symmetry 'spinflip' assumes that spin up/down is encoded with numerical values +1/-1.")
        maps = np.concatenate([maps, maps], axis=0)
        signs = np.concatenate([signs, -signs], axis=0)
        factors = np.concatenate([factors, spinflip_factor * factors])
    idx = _unique_order(maps, signs)
    return LatticeSymmetry(maps=maps[idx], signs=signs[idx], factor=factors[idx])


def get_orbit_2D_square(L, *args, translation_factor=1., reflection_factor=1., rotation_factor=1.,
                        spinflip_factor=1.):
    """symmetries.py:113-153: the symmetry group of the L x L square lattice for
    any choice of "reflection", "rotation", "translation" and "spinflip"."""
    kwargs = {"reflection": {"use": "reflection" in args, "factor": reflection_factor},
              "translation": {"use": "translation" in args, "factor": translation_factor},
              "rotation": {"use": "rotation" in args, "factor": rotation_factor},
              "spinflip": {"use": "spinflip" in args, "factor": spinflip_factor}}
    translation, reflection, rotation = get_maps_2D(L, **kwargs)
    maps, factors = _combine([translation["maps"], reflection["maps"], rotation["maps"]],
                             [translation["factor"], reflection["factor"], rotation["factor"]])
    return _finish(maps, factors, L * L, kwargs["spinflip"]["use"], spinflip_factor)


def get_maps_1D(L, **kwargs):
    """symmetries.py:156-181: reflection and translation maps of a chain."""
    idx = get_1D_index(L)
    reflection = {"maps": [idx], "factor": [1.]}
    translation = {"maps": [idx], "factor": [1.]}
    if kwargs["reflection"]["use"]:
        reflection["maps"].append(idx[::-1])
        reflection["factor"].append(kwargs["reflection"]["factor"])
    if kwargs["translation"]["use"]:
        f = kwargs["translation"]["factor"]
        for x in range(1, L):
            translation["maps"].append(np.roll(idx, x, axis=(0,)))
            translation["factor"].append(f ** x)
    return translation, reflection


def get_orbit_1D(L, *args, translation_factor=1., reflection_factor=1., spinflip_factor=1.):
    """symmetries.py:184-206: the symmetry group of a chain of L sites for any
    choice of "reflection", "translation" and "spinflip"."""
    kwargs = {"reflection": {"use": "reflection" in args, "factor": reflection_factor},
              "translation": {"use": "translation" in args, "factor": translation_factor},
              "spinflip": {"use": "spinflip" in args, "factor": spinflip_factor}}
    translation, reflection = get_maps_1D(L, **kwargs)
    maps, factors = _combine([translation["maps"], reflection["maps"]],
                             [translation["factor"], reflection["factor"]])
    return _finish(maps, factors, L, kwargs["spinflip"]["use"], spinflip_factor)


__all__ = ["LatticeSymmetry", "get_orbit_1D", "get_orbit_2D_square", "get_maps_1D", "get_maps_2D",
           "get_1D_index", "get_2D_index"]
