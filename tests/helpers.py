"""Shared builders for the parity tests: the same Hamiltonian / network on the
device (vmc_jax_amd) and in the FP64 oracle."""
import numpy as np

from oracle import operator as oop

WEIGHTS16 = np.array([0.23898957, 0.12614753, 0.19479055, 0.17325271, 0.14619853, 0.21392751,
                      0.19648707, 0.17103704, -0.15457255, 0.10954413, 0.13228065, -0.14935214,
                      -0.09963073, 0.17610707, 0.13386381, -0.14836467])
REF_ZZ = np.array([0.882762129306284, 0.8936168721790617, 0.9257753299594491, 0.9779836185039352,
                   1.0482156449061142, 1.1337654450614298, 1.231369697427413, 1.337354107391303,
                   1.447796176316155, 1.558696104640795, 1.666147269524912, 1.7664978782554912,
                   1.8564960156892512, 1.9334113379450693, 1.9951280521882777, 2.0402054805651546,
                   2.067904337137255, 2.078178742959828, 2.071635856483114, 2.049466698269522])


def bonds(shape):
    """Nearest-neighbour bonds of a periodic chain or square lattice; TFIM
    strings are added as ``for site: ZZ(bonds); X(site)``."""
    if len(shape) == 1:
        L = shape[0]
        return [(l, [(l + 1) % L]) for l in range(L)]
    Lx, Ly = shape
    out = []
    for i in range(Lx):
        for j in range(Ly):
            s = i * Ly + j
            out.append((s, [((i + 1) % Lx) * Ly + j, i * Ly + (j + 1) % Ly]))
    return out


def tfim_pair(shape, J, h):
    """(device operator, oracle operator) for H = J sum ZZ + h sum X."""
    from vmc_jax_amd import operator as dop
    H = dop.BranchFreeOperator()
    Ho = oop.BranchFreeOperator()
    for s, nbs in bonds(shape):
        for nb in nbs:
            H.add(dop.scal_opstr(J, (dop.Sz(s), dop.Sz(nb))))
            Ho.add(oop.scal_opstr(J, (oop.Sz(s), oop.Sz(nb))))
        H.add(dop.scal_opstr(h, (dop.Sx(s),)))
        Ho.add(oop.scal_opstr(h, (oop.Sx(s),)))
    return H, Ho


def xxz_pair(L, delta):
    """(device, oracle) XXZ chain: sum X X + Y Y + delta Z Z (separate strings)."""
    from vmc_jax_amd import operator as dop
    H = dop.BranchFreeOperator()
    Ho = oop.BranchFreeOperator()
    for l in range(L):
        r = (l + 1) % L
        for (a, b, c) in ((dop.Sx, oop.Sx, 1.0), (dop.Sy, oop.Sy, 1.0), (dop.Sz, oop.Sz, delta)):
            H.add(dop.scal_opstr(c, (a(l), a(r))))
            Ho.add(oop.scal_opstr(c, (b(l), b(r))))
    return H, Ho


def to_np(t):
    return t.detach().cpu().numpy()
