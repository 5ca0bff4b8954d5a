"""Derives the polynomial coefficients of the real log-cosh in common.hpp.

log cosh(x) = |x| - ln 2 + log1p(t),  t = exp(-2|x|) in (0, 1]

* exp(y) for y <= 0: y = n ln 2 + r with |r| <= ln 2 / 2; e^r is a degree-11
  polynomial (Chebyshev least-squares fit on [-ln 2 / 2, ln 2 / 2]).
* log1p(t) = 2 atanh(s), s = t / (2 + t) in [0, 1/3];
  atanh(s) = s + s^3 R(s^2) with R(z) = (atanh(sqrt z) - sqrt z) / z^{3/2},
  a degree-9 polynomial fit on z in [0, 1/9].

Run with ``python fit_log_cosh.py`` (needs mpmath); it prints the C++
initialisers and the fit errors.  ``--check`` evaluates the double-precision
evaluation scheme against mpmath at many points and prints the largest
absolute error of log cosh.
"""
import sys

import mpmath as mp

mp.mp.dps = 60
EXP_DEG = 11
ATANH_DEG = 9


def fits():
    ln2 = mp.log(2)
    pe, ee = mp.chebyfit(mp.exp, [-ln2 / 2, ln2 / 2], EXP_DEG + 1, error=True)

    def R(z):
        if z == 0:
            return mp.mpf(1) / 3
        r = mp.sqrt(z)
        return (mp.atanh(r) - r) / (z * r)

    pr, er = mp.chebyfit(R, [0, mp.mpf(1) / 9], ATANH_DEG + 1, error=True)
    # mpmath returns the highest power first
    return [float(c) for c in pe], ee, [float(c) for c in pr], er


def fma(a, b, c):
    """Fused multiply-add rounded once to double (the product is exact at 60 digits)."""
    return float(mp.mpf(a) * mp.mpf(b) + mp.mpf(c))


def lc_double(x, pe, pr):
    """The evaluation order of the device code, in IEEE double."""
    import math
    a = abs(x)
    y = -2.0 * min(a, 40.0)
    n = round(y * 1.4426950408889634)
    r = fma(-n, 6.93147180559945286227e-01, y)
    r = fma(-n, 2.31904681384629955842e-17, r)
    p = pe[0]
    for c in pe[1:]:
        p = fma(p, r, c)
    t = math.ldexp(p, int(n))
    s = t / (2.0 + t)
    z = s * s
    q = pr[0]
    for c in pr[1:]:
        q = fma(q, z, c)
    return (a - 0.69314718055994530942) + 2.0 * fma(s * z, q, s)


def main():
    pe, ee, pr, er = fits()
    print("// exp(r), |r| <= ln2/2, degree %d, fit error %.2e" % (EXP_DEG, ee))
    print("constexpr double LC_EXP[%d] = {%s};" % (len(pe), ", ".join("%.17g" % c for c in pe)))
    print("// R(z), z in [0, 1/9], degree %d, fit error %.2e" % (ATANH_DEG, er))
    print("constexpr double LC_ATANH[%d] = {%s};" % (len(pr), ", ".join("%.17g" % c for c in pr)))
    if "--check" in sys.argv:
        worst = 0.0
        for i in range(6001):
            x = -30.0 + 60.0 * i / 6000
            ref = mp.log(mp.cosh(mp.mpf(x)))
            worst = max(worst, abs(float(lc_double(x, pe, pr) - ref)))
        print("max |error| on [-30, 30]: %.3e" % worst)


if __name__ == "__main__":
    main()
