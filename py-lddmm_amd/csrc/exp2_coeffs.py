"""Coefficients of the float64 2^f polynomial used by fast_exp2(double) in
common.hpp, and a check of the evaluated error.

    python py-lddmm_amd/csrc/exp2_coeffs.py

Fits 2^f on [-1/2, 1/2] by Chebyshev interpolation of degree 10 in 50-digit
arithmetic (mpmath.chebyfit; near-minimax), rounds the monomial coefficients
to float64 and evaluates Horner's rule with a single rounding per step (as
the device's FMA does) on a dense grid, printing the coefficients (highest
degree first) and the largest relative error against 2^f.
"""
import mpmath as mp
import numpy as np

DEGREE = 10


def coefficients(degree=DEGREE):
    mp.mp.dps = 50
    poly, err = mp.chebyfit(lambda f: mp.power(2, f), [-0.5, 0.5], degree + 1, error=True)
    return [float(c) for c in poly], float(err)


def horner_fma(cs, f):
    p = mp.mpf(cs[0])
    for c in cs[1:]:
        p = mp.mpf(float(p * mp.mpf(f) + mp.mpf(c)))     # one rounding per step
    return p


def worst_relative_error(cs, points=20001):
    mp.mp.dps = 50
    return max(abs(float((horner_fma(cs, f) - mp.power(2, mp.mpf(f))) / mp.power(2, mp.mpf(f))))
               for f in np.linspace(-0.5, 0.5, points))


if __name__ == "__main__":
    cs, fit_err = coefficients()
    print("coefficients (highest degree first):")
    print(",\n".join(repr(c) for c in cs))
    print(f"fit error {fit_err:.3g}, evaluated worst relative error {worst_relative_error(cs):.3g}")
