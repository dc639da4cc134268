"""Fit the single-precision inverse normal CDF used by the path kernels.

Method (the functional form of M. Giles, "Approximating the erfinv function",
GPU Computing Gems Jade, 2011): with q = min(u, 1-u) and t = q (1-q),
w = -log(4t), the ratio f = Phi^{-1}(1-q) / (1-2q) is a smooth function of w
(and of t).  The kernels read f from a table of cubics, one per 1/16 of a binade
of t: the row index is bits 19..27 of t's FP32 encoding (the low 5 exponent
bits and the top 4 mantissa bits) and the variable is the remaining 19 mantissa
bits as an integer, so no logarithm and no branch is evaluated.  The smallest q
the kernels produce is 2^-33 (u = (s + 1/2) 2^-32 folded to the lower half), so
t spans [2^-33, 1/4]: 32 binades, biased exponents 94..125, 512 rows.

The coefficients are NOT taken from any library: they are fitted here by
weighted least squares (relative error) per row against scipy.special.ndtri
evaluated in FP64, and written to ../csrc/icdf_coeffs.h.

Usage: python fit_icdf.py
"""
import os

import numpy as np
from scipy.special import ndtri

TAB_SUB = 4          # 2^4 segments per binade of t
TAB_DEG = 3
TAB_XBITS = 23 - TAB_SUB


def q_of_w(w):
    return 0.5 * (1.0 - np.sqrt(-np.expm1(-w)))


def f_of_w(w):
    q = q_of_w(w)
    return -ndtri(q) / (1.0 - 2.0 * q)


def tab_exponent(row):
    """Biased FP32 exponent of t for a table row (the row's bits 4..8 are the
    low 5 bits of the exponent, which lies in 94..125)."""
    e5 = row >> TAB_SUB
    return e5 + 96 if e5 <= 29 else e5 + 64


def fit_table():
    """Row (E, j) covers t = 2^(E-127) (1 + (j + x 2^-19) / 16) for the integer
    x in [0, 2^19); its cubic is fitted in X = x 2^-19 in [0, 1) and stored
    scaled to x.  The point t = 1/4 (q = 1/2 to FP32 rounding, biased exponent
    125) is the constant f(0) = sqrt(pi/2), and rows no t in [2^-33, 1/4]
    reaches repeat it."""
    rows = 1 << (5 + TAB_SUB)
    tab = np.zeros((rows, TAB_DEG + 1))
    f0 = np.sqrt(np.pi / 2.0)
    worst = 0.0
    X = 0.5 - 0.5 * np.cos(np.linspace(0.0, np.pi, 513))     # Chebyshev-spaced, both ends
    X = np.minimum(X, 1.0 - 1e-9)
    for r in range(rows):
        E = tab_exponent(r)
        j = r & ((1 << TAB_SUB) - 1)
        if E >= 125:
            tab[r, 0] = f0
            continue
        t = 2.0 ** (E - 127) * (1.0 + (j + X) / (1 << TAB_SUB))
        f = f_of_w(-np.log(4.0 * t))
        a = np.polyfit(X, f, TAB_DEG, w=1.0 / f)[::-1]          # a_0 .. a_3 in X
        worst = max(worst, np.max(np.abs(np.polyval(a[::-1], X) / f - 1.0)))
        tab[r] = a * 2.0 ** (-TAB_XBITS * np.arange(TAB_DEG + 1))   # in x = X 2^19
    return tab, worst


def main():
    tab, err = fit_table()
    out = os.path.join(os.path.dirname(__file__), "..", "csrc", "icdf_coeffs.h")
    with open(out, "w") as f:
        f.write("/* Generated by gen/fit_icdf.py -- do not edit.\n")
        f.write(" * Piecewise-cubic table of f = Phi^-1(1-q)/(1-2q) in the bits of t = q(1-q):\n")
        f.write(" * row r = bits(t) >> 19 & 511 holds {c0, c1, c2, c3} of\n")
        f.write(" * f = c0 + x (c1 + x (c2 + x c3)), x = bits(t) & 0x7FFFF.\n")
        f.write(" * max relative fit error %.3e (FP64 evaluation). */\n" % err)
        f.write("#pragma once\n")
        f.write("#define FDMC_ICDF_TAB_ROWS %d\n" % len(tab))
        f.write("#define FDMC_ICDF_TAB { \\\n")
        for r in range(len(tab)):
            f.write("    " + ", ".join("%.9ef" % c for c in tab[r]) + (", \\\n" if r + 1 < len(tab) else " \\\n"))
        f.write("}\n")
    print("table err %.3e -> %s" % (err, out))


if __name__ == "__main__":
    main()
