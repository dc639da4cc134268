"""Fit the single-precision inverse normal CDF used by the path kernels.

Method (the functional form of M. Giles, "Approximating the erfinv function",
GPU Computing Gems Jade, 2011): with q = min(u, 1-u) and w = -log(4 q (1-q)),
the ratio f(w) = Phi^{-1}(1-q) / (1-2q) is a smooth function of w.  It is
approximated by one polynomial in (w - 1.5) for w < 7 and one in
(sqrt(w) - C_TAIL) for w >= 7.  The coefficients below are NOT taken from any
library: they are fitted here by weighted least squares (relative error) against
scipy.special.ndtri evaluated in FP64, and written to ../csrc/icdf_coeffs.h.

The smallest q the kernels produce is 2^-33 (u = (s + 0.5) 2^-32 folded to the
lower half), i.e. w <= 31 ln 2 = 21.49; the tail fit covers w in [7, 21.6].

The split sits at w = 7 (q = 4.6e-4) rather than lower because 64 consecutive
scrambled-Sobol points fill all 64 strata of [0, 1) in each coordinate: a
wavefront takes the tail branch whenever one of its lanes falls in a tail, which
at w = 7 is 6 % of wavefronts (43 % at w = 5), for two more terms in the central
polynomial.  Its centre w = 1.5 keeps the FP32 rounding of (u - U0) small where the
probability mass is: the largest error of the FP32 evaluation is 1.9e-7.

Usage: python fit_icdf.py
"""
import os

import numpy as np
from scipy.special import ndtri

W_SPLIT = 7.0
W_MAX = 21.6
DEG_CENTRAL = 11
DEG_TAIL = 8
C_CENTRAL = 1.5
C_TAIL = 0.5 * (np.sqrt(W_SPLIT) + np.sqrt(W_MAX))


def q_of_w(w):
    return 0.5 * (1.0 - np.sqrt(-np.expm1(-w)))


def f_of_w(w):
    q = q_of_w(w)
    return -ndtri(q) / (1.0 - 2.0 * q)


def fit():
    w1 = np.linspace(1e-6, W_SPLIT, 400001)
    f1 = f_of_w(w1)
    p1 = np.polyfit(w1 - C_CENTRAL, f1, DEG_CENTRAL, w=1.0 / f1)
    e1 = np.max(np.abs(np.polyval(p1, w1 - C_CENTRAL) / f1 - 1.0))
    w2 = np.linspace(W_SPLIT, W_MAX, 400001)
    f2 = f_of_w(w2)
    p2 = np.polyfit(np.sqrt(w2) - C_TAIL, f2, DEG_TAIL, w=1.0 / f2)
    e2 = np.max(np.abs(np.polyval(p2, np.sqrt(w2) - C_TAIL) / f2 - 1.0))
    return p1, e1, p2, e2


def u_form(p1):
    """The central polynomial as a polynomial in u = log2(q(1-q)), the quantity
    the device computes first: w - C_CENTRAL = -ln2 (u - U0), U0 = -C_CENTRAL/ln2 - 2,
    so sum_k c_k x^k = sum_k c_k (-ln2)^k (u - U0)^k (two operations fewer on the
    device, same polynomial)."""
    ln2 = np.log(2.0)
    u0 = -C_CENTRAL / ln2 - 2.0
    deg = len(p1) - 1
    d = np.array([c * (-ln2) ** (deg - i) for i, c in enumerate(p1)])   # Horner order
    u_split = -W_SPLIT / ln2 - 2.0
    return u0, u_split, d


TAB_SUB = 4          # 2^4 segments per binade of t
TAB_DEG = 3
TAB_XBITS = 23 - TAB_SUB


def tab_exponent(row):
    """Biased FP32 exponent of t for a table row: the row index is bits 19..27 of
    t's FP32 encoding (the low 5 exponent bits and the top 4 mantissa bits), and t
    lies in [2^-33, 2^-2], i.e. biased exponents 94..125."""
    e5 = row >> TAB_SUB
    return e5 + 96 if e5 <= 29 else e5 + 64


def fit_table():
    """Piecewise cubic f(t) on 1/16 of a binade of t = q(1-q) each, in the
    segment's integer mantissa x = bits(t) & (2^19 - 1): the device reads the row
    from the bits of t, with no logarithm and no branch.  Row (E, j) covers
    t = 2^(E-127) (1 + (j + x 2^-19) / 16); the point t = 1/4 (q = 1/2 to FP32
    rounding, biased exponent 125) is the constant f(0) = sqrt(pi/2), and rows no
    t in [2^-33, 1/4] reaches repeat it."""
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
        a = np.polyfit(X, f, TAB_DEG, w=1.0 / f)[::-1]          # a_0 .. a_3 in X in [0, 1)
        worst = max(worst, np.max(np.abs(np.polyval(a[::-1], X) / f - 1.0)))
        tab[r] = a * 2.0 ** (-TAB_XBITS * np.arange(TAB_DEG + 1))   # in x = X 2^19
    return tab, worst


def main():
    p1, e1, p2, e2 = fit()
    u0, u_split, d = u_form(p1)
    tab, e3 = fit_table()
    out = os.path.join(os.path.dirname(__file__), "..", "csrc", "icdf_coeffs.h")
    with open(out, "w") as f:
        f.write("/* Generated by gen/fit_icdf.py -- do not edit.\n")
        f.write(" * Least-squares fit of f(w) = Phi^-1(1-q)/(1-2q), w = -log(4q(1-q)).\n")
        f.write(" * max relative fit error: central %.3e, tail %.3e (FP64 evaluation). */\n" % (e1, e2))
        f.write("#pragma once\n")
        f.write("#define FDMC_ICDF_W_SPLIT %.1ff\n" % W_SPLIT)
        f.write("#define FDMC_ICDF_C_CENTRAL %.17gf\n" % C_CENTRAL)
        f.write("#define FDMC_ICDF_C_TAIL %.17gf\n" % C_TAIL)
        f.write("#define FDMC_ICDF_DEG_CENTRAL %d\n" % DEG_CENTRAL)
        f.write("#define FDMC_ICDF_DEG_TAIL %d\n" % DEG_TAIL)
        f.write("/* Horner order: highest degree first. */\n")
        f.write("#define FDMC_ICDF_CENTRAL_COEFFS {" + ", ".join("%.9ef" % c for c in p1) + "}\n")
        f.write("#define FDMC_ICDF_TAIL_COEFFS {" + ", ".join("%.9ef" % c for c in p2) + "}\n")
        f.write("/* The central polynomial in u = log2(q(1-q)): w < W_SPLIT <=> u > U_SPLIT,\n")
        f.write(" * f = sum_k d_k (u - U0)^k with d_k = c_k (-ln 2)^k. */\n")
        f.write("#define FDMC_ICDF_U0 %.17gf\n" % u0)
        f.write("#define FDMC_ICDF_U_SPLIT %.17gf\n" % u_split)
        f.write("#define FDMC_ICDF_CENTRAL_COEFFS_U {" + ", ".join("%.9ef" % c for c in d) + "}\n")
        f.write("/* Table form (piecewise cubic in the bits of t = q(1-q)); row r = bits(t) >> 19\n")
        f.write(" * & 511 holds {c0, c1, c2, c3} of f = c0 + x (c1 + x (c2 + x c3)), x = bits(t) & 0x7FFFF.\n")
        f.write(" * max relative fit error %.3e (FP64 evaluation). */\n" % e3)
        f.write("#define FDMC_ICDF_TAB_ROWS %d\n" % len(tab))
        f.write("#define FDMC_ICDF_TAB { \\\n")
        for r in range(len(tab)):
            f.write("    " + ", ".join("%.9ef" % c for c in tab[r]) + (", \\\n" if r + 1 < len(tab) else " \\\n"))
        f.write("}\n")
    print("central err %.3e tail err %.3e table err %.3e -> %s" % (e1, e2, e3, out))


if __name__ == "__main__":
    main()
