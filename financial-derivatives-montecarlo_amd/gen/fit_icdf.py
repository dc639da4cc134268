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


def main():
    p1, e1, p2, e2 = fit()
    u0, u_split, d = u_form(p1)
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
    print("central err %.3e tail err %.3e -> %s" % (e1, e2, out))


if __name__ == "__main__":
    main()
