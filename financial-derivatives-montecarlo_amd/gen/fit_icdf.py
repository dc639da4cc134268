"""Fit the single-precision inverse normal CDF used by the path kernels.

The kernels fold u = (s + 1/2) 2^-32 to the lower half, q = min(u, 1 - u),
computed from the integer s in two FP32 operations, and need
g(q) = -Phi^{-1}(q) >= 0 (the sign comes from s).  g is read from a table of
cubics indexed directly by q's FP32 encoding: one row per 1/16 of a binade of q,
the row index being bits 19..31 of q's encoding (the biased exponent and the top
4 mantissa bits) minus the first row's, the variable the remaining 19 mantissa
bits under the exponent of 1.0, i.e. the float 1 + x 2^-23 in [1, 1 + 2^-4)
(one AND-OR, no integer-to-float conversion).  q lies in [2^-33, 1/2] (2^-33
for s = 0; FP32 rounds q to exactly 1/2 for s within 64 of 2^31), biased
exponents 94..126: 33 binades and 528 rows, the last binade holding only
q = 1/2, where g = 0.  No logarithm and
no branch: a 16-byte LDS read, three FMAs and a multiply by the caller's scale.
The cubic's variable is not centred, so its FP32 evaluation loses a little to
cancellation: 2.1e-7 at most against the exact g (max(1, g)-weighted).

The coefficients are NOT taken from any library: they are fitted here by
weighted least squares per row against scipy.special.ndtri evaluated in FP64
(weights 1 / max(1, g), the error measure the kernels are held to:
|error| <= 4e-7 max(1, |z|)), and written to ../csrc/icdf_coeffs.h.

Usage: python fit_icdf.py
"""
import os

import numpy as np
from scipy.special import ndtri

TAB_SUB = 4           # 2^4 segments per binade of q
TAB_DEG = 3
E_MIN, E_MAX = 94, 126  # biased FP32 exponents of q


def fit_table():
    """Row r = (E - E_MIN) 16 + j covers q = 2^(E-127) (1 + (j + x 2^-19) / 16)
    for the integer x in [0, 2^19); its cubic is fitted in the kernels' variable
    V = 1 + x 2^-23 = 1 + X / 16, X = x 2^-19 in [0, 1).  The binade of
    E_MAX = 126 holds only q = 1/2: its rows are zero (g(1/2) = 0)."""
    rows = (E_MAX - E_MIN + 1) << TAB_SUB
    tab = np.zeros((rows, TAB_DEG + 1))
    worst = 0.0
    X = 0.5 - 0.5 * np.cos(np.linspace(0.0, np.pi, 513))     # Chebyshev-spaced, both ends
    X = np.minimum(X, 1.0 - 1e-9)
    for r in range(rows):
        E = E_MIN + (r >> TAB_SUB)
        j = r & ((1 << TAB_SUB) - 1)
        if E == E_MAX:
            continue
        q = 2.0 ** (E - 127) * (1.0 + (j + X) / (1 << TAB_SUB))
        g = -ndtri(q)
        w = 1.0 / np.maximum(1.0, g)
        V = 1.0 + X / (1 << TAB_SUB)
        a = np.polyfit(V, g, TAB_DEG, w=w)[::-1]                  # a_0 .. a_3 in V
        worst = max(worst, np.max(np.abs(np.polyval(a[::-1], V) - g) * w))
        tab[r] = a
    return tab, worst


def main():
    tab, err = fit_table()
    out = os.path.join(os.path.dirname(__file__), "..", "csrc", "icdf_coeffs.h")
    with open(out, "w") as f:
        f.write("/* Generated by gen/fit_icdf.py -- do not edit.\n")
        f.write(" * Piecewise-cubic table of g = -Phi^-1(q), q in [2^-33, 1/2], in the bits of q:\n")
        f.write(" * row r = (bits(q) >> 19) - FDMC_ICDF_ROW0 holds {c0, c1, c2, c3} of\n")
        f.write(" * g = c0 + v (c1 + v (c2 + v c3)), v = float(bits(q) & 0x7FFFF | bits(1.0f)).\n")
        f.write(" * max weighted fit error |g - p| / max(1, g) %.3e (FP64 evaluation). */\n" % err)
        f.write("#pragma once\n")
        f.write("#define FDMC_ICDF_TAB_ROWS %d\n" % len(tab))
        f.write("#define FDMC_ICDF_ROW0 %du   /* (biased exponent 94) x 16 */\n" % (E_MIN << TAB_SUB))
        f.write("#define FDMC_ICDF_TAB { \\\n")
        for r in range(len(tab)):
            f.write("    " + ", ".join("%.9ef" % c for c in tab[r]) + (", \\\n" if r + 1 < len(tab) else " \\\n"))
        f.write("}\n")
    print("table err %.3e -> %s" % (err, out))


if __name__ == "__main__":
    main()
