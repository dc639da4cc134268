"""Write the Joe-Kuo Sobol direction-number table as a C include.

Source data: S. Joe and F. Y. Kuo, "Constructing Sobol sequences with better
two-dimensional projections", SIAM J. Sci. Comput. 30 (2008) 2635-2654, file
``new-joe-kuo-6.21201`` (public).  The copy read here is the plain data archive
``scipy/stats/_sobol_direction_numbers.npz`` shipped with SciPy (arrays ``poly``
and ``vinit``; dimension 0 is van der Corput).  Only the data archive is read.

The emitted include holds, for dimensions 0 .. FDMC_JK_NDIM-1, the primitive
polynomial (with its leading and trailing 1 bits, as in the archive) and the
initial direction integers m_1..m_s.  Expanding them into 32 direction words per
dimension is done by the library and by the oracle at start-up (Bratley & Fox
recurrence), so the include stays small.

The reference bounds Sobol dimensions by ``D_MAX_SOBOL = 20000`` (models.hpp:9)
and takes its table from the JOEKUO6 set (methods_cuda.cu:764-775).

Usage: python gen_joe_kuo_table.py  (rewrites ../csrc/joe_kuo6_table.inc)
"""
import os

import numpy as np
import scipy.stats

NDIM = 20000


def main():
    path = os.path.join(os.path.dirname(scipy.stats.__file__), "_sobol_direction_numbers.npz")
    with np.load(path, allow_pickle=False) as z:
        poly = z["poly"].astype(np.int64)[:NDIM]
        vinit = z["vinit"].astype(np.int64)[:NDIM]
    offs = [0]
    mvals = []
    for d in range(NDIM):
        s = int(poly[d]).bit_length() - 1
        if d == 0:
            s = 0
        for j in range(s):
            m = int(vinit[d, j])
            assert m % 2 == 1 and m < (1 << (j + 1)), (d, j, m)
            mvals.append(m)
        offs.append(len(mvals))
    out = os.path.join(os.path.dirname(__file__), "..", "csrc", "joe_kuo6_table.inc")
    with open(out, "w") as f:
        f.write("/* Generated by gen/gen_joe_kuo_table.py -- do not edit.\n")
        f.write(" * Joe & Kuo new-joe-kuo-6.21201 direction numbers (public data),\n")
        f.write(" * dimensions 0..%d. Dimension 0 is van der Corput (degree 0).\n" % (NDIM - 1))
        f.write(" * fdmc_jk_poly[d]: primitive polynomial bits incl. leading/trailing 1.\n")
        f.write(" * fdmc_jk_m[fdmc_jk_off[d] .. fdmc_jk_off[d+1]): m_1..m_s. */\n")
        f.write("#define FDMC_JK_NDIM %d\n" % NDIM)
        f.write("static const unsigned int fdmc_jk_poly[%d] = {\n" % NDIM)
        for i in range(0, NDIM, 12):
            f.write(",".join(str(int(v)) for v in poly[i:i + 12]) + ",\n")
        f.write("};\n")
        f.write("static const unsigned int fdmc_jk_off[%d] = {\n" % (NDIM + 1))
        for i in range(0, NDIM + 1, 12):
            f.write(",".join(str(v) for v in offs[i:i + 12]) + ",\n")
        f.write("};\n")
        f.write("static const unsigned int fdmc_jk_m[%d] = {\n" % len(mvals))
        for i in range(0, len(mvals), 16):
            f.write(",".join(str(v) for v in mvals[i:i + 16]) + ",\n")
        f.write("};\n")
    print("wrote", out, "dims", NDIM, "m values", len(mvals))


if __name__ == "__main__":
    main()
