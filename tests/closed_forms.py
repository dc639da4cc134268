"""Independent closed forms used as pins by the tests (FP64, scipy quadrature).

heston_call: Heston (1993) European call with the characteristic function in
the rotation-free form of Albrecher, Mayer, Schoutens and Tistaert ("The little
Heston trap", Wilmott 2007), priced by Gil-Pelaez inversion.
"""
import cmath
import math

from scipy import integrate


def heston_call(S0, K, T, r, kappa, theta, xi, rho, v0):
    lnK = math.log(K)

    def cf(u):  # E[exp(i u ln S_T)]
        iu = 1j * u
        d = cmath.sqrt((rho * xi * iu - kappa) ** 2 + xi * xi * (iu + u * u))
        g = (kappa - rho * xi * iu - d) / (kappa - rho * xi * iu + d)
        e = cmath.exp(-d * T)
        C = r * iu * T + kappa * theta / (xi * xi) * ((kappa - rho * xi * iu - d) * T - 2.0 * cmath.log((1 - g * e) / (1 - g)))
        D = (kappa - rho * xi * iu - d) / (xi * xi) * (1 - e) / (1 - g * e)
        return cmath.exp(C + D * v0 + iu * math.log(S0))

    cf_mi = cf(-1j)  # E[S_T]

    def p1(u):
        return (cmath.exp(-1j * u * lnK) * cf(u - 1j) / (1j * u * cf_mi)).real

    def p2(u):
        return (cmath.exp(-1j * u * lnK) * cf(u) / (1j * u)).real

    P1 = 0.5 + integrate.quad(p1, 1e-10, 200, limit=400)[0] / math.pi
    P2 = 0.5 + integrate.quad(p2, 1e-10, 200, limit=400)[0] / math.pi
    return S0 * P1 - K * math.exp(-r * T) * P2


def pca_uniform(m, T):
    """Increment matrix M (dW = M z, row-major [date][component]) of the PCA of a
    Brownian motion on m uniform dates over [0, T], from the closed-form
    eigenpairs of the covariance h min(i, j): lambda_k = h / (4 sin^2((2k-1) pi /
    (2 (2m+1)))), v_k(i) = 2 / sqrt(2m+1) sin((2k-1) i pi / (2m+1)), components in
    descending eigenvalue order; rows of the cumulative matrix V sqrt(Lambda)
    differenced into increments."""
    import numpy as np

    h = T / m
    k = np.arange(1, m + 1)
    lam = h / (4.0 * np.sin((2 * k - 1) * np.pi / (2 * (2 * m + 1))) ** 2)
    V = 2.0 / np.sqrt(2 * m + 1) * np.sin(np.outer(k, 2 * k - 1) * np.pi / (2 * m + 1))
    Mc = V * np.sqrt(lam)
    M = Mc.copy()
    M[1:] -= Mc[:-1]
    return M
