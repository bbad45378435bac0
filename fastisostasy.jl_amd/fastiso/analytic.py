"""analytic_solution (src/analytic_solutions.jl:11-65): bedrock displacement of a cylindrical
load on an elastic plate over a viscous half-space (Bueler et al. 2007). t in seconds."""
import numpy as np
from scipy.special import j0, j1


def analytic_solution(r, t, c, solidearth, H0, R0, n_quad_support=5):
    x, w = np.polynomial.legendre.leggauss(n_quad_support)
    support = np.concatenate([[1e-14], 10.0 ** (-10 + 0.05 * np.arange(141)), [1.0]])
    a, b = support[:-1], support[1:]
    m = 2.0 / (b - a)
    p = -1.0 - m * a
    kappa = (x[None, :] - p[:, None]) / m[:, None]                       # (nint, nq)
    D = float(np.mean(solidearth.litho_rigidity))
    eta = float(np.mean(solidearth.effective_viscosity))
    beta = solidearth.rho_uppermantle * c.g + D * kappa ** 4
    r = np.atleast_1d(np.asarray(r, dtype=np.float64))
    out = np.empty_like(r)
    for k, rr in enumerate(r):
        if np.isinf(t):
            f = -j0(kappa * rr) * j1(kappa * R0) / beta
        else:
            f = (np.exp(-beta * t / (2 * eta * kappa)) - 1) * j0(kappa * rr) * j1(kappa * R0) / beta
        out[k] = np.sum(np.sum(f * w[None, :], axis=1) / m)
    return c.rho_ice * c.g * H0 * R0 * out
