"""Green's-function tables (host setup, FP64): elastic (Farrell 1972) and sea-surface kernels."""
import os

import numpy as np

_DATA = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data")


def gauss_legendre(n):
    """Gauss-Legendre nodes and weights on [-1, 1], ascending (FastGaussQuadrature.gausslegendre,
    used by get_quad_coeffs, src/utils.jl:118-121)."""
    x, w = np.polynomial.legendre.leggauss(int(n))
    return x.astype(np.float64), w.astype(np.float64)


def load_farrell_table():
    """Distance rm (m) and coefficient GE of Farrell (1972) table A3 (src/dataloaders.jl:257-273),
    from the copy the reference ships in src/input/elasticgreencoeffs_farrell1972.jld2."""
    d = np.loadtxt(os.path.join(_DATA, "elasticgreencoeffs_farrell1972.csv"), delimiter=",", skiprows=1)
    return d[:, 1].copy(), d[:, 2].copy()


def loadgreen(r, rm, ge):
    """get_loadgreen with a linear interpolant in rm (src/material.jl:299-332), vectorised."""
    r = np.asarray(r, dtype=np.float64)
    out = np.interp(r, rm, ge) / (r * 1e12 + (r == 0))
    out = np.where(r < 0.01, ge[0] / (rm[1] * 1e12), out)
    out = np.where(r > rm[-1], 0.0, out)
    return out


def elastic_green(domain, rm=None, ge=None, quad_precision=4):
    """get_elastic_green (src/material.jl:340-364) with quadrature2D (src/utils.jl:146-186):
    cell (i, j) integrates g(r) over [p dx, (p+1) dx] x [q dy, (q+1) dy], p = i - mx - 1,
    q = j - my - 1 (1-based i, j)."""
    if rm is None:
        rm, ge = load_farrell_table()
    xq, wq = gauss_legendre(quad_precision)
    nx, ny, dx, dy = domain.nx, domain.ny, domain.dx, domain.dy
    p = np.arange(1, nx + 1) - domain.mx - 1
    q = np.arange(1, ny + 1) - domain.my - 1
    x1, x2 = p * dx, (p + 1) * dx
    y1, y2 = q * dy, (q + 1) * dy
    mxl = 2.0 / (x2 - x1)
    pxl = -1.0 - mxl * x1
    myl = 2.0 / (y2 - y1)
    pyl = -1.0 - myl * y1
    xx = (xq[None, :] - pxl[:, None]) / mxl[:, None]          # (nx, nq)
    yy = (xq[None, :] - pyl[:, None]) / myl[:, None]          # (ny, nq)
    r = np.sqrt(xx[:, None, :, None] ** 2 + yy[None, :, None, :] ** 2)   # (nx, ny, nq, nq)
    f = loadgreen(r, rm, ge)
    w2 = wq[:, None] * wq[None, :]
    return np.einsum("ijab,ab->ij", f, w2) / (mxl[:, None] * myl[None, :])


def unbounded_dz_ssgreen(R, c):
    """src/sealevel.jl:217-219."""
    with np.errstate(divide="ignore"):
        return c.r_pole / (2.0 * c.mE * np.sin(R / (2.0 * c.r_pole)))


def geoid_green(domain, c):
    """get_dz_ss_green (src/sealevel.jl:208-215)."""
    g = unbounded_dz_ssgreen(domain.R, c)
    return np.minimum(g, unbounded_dz_ssgreen(domain.dx / 2, c))


def green_viscous(domain, rho, D):
    """green_viscous(domain, rho, D) (src/material.jl:370-380): the ELRA response to a point load,
    -L^2 / (2 pi D) kei(r / L) dx dy with the flexural length L = (D / (9.81 rho))^(1/4)
    (src/material.jl:396-399, g fixed at 9.81 there) and r = max(R, 1). besselkei(x) =
    Im K0(x e^{i pi/4}) is the Kelvin function kei (scipy.special.kei)."""
    from scipy.special import kei
    from .solidearth import get_flexural_lengthscale
    L = get_flexural_lengthscale(D, rho, 9.81)
    R = np.maximum(domain.R, 1.0)
    return -(L ** 2 / (2 * np.pi * D) * kei(R / L)) * (domain.dx * domain.dy)
