"""Pins of the oracle's building blocks against closed forms, brute force and library routines."""
import numpy as np
import pytest
import scipy.signal

import fastiso


@pytest.mark.parametrize("n", [2, 3, 4, 5, 6, 7, 8, 11, 12, 13, 35, 64, 96, 175, 350, 700, 1024])
def test_fft_matches_numpy(ref, n):
    rng = np.random.default_rng(n)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    np.testing.assert_allclose(ref.fft(x, -1), np.fft.fft(x), rtol=0, atol=1e-12 * np.sqrt(n) * n)
    np.testing.assert_allclose(ref.fft(x, +1), np.fft.ifft(x) * n, rtol=0, atol=1e-12 * np.sqrt(n) * n)


def test_derivatives_closed_form(ref):
    """test/test_derivatives.jl:15-26: u = X^2 Y^2 on RegionalDomain(3e6, 7)."""
    d = fastiso.RegionalDomain(3e6, 7)
    X, Y = d.X, d.Y
    u = X ** 2 * Y ** 2
    uxx, uyy, ux, uxy = ref.derivatives(u, d.dx, d.dy)
    inn = (slice(1, -1), slice(1, -1))
    for got, want in ((uxx, 2 * Y ** 2), (uyy, 2 * X ** 2), (uxy, 4 * X * Y)):
        g, w = got[inn], want[inn]
        assert np.linalg.norm(g - w) <= 1e-8 * max(np.linalg.norm(g), np.linalg.norm(w))


def test_derivative_edge_rules(ref):
    """CPU edge rules (src/derivatives.jl:66-76,171-191): one-sided second differences are exact
    on quadratics, first differences at the edges are the forward/backward quotients."""
    d = fastiso.RegionalDomain(2e6, 2e6, 12, 9)
    X, Y = d.X / 1e6, d.Y / 1e6
    u = 3 * X ** 2 - 2 * X * Y + 5 * Y ** 2 + X - 7
    uxx, uyy, ux, uxy = ref.derivatives(u, d.dx / 1e6, d.dy / 1e6)
    np.testing.assert_allclose(uxx, 6.0, atol=1e-9)
    np.testing.assert_allclose(uyy, 10.0, atol=1e-9)
    hx = d.dx / 1e6
    exp_ux = 6 * X - 2 * Y + 1
    exp_ux[0, :] = (u[1, :] - u[0, :]) / hx
    exp_ux[-1, :] = (u[-1, :] - u[-2, :]) / hx
    np.testing.assert_allclose(ux, exp_ux, atol=1e-9)
    # interior cross derivative of a quadratic is exact away from the edges
    np.testing.assert_allclose(uxy[1:-1, 1:-1], -2.0, atol=1e-9)


def test_composed_flexure_operator(ref):
    """u = cos(a x): Mxx = -D uxx, Myy = -D nu uxx, Mxy = 0, so the composed operator is
    -D lam^2 u with lam = (2 cos(a dx) - 2)/dx^2 at cells >= 2 from the x edges."""
    d = fastiso.RegionalDomain(3e6, 3e6, 40, 24)
    a = 2 * np.pi * 3 / (2 * d.Wx)
    u = np.cos(a * d.X)
    D, nu = 1e24, 0.28
    uxx, uyy, ux, uxy = ref.derivatives(u, d.dx, d.dy)
    Mxx, Myy, Mxy = -D * (nu * uyy + uxx), -D * (nu * uxx + uyy), -D * (1 - nu) * uxy
    bxx = ref.derivatives(Mxx, d.dx, d.dy)[0]
    byy = ref.derivatives(Myy, d.dx, d.dy)[1]
    bxy = ref.derivatives(Mxy, d.dx, d.dy)[3]
    rhs = bxx + (2 * bxy + byy)
    lam = (2 * np.cos(a * d.dx) - 2) / d.dx ** 2
    np.testing.assert_allclose(rhs[2:-2, :], (-D * lam ** 2 * u)[2:-2, :], rtol=0,
                               atol=1e-9 * D * lam ** 2)


@pytest.mark.parametrize("n", [8, 15, 24])
def test_padded_convolution_vs_direct(ref, n):
    """test/test_convolution.jl:2-10 (conv! == DSP.conv): brute-force direct convolution."""
    rng = np.random.default_rng(100 + n)
    k, x = rng.random((n, n)), rng.random((n, n))
    full = ref.conv_cropped(k, x)
    direct = scipy.signal.convolve2d(k, x, mode="full")
    # independent brute force
    bf = np.zeros((2 * n - 1, 2 * n - 1))
    for i in range(n):
        for j in range(n):
            bf[i:i + n, j:j + n] += k[i, j] * x
    np.testing.assert_allclose(full, bf, rtol=0, atol=1e-12 * n * n)
    np.testing.assert_allclose(direct, bf, rtol=0, atol=1e-12 * n * n)


def test_samesize_conv_crop_and_bc(ref):
    """Crop [i1:i2, j1:j2] of the (2n-1)^2 convolution after subtracting its corner mean
    (src/convolutions.jl:134-143, src/boundary_conditions.jl:144-148)."""
    n = 10
    rng = np.random.default_rng(7)
    k, x = rng.random((n, n)), rng.random((n, n))
    bf = np.zeros((2 * n - 1, 2 * n - 1))
    for i in range(n):
        for j in range(n):
            bf[i:i + n, j:j + n] += k[i, j] * x
    bf -= (bf[0, 0] + bf[-1, 0] + bf[0, -1] + bf[-1, -1]) / 4 - 0.5
    i1, i2 = n // 2, 2 * n - 1 - n // 2  # even n, 1-based
    want = bf[i1 - 1:i2, i1 - 1:i2]
    np.testing.assert_allclose(ref.samesize_conv(k, x, 0.5), want, rtol=0, atol=1e-12)
    # a delta kernel at the centre (mx, my) reproduces the input, up to the corner offset
    dk = np.zeros((n, n))
    dk[n // 2 - 1, n // 2 - 1] = 1.0
    out = ref.samesize_conv(dk, x, 0.0)
    np.testing.assert_allclose(out, x, atol=1e-12)


def test_nextfastfft(ref):
    for n, want in ((127, 128), (699, 700), (191, 192), (2047, 2048), (121, 125), (143, 144)):
        assert ref.lib().ref_nextfastfft(n) == want == fastiso.nextfastfft(n)


def test_gauss_legendre_closed_form(ref):
    x, w = ref.gauss_legendre(4)
    a, b = np.sqrt(3 / 7 - 2 / 7 * np.sqrt(6 / 5)), np.sqrt(3 / 7 + 2 / 7 * np.sqrt(6 / 5))
    np.testing.assert_allclose(x, [-b, -a, a, b], atol=1e-15)
    wa, wb = (18 + np.sqrt(30)) / 36, (18 - np.sqrt(30)) / 36
    np.testing.assert_allclose(w, [wb, wa, wa, wb], atol=1e-15)
    x5, w5 = ref.gauss_legendre(5)
    c = np.sqrt(5 - 2 * np.sqrt(10 / 7)) / 3
    e = np.sqrt(5 + 2 * np.sqrt(10 / 7)) / 3
    np.testing.assert_allclose(x5, [-e, -c, 0, c, e], atol=1e-15)
    np.testing.assert_allclose(w5[2], 128 / 225, atol=1e-15)
    for n in (1, 2, 3, 8, 16):
        xr, wr = ref.gauss_legendre(n)
        xn, wn = np.polynomial.legendre.leggauss(n)
        np.testing.assert_allclose(xr, xn, atol=1e-14)
        np.testing.assert_allclose(wr, wn, atol=1e-14)


def test_pseudodiff_symbol(ref):
    """kappa = sqrt((pi/Wx f_i)^2 + (pi/Wy f_j)^2) with f = fftint, DC set to 1e-3 * mean of the
    first two modes (src/derivatives.jl:201-222, src/domain.jl:151-155)."""
    nx, ny, Wx, Wy = 12, 9, 3e6, 2e6
    p = ref.pseudodiff(nx, ny, Wx, Wy)
    fx = np.abs(np.fft.fftfreq(nx) * nx)
    fy = np.abs(np.fft.fftfreq(ny) * ny)
    want = np.sqrt((np.pi / Wx * fx[:, None]) ** 2 + (np.pi / Wy * fy[None, :]) ** 2)
    want[0, 0] = 1e-3 * (want[0, 1] + want[1, 0]) / 2
    np.testing.assert_allclose(p, want, rtol=1e-14)
    np.testing.assert_allclose(fastiso.RegionalDomain(Wx, Wy, nx, ny).pseudodiff, want, rtol=1e-14)


def test_geoid_green_closed_form(ref):
    d = fastiso.RegionalDomain(3e6, 3e6, 16, 16)
    c = fastiso.PhysicalConstants()
    G = ref.geoid_green(16, 16, 3e6, 3e6)
    unb = lambda R: c.r_pole / (2 * c.mE * np.sin(R / (2 * c.r_pole)))
    with np.errstate(divide="ignore"):
        want = np.minimum(unb(d.R), unb(d.dx / 2))
    np.testing.assert_allclose(G, want, rtol=1e-14)
    np.testing.assert_allclose(fastiso.geoid_green(d, c), want, rtol=1e-14)


def test_elastic_green_two_implementations_and_convergence(ref):
    rm, ge = fastiso.load_farrell_table()
    assert len(rm) == 42 and rm[0] == 0 and rm[1] == 11.0
    d = fastiso.RegionalDomain(3e6, 3e6, 24, 24)
    Gr = ref.elastic_green(24, 24, 3e6, 3e6, rm, ge, 4)
    Gp = fastiso.elastic_green(d, rm, ge, quad_precision=4)
    np.testing.assert_allclose(Gp, Gr, rtol=1e-12, atol=1e-30)
    # quadrature convergence: GL8 within 2 % of GL4 away from the singular centre cell
    G8 = ref.elastic_green(24, 24, 3e6, 3e6, rm, ge, 8)
    far = d.R > 3 * d.dx
    np.testing.assert_allclose(Gr[far], G8[far], rtol=2e-2)
    # far field beyond the table end is zero; the response is a depression (GE < 0 near field)
    assert Gr[12, 12] < 0


def _cell_integral_r2(x1, x2, y1, y2):
    """Closed form of the integral of x^2 + y^2 over [x1, x2] x [y1, y2]."""
    return (x2 ** 3 - x1 ** 3) / 3 * (y2 - y1) + (x2 - x1) * (y2 ** 3 - y1 ** 3) / 3


@pytest.mark.parametrize("n", [24, 25])
def test_elastic_green_cells_closed_form(ref, n):
    """get_elastic_green / get_loadgreen / quadrature2D (src/material.jl:299-364,
    src/utils.jl:146-186) on a synthetic table GE = c r^3 (dense, so its linear interpolant is
    r^3 to ~1e-7): the integrand GE(r) / (r 1e12) = c (x^2 + y^2) / 1e12 is a polynomial that
    4-point Gauss-Legendre integrates exactly, so every cell inside the table's range equals
    c / 1e12 times the closed-form integral over [p dx, (p+1) dx] x [q dy, (q+1) dy] with
    p = i - mx - 1, q = j - my - 1 (1-based i, j; mx = nx div 2). A one-cell shift, a lost
    1e12 or a wrong cell size fails; cells beyond rm[end] are zero."""
    W = 3e6
    d = fastiso.RegionalDomain(W, W, n, n)
    r_end = 2.2e6
    rm = np.linspace(0.0, r_end, 400001)
    c = 7.0e-6
    ge = c * rm ** 3
    for G in (fastiso.elastic_green(d, rm, ge, quad_precision=4), ref.elastic_green(n, n, W, W, rm, ge, 4)):
        p = np.arange(1, n + 1) - n // 2 - 1
        x1, x2 = p * d.dx, (p + 1) * d.dx
        want = c / 1e12 * _cell_integral_r2(x1[:, None], x2[:, None], x1[None, :], x2[None, :])
        # distance of the cell's farthest / nearest corner from the origin
        far = np.hypot(np.maximum(abs(x1), abs(x2))[:, None], np.maximum(abs(x1), abs(x2))[None, :])
        near = np.hypot(np.where(p >= 0, x1, -x2)[:, None], np.where(p >= 0, x1, -x2)[None, :])
        inside, outside = far < r_end, near > r_end
        assert inside.sum() > n and outside.sum() > n
        np.testing.assert_allclose(G[inside], want[inside], rtol=2e-6)
        assert np.all(G[outside] == 0.0)
        straddle = ~inside & ~outside
        assert np.all((G[straddle] >= 0) & (G[straddle] <= want[straddle] * (1 + 1e-6)))
        # the planted mistakes this test exists for
        assert not np.allclose(np.roll(G, 1, axis=0)[inside], want[inside], rtol=1e-3)
        assert not np.allclose(1e12 * G[inside], want[inside], rtol=1e-3)


def test_elastic_green_small_distance_branch(ref):
    """get_loadgreen below r = 0.01 m returns GE[1] / (rm[2] 1e12) (src/material.jl:323-324): on
    a 4-cm domain (5-mm cells) every node of the four centre cells lies below 0.01 m, so those
    cells equal that constant times the cell area; doubling GE[1] doubles them."""
    n, W = 8, 0.02
    d = fastiso.RegionalDomain(W, W, n, n)
    rm = np.array([0.0, 0.5, 1.0, 2.0])
    ge = np.array([-3.0, -2.0, -1.0, 0.5])
    p = np.arange(1, n + 1) - n // 2 - 1
    x1, x2 = p * d.dx, (p + 1) * d.dx
    far = np.hypot(np.maximum(abs(x1), abs(x2))[:, None], np.maximum(abs(x1), abs(x2))[None, :])
    small = far < 0.01
    assert small.sum() >= 4
    ge2 = ge.copy()
    ge2[0] = -6.0
    for table in (lambda g: fastiso.elastic_green(d, rm, g, quad_precision=4),
                  lambda g: ref.elastic_green(n, n, W, W, rm, g, 4)):
        G, G2 = table(ge), table(ge2)
        np.testing.assert_allclose(G[small], ge[0] / (rm[1] * 1e12) * d.dx * d.dy, rtol=1e-12)
        np.testing.assert_allclose(G2[small], 2 * G[small], rtol=1e-12)
        assert np.all(G[~small] != G[small][0])


def test_farrell_table_columns():
    """The shipped table: rm is the arc length of the angular distance on a 6371-km sphere
    (Farrell 1972, table A3; the reference reads rm and GE as columns 2 and 3,
    src/dataloaders.jl:257-273), GE is negative near the load and its published first rows."""
    d = np.loadtxt(fastiso.greens.os.path.join(fastiso.greens._DATA, "elasticgreencoeffs_farrell1972.csv"),
                   delimiter=",", skiprows=1)
    theta, rm, ge = d[:, 0], d[:, 1], d[:, 2]
    arc = np.deg2rad(theta) * 6.371e6
    np.testing.assert_allclose(rm[3:], arc[3:], rtol=3e-4)
    np.testing.assert_allclose(rm[1:3], arc[1:3], rtol=2e-2)
    assert rm[0] == 0 and theta[0] == 0 and np.all(np.diff(rm) > 0)
    assert ge[0] == -33.6488 and ge[1] == -33.64 and ge[2] == -33.56


def test_farrell_csv_equals_reference_jld2_bytes():
    """The CSV holds the float64 arrays of the reference's src/input/elasticgreencoeffs_farrell1972.jld2
    (read as data). Runs where the reference is mounted."""
    import os
    src = "/root/reference/src/input/elasticgreencoeffs_farrell1972.jld2"
    if not os.path.exists(src):
        pytest.skip("reference not mounted")
    raw = open(src, "rb").read()
    d = np.loadtxt(fastiso.greens.os.path.join(fastiso.greens._DATA, "elasticgreencoeffs_farrell1972.csv"),
                   delimiter=",", skiprows=1)
    for k, off in enumerate((622, 1024, 1426)):
        col = np.frombuffer(raw[off:off + 42 * 8], dtype="<f8")
        assert np.array_equal(col, d[:, k])


def test_solidearth_two_implementations(ref):
    d = fastiso.RegionalDomain(3e6, 3e6, 16, 16)
    rng = np.random.default_rng(3)
    h = 80e3 + 40e3 * rng.random((16, 16))
    lb = [h, 300e3 + 0 * h, 600e3 + 0 * h]
    lv = [1e20 + 0 * h, 10 ** (19 + rng.random((16, 16))), 1e21 + 0 * h]
    se = fastiso.SolidEarth(d, layer_boundaries=lb, layer_viscosities=lv)
    D, eta, R, spinv = ref.solidearth(16, 16, 3e6, 3e6, lb, lv)
    np.testing.assert_allclose(se.litho_rigidity, D, rtol=1e-14)
    np.testing.assert_allclose(se.effective_viscosity, eta, rtol=1e-14)
    np.testing.assert_allclose(se.pseudodiff_scaling, R, rtol=1e-12)
    np.testing.assert_allclose(se.scaled_pseudodiff_inv, spinv, rtol=1e-12)
    # closed forms: D = E h^3 / (12 (1 - nu^2)); compressible eta = 1.5/1.28 * eta_halfspace;
    # a single layer has R = 1
    np.testing.assert_allclose(D, 6.6e10 * h ** 3 / (12 * (1 - 0.28 ** 2)), rtol=1e-14)
    np.testing.assert_allclose(eta, 1e21 * 1.5 / 1.28, rtol=1e-14)
    se1 = fastiso.SolidEarth(d, layer_boundaries=[88e3], layer_viscosities=[1e21])
    np.testing.assert_allclose(se1.pseudodiff_scaling, 1.0)
    # channel scaling tends to 1 for a channel as viscous as the half-space (visc_ratio = 1)
    se2 = fastiso.SolidEarth(d, layer_boundaries=[88e3, 400e3], layer_viscosities=[1e21, 1e21])
    np.testing.assert_allclose(se2.pseudodiff_scaling, 1.0, rtol=1e-12)


def test_channel_scaling_limits(ref):
    """R(kappa) of one low-viscosity channel over the half-space (channel_scaling,
    src/material.jl:232-262) against what the formula's mathematics fixes, for the host set-up and
    the oracle's restatement: with x = kappa T and nu = visc_ratio, R = 1 + (2/3)(nu - 1/nu) x^3
    + O(x^4) for a thin channel (its Taylor series: numerator and denominator agree to second
    order), R -> nu for a thick one (cosh ~ sinh ~ e^x / 2), and R = 1 at nu = 1. A dropped or
    sign-flipped term changes the x, x^2 or x^3 coefficient or the thick limit."""
    from fastiso.solidearth import channel_scaling
    for nu in (0.1, 0.01, 3.0):
        x = np.array([1e-3, 3e-3, 1e-2])
        R = channel_scaling(x, 1.0, nu)
        np.testing.assert_allclose((R - 1) / x ** 3, 2.0 / 3.0 * (nu - 1 / nu), rtol=0.05)
        xl = np.array([15.0, 20.0, 30.0])
        np.testing.assert_allclose(channel_scaling(xl, 1.0, nu), nu, rtol=1e-6)
        np.testing.assert_allclose(channel_scaling(np.array([0.5, 2.0]), 1.0, 1.0), 1.0, rtol=1e-14)
    # the oracle's two-layer set-up: a channel of viscosity eta_c from h to h + T over the
    # half-space eta_h (visc_ratio = eta_c / eta_h; layer boundaries uniform)
    n, W = 16, 3e6
    d = fastiso.RegionalDomain(W, W, n, n)
    kappa = d.pseudodiff
    eta_c, eta_h = 1e19, 1e21
    nu = eta_c / eta_h
    for T, check in ((200.0, "thin"), (6e6, "thick")):
        lb = [np.full((n, n), 88e3), np.full((n, n), 88e3 + T)]
        lv = [np.full((n, n), eta_c), np.full((n, n), eta_h)]
        _, _, R, _ = ref.solidearth(n, n, W, W, lb, lv)
        x = kappa * T
        if check == "thin":
            assert x.max() < 1e-2
            ok = x > 1e-5   # (R - 1) well above rounding; the DC mode's kappa is 1e-3 of its neighbours'
            assert ok.sum() == n * n - 1
            np.testing.assert_allclose((R[ok] - 1) / x[ok] ** 3, 2.0 / 3.0 * (nu - 1 / nu), rtol=0.05)
        else:
            big = x > 15
            assert big.sum() > n * n // 2
            np.testing.assert_allclose(R[big], nu, rtol=1e-6)


def test_bs3_order_and_adaptivity(ref):
    lam, T = -0.5, 2.0
    errs = []
    for n in (20, 40, 80):
        u, _ = ref.bs3_linear(lam, 1.0, T, dt_fixed=T / n)
        errs.append(abs(u - np.exp(lam * T)))
    assert 7.0 < errs[0] / errs[1] < 9.0 and 7.0 < errs[1] / errs[2] < 9.0   # third order
    # one step of the tableau against the exact stability polynomial 1 + z + z^2/2 + z^3/6
    for z in (-0.3, -1.0, -2.0):
        u1, _ = ref.bs3_linear(z, 1.0, 1.0, dt_fixed=1.0)
        assert abs(u1 - (1 + z + z * z / 2 + z ** 3 / 6)) < 1e-15
    u, nst = ref.bs3_linear(lam, 1.0, 20.0, reltol=1e-8, abstol=1e-10)
    assert abs(u - np.exp(lam * 20)) < 1e-8 and nst > 10


def test_analytic_two_implementations(ref):
    d = fastiso.RegionalDomain(3e6, 3e6, 16, 16)
    se = fastiso.SolidEarth(d, layer_boundaries=[88e3], layer_viscosities=[1e21], rho_uppermantle=3300.0,
                            rho_litho=0.0)
    c = fastiso.PhysicalConstants()
    from fastiso.analytic import analytic_solution
    D, eta = float(se.litho_rigidity.mean()), float(se.effective_viscosity.mean())
    for t_yr in (100.0, 5e3, 50e3, np.inf):
        ts = t_yr * c.seconds_per_year
        for r in (0.0, 5e5, 9.5e5, 1.5e6):
            a = ref.analytic_solution(r, ts, c.rho_ice, c.g, 1e3, 1e6, 3300.0, D, eta)
            b = analytic_solution(r, ts, c, se, 1e3, 1e6)[0]
            assert abs(a - b) <= 1e-9 * max(1.0, abs(a))
    # equilibrium at the disc centre approaches the local isostatic depth rho_i H / rho_m
    ueq = ref.analytic_solution(0.0, np.inf, c.rho_ice, c.g, 1e3, 1e6, 3300.0, D, eta)
    assert -910 / 3300 * 1e3 * 1.02 < ueq < -910 / 3300 * 1e3 * 0.9
