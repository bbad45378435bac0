"""Seeded synthetic workloads for the benchmark and parity configurations.

disc_benchmark  64x64 homogeneous Earth (eta 1e21 Pa s, 88 km lithosphere, rho_mantle 3300,
                rho_litho 0), 1000 km x 1000 m ice disc applied over the first year, 0-50 kyr
                (docs/src/examples/benchmark_analytic.jl:11-43); compared with analytic_solution.
lv_elva         N x N laterally variable Earth: log10 viscosity a Gaussian random field
                (correlation length 500 km) mapped to [19, 23], lithosphere thickness a smooth
                random field (1000 km) mapped to [50, 250] km, 3-6 Gaussian ice domes (0-4000 m)
                scaled by a sawtooth glacial cycle over 0-120 kyr with knots every 1 kyr; elastic,
                viscous and sea-surface responses on. N = 350 over a 6300 km square is dx = 18 km.
"""
import numpy as np

from .bcs import BoundaryConditions, SeparableIceThickness, TimeInterpolatedIceThickness
from .domain import PhysicalConstants, RegionalDomain
from .sealevel import LaterallyVariableSeaSurface, RegionalSeaLevel
from .simulation import NativeOutput, SolverOptions
from .solidearth import SolidEarth
from .utils import uniform_ice_cylinder


def gaussian_random_field(domain, seed, corr_length):
    """White N(0,1) noise filtered by exp(-kappa^2 l^2 / 4) in Fourier space, min-max scaled to
    [0, 1]."""
    rng = np.random.default_rng(seed)
    w = rng.standard_normal((domain.nx, domain.ny))
    kx = 2 * np.pi * np.fft.fftfreq(domain.nx, d=domain.dx)
    ky = 2 * np.pi * np.fft.fftfreq(domain.ny, d=domain.dy)
    k2 = kx[:, None] ** 2 + ky[None, :] ** 2
    f = np.real(np.fft.ifft2(np.fft.fft2(w) * np.exp(-k2 * corr_length ** 2 / 4)))
    return (f - f.min()) / (f.max() - f.min())


def ice_domes(domain, seed, inner=0.7, hmax=4000.0):
    """3-6 Gaussian domes: centres uniform in the inner 70 % of the domain, widths 200-600 km,
    peaks 1000-4000 m, total clipped to [0, hmax]."""
    rng = np.random.default_rng(seed)
    n = int(rng.integers(3, 7))
    H = np.zeros((domain.nx, domain.ny))
    for _ in range(n):
        cx = rng.uniform(-inner * domain.Wx, inner * domain.Wx)
        cy = rng.uniform(-inner * domain.Wy, inner * domain.Wy)
        s = rng.uniform(200e3, 600e3)
        a = rng.uniform(1000.0, 4000.0)
        H += a * np.exp(-((domain.X - cx) ** 2 + (domain.Y - cy) ** 2) / (2 * s ** 2))
    return np.clip(H, 0.0, hmax)


def sawtooth(t_end=120e3, t_peak=100e3, dt=1e3):
    """Glacial-cycle amplitude: 0 -> 1 linearly until t_peak, 1 -> 0 until t_end."""
    t = np.arange(0.0, t_end + dt / 2, dt)
    a = np.where(t <= t_peak, t / t_peak, (t_end - t) / (t_end - t_peak))
    return t, a


def disc_benchmark(n=64, W=3e6, t_end=50e3, t_out=None):
    domain = RegionalDomain(W, W, n, n)
    H1 = uniform_ice_cylinder(domain, 1e6, 1e3)
    it = TimeInterpolatedIceThickness([0.0, 1.0, t_end], [domain.zeros(), H1, H1], domain)
    bcs = BoundaryConditions(domain, ice_thickness=it)
    se = SolidEarth(domain, layer_boundaries=[88e3], layer_viscosities=[1e21], rho_uppermantle=3300.0,
                    rho_litho=0.0)
    sl = RegionalSeaLevel()
    if t_out is None:
        t_out = np.arange(1e3, t_end + 1, 1e3)
    nout = NativeOutput(t=t_out, vars=["u"])
    return dict(domain=domain, bcs=bcs, sealevel=sl, solidearth=se, t_span=(0.0, t_end), nout=nout,
                opts=SolverOptions(), c=PhysicalConstants())


def lv_elva(n=350, W=3.15e6, t_end=120e3, seeds=(1001, 1002, 1003), separable=True, t_out=None):
    domain = RegionalDomain(W, W, n, n)
    logeta = 19.0 + 4.0 * gaussian_random_field(domain, seeds[0], 500e3)
    h = 50e3 + 200e3 * gaussian_random_field(domain, seeds[1], 1000e3)
    se = SolidEarth(domain, layer_boundaries=[h], layer_viscosities=[10.0 ** logeta])
    Hd = ice_domes(domain, seeds[2])
    tk, ak = sawtooth(t_end=max(t_end, 120e3))
    if separable:
        it = SeparableIceThickness(tk, ak, Hd)
    else:
        it = TimeInterpolatedIceThickness(tk, [a * Hd for a in ak], domain)
    bcs = BoundaryConditions(domain, ice_thickness=it)
    sl = RegionalSeaLevel(surface=LaterallyVariableSeaSurface())
    if t_out is None:
        t_out = np.arange(1e3, t_end + 1, 1e3)
    nout = NativeOutput(t=t_out, vars=["u"])
    return dict(domain=domain, bcs=bcs, sealevel=sl, solidearth=se, t_span=(0.0, t_end), nout=nout,
                opts=SolverOptions(), c=PhysicalConstants(), logeta=logeta, h=h, H_dome=Hd)


def elra(n=64, W=3e6, t_end=50e3, tau=3e3, lithosphere="rigid", t_out=None, interactive=False):
    """ELRA (docs/src/examples/alternative_models.jl:10-33,44-70): a 1000 m ice cylinder of
    radius 1000 km switched on over the first year, RelaxedMantle, 0-50 kyr. tau="weak" is the
    2D relaxation time get_relaxation_time_weaker of a viscosity 10^(21 - g) with g a Gaussian of
    peak 1 and width W/4 at the centre; lithosphere "rigid" (the example) or "constant" (88 km,
    elastic response on); interactive adds the laterally variable, interactive sea level."""
    from .sealevel import InteractiveSealevelLoad
    from .solidearth import (LaterallyConstantLithosphere, RelaxedMantle, RigidLithosphere,
                             get_relaxation_time_weaker)
    domain = RegionalDomain(W, W, n, n)
    H1 = uniform_ice_cylinder(domain, 1e6, 1e3)
    it = TimeInterpolatedIceThickness([0.0, 1.0, t_end], [domain.zeros(), H1, H1], domain)
    bcs = BoundaryConditions(domain, ice_thickness=it)
    if isinstance(tau, str):
        s2 = (W / 4) ** 2
        g = np.exp(-0.5 * (domain.X ** 2 + domain.Y ** 2) / s2)
        tau = get_relaxation_time_weaker(10.0 ** (21.0 - g / g.max()))
    litho = RigidLithosphere() if lithosphere == "rigid" else LaterallyConstantLithosphere()
    se = SolidEarth(domain, tau=tau, lithosphere=litho, mantle=RelaxedMantle())
    sl = RegionalSeaLevel()
    if interactive:
        sl = RegionalSeaLevel(surface=LaterallyVariableSeaSurface(), load=InteractiveSealevelLoad())
    if t_out is None:
        t_out = [1e2, 3e2, 1e3, 3e3, 1e4, 2e4, 5e4]
    nout = NativeOutput(t=[x for x in t_out if x <= t_end], vars=["u"])
    return dict(domain=domain, bcs=bcs, sealevel=sl, solidearth=se, t_span=(0.0, t_end), nout=nout,
                opts=SolverOptions(), c=PhysicalConstants(), H1=H1, tau=tau)


def ensemble(nmem=256, W=3e6, t_end=20e3, sizes=(64, 96, 128, 192, 256), member_seed=2000, eta_seed=3000,
             litho_seed=4000, dome_seed=1003, separable=True):
    """Uneven ensemble: member m has side n_m drawn uniformly from sizes (seed member_seed + m), its
    own log10 viscosity and lithosphere fields as in lv_elva (seeds eta_seed + m, litho_seed + m),
    and the shared dome field (dome_seed) evaluated on its grid. All members span the same
    2W x 2W square."""
    out = []
    for m in range(nmem):
        n = int(np.random.default_rng(member_seed + m).choice(sizes))
        w = lv_elva(n=n, W=W, t_end=t_end, seeds=(eta_seed + m, litho_seed + m, dome_seed), separable=separable,
                    t_out=[])
        w["member"] = m
        out.append(w)
    return out


def elra_ensemble(nmem=256, W=3e6, t_end=50e3, sizes=(64, 96, 128, 192, 256), member_seed=2000):
    """Uneven ensemble of ELRA members: member m has side n_m drawn uniformly from sizes (seed
    member_seed + m) and the 2D relaxation time get_relaxation_time_weaker of a viscosity
    10^(21 - a_m g), g the Gaussian of elra(tau="weak") and a_m uniform in [0.5, 1.5] from the
    same seed; the ice cylinder of elra, a laterally constant 88 km lithosphere with the elastic
    response, and the laterally variable sea surface."""
    from .solidearth import get_relaxation_time_weaker
    out = []
    for m in range(nmem):
        rng = np.random.default_rng(member_seed + m)
        n = int(rng.choice(sizes))
        amp = float(rng.uniform(0.5, 1.5))
        d = RegionalDomain(W, W, n, n)
        g = np.exp(-0.5 * (d.X ** 2 + d.Y ** 2) / (W / 4) ** 2)
        tau = get_relaxation_time_weaker(10.0 ** (21.0 - amp * g / g.max()))
        w = elra(n=n, W=W, t_end=t_end, tau=tau, lithosphere="constant", t_out=[])
        w["sealevel"] = RegionalSeaLevel(surface=LaterallyVariableSeaSurface())
        w["member"] = m
        out.append(w)
    return out
