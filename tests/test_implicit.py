"""Implicit step of a Maxwell mantle under a rigid or laterally constant lithosphere
(src/deformation.jl:52-89), run by Euler with dt = dt_min (src/simulation.jl:244-253,
docs/src/examples/benchmark_analytic.jl:66-80).

Each call of update_diagnostics! replaces u by one trapezoidal step of dt_min on the n x n
spectrum and leaves dudt at zero, so Euler's u = uprev + dt * 0 followed by the FSAL call
advances one implicit step per integrator step (and one more at each output callback).

Pins of the oracle:
* a cosine load on a flexure-free Earth: the mode relaxes as u*(1 - e^{-t/tau}) with the
  isostatic amplitude u* = -rho_ice A / rho_mantle and the half-space relaxation time
  tau = 2 eta kappa / (rho_mantle g) (the trapezoidal rule is second order, error ~ N (dt/tau)^3);
* the disc benchmark of the reference's docs against analytic_solution, at the tolerance of the
  adaptive solver's test.
The HIP engine is compared with the oracle call by call and over integrations.
"""
import numpy as np
import pytest

import fastiso
from fastiso import workloads
from fastiso.analytic import analytic_solution

from _cases import hip_sim, sim_inputs


def cosine_implicit_case(n=16, k=2, A=300.0, H0=1000.0, eta=1e21, dt_min=100.0, t_end=5000.0):
    d = fastiso.RegionalDomain(3e6, 3e6 * 12 / 16, n, 12)
    se = fastiso.SolidEarth(d, lithosphere=fastiso.RigidLithosphere(), layer_boundaries=[np.zeros((n, 12))],
                            layer_viscosities=[eta])
    H1 = H0 + A * np.cos(2 * np.pi * k * d.X / (2 * d.Wx))
    it = fastiso.TimeInterpolatedIceThickness([0.0, 1.0, t_end], [d.zeros(), H1, H1])
    bcs = fastiso.BoundaryConditions(d, ice_thickness=it)
    opts = fastiso.SolverOptions(diffeq=fastiso.DiffEqOptions(alg="Euler", dt_min=dt_min))
    return dict(domain=d, bcs=bcs, sealevel=fastiso.RegionalSeaLevel(), solidearth=se, t_span=(0.0, t_end),
                nout=fastiso.NativeOutput(), opts=opts, c=fastiso.PhysicalConstants(), k=k, A=A)


def _mode(u, d, k):
    return 2.0 / u.size * np.sum(u * np.cos(2 * np.pi * k * d.X / (2 * d.Wx)))


def test_oracle_implicit_cosine_mode_relaxation():
    from oracle import pyref
    w = cosine_implicit_case()
    d, c, se = w["domain"], w["c"], w["solidearth"]
    ref = pyref.RefEngine.from_inputs(sim_inputs(w).inputs)
    ref.integrator_init()
    kap = 2 * np.pi * w["k"] / (2 * d.Wx)
    eta = float(se.effective_viscosity[0, 0])
    tau = 2 * eta * kap / (se.rho_uppermantle * c.g) / c.seconds_per_year
    ustar = -c.rho_ice * w["A"] / se.rho_uppermantle
    for t in (1000.0, 3000.0, 5000.0):
        ref.integrate(t)
        got = _mode(ref.get_field("u"), d, w["k"])
        want = ustar * (1 - np.exp(-t / tau))
        assert abs(got - want) <= 1e-4 * abs(ustar), (t, got, want)
    assert ref.stats()["n_accept"] == 50
    np.testing.assert_array_equal(ref.get_field("dudt"), 0.0)


def test_dt_min_needs_euler_and_the_implicit_step_needs_dt_min():
    from oracle import pyref
    w = cosine_implicit_case()
    w["opts"] = fastiso.SolverOptions(diffeq=fastiso.DiffEqOptions(alg="BS3", dt_min=100.0))
    with pytest.raises(ValueError, match="Euler"):
        sim_inputs(w)
    w["opts"] = fastiso.SolverOptions()
    with pytest.raises(RuntimeError, match="dt_min"):
        pyref.RefEngine.from_inputs(sim_inputs(w).inputs)


def disc_implicit_case(n=64, t_end=50e3, lithosphere="rigid"):
    w = workloads.disc_benchmark(n=n, t_end=t_end, t_out=np.arange(5e3, t_end + 1, 5e3))
    d = w["domain"]
    litho = fastiso.RigidLithosphere() if lithosphere == "rigid" else fastiso.LaterallyConstantLithosphere()
    w["solidearth"] = fastiso.SolidEarth(d, lithosphere=litho, layer_boundaries=[88e3], layer_viscosities=[1e21],
                                         rho_uppermantle=3300.0, rho_litho=0.0)
    w["opts"] = fastiso.SolverOptions(diffeq=fastiso.DiffEqOptions(alg="Euler", dt_min=100.0))
    return w


def test_oracle_implicit_disc_vs_analytic():
    """docs/src/examples/benchmark_analytic.jl:71-80 (RigidLithosphere, Euler, dt_min = 100)."""
    from oracle import pyref
    w = disc_implicit_case()
    d, c, se = w["domain"], w["c"], w["solidearth"]
    ref = pyref.RefEngine.from_inputs(sim_inputs(w).inputs)
    out = []
    ref.integrator_init()
    ref.integrate(w["t_span"][1], lambda t, k: out.append((t, ref.get_field("u"))))
    assert [t for t, _ in out] == list(np.arange(5e3, 50e3 + 1, 5e3))
    j = d.my - 1
    for t, u in out:
        ua = analytic_solution(np.abs(d.x), t * c.seconds_per_year, c, se, 1e3, 1e6)
        assert np.abs(u[:, j] - ua).max() < 0.05 * np.abs(ua).max() + 5.0, t


@pytest.mark.gpu
@pytest.mark.parametrize("lithosphere", ["rigid", "constant"])
def test_gpu_implicit_step_vs_oracle(lithosphere):
    from oracle import pyref
    w = disc_implicit_case(n=48, t_end=20e3, lithosphere=lithosphere)
    sim = hip_sim(w)
    ref = pyref.RefEngine.from_inputs(sim.inputs)
    d = w["domain"]
    u = -50 * np.exp(-(d.R / 1e6) ** 2)
    for t in (0.0, 100.0, 200.0, 300.0):   # each call replaces u by one implicit step
        ug = sim.engine.rhs(t, u)[1]
        ur = ref.rhs(t, u)[1]
        assert np.abs(ug - ur).max() <= 2e-6 * np.abs(ur).max(), t
        u = ur
    sim2, ref2 = hip_sim(w), pyref.RefEngine.from_inputs(sim.inputs)
    og, orf = [], []
    sim2.engine.integrator_init()
    ref2.integrator_init()
    sim2.engine.integrate(w["t_span"][1], lambda t, k: og.append((t, sim2.engine.get_field("u"))))
    ref2.integrate(w["t_span"][1], lambda t, k: orf.append((t, ref2.get_field("u"))))
    assert [t for t, _ in og] == [t for t, _ in orf]
    for (t, a), (_, b) in zip(og, orf):
        assert np.abs(a - b).max() <= 2e-5 * np.abs(b).max(), t
    assert sim2.engine.stats()["n_accept"] == ref2.stats()["n_accept"]


@pytest.mark.gpu
def test_gpu_implicit_step_with_sea_level_and_distortion_vs_oracle():
    """Laterally constant lithosphere with the elastic response, interactive sea level on a
    marine margin and map distortion (K^2 on the load of the implicit step)."""
    from oracle import pyref
    w = disc_implicit_case(n=40, t_end=20e3, lithosphere="constant")
    d0 = w["domain"]
    r = d0.R / d0.R.max()
    d = fastiso.RegionalDomain(d0.Wx, d0.Wy, d0.nx, d0.ny, K=0.95 + 0.1 * r ** 2)
    w["domain"] = d
    H1 = fastiso.uniform_ice_cylinder(d, 1e6, 1e3)
    w["bcs"] = fastiso.BoundaryConditions(
        d, ice_thickness=fastiso.TimeInterpolatedIceThickness([0.0, 1.0, 20e3], [d.zeros(), H1, H1]))
    w["solidearth"] = fastiso.SolidEarth(d, lithosphere=fastiso.LaterallyConstantLithosphere(),
                                         layer_boundaries=[88e3], layer_viscosities=[1e21], maskactive=r < 0.55)
    w["sealevel"] = fastiso.RegionalSeaLevel(surface=fastiso.LaterallyVariableSeaSurface(),
                                             load=fastiso.InteractiveSealevelLoad())
    w["z_b_ref"] = 500.0 - 1500.0 * r
    sim = hip_sim(w)
    ref = pyref.RefEngine.from_inputs(sim.inputs)
    sim.engine.integrator_init()
    ref.integrator_init()
    for t in (500.0, 2000.0):
        sim.engine.integrate(t)
        ref.integrate(t)
        a, b = sim.engine.get_field("u"), ref.get_field("u")
        assert np.abs(a - b).max() <= 2e-5 * np.abs(b).max(), t
    assert np.abs(ref.get_field("canom_seawater")).max() > 0


@pytest.mark.gpu
def test_gpu_implicit_checkpoint_restart_is_bitwise():
    w = disc_implicit_case(n=32, t_end=20e3)
    e = hip_sim(w).engine
    e.integrator_init()
    e.integrate(3000.0)
    blob = e.checkpoint()
    e.integrate(7000.0)
    ua, sa = e.get_field("u"), e.stats()
    e2 = hip_sim(w).engine
    e2.restore(blob)
    e2.integrate(7000.0)
    np.testing.assert_array_equal(e2.get_field("u"), ua)
    assert e2.stats()["n_accept"] == sa["n_accept"]


@pytest.mark.gpu
@pytest.mark.parametrize("nx,ny", [(40, 30), (35, 28)])
def test_gpu_implicit_non_square_grid_vs_oracle(nx, ny):
    from oracle import pyref
    d = fastiso.RegionalDomain(3e6, 3e6 * ny / nx, nx, ny)
    H1 = fastiso.uniform_ice_cylinder(d, 1e6, 1e3)
    it = fastiso.TimeInterpolatedIceThickness([0.0, 1.0, 20e3], [d.zeros(), H1, H1])
    se = fastiso.SolidEarth(d, lithosphere=fastiso.LaterallyConstantLithosphere(), layer_boundaries=[88e3],
                            layer_viscosities=[1e21])
    w = dict(domain=d, bcs=fastiso.BoundaryConditions(d, ice_thickness=it),
             sealevel=fastiso.RegionalSeaLevel(surface=fastiso.LaterallyVariableSeaSurface()), solidearth=se,
             t_span=(0.0, 20e3), nout=fastiso.NativeOutput(t=[1000.0, 2000.0]),
             opts=fastiso.SolverOptions(diffeq=fastiso.DiffEqOptions(alg="Euler", dt_min=100.0)),
             c=fastiso.PhysicalConstants())
    sim = hip_sim(w)
    ref = pyref.RefEngine.from_inputs(sim.inputs)
    sim.engine.integrator_init()
    ref.integrator_init()
    for t in (1000.0, 2000.0):
        sim.engine.integrate(t)
        ref.integrate(t)
        a, b = sim.engine.get_field("u"), ref.get_field("u")
        assert np.abs(a - b).max() <= 2e-5 * np.abs(b).max(), t


def test_oracle_set_time_is_refused_under_the_implicit_step():
    from oracle import pyref
    ref = pyref.RefEngine.from_inputs(sim_inputs(cosine_implicit_case()).inputs)
    ref.integrator_init()
    with pytest.raises(RuntimeError, match="implicit"):
        ref.set_time(500.0)


@pytest.mark.gpu
def test_gpu_implicit_bench_workload_n350_vs_oracle():
    """The N=350 implicit-step bench workload (the docs' disc on a 350^2 grid, Euler with
    dt_min = 100 yr) over 2 kyr: fixed steps, so both back-ends take the same 20 steps."""
    from oracle import pyref
    w = disc_implicit_case(n=350, t_end=1e6)
    w["nout"] = fastiso.NativeOutput(t=[1000.0, 2000.0])
    sim = hip_sim(w)
    ref = pyref.RefEngine.from_inputs(sim.inputs)
    og, orf = [], []
    sim.engine.integrator_init()
    ref.integrator_init()
    sim.engine.integrate(2000.0, lambda t, k: og.append((t, sim.engine.get_field("u"))))
    ref.integrate(2000.0, lambda t, k: orf.append((t, ref.get_field("u"))))
    for (t, a), (_, b) in zip(og, orf):
        e = np.abs(a - b).max() / np.abs(b).max()
        print(f"t={t} u max-rel {e:.2e}")
        assert e <= 2e-5, (t, e)
    assert sim.engine.stats()["n_accept"] == ref.stats()["n_accept"]
