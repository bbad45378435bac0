"""RHS-level and trajectory pins of the oracle."""
import numpy as np
import pytest

import fastiso
from fastiso import workloads
from fastiso.analytic import analytic_solution

from _cases import corner_mean, cosine_case, lv_case, sim_inputs


def _expected_cosine_dudt(w, u_amp):
    d, c, se = w["domain"], w["c"], w["solidearth"]
    eta = w["eta"] * 1.5 / 1.28
    fx = -c.g * c.rho_ice * w["A"] * np.cos(2 * np.pi * w["k"] * d.X / (2 * d.Wx)) * (d.Wx / (np.pi * w["k"]))
    fy = -c.g * se.rho_uppermantle * u_amp * np.cos(2 * np.pi * w["l"] * d.Y / (2 * d.Wy)) * (d.Wy / (np.pi * w["l"]))
    F = c.seconds_per_year / (2 * eta) * (fx + fy)
    return F - corner_mean(F)


def test_rhs_spectral_closed_form(ref):
    """Without flexure, dudt = spy/(2 eta) IFFT(FFT(-g sigma)/kappa) minus its corner mean:
    cosine modes of the ice and mantle loads are scaled by 1/kappa_k exactly
    (src/deformation.jl:91-103,110-131)."""
    w = cosine_case()
    e = ref.RefEngine.from_inputs(sim_inputs(w).inputs)
    d = w["domain"]
    u = 7.0 * np.cos(2 * np.pi * w["l"] * d.Y / (2 * d.Wy)) + 3.0
    du, ubc = e.rhs(1.0, u)
    want = _expected_cosine_dudt(w, 7.0)
    np.testing.assert_allclose(du, want, rtol=0, atol=1e-10 * np.abs(want).max())
    np.testing.assert_allclose(ubc, u - corner_mean(u), atol=1e-12)   # the stage vector is mutated
    assert abs(corner_mean(du)) < 1e-9 * np.abs(du).max()


def test_rhs_linearity_and_stale_loads(ref):
    """The sparse update reads the previous call's column anomalies (src/simulation.jl:355-385):
    the first call at t0 sees zero loads, so ue stays 0; a later trigger sees the loads of the call
    before it."""
    w = lv_case(n=16)
    e = ref.RefEngine.from_inputs(sim_inputs(w).inputs)
    u0 = np.zeros((16, 16))
    e.rhs(0.0, u0)
    assert np.all(e.get_field("ue") == 0) and e.stats()["count_sparse_updates"] == 1
    e.rhs(500.0, u0)   # t/10 >= 1: triggers with the loads of the call at t = 0 (zero ice)
    assert np.all(e.get_field("ue") == 0)
    load_500 = e.get_field("canom_load").copy()
    e.rhs(600.0, u0)   # triggers again, now with the loads of t = 500
    ue = e.get_field("ue")
    assert np.abs(ue).max() > 0
    # ue is linear in the load: halving the load halves ue (fresh context, same sequence)
    e2 = ref.RefEngine.from_inputs(sim_inputs(w).inputs)
    e2.rhs(0.0, u0)
    e2.rhs(500.0, u0)
    np.testing.assert_allclose(e2.get_field("canom_load"), load_500)


@pytest.fixture(scope="module")
def c0_run():
    from oracle import pyref
    w = workloads.disc_benchmark(n=64)
    e = pyref.RefEngine.from_inputs(sim_inputs(w).inputs)
    e.integrator_init()
    outs = []
    e.integrate(w["t_span"][1], lambda t, k: outs.append((t, e.get_field("u"))))
    return w, e, outs


def test_disc_benchmark_vs_analytic(c0_run):
    """64x64 disc load (docs/src/examples/benchmark_analytic.jl) against analytic_solution
    (src/analytic_solutions.jl:11-65) on the x transect through the centre."""
    w, e, outs = c0_run
    d, c, se = w["domain"], w["c"], w["solidearth"]
    assert len(outs) == 50 and outs[-1][0] == 50e3
    j = d.my - 1
    for t, u in outs:
        ua = analytic_solution(np.abs(d.x), t * c.seconds_per_year, c, se, 1e3, 1e6)
        err = np.abs(u[:, j] - ua)
        assert err.max() < 0.05 * np.abs(ua).max() + 5.0, (t, err.max())
        if t >= 10e3:
            assert abs(u[d.mx - 1, j] - ua[d.mx - 1]) < 0.5, t
    st = e.stats()
    assert st["n_accept"] > 50 and st["t"] == 50e3


def test_disc_benchmark_mass_sign(c0_run):
    w, e, outs = c0_run
    u = outs[-1][1]
    assert u.min() < -250 and u.min() > -300       # ~ rho_i/rho_m * 1 km depression


def test_oracle_checkpoint_restart_is_bitwise(ref):
    from _cases import lv_case, sim_inputs
    w = lv_case(n=24, t_end=300.0)
    inp = sim_inputs(w).inputs
    e = ref.RefEngine.from_inputs(inp)
    e.integrator_init()
    e.integrate(120.0)
    blob = e.checkpoint()
    e.integrate(300.0)
    e2 = ref.RefEngine.from_inputs(inp)
    e2.restore(blob)
    e2.integrate(300.0)
    np.testing.assert_array_equal(e2.get_field("u"), e.get_field("u"))
    np.testing.assert_array_equal(e2.get_field("ue"), e.get_field("ue"))
    assert e2.stats()["n_accept"] == e.stats()["n_accept"]
    bad = bytearray(blob)
    bad[0] ^= 0xFF
    import pytest
    with pytest.raises(RuntimeError):
        e2.restore(bytes(bad))
