"""RKC2 integration of the right-hand side (fi_integrate_rkc), the macro-step predictor of the
surrogate-with-fallback mode. It has no counterpart in the reference; it is pinned on the
mathematics of the method and compared with the plain BS3 solver:

* a cosine load on a flexure-free Earth: after the load is switched on, the forced mode obeys
  m' = -(m - m*)/tau with m* = -rho_ice A / rho_mantle and tau = 2 eta kappa / (rho_mantle g), so n
  RKC sub-steps of h give m* + (m0 - m*) R_s(-h/tau)^n with the published stability polynomial
  R_s(z) = a_s + b_s T_s(w0 + w1 z) (Sommeijer, Shampine & Verwer 1997);
* the power-iteration estimate of the spectral radius equals the fastest mode's rate, which on a
  flexure-free Earth is the longest wavelength's g rho_mantle / (2 eta kappa_min);
* the stage count is the smallest s with beta(s) >= 1.2 rho h;
* on a laterally variable case the RKC result approaches the plain BS3 trajectory as the sub-step
  shrinks and stays within millimetres of it.
"""
import numpy as np
import pytest

import fastiso
from fastiso import workloads

from _cases import DRYRUN, hip_sim, sim_inputs

EPS = 2.0 / 13.0


def cheb(s, x):
    """T_s(x), T_s'(x), T_s''(x) by the three-term recurrences."""
    T0, T1 = 1.0, x
    d0, d1 = 0.0, 1.0
    e0, e1 = 0.0, 0.0
    for _ in range(2, s + 1):
        T0, T1 = T1, 2 * x * T1 - T0
        d0, d1 = d1, 2 * T0 + 2 * x * d1 - d0
        e0, e1 = e1, 4 * d0 + 2 * x * e1 - e0
    return T1, d1, e1


def stability_polynomial(s, z):
    w0 = 1 + EPS / s ** 2
    Ts, dTs, ddTs = cheb(s, w0)
    w1 = dTs / ddTs
    b = ddTs / dTs ** 2
    a = 1 - b * Ts
    return a + b * cheb(s, w0 + w1 * z)[0]


def beta(s):
    w0 = 1 + EPS / s ** 2
    _, d, dd = cheb(s, w0)
    return (w0 + 1) * dd / d


def test_stability_polynomial_is_second_order_and_damped():
    """Checks of the test's own closed form: R_s(z) = 1 + z + z^2/2 + O(z^3), and |R_s| <= 1 on
    [-beta(s), 0] with beta(s) close to the published 2/3 (s^2 - 1)(1 - 2 eps / 15)."""
    for s in (2, 5, 17, 60):
        for z in (-1e-2, -2e-3):
            assert abs(stability_polynomial(s, z) - (1 + z + z * z / 2)) < 2 * abs(z) ** 3
        zz = np.linspace(-beta(s), 0, 801)
        assert max(abs(stability_polynomial(s, z)) for z in zz) <= 1 + 1e-9
        assert abs(beta(s) / (2 / 3 * (s * s - 1) * (1 - 2 * EPS / 15)) - 1) < 0.05


def cosine_rkc_case(n=16, ny=12, k=2, A=300.0, H0=1000.0, eta=1e21, t_end=20000.0):
    d = fastiso.RegionalDomain(3e6, 3e6 * ny / n, n, ny)
    se = fastiso.SolidEarth(d, layer_boundaries=[np.zeros((n, ny))], layer_viscosities=[eta], rho_litho=0.0)
    H1 = H0 + A * np.cos(2 * np.pi * k * d.X / (2 * d.Wx))
    it = fastiso.TimeInterpolatedIceThickness([0.0, 1.0, t_end], [d.zeros(), H1, H1])
    bcs = fastiso.BoundaryConditions(d, ice_thickness=it)
    return dict(domain=d, bcs=bcs, sealevel=fastiso.RegionalSeaLevel(), solidearth=se, t_span=(0.0, t_end),
                nout=fastiso.NativeOutput(), opts=fastiso.SolverOptions(), c=fastiso.PhysicalConstants(), k=k, A=A)


def mode(u, d, k):
    return 2.0 / u.size * np.sum(u * np.cos(2 * np.pi * k * d.X / (2 * d.Wx)))


def cosine_rkc_run(make, rtol):
    w = cosine_rkc_case()
    d, c, se = w["domain"], w["c"], w["solidearth"]
    e = make(w)
    e.integrator_init()
    e.integrate(2.0)
    eta = float(se.effective_viscosity[0, 0])
    kap = 2 * np.pi * w["k"] / (2 * d.Wx)
    tau = 2 * eta * kap / (se.rho_uppermantle * c.g) / c.seconds_per_year
    mstar = -c.rho_ice * w["A"] / se.rho_uppermantle
    m = mode(e.get_field("u"), d, w["k"])
    # the fastest mode: the longest wavelength (no flexure), kappa = pi / max(Wx, Wy)
    rate_max = se.rho_uppermantle * c.g / (2 * eta * np.pi / max(d.Wx, d.Wy)) * c.seconds_per_year
    t = 2.0
    for t_next, dt_max in ((1002.0, 100.0), (3002.0, 400.0)):
        info = e.integrate_rkc(t_next, dt_max)
        assert abs(info["rho"] / rate_max - 1) < 1e-4, (info, rate_max)
        s, h, nsub = info["stages"], info["h"], info["substeps"]
        assert nsub == int(np.ceil((t_next - t) / dt_max - 1e-12)) and abs(h * nsub - (t_next - t)) < 1e-9
        assert beta(s) >= 1.2 * info["rho"] * h and (s == 2 or beta(s - 1) < 1.2 * info["rho"] * h)
        m = mstar + (m - mstar) * stability_polynomial(s, -h / tau) ** nsub
        got = mode(e.get_field("u"), d, w["k"])
        assert abs(got - m) <= rtol * abs(mstar), (t_next, got, m)
        assert e.stats()["t"] == t_next
        t = t_next
    return e


def test_oracle_rkc_cosine_mode_closed_form():
    from oracle import pyref
    cosine_rkc_run(lambda w: pyref.RefEngine.from_inputs(sim_inputs(w).inputs), 1e-9)


def lv_rkc_case(n=48):
    return workloads.lv_elva(n=n, W=n * 18e3 / 2, t_end=5000.0, seeds=(71, 72, 73), t_out=[])


def test_oracle_rkc_converges_to_bs3_on_lv_case():
    """LV-ELVA with dx = 18 km: RKC from the BS3 state at 500 yr to 700 yr against BS3 itself
    (reltol 1e-5) and against RKC with 6.25-yr sub-steps. The error falls as the sub-step shrinks,
    though not at a clean second-order rate: the sparse elastic / sea-surface update samples the
    loads at the first right-hand-side call past each 10-yr mark, and those calls fall at
    different times for different sub-steps. It stays within 0.3% of the increment."""
    from oracle import pyref
    w = lv_rkc_case()
    inp = sim_inputs(w).inputs
    plain = pyref.RefEngine.from_inputs(inp)
    plain.integrator_init()
    plain.integrate(500.0)
    u0 = plain.get_field("u")
    plain.integrate(700.0)
    up = plain.get_field("u")
    res = {}
    for dt_max in (100.0, 25.0, 6.25):
        e = pyref.RefEngine.from_inputs(inp)
        e.integrator_init()
        e.set_field("u", u0)
        e.set_time(500.0)
        e.integrate_rkc(700.0, dt_max)
        res[dt_max] = e.get_field("u")
    du = np.abs(up - u0).max()
    to_bs3 = [np.abs(res[h] - up).max() for h in (100.0, 25.0, 6.25)]
    to_fine = [np.abs(res[h] - res[6.25]).max() for h in (100.0, 25.0)]
    assert max(to_bs3) < 3e-3 * du, (to_bs3, du)
    assert to_bs3[1] < to_bs3[0] and to_fine[1] < 0.5 * to_fine[0], (to_bs3, to_fine)


def test_oracle_rkc_refuses_the_implicit_step_and_past_targets():
    from oracle import pyref
    from _cases import lv_case
    from test_implicit import cosine_implicit_case
    e = pyref.RefEngine.from_inputs(sim_inputs(lv_case(n=16, t_end=100.0)).inputs)
    e.integrator_init()
    e.integrate(10.0)
    with pytest.raises(RuntimeError, match="error -1:"):
        e.integrate_rkc(5.0, 10.0)
    imp = pyref.RefEngine.from_inputs(sim_inputs(cosine_implicit_case()).inputs)
    imp.integrator_init()
    with pytest.raises(RuntimeError, match="error -2:"):
        imp.integrate_rkc(100.0, 10.0)


# ---- GPU ----------------------------------------------------------------------------------------
@pytest.mark.gpu
def test_gpu_rkc_cosine_mode_closed_form():
    e = cosine_rkc_run(lambda w: hip_sim(w).engine, 1e-9 if DRYRUN else 2e-5)
    if not DRYRUN:
        e.close()


@pytest.mark.gpu
def test_gpu_rkc_matches_oracle_rkc_and_bs3_on_lv_case():
    """The same RKC sub-steps and stages (the oracle's spectral radius passed to both) from the
    same state: the HIP result is within FP32-transform rounding of the oracle's, and both are
    within millimetres of the plain BS3 solver."""
    from oracle import pyref
    w = lv_rkc_case(n=64)
    inp = sim_inputs(w).inputs
    plain = pyref.RefEngine.from_inputs(inp)
    plain.integrator_init()
    plain.integrate(400.0)
    u0 = plain.get_field("u")
    plain.integrate(600.0)
    up = plain.get_field("u")
    ref = pyref.RefEngine.from_inputs(inp)
    ref.integrator_init()
    ref.set_field("u", u0)
    ref.set_time(400.0)
    ir = ref.integrate_rkc(600.0, 50.0)
    hs = hip_sim(w)
    e = hs.engine
    e.integrator_init()
    e.set_field("u", u0)
    e.set_time(400.0)
    ih = e.integrate_rkc(600.0, 50.0, rho=ir["rho"])
    assert (ih["stages"], ih["substeps"]) == (ir["stages"], ir["substeps"])
    uh, ur = e.get_field("u"), ref.get_field("u")
    du = np.abs(up - u0).max()
    assert np.abs(uh - ur).max() <= 1e-4 * np.abs(ur).max() + 1e-4, (np.abs(uh - ur).max(), np.abs(ur).max())
    assert np.abs(uh - up).max() <= 5e-3 * du
    # the engine's own estimate of the spectral radius agrees with the oracle's
    e2 = hip_sim(w).engine
    e2.integrator_init()
    e2.set_field("u", u0)
    e2.set_time(400.0)
    i2 = e2.integrate_rkc(450.0, 50.0)
    assert abs(i2["rho"] / ir["rho"] - 1) < 1e-3, (i2["rho"], ir["rho"])
    # BS3 continues from each RKC state (its first-same-as-last derivative re-evaluated) and
    # matches the oracle's continuation. The stage recursion can end in the scratch state buffer
    # (which of the three depends on the stage count's parity and the starting buffer); the
    # integration loop must find the state in one of the two buffers it alternates between.
    # Stage counts of both parities are covered.
    parities = set()
    t = 600.0
    for span in (50.0, 51.0, 52.0, 53.0, 54.0):
        ir2 = ref.integrate_rkc(t + span, span)
        ih2 = e.integrate_rkc(t + span, span, rho=ir["rho"])
        assert ih2["stages"] == ir2["stages"] and ih2["substeps"] == 1
        parities.add(ih2["stages"] % 2)
        t += span + 5.0
        e.integrate(t)
        ref.integrate(t)
        assert e.stats()["t"] == t
        assert np.abs(e.get_field("u") - ref.get_field("u")).max() <= 2e-4 * np.abs(ur).max() + 1e-4, t
    assert parities == {0, 1}
    if not DRYRUN:
        e.close()
        e2.close()


@pytest.mark.gpu
def test_gpu_rkc_graph_replay_equals_direct_launches():
    """The chain of an fi_integrate_rkc call is captured into a HIP graph on its first call with a
    given sub-step count, stage count, sub-step and starting buffer, and replayed by later calls
    (the stage times are a device table written once per call). Two contexts from the same state,
    one launching the chain directly (FASTISO_RKC_GRAPH=0), advance through three macro steps
    with a BS3 step in between (the starting buffer changes): equal bit for bit, with the graph
    captured, replayed and captured again for the other starting buffer."""
    import os
    from oracle import pyref
    w = lv_rkc_case(n=48)
    inp = sim_inputs(w).inputs
    plain = pyref.RefEngine.from_inputs(inp)
    plain.integrator_init()
    plain.integrate(100.0)
    u0 = plain.get_field("u")

    def fresh():
        e = hip_sim(w).engine
        e.integrator_init()
        e.set_field("u", u0)
        e.set_time(100.0)
        return e

    a, b = fresh(), fresh()
    graphs = []
    os.environ["FASTISO_RKC_GRAPH"] = "0"
    try:
        ib = b.integrate_rkc(200.0, 50.0, rho=120.0)
    finally:
        del os.environ["FASTISO_RKC_GRAPH"]
    assert ib["graph"] == 0
    ia = a.integrate_rkc(200.0, 50.0, rho=120.0)
    graphs.append(ia["graph"])
    np.testing.assert_array_equal(a.get_field("u"), b.get_field("u"))
    for t in (300.0, 300.5, 400.5):
        for e in (a, b):
            if t == 300.5:
                e.integrate(t)          # one BS3 stretch: the state moves to the other buffer
            else:
                info = e.integrate_rkc(t, 50.0, rho=120.0)
                if e is a:
                    graphs.append(info["graph"])
                else:
                    assert info["graph"] == 0
        np.testing.assert_array_equal(a.get_field("u"), b.get_field("u"), err_msg=f"t={t}")
    if not DRYRUN:
        # three calls of the same counts and sub-step from at most two starting buffers
        assert graphs[0] == 2 and 1 in graphs[1:] and set(graphs) <= {1, 2}, graphs
