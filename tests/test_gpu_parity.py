"""Parity of the HIP engine (through its C ABI) against the FP64 oracle and closed forms.

Tolerances: the engine runs the FFTs and the second stencil sweep in FP32 (FP64 state, first
stencil sweep, loads and error norm), so single RHS evaluations are compared at
max |d - d_ref| <= 2e-6 max |d_ref| and relative L2 <= 1e-6; trajectories at the output times
within 1e-5 (disc) or 1e-4 (laterally variable) of max |u| (the step-size sequences of the two back-ends differ once their error
estimates differ in the last FP32 digits).
"""
import numpy as np
import pytest

import fastiso
from fastiso import workloads
from fastiso.analytic import analytic_solution

from _cases import DRYRUN, corner_mean, cosine_case, hip_sim, lv_case, sim_inputs

pytestmark = pytest.mark.gpu

RHS_MAX = 2e-6
RHS_L2 = 1e-6


def rel(a, b):
    return np.abs(a - b).max() / np.abs(b).max(), np.linalg.norm(a - b) / np.linalg.norm(b)


def pair(w):
    from oracle import pyref
    sim = hip_sim(w)
    return sim, pyref.RefEngine.from_inputs(sim.inputs)


def test_rhs_cosine_closed_form():
    w = cosine_case()
    sim, ref = pair(w)
    d, c, se = w["domain"], w["c"], w["solidearth"]
    u = 7.0 * np.cos(2 * np.pi * w["l"] * d.Y / (2 * d.Wy)) + 3.0
    du, ubc = sim.engine.rhs(1.0, u)
    eta = w["eta"] * 1.5 / 1.28
    F = c.seconds_per_year / (2 * eta) * (
        -c.g * c.rho_ice * w["A"] * np.cos(2 * np.pi * w["k"] * d.X / (2 * d.Wx)) * (d.Wx / (np.pi * w["k"]))
        - c.g * se.rho_uppermantle * 7.0 * np.cos(2 * np.pi * w["l"] * d.Y / (2 * d.Wy)) * (d.Wy / (np.pi * w["l"])))
    want = F - corner_mean(F)
    m, l2 = rel(du, want)
    assert m < RHS_MAX and l2 < RHS_L2, (m, l2)
    np.testing.assert_allclose(ubc, u - corner_mean(u), atol=1e-12)
    np.testing.assert_allclose(sim.engine.get_field("u"), ubc, atol=1e-12)


# odd sizes end with a one-row item and a three-row moment window at the bottom edge
@pytest.mark.parametrize("n,seed,surface", [(32, 11, True), (64, 5, True), (48, 7, False), (96, 3, True),
                                            (45, 13, True), (35, 17, False)])
def test_rhs_sequence_vs_oracle(n, seed, surface):
    """A sequence of RHS calls through two sparse updates: the first at t0 reads zero loads, the
    second (t = 600 after t = 500) reads the loads of the call before it."""
    w = lv_case(n=n, seed=seed, sea_surface=surface)
    sim, ref = pair(w)
    rng = np.random.default_rng(seed)
    u = -50 * np.exp(-(w["domain"].R / 1e6) ** 2) + 0.01 * rng.standard_normal((n, n))
    for t in (0.0, 500.0, 600.0, 605.0):
        dg, ug = sim.engine.rhs(t, u)
        dr, ur = ref.rhs(t, u)
        m, l2 = rel(dg, dr)
        assert m < RHS_MAX and l2 < RHS_L2, (t, m, l2)
        np.testing.assert_allclose(ug, ur, atol=1e-9)
    # triggers at t = 0 (count 0), 500 (50 >= 1), 600 (60 >= 2) and 605 (60.5 >= 3)
    assert sim.engine.stats()["count_sparse_updates"] == ref.stats()["count_sparse_updates"] == 4
    for f in ("ue", "dz_ss", "z_ss"):
        a, b = sim.engine.get_field(f), ref.get_field(f)
        if np.abs(b).max() == 0:
            assert np.abs(a).max() == 0
            continue
        m, l2 = rel(a, b)
        assert m < 5e-6 and l2 < 5e-6, (f, m, l2)
    for f in ("canom_full", "canom_load", "H_af", "z_b"):
        a, b = sim.engine.get_field(f), ref.get_field(f)
        assert np.abs(a - b).max() <= 1e-5 * max(np.abs(b).max(), 1.0), f


# Grids wider than 512 cells keep the plate moments in FP32 (one workgroup's LDS holds four
# moment rows per field only at single precision there); they are held to the same bounds.
@pytest.mark.parametrize("n", [600, 1024])
def test_rhs_large_grid_vs_oracle(n):
    w = lv_case(n=n, seed=21)
    sim, ref = pair(w)
    rng = np.random.default_rng(n)
    u = -50 * np.exp(-(w["domain"].R / 1e6) ** 2) + 0.01 * rng.standard_normal((n, n))
    for t in (0.0, 500.0, 600.0):   # the call at 600 convolves the nonzero loads of t = 500
        dg, ug = sim.engine.rhs(t, u)
        dr, ur = ref.rhs(t, u)
        m, l2 = rel(dg, dr)
        print(f"n={n} t={t} dudt max-rel {m:.3e} L2-rel {l2:.3e}")
        assert m < RHS_MAX and l2 < RHS_L2, (t, m, l2)
        np.testing.assert_allclose(ug, ur, atol=1e-9)
    for f in ("ue", "dz_ss"):
        m, l2 = rel(sim.engine.get_field(f), ref.get_field(f))
        print(f"n={n} {f} max-rel {m:.3e} L2-rel {l2:.3e}")
        assert m < 5e-6 and l2 < 5e-6, (f, m, l2)


def test_rhs_snapshots_equal_separable():
    """TimeInterpolatedIceThickness snapshots (src/interpolations.jl:51-74) and the separable
    form a(t) H give the same RHS, each against the FP64 oracle fed the same snapshot inputs, at a
    knot (t = 500) and between knots (t = 1500.5)."""
    from oracle import pyref
    a = lv_case(n=32, separable=True)
    b = lv_case(n=32, separable=False)
    sa, sb = hip_sim(a), hip_sim(b)
    rb = pyref.RefEngine.from_inputs(sb.inputs)
    u = np.zeros((32, 32))
    for t in (500.0, 1500.5):
        da, _ = sa.engine.rhs(t, u)
        db, _ = sb.engine.rhs(t, u)
        dr, _ = rb.rhs(t, u)
        m, l2 = rel(da, db)
        assert m < 1e-6 and l2 < 1e-6
        for d in (da, db):
            m, l2 = rel(d, dr)
            assert m < RHS_MAX and l2 < RHS_L2, (t, m, l2)


@pytest.mark.skipif(DRYRUN, reason="device pointers need the GPU")
def test_rhs_device_pointers_match_host():
    """fi_rhs_device on device-resident FP64 arrays equals fi_rhs on host arrays. Device memory
    comes from the same HIP runtime the engine links (libamdhip64 of /opt/rocm)."""
    import ctypes as C
    hip = C.CDLL("libamdhip64.so")
    w = lv_case(n=32)
    sim = sim_inputs(w, engine="hip")
    u = -20 * np.exp(-(w["domain"].R / 8e5) ** 2)
    dh, uh = sim.engine.rhs(0.0, u)
    sim2 = sim_inputs(w, engine="hip")
    nb = 32 * 32 * 8
    ud, dd = C.c_void_p(), C.c_void_p()
    assert hip.hipMalloc(C.byref(ud), C.c_size_t(nb)) == 0 and hip.hipMalloc(C.byref(dd), C.c_size_t(nb)) == 0
    try:
        uf = np.ascontiguousarray(np.asarray(u, dtype=np.float64).ravel(order="F"))
        assert hip.hipMemcpy(ud, uf.ctypes.data_as(C.c_void_p), C.c_size_t(nb), 1) == 0   # H2D
        st = sim2.engine.lib.fi_rhs_device(sim2.engine.h, 0.0, ud, dd)
        assert st == 0
        dout, uout = np.empty(32 * 32), np.empty(32 * 32)
        assert hip.hipMemcpy(dout.ctypes.data_as(C.c_void_p), dd, C.c_size_t(nb), 2) == 0  # D2H
        assert hip.hipMemcpy(uout.ctypes.data_as(C.c_void_p), ud, C.c_size_t(nb), 2) == 0
    finally:
        hip.hipFree(ud)
        hip.hipFree(dd)
    np.testing.assert_allclose(dout.reshape((32, 32), order="F"), dh, rtol=0, atol=0)
    np.testing.assert_allclose(uout.reshape((32, 32), order="F"), uh, rtol=0, atol=1e-12)


def test_time_out_of_range_is_an_error():
    w = lv_case(n=32, t_end=1000.0)
    sim = hip_sim(w)
    with pytest.raises(Exception) as e:
        sim.engine.rhs(200e3, np.zeros((32, 32)))
    assert getattr(e.value, "code", -3) == -3 and "out of range" in str(e.value)


@pytest.fixture(scope="module")
def disc_runs():
    from oracle import pyref
    w = workloads.disc_benchmark(n=64)
    sim = hip_sim(w)
    ref = pyref.RefEngine.from_inputs(sim.inputs)
    og, orf = [], []
    sim.engine.integrator_init()
    sim.engine.integrate(w["t_span"][1], lambda t, k: og.append((t, sim.engine.get_field("u"))))
    ref.integrator_init()
    ref.integrate(w["t_span"][1], lambda t, k: orf.append((t, ref.get_field("u"))))
    return w, sim, ref, og, orf


def test_disc_trajectory_vs_oracle(disc_runs):
    w, sim, ref, og, orf = disc_runs
    assert [t for t, _ in og] == [t for t, _ in orf] == list(np.arange(1e3, 50e3 + 1, 1e3))
    for (t, a), (_, b) in zip(og, orf):
        assert np.abs(a - b).max() <= 1e-5 * np.abs(b).max(), t


def test_disc_trajectory_vs_analytic(disc_runs):
    """64x64 disc benchmark (docs/src/examples/benchmark_analytic.jl) against
    analytic_solution (src/analytic_solutions.jl:11-65) on the centre transect. The FP64 oracle's
    own transect error is the discretisation error of the 93.75-km grid (largest at the disc edge:
    11.4 m of 287 m at 50 kyr, 7.4 m of 88 m at 1 kyr); the device run may exceed it by at most
    1e-3 of the analytic maximum, and the centre is within 0.5 m from 10 kyr on."""
    w, sim, ref, og, orf = disc_runs
    d, c, se = w["domain"], w["c"], w["solidearth"]
    j = d.my - 1
    for (t, u), (_, ur) in zip(og, orf):
        ua = analytic_solution(np.abs(d.x), t * c.seconds_per_year, c, se, 1e3, 1e6)
        e_ref = np.abs(ur[:, j] - ua).max()
        assert e_ref < 11.5 and e_ref < 0.085 * np.abs(ua).max(), (t, e_ref)
        assert np.abs(u[:, j] - ua).max() <= e_ref + 1e-3 * np.abs(ua).max(), t
        if t >= 10e3:
            assert abs(u[d.mx - 1, j] - ua[d.mx - 1]) < 0.5, t


def test_lv_trajectory_vs_oracle():
    from oracle import pyref
    w = lv_case(n=32, t_end=200.0)
    sim = hip_sim(w)
    ref = pyref.RefEngine.from_inputs(sim.inputs)
    og, orf = [], []
    sim.engine.integrator_init()
    sim.engine.integrate(200.0, lambda t, k: og.append((t, sim.engine.get_field("u"), sim.engine.get_field("ue"))))
    ref.integrator_init()
    ref.integrate(200.0, lambda t, k: orf.append((t, ref.get_field("u"), ref.get_field("ue"))))
    assert len(og) == len(orf) == 4
    for (t, a, ea), (_, b, eb) in zip(og, orf):
        assert np.abs(a - b).max() <= 1e-4 * np.abs(b).max() + 1e-6, t
        assert np.abs(ea - eb).max() <= 1e-4 * np.abs(eb).max() + 1e-6, t
    sg, sr = sim.engine.stats(), ref.stats()
    assert abs(sg["n_accept"] - sr["n_accept"]) <= 0.05 * sr["n_accept"] + 3
    assert sg["count_sparse_updates"] == sr["count_sparse_updates"]


def test_set_time_jump_vs_oracle():
    """fi_set_field(u) + fi_set_time (the surrogate's macro step) then integration, both engines."""
    from oracle import pyref
    w = lv_case(n=32, t_end=20e3)
    sim = hip_sim(w)
    ref = pyref.RefEngine.from_inputs(sim.inputs)
    d = w["domain"]
    u1 = -30.0 * np.exp(-(d.R / 8e5) ** 2)
    outs = []
    for e in (sim.engine, ref):
        e.integrator_init()
        e.integrate(100.0)
        e.set_field("u", u1)
        e.set_time(5000.0)
        ue_jump = e.get_field("ue")
        e.integrate(5100.0)
        outs.append((e.get_field("u"), ue_jump, e.stats()))
    (a, ea, sa), (b, eb, sb) = outs
    assert sa["t"] == sb["t"] == 5100.0
    assert sa["count_sparse_updates"] == sb["count_sparse_updates"]
    assert np.abs(ea).max() > 0 and np.abs(ea - eb).max() <= 1e-5 * np.abs(eb).max()
    assert np.abs(a - b).max() <= 1e-4 * np.abs(b).max() + 1e-6


def test_interactive_sealevel_rhs_vs_oracle():
    """InteractiveSealevelLoad (src/loads.jl:56-61): a marine margin, the active-load mask on 40 % of
    the cells; the sea-water column anomaly enters the load after each sparse update."""
    import fastiso
    from oracle import pyref
    n = 48
    w = lv_case(n=n, seed=9, interactive=True)
    d = w["domain"]
    r = d.R / d.R.max()
    w["z_b_ref"] = 400.0 - 1400.0 * r                       # land inside, ocean outside
    se = w["solidearth"]
    mask = r < 0.55
    assert 0.2 < mask.mean() < 0.6
    w["solidearth"] = fastiso.SolidEarth(d, layer_boundaries=[w["h"]], layer_viscosities=[10.0 ** w["logeta"]],
                                         maskactive=mask)
    sim = hip_sim(w)
    ref = pyref.RefEngine.from_inputs(sim.inputs)
    rng = np.random.default_rng(4)
    u = -80 * np.exp(-(d.R / 1e6) ** 2) + 0.01 * rng.standard_normal((n, n))
    for t in (0.0, 500.0, 600.0, 605.0):
        dg, _ = sim.engine.rhs(t, u)
        dr, _ = ref.rhs(t, u)
        m, l2 = rel(dg, dr)
        assert m < RHS_MAX and l2 < RHS_L2, (t, m, l2)
    sw_g, sw_r = sim.engine.get_field("canom_seawater"), ref.get_field("canom_seawater")
    assert np.abs(sw_r).max() > 0
    assert np.abs(sw_g - sw_r).max() <= 1e-5 * np.abs(sw_r).max()
    for f in ("maskocean", "maskgrounded"):
        np.testing.assert_array_equal(sim.engine.get_field(f), ref.get_field(f))


def test_checkpoint_restart_is_bitwise():
    """fi_checkpoint_save at t1 (after sparse updates), then integrate to t2; a fresh context that
    restores the checkpoint and integrates to t2 ends bit for bit equal."""
    w = lv_case(n=32, t_end=400.0)
    sim = hip_sim(w)
    e = sim.engine
    e.integrator_init()
    e.integrate(150.0)
    blob = e.checkpoint()
    e.integrate(400.0)
    ua, uea, sa = e.get_field("u"), e.get_field("ue"), e.stats()
    sim2 = hip_sim(w)
    e2 = sim2.engine
    e2.restore(blob)
    e2.integrate(400.0)
    np.testing.assert_array_equal(e2.get_field("u"), ua)
    np.testing.assert_array_equal(e2.get_field("ue"), uea)
    sb = e2.stats()
    for k in ("t", "n_accept", "n_reject", "count_sparse_updates", "n_rhs"):
        assert sa[k] == sb[k], k
    # rollback on the same context
    e2.restore(blob)
    assert e2.stats()["t"] == 150.0


@pytest.mark.gpu
def test_n350_workload_fields_vs_oracle():
    """The N=350 workload itself (the bench configuration) on the device and on the FP64 oracle
    over 0-20 yr (two sparse updates): per-field max and relative-L2 errors of u, ue, dz_ss, z_ss,
    H_af and the full column anomaly, printed and bounded. This early in the glacial cycle |u| is
    below 1 mm, so the step control works at its absolute tolerance: each of the ~850 steps is
    held to abstol + reltol |u| ~ 1e-6 m, and the two back-ends' step sequences differ (FP32
    error estimates), so u is bounded by the accumulated local tolerance (5e-4 m; measured 1.3e-4 m
    at 20 yr). The sparse update convolves the loads of the call before it, sampled at times that
    differ by a fraction of a step while the load grows as t: ue and dz_ss are bounded at 1e-3 of
    their maximum (measured 2.3e-4). H_af follows the forcing exactly."""
    from oracle import pyref
    w = workloads.lv_elva(n=350, t_end=120e3, t_out=[10.0, 20.0])
    sim = hip_sim(w)
    ref = pyref.RefEngine.from_inputs(sim.inputs)
    fields = ("u", "ue", "dz_ss", "z_ss", "H_af", "canom_full")
    og, orf = [], []
    sim.engine.integrator_init()
    sim.engine.integrate(20.0, lambda t, k: og.append((t, {f: sim.engine.get_field(f) for f in fields})))
    ref.integrator_init()
    ref.integrate(20.0, lambda t, k: orf.append((t, {f: ref.get_field(f) for f in fields})))
    assert [t for t, _ in og] == [t for t, _ in orf] == [10.0, 20.0]
    atol = {"u": 5e-4, "ue": 1e-3 * 0.03, "dz_ss": 1e-3 * 0.025, "z_ss": 1e-3 * 0.025, "H_af": 1e-5,
            "canom_full": 1.0}
    for (t, a), (_, b) in zip(og, orf):
        for f in fields:
            den = max(np.abs(b[f]).max(), 1e-30)
            eabs = np.abs(a[f] - b[f]).max()
            el2 = np.linalg.norm(a[f] - b[f]) / max(np.linalg.norm(b[f]), 1e-30)
            print(f"t={t:5.1f} {f:10s} max|b|={den:10.4g} abs max {eabs:9.2e} rel max {eabs / den:9.2e} "
                  f"rel L2 {el2:9.2e}")
            assert eabs <= atol[f], (t, f, eabs)
    # step counts are not compared here: in this abstol-dominated start the two back-ends' step
    # sequences differ (971 and 848 accepted steps to 20 yr); see the next test for the regime of
    # the bench windows
    sg, sr = sim.engine.stats(), ref.stats()
    print("accepted steps", sg["n_accept"], sr["n_accept"])


@pytest.mark.gpu
def test_n350_bench_window_from_the_same_state():
    """One 10-yr window of the N=350 workload at t = 1000 yr (the regime of the bench windows)
    from the same state on both back-ends (fi_set_field(u) + fi_set_time): the FP32-transform
    engine takes the FP64 oracle's step count (524 and 526 when measured) and ends within 1e-4 of
    max |u|."""
    from oracle import pyref
    w = workloads.lv_elva(n=350, t_end=120e3, t_out=[])
    sim = hip_sim(w)
    e = sim.engine
    e.integrator_init()
    e.integrate(1000.0)
    u0 = e.get_field("u")
    ref = pyref.RefEngine.from_inputs(sim.inputs)
    ref.integrator_init()
    for eng in (ref, e):
        eng.set_field("u", u0)
        eng.set_time(1000.0)
    a0, b0 = e.stats()["n_accept"], ref.stats()["n_accept"]
    e.integrate(1010.0)
    ref.integrate(1010.0)
    na, nb = e.stats()["n_accept"] - a0, ref.stats()["n_accept"] - b0
    ua, ub = e.get_field("u"), ref.get_field("u")
    err = np.abs(ua - ub).max() / np.abs(ub).max()
    print(f"window steps engine {na} oracle {nb}; u max-rel {err:.2e}")
    assert abs(na - nb) <= 0.02 * nb + 2
    assert err <= 1e-4
