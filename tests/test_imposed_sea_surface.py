"""ImposedSeaSurface (src/sealevel.jl:132-134): every sparse update sets dz_ss to the
TimeInterpolation2D of the given snapshots at the update's time (update_dz_ss!,
src/sealevel.jl:188-191), then z_ss = z_ss_ref + dz_ss + z_bsl (src/sealevel.jl:228-231).

Pins (oracle, CPU):
* Hand-computed interpolation values: snapshots c_k F at t = 0, 4000, 12000 yr with
  c = (0, 2, -2); an RHS call at t triggers one sparse update at t, after which dz_ss must be
  0, 1.5 F, -F and -1.005 F at t = 0, 5000, 10000, 10010 (a knot hit, interior points of the
  second bracket). A wrong bracket, time or slope changes these numbers.
* Self-consistency: an imposed surface fed with the dz_ss that a laterally variable run produced
  at its sparse-update times reproduces that run's dudt, z_ss and masks bit for bit (knot hits
  return the snapshot itself), with the interactive sea-level load on, so dz_ss reaches the
  load.
* Out of range: FI_ERR_TIME_RANGE without flat_bc; the end snapshot with it.
* fi_init refuses an imposed surface without snapshots.
HIP engine (GPU): the same RHS sequence against the oracle, at the tolerances of the other
sea-level parity tests (the device stores dz_ss and z_ss in FP32), and the range error.
"""
import numpy as np
import pytest

import fastiso
from fastiso import _abi

from _cases import hip_sim, lv_case, sim_inputs

TIMES = (0.0, 5000.0, 10000.0, 10010.0)   # each RHS call triggers one sparse update
T_SNAP = (0.0, 4000.0, 12000.0)
C_SNAP = (0.0, 2.0, -2.0)
C_EXPECTED = (0.0, 1.5, -1.0, -1.005)      # 2 + (-4) (t - 4000) / 8000 inside [4000, 12000]


def field(d):
    return 0.8 * np.exp(-((d.X - 4e5) ** 2 + d.Y ** 2) / (1.2e6) ** 2) - 0.3 * np.cos(d.Y / 9e5)


def margin_case(surface, n=32):
    """The marine-margin case of the interactive sea-level test: land inside, ocean outside,
    the active-load mask on the inner cells, so z_ss reaches the load."""
    w = lv_case(n=n, seed=9, interactive=True, t_end=20000.0)
    d = w["domain"]
    r = d.R / d.R.max()
    w["z_b_ref"] = 400.0 - 1400.0 * r
    w["solidearth"] = fastiso.SolidEarth(d, layer_boundaries=[w["h"]], layer_viscosities=[10.0 ** w["logeta"]],
                                         maskactive=r < 0.55)
    w["sealevel"] = fastiso.RegionalSeaLevel(surface=surface, load=fastiso.InteractiveSealevelLoad())
    return w


def start_u(d):
    rng = np.random.default_rng(4)
    return -80 * np.exp(-(d.R / 1e6) ** 2) + 0.01 * rng.standard_normal((d.nx, d.ny))


def oracle(w):
    from oracle import pyref
    return pyref.RefEngine.from_inputs(sim_inputs(w).inputs)


def imposed(d, flat=False, times=T_SNAP, coefs=C_SNAP):
    F = field(d)
    return fastiso.ImposedSeaSurface(list(times), [c * F for c in coefs], flat_bc=flat)


def test_oracle_imposed_dz_ss_hand_values():
    w = margin_case(fastiso.LaterallyConstantSeaSurface())
    d = w["domain"]
    w["sealevel"] = fastiso.RegionalSeaLevel(surface=imposed(d), load=fastiso.InteractiveSealevelLoad())
    ref = oracle(w)
    lc = oracle(margin_case(fastiso.LaterallyConstantSeaSurface()))
    F = field(d)
    u = start_u(d)
    for t, cexp in zip(TIMES, C_EXPECTED):
        ref.rhs(t, u)
        lc.rhs(t, u)
        dz = ref.get_field("dz_ss")
        np.testing.assert_allclose(dz, cexp * F, rtol=0, atol=1e-13)
        # z_ss = z_ss_ref + dz_ss + z_bsl; the laterally constant run has dz_ss = 0
        assert ref.stats()["z_bsl"] == lc.stats()["z_bsl"] == 0.0
        np.testing.assert_allclose(ref.get_field("z_ss") - lc.get_field("z_ss"), dz, rtol=0, atol=1e-12)
    assert ref.stats()["count_sparse_updates"] == len(TIMES)


def test_oracle_imposed_surface_reproduces_the_variable_surface_run():
    w = margin_case(fastiso.LaterallyVariableSeaSurface())
    d = w["domain"]
    lv = oracle(w)
    u = start_u(d)
    snaps, out_lv = [], []
    for t in TIMES:
        du, _ = lv.rhs(t, u)
        snaps.append(lv.get_field("dz_ss"))
        out_lv.append((du, lv.get_field("z_ss"), lv.get_field("maskocean"), lv.get_field("canom_seawater")))
    assert np.abs(snaps[-1]).max() > 1e-3
    w2 = margin_case(fastiso.ImposedSeaSurface(list(TIMES), snaps))
    im = oracle(w2)
    for t, (du_lv, zss_lv, mo_lv, sw_lv) in zip(TIMES, out_lv):
        du, _ = im.rhs(t, u)
        np.testing.assert_array_equal(du, du_lv)
        np.testing.assert_array_equal(im.get_field("z_ss"), zss_lv)
        np.testing.assert_array_equal(im.get_field("maskocean"), mo_lv)
        np.testing.assert_array_equal(im.get_field("canom_seawater"), sw_lv)
    assert np.abs(out_lv[-1][3]).max() > 0


def test_oracle_imposed_surface_range():
    w = margin_case(fastiso.LaterallyConstantSeaSurface())
    d = w["domain"]
    u = start_u(d)
    w["sealevel"] = fastiso.RegionalSeaLevel(surface=imposed(d, times=(0.0, 6000.0), coefs=(0.0, 1.0)),
                                             load=fastiso.InteractiveSealevelLoad())
    ref = oracle(w)
    ref.rhs(0.0, u)
    with pytest.raises(RuntimeError, match="TimeInterpolation2D"):
        ref.rhs(10000.0, u)
    w["sealevel"] = fastiso.RegionalSeaLevel(surface=imposed(d, flat=True, times=(0.0, 6000.0), coefs=(0.0, 1.0)),
                                             load=fastiso.InteractiveSealevelLoad())
    ref = oracle(w)
    ref.rhs(0.0, u)
    ref.rhs(10000.0, u)
    np.testing.assert_array_equal(ref.get_field("dz_ss"), field(d))


def test_oracle_imposed_surface_needs_snapshots():
    from oracle import pyref
    w = margin_case(fastiso.LaterallyConstantSeaSurface())
    inputs = sim_inputs(w).inputs
    inputs["cfg"].sea_surface = _abi.FI_IMPOSED_SEA_SURFACE
    with pytest.raises(RuntimeError, match="snapshots"):
        pyref.RefEngine.from_inputs(inputs)


def test_imposed_surface_host_checks():
    with pytest.raises(ValueError):
        fastiso.ImposedSeaSurface([0.0, 1.0], [np.zeros((4, 4))])
    with pytest.raises(ValueError):
        fastiso.ImposedSeaSurface([1.0, 0.0], [np.zeros((4, 4))] * 2)


@pytest.mark.gpu
def test_gpu_imposed_surface_vs_oracle():
    from oracle import pyref
    w = margin_case(fastiso.LaterallyConstantSeaSurface(), n=48)
    d = w["domain"]
    w["sealevel"] = fastiso.RegionalSeaLevel(surface=imposed(d), load=fastiso.InteractiveSealevelLoad())
    sim = hip_sim(w)
    ref = pyref.RefEngine.from_inputs(sim.inputs)
    u = start_u(d)
    for t, cexp in zip(TIMES, C_EXPECTED):
        dg, _ = sim.engine.rhs(t, u)
        dr, _ = ref.rhs(t, u)
        m = np.abs(dg - dr).max() / np.abs(dr).max()
        assert m < 2e-6, (t, m)
        F = field(d)
        # dz_ss is stored in FP32 on the device
        assert np.abs(sim.engine.get_field("dz_ss") - cexp * F).max() <= 1e-6 * np.abs(F).max(), t
        zg, zr = sim.engine.get_field("z_ss"), ref.get_field("z_ss")
        assert np.abs(zg - zr).max() <= 1e-4 + 1e-6 * np.abs(zr).max(), t
        for f in ("maskocean", "maskgrounded"):
            np.testing.assert_array_equal(sim.engine.get_field(f), ref.get_field(f))
    sw_g, sw_r = sim.engine.get_field("canom_seawater"), ref.get_field("canom_seawater")
    assert np.abs(sw_r).max() > 0
    assert np.abs(sw_g - sw_r).max() <= 1e-5 * np.abs(sw_r).max()
    sim.close()


@pytest.mark.gpu
def test_gpu_imposed_surface_range_error_and_integration():
    from oracle import pyref
    w = margin_case(fastiso.LaterallyConstantSeaSurface())
    d = w["domain"]
    u = start_u(d)
    w["sealevel"] = fastiso.RegionalSeaLevel(surface=imposed(d, times=(0.0, 6000.0), coefs=(0.0, 1.0)),
                                             load=fastiso.InteractiveSealevelLoad())
    sim = hip_sim(w)
    sim.engine.rhs(0.0, u)
    with pytest.raises(RuntimeError, match="TimeInterpolation2D"):
        sim.engine.rhs(10000.0, u)
    sim.close()
    # an integration over 0-200 yr against the oracle's (sparse updates every 10 yr)
    w["sealevel"] = fastiso.RegionalSeaLevel(surface=imposed(d, times=(0.0, 100.0, 300.0), coefs=(0.0, 3.0, -1.0)),
                                             load=fastiso.InteractiveSealevelLoad())
    sim = hip_sim(w)
    ref = pyref.RefEngine.from_inputs(sim.inputs)
    for e in (sim.engine, ref):
        e.integrator_init()
        e.integrate(200.0)
    ug, ur = sim.engine.get_field("u"), ref.get_field("u")
    assert np.abs(ug - ur).max() <= 1e-4 * np.abs(ur).max() + 1e-6
    assert abs(sim.engine.stats()["count_sparse_updates"] - ref.stats()["count_sparse_updates"]) == 0
    sim.close()
