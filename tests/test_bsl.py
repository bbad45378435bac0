"""Barystatic sea-level rules (src/barystatic_sealevel.jl:215-235) and the internal/external update
(src/sealevel.jl:242-262): the host mirror, the oracle and the HIP engine.

Pins:
* TimeInterpolation0D against hand-computed values: knot hits, interior points, flat and
  non-flat ends.
* The reference's own test (test/test_barystatic_sea_level.jl): 45e6 km^3 of ice volume taken
  out in 10 000 steps lowers the BSL below -120 m and shrinks the ocean surface. The reference
  downloads its ocean-surface table; the synthetic hypsometry here, A(z) = A_pd (1 + z / 400 m),
  keeps the same monotonicity.
* Oracle: z_bsl follows z_k = z_{k-1} - delta_V_k / A at every sparse update; it falls while the
  ice grows; and with a laterally constant sea surface z_ss shifts by exactly z_bsl against a run
  with a constant BSL.
* HIP engine: the same RHS sequence as the oracle, with its dudt, z_bsl, ocean surface and sea
  surface compared to the tolerances of the other RHS parity tests (ΔV is summed from FP32
  fields on the device, so z_bsl agrees to ~1e-6 relative).
"""
import numpy as np
import pytest

import fastiso
from fastiso.sealevel import (A_OCEAN_PD, ConstantOceanSurfaceBSL, ImposedBSL, PiecewiseConstantBSL, ReferenceBSL,
                              interpolate0d, update_bsl)

from _cases import hip_sim, lv_case, sim_inputs

TIMES = (0.0, 5000.0, 10000.0, 10010.0)   # each call triggers one sparse update


def hypsometry():
    z = np.linspace(-400.0, 100.0, 501)
    return z, A_OCEAN_PD * (1.0 + z / 400.0 + 1e-3)


def test_interpolate0d_rules():
    x, y = [0.0, 10.0, 20.0], [1.0, 3.0, 2.0]
    assert interpolate0d(x, y, 10.0) == 3.0
    assert interpolate0d(x, y, 5.0) == 2.0
    assert interpolate0d(x, y, 15.0) == 2.5
    assert interpolate0d(x, y, -1.0, flat_bc=True) == 1.0
    assert interpolate0d(x, y, 25.0, flat_bc=True) == 2.0
    with pytest.raises(ValueError):
        interpolate0d(x, y, 25.0)


def test_piecewise_constant_bsl_falls_as_ice_grows():
    z, A = hypsometry()
    bsl = PiecewiseConstantBSL(ReferenceBSL(0.0, z_vec=z, A_vec=A))
    A0 = bsl.A
    Vice = 45e6 * 1e9
    nV = 10000
    for _ in range(nV):
        update_bsl(bsl, -Vice / nV, 0.0)
    assert bsl.z < -120.0
    assert bsl.A < A0
    # constant ocean surface: the same volume lowers z by exactly V / A
    c = ConstantOceanSurfaceBSL()
    for _ in range(nV):
        update_bsl(c, -Vice / nV, 0.0)
    assert abs(c.z + Vice / A_OCEAN_PD) < 1e-9 * Vice / A_OCEAN_PD
    # the shrinking surface of the piecewise-constant rule makes the fall larger
    assert bsl.z < c.z


def case(bsl, surface=False, update=None):
    w = lv_case(n=32, seed=5, sea_surface=surface, t_end=20000.0)
    kw = dict(bsl=bsl)
    if update is not None:
        kw["update_bsl"] = update
    if surface:
        kw["surface"] = fastiso.LaterallyVariableSeaSurface()
    w["sealevel"] = fastiso.RegionalSeaLevel(**kw)
    return w


def u0(w):
    rng = np.random.default_rng(3)
    d = w["domain"]
    return -30 * np.exp(-(d.R / 1e6) ** 2) + 0.01 * rng.standard_normal((d.nx, d.ny))


def oracle_run(w):
    from oracle import pyref
    sim = sim_inputs(w)
    ref = pyref.RefEngine.from_inputs(sim.inputs)
    u = u0(w)
    out = []
    for t in TIMES:
        d, _ = ref.rhs(t, u)
        st = ref.stats()
        out.append((d, st["z_bsl"], st["A_ocean"], st["delta_V"], ref.get_field("z_ss")))
    return out


def test_oracle_constant_ocean_surface_recurrence():
    runs = oracle_run(case(ConstantOceanSurfaceBSL()))
    const = oracle_run(case(fastiso.ConstantBSL()))
    z_prev = 0.0
    for k, (_, z, A, dV, zss) in enumerate(runs):
        assert A == A_OCEAN_PD
        assert abs(z - (z_prev - dV / A)) <= 1e-12 * max(abs(z), 1e-6)
        # laterally constant sea surface: z_ss = z_ss_ref + z_bsl
        np.testing.assert_allclose(zss - const[k][4], z, rtol=0, atol=1e-9)
        z_prev = z
    # the ice grows over the first 10 kyr of the cycle: the volume above floatation rises
    assert runs[1][3] > 0 and runs[2][3] > 0
    assert runs[2][1] < runs[1][1] < 0.0


def test_oracle_piecewise_constant_and_imposed_rules():
    z, A = hypsometry()
    runs = oracle_run(case(PiecewiseConstantBSL(ReferenceBSL(0.0, z_vec=z, A_vec=A))))
    A_prev = interpolate0d(z, A, 0.0)
    z_prev = 0.0
    for _, zb, Ab, dV, _ in runs:
        assert abs(zb - (z_prev - dV / A_prev)) <= 1e-12 * max(abs(zb), 1e-6)
        assert abs(Ab - interpolate0d(z, A, zb)) <= 1e-9 * Ab
        z_prev, A_prev = zb, Ab
    tv, zv = np.array([0.0, 8000.0, 12000.0]), np.array([0.0, -4.0, -2.0])
    runs = oracle_run(case(ImposedBSL(tv, zv)))
    for t, (_, zb, _, _, _) in zip(TIMES, runs):
        assert zb == interpolate0d(tv, zv, t)
    # an external update leaves the BSL where it was
    runs = oracle_run(case(ConstantOceanSurfaceBSL(), update=fastiso.ExternalBSLUpdate()))
    assert all(r[1] == 0.0 for r in runs) and runs[1][3] > 0


def test_oracle_bsl_table_out_of_range_is_an_error():
    from oracle import pyref
    tv, zv = np.array([0.0, 6000.0]), np.array([0.0, -1.0])
    w = case(ImposedBSL(tv, zv))
    ref = pyref.RefEngine.from_inputs(sim_inputs(w).inputs)
    u = u0(w)
    ref.rhs(0.0, u)
    with pytest.raises(RuntimeError, match="TimeInterpolation0D"):
        ref.rhs(10000.0, u)


@pytest.mark.gpu
@pytest.mark.parametrize("surface", [False, True])
def test_gpu_piecewise_constant_bsl_vs_oracle(surface):
    from oracle import pyref
    z, A = hypsometry()
    w = case(PiecewiseConstantBSL(ReferenceBSL(0.0, z_vec=z, A_vec=A)), surface=surface)
    sim = hip_sim(w)
    ref = pyref.RefEngine.from_inputs(sim.inputs)
    u = u0(w)
    for t in TIMES:
        dg, _ = sim.engine.rhs(t, u)
        dr, _ = ref.rhs(t, u)
        m = np.abs(dg - dr).max() / np.abs(dr).max()
        assert m < 2e-6, (t, m)
        sg, sr = sim.engine.stats(), ref.stats()
        assert abs(sg["z_bsl"] - sr["z_bsl"]) <= 1e-5 * max(abs(sr["z_bsl"]), 1e-3), (t, sg["z_bsl"], sr["z_bsl"])
        assert abs(sg["A_ocean"] - sr["A_ocean"]) <= 1e-8 * sr["A_ocean"]
        zg, zr = sim.engine.get_field("z_ss"), ref.get_field("z_ss")
        assert np.abs(zg - zr).max() <= 1e-4 + 1e-6 * np.abs(zr).max(), t
    assert sr["z_bsl"] < -0.05


@pytest.mark.gpu
def test_gpu_bsl_table_out_of_range_is_an_error():
    tv, zv = np.array([0.0, 6000.0]), np.array([0.0, -1.0])
    w = case(ImposedBSL(tv, zv))
    sim = hip_sim(w)
    u = u0(w)
    sim.engine.rhs(0.0, u)
    with pytest.raises(RuntimeError, match="TimeInterpolation0D"):
        sim.engine.rhs(10000.0, u)
