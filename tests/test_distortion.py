"""Map distortion K != 1 (RegionalDomain.K, src/domain.jl:140-147,167): grid steps K dx, K dy per
cell, the RHS factor K / (2 eta) (src/deformation.jl:95), K^2 on the elastic convolution input
(src/deformation.jl:177), and the cell area dx dy K^2 in the geoid mass anomaly and the volume
contributions (src/sealevel.jl:193-201,273-331).

Pins of the oracle:
* The stencils with per-cell steps, on polynomials whose differences are exact. With
  u = a x^2 + b x y + c y^2 these hold at every cell, edges included:
  * uxx = 2a / K^2 and uyy = 2c / K^2 at the output cell;
  * ux = (a (x_l + x_r) + b y) / K, with (l, r) the two cells of the difference;
  * uxy = (ux(row above) - ux(row below)) / (span * K dy).
  Using a neighbour's K, or K instead of K^2, fails one of them.
* A constant K = k0 is an Earth with rigidity D / k0^4 and viscosity eta / k0 on an undistorted
  grid. With the elastic response off (zero Green kernel) and a constant sea surface, both give
  the same RHS.

The HIP engine is compared with the oracle on a smooth, variable K with the interactive sea level,
a laterally variable sea surface and a BSL rule that reads the cell areas.
"""
import numpy as np
import pytest

import fastiso
from fastiso.sealevel import ConstantOceanSurfaceBSL

from _cases import hip_sim, lv_case, sim_inputs


def smooth_K(d, amp=0.06):
    r = d.R / d.R.max()
    return 1.0 - amp + 2 * amp * r ** 2 + 0.01 * np.cos(3 * d.X / d.Wx) * np.sin(2 * d.Y / d.Wy)


def test_oracle_per_cell_derivatives_closed_form():
    from oracle import pyref
    nx, ny, dx, dy = 9, 7, 3.0, 2.0
    rng = np.random.default_rng(2)
    K = 0.8 + 0.4 * rng.random((nx, ny))
    x = 5.0 + dx * np.arange(nx)
    y = -4.0 + dy * np.arange(ny)
    X, Y = np.meshgrid(x, y, indexing="ij")
    a, b, c = 0.7, -1.3, 0.4
    u = a * X ** 2 + b * X * Y + c * Y ** 2
    uxx, uyy, ux, uxy = pyref.derivatives_k(u, dx, dy, K)
    np.testing.assert_allclose(uxx, 2 * a / K ** 2, rtol=1e-9, atol=1e-9)
    np.testing.assert_allclose(uyy, 2 * c / K ** 2, rtol=1e-9, atol=1e-9)
    xl = np.concatenate([[x[0]], x[:-2], [x[-2]]])    # the two cells of each x difference
    xr = np.concatenate([[x[1]], x[2:], [x[-1]]])
    want_ux = (a * (xl + xr)[:, None] + b * Y) / K
    np.testing.assert_allclose(ux, want_ux, rtol=1e-9, atol=1e-9)
    jb = np.concatenate([[0], np.arange(ny - 2), [ny - 2]])   # rows of each y difference
    ja = np.concatenate([[1], np.arange(2, ny), [ny - 1]])
    span = (ja - jb)[None, :]
    want_uxy = (want_ux[:, ja] - want_ux[:, jb]) / (span * K * dy)
    np.testing.assert_allclose(uxy, want_uxy, rtol=1e-9, atol=1e-9)


def test_oracle_constant_K_is_a_rescaled_earth():
    from oracle import pyref
    k0 = 1.07
    w = lv_case(n=32, seed=7, sea_surface=False)
    inp = sim_inputs(w).inputs
    d = w["domain"]
    u = -40 * np.exp(-(d.R / 1e6) ** 2)
    # the elastic response off (zero Green kernel): K^2 on its input then changes nothing
    f0 = dict(inp["fields"], green_elastic=np.zeros((d.nx, d.ny)))
    inp = dict(inp, fields=f0)
    a = dict(inp, fields=dict(f0, distortion=np.full((d.nx, d.ny), k0)))
    b = dict(inp, fields=dict(f0, litho_rigidity=f0["litho_rigidity"] / k0 ** 4,
                              effective_viscosity=f0["effective_viscosity"] / k0))
    ea, eb = pyref.RefEngine.from_inputs(a), pyref.RefEngine.from_inputs(b)
    for t in (0.0, 500.0, 600.0):
        da, _ = ea.rhs(t, u)
        db, _ = eb.rhs(t, u)
        assert np.abs(da - db).max() <= 1e-10 * np.abs(db).max(), t
    # the distortion itself changes the answer
    e1 = pyref.RefEngine.from_inputs(inp)
    d1, _ = e1.rhs(0.0, u)
    assert np.abs(d1 - ea.rhs(0.0, u)[0]).max() > 1e-3 * np.abs(d1).max()


def test_domain_distortion_arguments():
    with pytest.raises(NotImplementedError):
        fastiso.RegionalDomain(3e6, 3e6, 16, 16, correct_distortion=True)
    with pytest.raises(ValueError):
        fastiso.RegionalDomain(3e6, 3e6, 16, 16, K=-np.ones((16, 16)))
    K = np.full((16, 16), 1.1)
    d = fastiso.RegionalDomain(3e6, 3e6, 16, 16, K=K)
    assert d.correct_distortion
    np.testing.assert_allclose(d.A, d.dx * d.dy * K ** 2)
    np.testing.assert_allclose(d.Dx, 1.1 * d.dx)


def distorted_case(n=48, seed=9):
    w = lv_case(n=n, seed=seed, interactive=True)
    d0 = w["domain"]
    d = fastiso.RegionalDomain(d0.Wx, d0.Wy, d0.nx, d0.ny, K=smooth_K(d0))
    w["domain"] = d
    r = d.R / d.R.max()
    w["z_b_ref"] = 400.0 - 1400.0 * r
    w["solidearth"] = fastiso.SolidEarth(d, layer_boundaries=[w["h"]], layer_viscosities=[10.0 ** w["logeta"]],
                                         maskactive=r < 0.55)
    w["sealevel"] = fastiso.RegionalSeaLevel(surface=fastiso.LaterallyVariableSeaSurface(),
                                             load=fastiso.InteractiveSealevelLoad(), bsl=ConstantOceanSurfaceBSL())
    w["bcs"] = fastiso.BoundaryConditions(d, ice_thickness=w["bcs"].ice_thickness)
    return w


def test_oracle_distortion_moves_the_sparse_diagnostics():
    from oracle import pyref
    w = distorted_case()
    ref = pyref.RefEngine.from_inputs(sim_inputs(w).inputs)
    d = w["domain"]
    u = -60 * np.exp(-(d.R / 1e6) ** 2)
    for t in (0.0, 5000.0, 5010.0):   # a sparse update convolves the loads of the call before it
        ref.rhs(t, u)
    assert np.abs(ref.get_field("ue")).max() > 0 and ref.stats()["z_bsl"] != 0.0


@pytest.mark.gpu
def test_gpu_distortion_rhs_vs_oracle():
    from oracle import pyref
    w = distorted_case()
    sim = hip_sim(w)
    ref = pyref.RefEngine.from_inputs(sim.inputs)
    d = w["domain"]
    rng = np.random.default_rng(6)
    u = -60 * np.exp(-(d.R / 1e6) ** 2) + 0.01 * rng.standard_normal((d.nx, d.ny))
    for t in (0.0, 5000.0, 5010.0, 5015.0):
        dg, _ = sim.engine.rhs(t, u)
        dr, _ = ref.rhs(t, u)
        m = np.abs(dg - dr).max() / np.abs(dr).max()
        l2 = np.linalg.norm(dg - dr) / np.linalg.norm(dr)
        assert m < 2e-6 and l2 < 1e-6, (t, m, l2)
    for f in ("ue", "dz_ss"):
        a, b = sim.engine.get_field(f), ref.get_field(f)
        assert np.abs(a - b).max() <= 5e-6 * np.abs(b).max(), f
    sg, sr = sim.engine.stats(), ref.stats()
    assert abs(sg["z_bsl"] - sr["z_bsl"]) <= 1e-5 * abs(sr["z_bsl"])


@pytest.mark.gpu
def test_gpu_distortion_large_grid_rhs_vs_oracle():
    """Grids wider than 512 cells keep the plate moments in FP32 and 1/K in registers: the same
    distorted case at n = 600 against the oracle, to the bounds of the undistorted large grids."""
    from oracle import pyref
    w = distorted_case(n=600, seed=3)
    sim = hip_sim(w)
    ref = pyref.RefEngine.from_inputs(sim.inputs)
    d = w["domain"]
    rng = np.random.default_rng(8)
    u = -60 * np.exp(-(d.R / 1e6) ** 2) + 0.01 * rng.standard_normal((d.nx, d.ny))
    for t in (0.0, 5000.0, 5010.0):
        dg, _ = sim.engine.rhs(t, u)
        dr, _ = ref.rhs(t, u)
        m = np.abs(dg - dr).max() / np.abs(dr).max()
        l2 = np.linalg.norm(dg - dr) / np.linalg.norm(dr)
        print(f"t={t} dudt max-rel {m:.3e} L2-rel {l2:.3e}")
        assert m < 2e-6 and l2 < 1e-6, (t, m, l2)
    for f in ("ue", "dz_ss"):
        a, b = sim.engine.get_field(f), ref.get_field(f)
        assert np.abs(a - b).max() <= 5e-6 * np.abs(b).max(), f
    sim.engine.close()


@pytest.mark.gpu
def test_gpu_distortion_needs_six_rows():
    w = lv_case(n=16, seed=3)
    d0 = w["domain"]
    d = fastiso.RegionalDomain(d0.Wx, d0.Wy * 5 / 16, d0.nx, 5, K=np.full((d0.nx, 5), 1.05))
    w["domain"] = d
    w["solidearth"] = fastiso.SolidEarth(d, layer_boundaries=[np.full((d.nx, d.ny), 88e3)],
                                         layer_viscosities=[np.full((d.nx, d.ny), 1e21)])
    w["bcs"] = fastiso.BoundaryConditions(d)
    with pytest.raises(RuntimeError, match="distortion"):
        hip_sim(w)
