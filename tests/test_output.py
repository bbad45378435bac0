"""Output path: NativeOutput / NetcdfOutput snapshots at the output times and the thin-plate
horizontal displacement (src/io.jl:1-353, src/simulation.jl:190-225, src/deformation.jl:141-164).

Pins:
* u_x, u_y of the oracle on a quadratic u, where the CPU difference rules are exact at every
  cell (edges included): u_x = -h/2 (a (x_l + x_r) + b y), u_y = -h/2 (b x + c (y_b + y_a)),
  with (l, r) and (b, a) the two cells of each difference. A wrong sign, h instead of h/2 or a
  swapped axis fails it.
* The NetCDF file read back equals the native snapshots of the same run, in the reference's
  layout (Julia [x, y, t] is (t, y, x) on disk), with the parameters mapped as io_dict maps them
  (log10 viscosity, thickness in km) and the crop applied.
"""
import os

import numpy as np
import pytest

import fastiso
from fastiso import workloads

from _cases import hip_sim, lv_case, sim_inputs


def _quadratic_case(n=12):
    w = workloads.disc_benchmark(n=n, t_end=2e3, t_out=[1e3, 2e3])
    d = w["domain"]
    a, b, c = 2e-9, -3e-9, 1.5e-9
    u = a * d.X ** 2 + b * d.X * d.Y + c * d.Y ** 2
    return w, d, u, (a, b, c)


def test_oracle_thinplate_horizontal_displacement_closed_form():
    from oracle import pyref
    w, d, u, (a, b, c) = _quadratic_case()
    rng = np.random.default_rng(3)
    h = 50e3 + 100e3 * rng.random((d.nx, d.ny))
    inp = sim_inputs(w).inputs
    inp = dict(inp, fields=dict(inp["fields"], litho_thickness=h))
    ref = pyref.RefEngine.from_inputs(inp)
    ref.set_field("u", u)
    x, y = d.x, d.y
    xl = np.concatenate([[x[0]], x[:-2], [x[-2]]])
    xr = np.concatenate([[x[1]], x[2:], [x[-1]]])
    yb = np.concatenate([[y[0]], y[:-2], [y[-2]]])
    ya = np.concatenate([[y[1]], y[2:], [y[-1]]])
    want_x = -h / 2 * (a * (xl + xr)[:, None] + b * d.Y)
    want_y = -h / 2 * (b * d.X + c * (yb + ya)[None, :])
    np.testing.assert_allclose(ref.get_field("u_x"), want_x, rtol=1e-9, atol=1e-12)
    np.testing.assert_allclose(ref.get_field("u_y"), want_y, rtol=1e-9, atol=1e-12)


def _run_with_outputs(w, make_engine, path, pad=1):
    d = w["domain"]
    se = w["solidearth"]
    nout = fastiso.NativeOutput(t=[500.0, 1000.0, 2000.0], vars=["u", "ue", "u_x", "dz_ss"])
    ncout = fastiso.NetcdfOutput(d, [1000.0, 2000.0], path, vars3D=["u", "ue", "u_x", "dz_ss"], vars1D=["z_bsl"],
                                 params2D=["effective_viscosity", "litho_thickness"],
                                 output_crop=fastiso.PaddedOutputCrop(pad), solidearth=se)
    sim = fastiso.Simulation(d, w["bcs"], w["sealevel"], se, (0.0, 2000.0), nout=nout, ncout=ncout,
                             opts=w["opts"], c=w["c"], engine=None)
    sim.engine = make_engine(sim.inputs)
    fastiso.run(sim)
    return sim, nout, ncout


def _check_file(path, nout, se, pad):
    from scipy.io import netcdf_file
    with netcdf_file(path, "r", mmap=False) as nc:
        assert nc.variables["u"].dimensions == ("t", "y", "x")
        assert nc.variables["u"].units == b"m" and nc.variables["t"][:].tolist() == [1000.0, 2000.0]
        assert nc.variables["effective_viscosity"].dimensions == ("y", "x")
    sl = slice(pad, -pad)
    for v in ("u", "ue", "u_x", "dz_ss"):
        F = fastiso.read_nc(path, v)
        assert F.shape[2] == 2
        for k, snap in enumerate(nout.vals[v][1:]):    # nout also holds t = 500
            np.testing.assert_array_equal(F[:, :, k], snap[sl, sl].astype(np.float32))
    np.testing.assert_allclose(fastiso.read_nc(path, "effective_viscosity"),
                               np.log10(se.effective_viscosity[sl, sl]), rtol=1e-6)
    np.testing.assert_allclose(fastiso.read_nc(path, "litho_thickness"), 1e-3 * se.litho_thickness[sl, sl],
                               rtol=1e-6)
    assert fastiso.read_nc(path, "z_bsl").shape == (2,)


def test_netcdf_output_matches_native_snapshots_oracle(tmp_path):
    from oracle import pyref
    w = workloads.disc_benchmark(n=16, t_end=2e3)
    path = str(tmp_path / "out.nc")
    sim, nout, ncout = _run_with_outputs(w, pyref.RefEngine.from_inputs, path)
    assert nout.t_saved == [500.0, 1000.0, 2000.0] and ncout.k == 3
    assert np.abs(nout.vals["u_x"][-1]).max() > 0
    _check_file(path, nout, w["solidearth"], 1)


def test_netcdf_output_needs_nc_suffix(tmp_path):
    d = fastiso.RegionalDomain(1e6, 1e6, 8, 8)
    with pytest.warns(UserWarning):
        o = fastiso.NetcdfOutput(d, [1.0], str(tmp_path / "out.txt"))
    assert not o.active and not os.path.exists(tmp_path / "out.txt")
    with pytest.raises(KeyError):
        fastiso.NetcdfOutput(d, [1.0], str(tmp_path / "o.nc"), vars3D=["nope"])


def test_crops():
    X = np.arange(30.0).reshape(5, 6)
    np.testing.assert_array_equal(fastiso.crop(X, fastiso.PaddedOutputCrop(1)), X[1:4, 1:5])
    np.testing.assert_array_equal(fastiso.crop(X, fastiso.AsymetricOutputCrop(0, 1, 2, 0)), X[0:4, 2:6])
    np.testing.assert_array_equal(fastiso.crop(np.arange(6.0), fastiso.PaddedOutputCrop(2)), [2.0, 3.0])


@pytest.mark.gpu
@pytest.mark.parametrize("distorted", [False, True])
def test_gpu_horizontal_displacement_vs_oracle(distorted):
    from oracle import pyref
    w = lv_case(n=48, seed=21)
    if distorted:
        d0 = w["domain"]
        r = d0.R / d0.R.max()
        d = fastiso.RegionalDomain(d0.Wx, d0.Wy, d0.nx, d0.ny, K=1.0 - 0.05 + 0.1 * r ** 2)
        w["domain"] = d
        w["solidearth"] = fastiso.SolidEarth(d, layer_boundaries=[w["h"]], layer_viscosities=[10.0 ** w["logeta"]])
        w["bcs"] = fastiso.BoundaryConditions(d, ice_thickness=w["bcs"].ice_thickness)
    sim = hip_sim(w)
    ref = pyref.RefEngine.from_inputs(sim.inputs)
    d = w["domain"]
    u = -60 * np.exp(-(d.R / 1e6) ** 2)
    for t in (0.0, 500.0, 510.0):
        sim.engine.rhs(t, u)
        ref.rhs(t, u)
    assert np.abs(ref.get_field("ue")).max() > 0
    for f in ("u_x", "u_y"):
        a, b = sim.engine.get_field(f), ref.get_field(f)
        assert np.abs(a - b).max() <= 2e-6 * np.abs(b).max(), f


@pytest.mark.gpu
def test_gpu_netcdf_output_matches_native_snapshots(tmp_path):
    from _cases import DRYRUN
    w = workloads.disc_benchmark(n=32, t_end=2e3)
    path = str(tmp_path / "out_gpu.nc")
    if DRYRUN:
        from oracle import pyref
        make = pyref.RefEngine.from_inputs
    else:
        def make(inputs):
            return fastiso.simulation.push_inputs(fastiso.Engine(inputs["cfg"]), inputs)
    sim, nout, ncout = _run_with_outputs(w, make, path, pad=2)
    assert nout.t_saved == [500.0, 1000.0, 2000.0]
    _check_file(path, nout, w["solidearth"], 2)
