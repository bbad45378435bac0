"""Small seeded cases shared by the oracle pins and the GPU parity tests."""
import os

import numpy as np

import fastiso
from fastiso import workloads


# FASTISO_TEST_DRYRUN=1 runs the GPU tests on the CPU with the oracle standing in for the HIP
# engine: it checks the tests' own expectations before GPU time is spent; it proves nothing
# about the engine.
DRYRUN = os.environ.get("FASTISO_TEST_DRYRUN") == "1"


def hip_sim(w):
    sim = sim_inputs(w, engine=None if DRYRUN else "hip")
    if DRYRUN:
        from oracle import pyref
        sim.engine = pyref.RefEngine.from_inputs(sim.inputs)
    return sim


def sim_inputs(w, engine=None):
    return fastiso.Simulation(w["domain"], w["bcs"], w["sealevel"], w["solidearth"], w["t_span"],
                              nout=w["nout"], opts=w.get("opts"), c=w.get("c"), engine=engine,
                              z_b_ref=w.get("z_b_ref"))


def cosine_case(nx=16, ny=12, W=3e6, k=2, l=1, A=300.0, H0=1000.0, eta=1e21):
    """Flexure-free Earth (zero lithosphere thickness), constant viscosity, ice H0 + A cos(k x)
    switched on over the first year: the RHS is a closed form in the Fourier modes k, l."""
    d = fastiso.RegionalDomain(W, W * ny / nx, nx, ny)
    se = fastiso.SolidEarth(d, layer_boundaries=[np.zeros((nx, ny))], layer_viscosities=[eta], rho_litho=500.0)
    H1 = H0 + A * np.cos(2 * np.pi * k * d.X / (2 * d.Wx))
    it = fastiso.TimeInterpolatedIceThickness([0.0, 1.0, 2.0], [d.zeros(), H1, H1])
    bcs = fastiso.BoundaryConditions(d, ice_thickness=it)
    return dict(domain=d, bcs=bcs, sealevel=fastiso.RegionalSeaLevel(), solidearth=se, t_span=(0.0, 2.0),
                nout=fastiso.NativeOutput(), opts=fastiso.SolverOptions(), c=fastiso.PhysicalConstants(),
                k=k, l=l, A=A, H0=H0, eta=eta)


def lv_case(n=32, seed=11, sea_surface=True, interactive=False, separable=True, t_end=2000.0, W=3e6,
            z_b_ref=None):
    """Small laterally variable case with the same generators as the N=350 workload."""
    w = workloads.lv_elva(n=n, W=W, t_end=t_end, seeds=(seed, seed + 1, seed + 2), separable=separable,
                          t_out=np.linspace(t_end / 4, t_end, 4))
    if not sea_surface:
        w["sealevel"] = fastiso.RegionalSeaLevel()
    if interactive:
        w["sealevel"] = fastiso.RegionalSeaLevel(surface=fastiso.LaterallyVariableSeaSurface(),
                                                 load=fastiso.InteractiveSealevelLoad())
    w["z_b_ref"] = z_b_ref
    return w


def corner_mean(F):
    return (F[0, 0] + F[-1, 0] + F[0, -1] + F[-1, -1]) / 4
