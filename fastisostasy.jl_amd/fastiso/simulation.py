"""Simulation / run! / init_integrator / step! / update_diagnostics! over the HIP engine.

Host-side mirror of src/simulation.jl: the setup (GIATools Green kernels, solid-Earth fields,
reference state inputs) is computed here in FP64 numpy as the reference computes it on the CPU;
every right-hand side, sparse update and integration step runs in libfastiso_hip.so.
"""
import time
from dataclasses import dataclass, field

import numpy as np

from . import _abi
from .bcs import ExternallyUpdatedIceThickness, SeparableIceThickness, TimeInterpolatedIceThickness
from .domain import PhysicalConstants
from .engine import Engine
from .greens import elastic_green, geoid_green, green_viscous
from .sealevel import (GoelzerAdjustmentContribution, GoelzerDensityContribution, GoelzerVolumeContribution,
                       ImposedSeaSurface, InteractiveSealevelLoad, LaterallyVariableSeaSurface, bsl_inputs)
from .solidearth import (CompressibleLithosphereColumn, LaterallyConstantLithosphere, MaxwellMantle,
                         RelaxedMantle, RigidLithosphere, RigidMantle)


@dataclass
class DiffEqOptions:
    """src/simulation.jl:11-15 (BS3, reltol 1f-5); abstol is OrdinaryDiffEq's default 1e-6."""
    alg: str = "BS3"
    reltol: float = 1e-5
    abstol: float = 1e-6
    dt_min: float = None


@dataclass
class SolverOptions:
    """src/simulation.jl:22-26."""
    diffeq: DiffEqOptions = field(default_factory=DiffEqOptions)
    dt_sparse_diagnostics: float = 10.0
    verbose: bool = False


class NativeOutput:
    """In-memory snapshots at output times (src/io.jl:336-353)."""
    def __init__(self, t=(), vars=()):
        self.t = [float(x) for x in t]
        self.vars = list(vars)
        self.vals = {v: [] for v in self.vars}
        self.t_saved = []
        self.computation_time = 0.0
        self.k = 1


class Timer:
    """Timer (src/simulation.jl:28-52): t_vec and t_computation hold, for every right-hand-side
    call past t_span[0], its time (yr) and the seconds since the first such call, recorded on the
    device (fi_set_timer_history); they are read from the engine on access. n_total counts every
    call; the first `capacity` are kept."""
    def __init__(self, sim, capacity):
        self._sim, self.capacity = sim, int(capacity)
        self.t_span = sim.t_span

    @property
    def t(self):
        return self._sim.engine.stats()["t"]

    def _read(self):
        if self._sim.engine is None or self.capacity == 0:
            return np.zeros(0), np.zeros(0), 0
        return self._sim.engine.timer_history(self.capacity)

    @property
    def t_vec(self):
        return self._read()[0]

    @property
    def t_computation(self):
        return self._read()[1]

    @property
    def n_total(self):
        return self._read()[2]


def build_config(domain, c, solidearth, sealevel, bcs, opts, device=0):
    """The POD fi_config for a set of reference-style model objects."""
    cfg = _abi.FiConfig()
    cfg.nx, cfg.ny = domain.nx, domain.ny
    cfg.Wx, cfg.Wy = domain.Wx, domain.Wy
    cfg.g, cfg.seconds_per_year = c.g, c.seconds_per_year
    cfg.rho_ice, cfg.rho_water, cfg.rho_seawater = c.rho_ice, c.rho_water, c.rho_seawater
    cfg.r_pole, cfg.mE = c.r_pole, c.mE
    cfg.litho_poissonratio = solidearth.litho_poissonratio
    cfg.rho_uppermantle, cfg.rho_litho = solidearth.rho_uppermantle, solidearth.rho_litho
    li = solidearth.lithosphere
    cfg.lithosphere = (_abi.FI_RIGID_LITHOSPHERE if isinstance(li, RigidLithosphere) else
                       _abi.FI_LATERALLY_CONSTANT_LITHOSPHERE if isinstance(li, LaterallyConstantLithosphere)
                       else _abi.FI_LATERALLY_VARIABLE_LITHOSPHERE)
    ma = solidearth.mantle
    cfg.mantle = (_abi.FI_RIGID_MANTLE if isinstance(ma, RigidMantle) else
                  _abi.FI_MAXWELL_MANTLE if isinstance(ma, MaxwellMantle) else _abi.FI_RELAXED_MANTLE)
    cfg.lithosphere_column = (_abi.FI_COMPRESSIBLE_LITHOSPHERE_COLUMN
                              if isinstance(solidearth.lithosphere_column, CompressibleLithosphereColumn)
                              else _abi.FI_INCOMPRESSIBLE_LITHOSPHERE_COLUMN)
    cfg.sea_surface = (_abi.FI_LATERALLY_VARIABLE_SEA_SURFACE
                       if isinstance(sealevel.surface, LaterallyVariableSeaSurface)
                       else _abi.FI_IMPOSED_SEA_SURFACE if isinstance(sealevel.surface, ImposedSeaSurface)
                       else _abi.FI_LATERALLY_CONSTANT_SEA_SURFACE)
    cfg.sealevel_load = (_abi.FI_INTERACTIVE_SEALEVEL_LOAD if isinstance(sealevel.load, InteractiveSealevelLoad)
                         else _abi.FI_NO_SEALEVEL_LOAD)
    cfg.volume_contribution = int(isinstance(sealevel.volume_contribution, GoelzerVolumeContribution))
    cfg.density_contribution = int(isinstance(sealevel.density_contribution, GoelzerDensityContribution))
    cfg.adjustment_contribution = int(isinstance(sealevel.adjustment_contribution, GoelzerAdjustmentContribution))
    cfg.bsl_z = sealevel.bsl.z
    cfg.viscous_x_border = bcs.viscous_displacement.x_border
    cfg.elastic_x_border = bcs.elastic_displacement.x_border
    cfg.ssh_x_border = bcs.sea_surface_perturbation.x_border
    cfg.dt_sparse = opts.dt_sparse_diagnostics
    cfg.reltol = opts.diffeq.reltol
    cfg.abstol = opts.diffeq.abstol
    # dt_min is only compatible with Euler (src/simulation.jl:244-253); it drives the implicit step
    if opts.diffeq.dt_min is not None:
        if str(opts.diffeq.alg).lower() != "euler":
            raise ValueError("The `dt_min` option is only compatible with the Euler algorithm.")
        cfg.dt_min = float(opts.diffeq.dt_min)
    cfg.device = device
    return cfg


class Simulation:
    """Simulation(domain, bcs, sealevel, solidearth, t_span; opts, u_ref, ue_ref, dz_ss_ref, z_b_ref,
    nout, c) (src/simulation.jl:77-156).

    engine="hip" (default) creates the device context; engine=None only assembles the inputs
    (config, fields, forcing), which is what the CPU-side tests inspect."""

    def __init__(self, domain, bcs, sealevel, solidearth, t_span, opts=None, u_ref=None, ue_ref=None,
                 dz_ss_ref=None, z_b_ref=None, nout=None, ncout=None, c=None, engine="hip", device=0,
                 quad_precision=4, timer_capacity=1 << 20):
        self.domain, self.bcs, self.sealevel, self.solidearth = domain, bcs, sealevel, solidearth
        self.t_span = (float(t_span[0]), float(t_span[1]))
        self.opts = opts or SolverOptions()
        self.c = c or PhysicalConstants()
        self.nout = nout or NativeOutput()
        self.ncout = ncout
        if isinstance(solidearth.mantle, MaxwellMantle) and not hasattr(solidearth, "litho_rigidity"):
            raise ValueError("solidearth incomplete")
        cfg = build_config(domain, self.c, solidearth, sealevel, bcs, self.opts, device)
        fields = {
            "litho_rigidity": solidearth.litho_rigidity,
            "effective_viscosity": solidearth.effective_viscosity,
            "scaled_pseudodiff_inv": solidearth.scaled_pseudodiff_inv,
            "maskactive": solidearth.maskactive.astype(np.float64),
            "litho_thickness": solidearth.litho_thickness,
        }
        if np.any(domain.K != 1.0):
            fields["distortion"] = domain.K
        # GIATools (src/tools.jl:49-98): Green kernels of the elastic and sea-surface responses
        if not isinstance(solidearth.lithosphere, RigidLithosphere):
            fields["green_elastic"] = elastic_green(domain, quad_precision=quad_precision)
        if isinstance(sealevel.surface, LaterallyVariableSeaSurface):
            fields["green_geoid"] = geoid_green(domain, self.c)
        if isinstance(solidearth.mantle, RelaxedMantle):
            # ELRA (src/tools.jl:55-60): the viscous Green kernel of the mean rigidity, and tau
            fields["green_viscous"] = green_viscous(domain, solidearth.rho_uppermantle,
                                                    float(np.mean(solidearth.litho_rigidity)))
            fields["tau"] = np.broadcast_to(np.asarray(solidearth.tau, dtype=np.float64), (domain.nx, domain.ny))
        for name, val in (("ref_u", u_ref), ("ref_ue", ue_ref), ("ref_dz_ss", dz_ss_ref), ("ref_z_b", z_b_ref)):
            if val is not None:
                fields[name] = np.broadcast_to(np.asarray(val, dtype=np.float64), (domain.nx, domain.ny))
        it = bcs.ice_thickness
        if isinstance(it, TimeInterpolatedIceThickness):
            ice = ("snapshots", it.t_vec, it.H_vec, it.flat_bc)
        elif isinstance(it, SeparableIceThickness):
            ice = ("separable", it.t_vec, it.a_vec, it.H_shape, it.flat_bc)
        elif isinstance(it, ExternallyUpdatedIceThickness):
            ice = ("external",)
        else:
            raise TypeError("unknown ice thickness forcing")
        self.inputs = {"cfg": cfg, "fields": fields, "ice": ice, "t_span": self.t_span,
                       "tout": sorted(set(self.nout.t) | set(self.ncout.t if self.ncout else ())),
                       "bsl": bsl_inputs(sealevel.bsl, sealevel.update_bsl),
                       "sea_surface": ((sealevel.surface.t_vec, sealevel.surface.dz_ss_vec, sealevel.surface.flat_bc)
                                       if isinstance(sealevel.surface, ImposedSeaSurface) else None),
                       "timer_capacity": int(timer_capacity)}
        self.timer = Timer(self, timer_capacity)
        self.engine = None
        if engine == "hip":
            self.engine = Engine(cfg)
            push_inputs(self.engine, self.inputs)

    @property
    def now(self):
        return _StateView(self.engine)

    def stats(self):
        return self.engine.stats()

    def close(self):
        """Free the device context now (instead of at garbage collection)."""
        if self.engine is not None:
            self.engine.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
        return False


def push_inputs(eng, inputs, timer=True):
    """Load assembled inputs into an engine context and run fi_init (timer=False: no Timer
    history, as for ensemble members)."""
    for name, X in inputs["fields"].items():
        eng.set_field(name, X)
    ice = inputs["ice"]
    if ice[0] == "snapshots":
        eng.set_ice_snapshots(ice[1], ice[2], ice[3])
    elif ice[0] == "separable":
        eng.set_ice_separable(ice[1], ice[2], ice[3], ice[4])
    if inputs.get("sea_surface") is not None:
        eng.set_sea_surface_snapshots(*inputs["sea_surface"])
    if timer and inputs.get("timer_capacity"):
        eng.set_timer_history(inputs["timer_capacity"])
    if "bsl" in inputs:
        eng.set_bsl(*inputs["bsl"])
    eng.set_output_times(inputs["tout"])
    eng.init(*inputs["t_span"])
    return eng


class _StateView:
    """sim.now.<field> reads the current state from the device (src/state.jl:49-77)."""
    def __init__(self, eng):
        object.__setattr__(self, "_e", eng)

    def __getattr__(self, name):
        return self._e.get_field(name)

    def __setattr__(self, name, val):
        self._e.set_field(name, val)


def update_diagnostics(dudt, u, sim, t):
    """The ODE right-hand side (src/simulation.jl:335-393). dudt is written, u is mutated by the
    viscous corner BC exactly as the reference mutates the solver's stage vector."""
    du, ubc = sim.engine.rhs(t, u)
    dudt[...] = du
    u[...] = ubc


def _output_callback(sim):
    """nout_affect! / nc_affect! under their conditions (src/simulation.jl:190-225): the engine
    stops at the union of both output time lists; each output takes the times it asked for."""
    nout, ncout = sim.nout, sim.ncout
    t_start = time.perf_counter()

    def cb(t, k):
        if nout.k <= len(nout.t) and t >= nout.t[nout.k - 1]:
            for v in nout.vars:
                nout.vals[v].append(sim.engine.get_field(v))
            nout.t_saved.append(t)
            nout.k += 1
            nout.computation_time = time.perf_counter() - t_start
            if sim.opts.verbose:
                print(f"Saving native output at simulation year {t}...")
        if ncout is not None and ncout.k <= len(ncout.t) and t >= ncout.t[ncout.k - 1]:
            if ncout.active:
                if sim.opts.verbose:
                    print(f"Saving nc output at index {ncout.k}, sim year {t}...")
                ncout.write(sim.engine)
            ncout.k += 1
    return cb


def run(sim):
    """run!(sim) (src/simulation.jl:236-262): init_problem!, then BS3 over t_span with tstops at the
    output times and the output callback."""
    sim.engine.integrator_init()
    sim.engine.integrate(sim.t_span[1], _output_callback(sim))
    return None


class Integrator:
    """init_integrator(sim) + step!(integrator, dt, true) (src/simulation.jl:293-322)."""
    def __init__(self, sim):
        self.sim = sim
        sim.engine.integrator_init()
        self._cb = _output_callback(sim)

    @property
    def t(self):
        return self.sim.engine.stats()["t"]

    def step(self, dt, force_dt=True):
        if not force_dt:
            raise NotImplementedError("step! is only defined with force_dt = true")
        self.sim.engine.integrate(self.t + float(dt), self._cb)


def init_integrator(sim):
    return Integrator(sim)


def step(integrator, dt, force_dt=True):
    integrator.step(dt, force_dt)
