"""fastiso: MI355X-native LV-ELVA deformation engine behind the FastIsostasy.jl API.

The Python layer mirrors the reference's user API (RegionalDomain, SolidEarth, BoundaryConditions,
RegionalSeaLevel, Simulation, run!, init_integrator, step!) and performs the host-side FP64 setup;
all time stepping runs in the HIP library libfastiso_hip.so through its C ABI.
"""
from .domain import PhysicalConstants, RegionalDomain, SECONDS_PER_YEAR, meshgrid, nextfastfft
from .bcs import (BoundaryConditions, CornerBC, ExtendedBCSpace, ExternallyUpdatedIceThickness,
                  RegularBCSpace, SeparableIceThickness, TimeInterpolatedIceThickness)
from .sealevel import (A_OCEAN_PD, ConstantBSL, ConstantOceanSurfaceBSL, ExternalBSLUpdate,
                       GoelzerAdjustmentContribution, GoelzerDensityContribution, GoelzerVolumeContribution,
                       ImposedBSL, ImposedSeaSurface, InteractiveSealevelLoad, InternalBSLUpdate, LaterallyConstantSeaSurface,
                       LaterallyVariableSeaSurface, NoAdjustmentContribution, NoDensityContribution,
                       NoSealevelLoad, NoVolumeContribution, PiecewiseConstantBSL, ReferenceBSL,
                       RegionalSeaLevel, interpolate0d, update_bsl)
from .solidearth import (CompressibleLithosphereColumn, CompressibleMantle, IncompressibleLithosphereColumn,
                         IncompressibleMantle, LaterallyConstantLithosphere, LaterallyVariableLithosphere,
                         MaxwellMantle, NoCalibration, RelaxedMantle, RigidLithosphere, RigidMantle,
                         SeakonCalibration, SolidEarth, get_flexural_lengthscale, get_relaxation_time,
                         get_relaxation_time_stronger, get_relaxation_time_weaker)
from .greens import elastic_green, gauss_legendre, geoid_green, green_viscous, load_farrell_table
from .simulation import (DiffEqOptions, Integrator, NativeOutput, Simulation, SolverOptions, init_integrator,
                         push_inputs, run, step, update_diagnostics)
from .utils import uniform_ice_cylinder
from .io import AsymetricOutputCrop, NetcdfOutput, PaddedOutputCrop, crop, read_nc
from .engine import Engine, EngineError

# Julia-style aliases of the mutating entry points
run_ = run
step_ = step
update_diagnostics_ = update_diagnostics
from .ensemble import Ensemble
