"""SolidEarth (host setup, FP64): rigidity, effective viscosity and the Fourier scaling R."""
import numpy as np


class RigidLithosphere:                pass   # src/solidearth.jl:20
class LaterallyConstantLithosphere:    pass   # src/solidearth.jl:28
class LaterallyVariableLithosphere:    pass   # src/solidearth.jl:36
class RigidMantle:                     pass   # src/solidearth.jl:57
class RelaxedMantle:                   pass   # src/solidearth.jl:66
class MaxwellMantle:                   pass   # src/solidearth.jl:75
class CompressibleLithosphereColumn:   pass   # src/solidearth.jl:102
class IncompressibleLithosphereColumn: pass   # src/solidearth.jl:109
class CompressibleMantle:              pass   # src/material.jl:34
class IncompressibleMantle:            pass   # src/material.jl:39
class NoCalibration:                   pass   # src/material.jl:66
class FreqDomainViscosityLumping:      pass   # src/material.jl:108


class SeakonCalibration:
    """src/material.jl:71-87."""
    def __init__(self, ref_viscosity=1e21):
        self.ref_viscosity = ref_viscosity


DEFAULT_RHO_LITHO = 3.2e3
DEFAULT_LITHO_YOUNGMODULUS = 6.6e10
DEFAULT_LITHO_POISSONRATIO = 0.28
DEFAULT_LITHO_THICKNESS = 88e3
DEFAULT_RHO_UPPERMANTLE = 3.4e3
DEFAULT_MANTLE_POISSONRATIO = 0.28
DEFAULT_MANTLE_TAU = 855.0


def get_rigidity(t, E, nu):
    """src/material.jl:10."""
    return (E * t ** 3) / (12 * (1 - nu ** 2))


def get_relaxation_time(eta, m, p):
    """Relaxation time (yr) of the relaxed mantle from a viscosity map, 10^(m log10(eta) - p)
    (src/material.jl:406)."""
    return 10.0 ** (np.log10(eta) * m - p)


def get_relaxation_time_weaker(eta):
    """src/material.jl:413 (Van Calcar et al., weaker mantle)."""
    return get_relaxation_time(eta, 0.35, 4.63)


def get_relaxation_time_stronger(eta):
    """src/material.jl:420 (Van Calcar et al., stronger mantle)."""
    return get_relaxation_time(eta, 0.20, 1.41)


def get_flexural_lengthscale(litho_rigidity, rho_uppermantle, g):
    """(D / (rho g))^(1/4) (src/material.jl:396-399)."""
    return (litho_rigidity / (rho_uppermantle * g)) ** 0.25


def channel_scaling(kappa, channel_thickness, visc_ratio):
    """src/material.jl:232-262, in FP64."""
    C = np.cosh(channel_thickness * kappa)
    S = np.sinh(channel_thickness * kappa)
    num = np.zeros_like(kappa)
    den = np.zeros_like(kappa)
    num += 2 * visc_ratio * C * S
    num += (1 - visc_ratio ** 2) * channel_thickness ** 2 * kappa ** 2
    num += visc_ratio ** 2 * S ** 2 + C ** 2
    den += (visc_ratio + 1 / visc_ratio) * C * S
    den += (visc_ratio - 1 / visc_ratio) * channel_thickness * kappa
    den += S ** 2 + C ** 2
    return num / den


def _matrify(v, nx, ny):
    """Layer list of scalars or (nx, ny) arrays -> (nx, ny, nl) (src/utils.jl:60-70)."""
    layers = [np.full((nx, ny), float(x)) if np.ndim(x) == 0 else np.asarray(x, dtype=np.float64)
              for x in v]
    return np.stack(layers, axis=2)


class SolidEarth:
    """SolidEarth(domain; ...) (src/solidearth.jl:142-233) with FreqDomainViscosityLumping
    (src/material.jl:154-178): effective_viscosity is the half-space viscosity, the channels enter
    through R(kappa) = prod channel_scaling(kappa, mean thickness, mean viscosity ratio)."""

    def __init__(self, domain, lithosphere=None, mantle=None, calibration=None, compressibility=None,
                 lumping=None, lithosphere_column=None, maskactive=None,
                 layer_boundaries=(88e3, 400e3), layer_viscosities=(1e19, 1e21),
                 litho_youngmodulus=DEFAULT_LITHO_YOUNGMODULUS, litho_poissonratio=DEFAULT_LITHO_POISSONRATIO,
                 mantle_poissonratio=DEFAULT_MANTLE_POISSONRATIO, tau=DEFAULT_MANTLE_TAU, scale_elralength=1.0,
                 rho_uppermantle=DEFAULT_RHO_UPPERMANTLE, rho_litho=DEFAULT_RHO_LITHO):
        self.lithosphere = lithosphere or LaterallyVariableLithosphere()
        self.mantle = mantle or MaxwellMantle()
        self.calibration = calibration or NoCalibration()
        self.compressibility = compressibility or CompressibleMantle()
        self.lumping = lumping or FreqDomainViscosityLumping()
        self.lithosphere_column = lithosphere_column or IncompressibleLithosphereColumn()
        if not isinstance(self.lumping, FreqDomainViscosityLumping):
            raise NotImplementedError("only FreqDomainViscosityLumping (the reference default) is supported")
        nx, ny = domain.nx, domain.ny
        self.maskactive = (np.ones((nx, ny), dtype=bool) if maskactive is None
                           else np.asarray(maskactive, dtype=bool))
        lb = _matrify(layer_boundaries, nx, ny)
        lv = _matrify(layer_viscosities, nx, ny)
        self.layer_boundaries, self.layer_viscosities = lb, lv
        self.litho_thickness = lb[:, :, 0].copy()
        self.litho_rigidity = get_rigidity(self.litho_thickness, litho_youngmodulus, litho_poissonratio)
        eta = lv[:, :, -1].copy()
        R = np.ones((nx, ny))
        nl = lv.shape[2]
        for l in range(1, nl):
            channel_thickness = lb[:, :, nl - l] - lb[:, :, nl - l - 1]
            visc_ratio = lv[:, :, nl - l - 1] / eta
            R *= channel_scaling(domain.pseudodiff, np.mean(channel_thickness),
                                 np.mean(visc_ratio[self.maskactive]))
            if not np.all(np.isfinite(R)):
                raise ValueError("The scaling factor R is too large.")
        if isinstance(self.compressibility, CompressibleMantle):
            eta = eta * ((1 + 0.5) / (1 + mantle_poissonratio))        # src/material.jl:46-51
        if isinstance(self.calibration, SeakonCalibration):
            eta = eta * np.exp(np.log10(self.calibration.ref_viscosity / eta))  # src/material.jl:83-87
        self.effective_viscosity = eta
        self.pseudodiff_scaling = R
        self.scaled_pseudodiff_inv = 1.0 / (R * domain.pseudodiff)       # src/solidearth.jl:222
        self.litho_poissonratio = float(litho_poissonratio)
        self.mantle_poissonratio = float(mantle_poissonratio)
        self.litho_youngmodulus = float(litho_youngmodulus)
        self.litho_shearmodulus = litho_youngmodulus / (2 * (1 + litho_poissonratio))
        self.rho_uppermantle = float(rho_uppermantle)
        self.rho_litho = float(rho_litho)
        self.tau = np.full((nx, ny), float(tau)) if np.ndim(tau) == 0 else np.asarray(tau, dtype=np.float64)
        self.scale_elralength = scale_elralength
