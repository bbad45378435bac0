"""RegionalSeaLevel model choices (src/sealevel.jl:106-166, src/barystatic_sealevel.jl)."""
import numpy as np


class LaterallyConstantSeaSurface:  pass   # src/sealevel.jl:122
class LaterallyVariableSeaSurface:  pass   # src/sealevel.jl:130


class ImposedSeaSurface:
    """ImposedSeaSurface(dz_ss_itp) (src/sealevel.jl:132-134): every sparse update sets dz_ss to
    the piecewise-linear interpolation of the snapshots dz_ss_vec at times t_vec (yr)
    (update_dz_ss!, src/sealevel.jl:188-191; TimeInterpolation2D, src/interpolations.jl:51-74)."""
    def __init__(self, t_vec, dz_ss_vec, flat_bc=False):
        self.t_vec = np.asarray(t_vec, dtype=np.float64)
        self.dz_ss_vec = [np.asarray(x, dtype=np.float64) for x in dz_ss_vec]
        if len(self.t_vec) != len(self.dz_ss_vec) or len(self.t_vec) < 1:
            raise ValueError("t_vec and dz_ss_vec must have the same, non-zero length")
        if np.any(np.diff(self.t_vec) <= 0):
            raise ValueError("t_vec must increase")
        self.flat_bc = bool(flat_bc)
class NoSealevelLoad:               pass   # src/loads.jl:19
class InteractiveSealevelLoad:      pass   # src/loads.jl:26
class InternalBSLUpdate:            pass   # src/barystatic_sealevel.jl:60
class ExternalBSLUpdate:            pass   # src/barystatic_sealevel.jl:67
class NoVolumeContribution:         pass   # src/barystatic_sealevel.jl:38
class GoelzerVolumeContribution:    pass   # src/barystatic_sealevel.jl:45
class NoDensityContribution:        pass   # src/barystatic_sealevel.jl:70
class GoelzerDensityContribution:   pass   # src/barystatic_sealevel.jl:77
class NoAdjustmentContribution:     pass   # src/barystatic_sealevel.jl:94
class GoelzerAdjustmentContribution: pass  # src/barystatic_sealevel.jl:101


A_OCEAN_PD = 3.625e14   # ocean surface (m^2), src/barystatic_sealevel.jl:1


class ReferenceBSL:
    """ReferenceBSL(z, A, z_vec, A_vec; flat_bc) (src/barystatic_sealevel.jl:27-41). The reference
    downloads the ocean-surface table (z_vec, A_vec) and rescales it to A_OCEAN_PD at z = 0; no
    download is possible here, so the caller passes the table (or none: A = A_OCEAN_PD, and only
    the rules that never read the table can use it)."""
    def __init__(self, z=0.0, A=None, z_vec=None, A_vec=None, flat_bc=False):
        self.z = float(z)
        self.z_vec = None if z_vec is None else np.asarray(z_vec, dtype=np.float64)
        self.A_vec = None if A_vec is None else np.asarray(A_vec, dtype=np.float64)
        self.flat_bc = bool(flat_bc)
        if A is None:
            A = interpolate0d(self.z_vec, self.A_vec, self.z, self.flat_bc) if self.z_vec is not None else A_OCEAN_PD
        self.A = float(A)


def interpolate0d(x, y, t, flat_bc=False):
    """interpolate(t, ::TimeInterpolation0D) (src/interpolations.jl:28-49)."""
    x, y = np.asarray(x, dtype=np.float64), np.asarray(y, dtype=np.float64)
    if t < x[0] or t > x[-1]:
        if not flat_bc:
            raise ValueError("TimeInterpolation0D out of range.")
        return float(y[0] if t < x[0] else y[-1])
    i = int(np.searchsorted(x, t, side="left"))
    if x[i] == t:
        return float(y[i])
    return float(y[i - 1] + (y[i] - y[i - 1]) / (x[i] - x[i - 1]) * (t - x[i - 1]))


class ConstantBSL:
    """BSL held at its reference value (src/barystatic_sealevel.jl:100-110, 215-218)."""
    def __init__(self, z=0.0, ref=None):
        self.ref = ref or ReferenceBSL(z)
        self.z, self.A = self.ref.z, self.ref.A


class ConstantOceanSurfaceBSL:
    """z += -delta_V / A with A = ref.A (src/barystatic_sealevel.jl:123-129, 220-224)."""
    def __init__(self, ref=None):
        self.ref = ref or ReferenceBSL()
        self.z, self.A = self.ref.z, self.ref.A


class PiecewiseConstantBSL:
    """z += -delta_V / A, then A = ref.A_itp(z) (src/barystatic_sealevel.jl:142-149, 226-230).
    Needs a ReferenceBSL with its ocean-surface table."""
    def __init__(self, ref):
        if ref is None or ref.z_vec is None:
            raise ValueError("PiecewiseConstantBSL needs ReferenceBSL(z_vec=..., A_vec=...)")
        self.ref = ref
        self.z, self.A = ref.z, ref.A


class ImposedBSL:
    """z = z_itp(t), a TimeInterpolation0D over (t_vec, z_vec) (src/barystatic_sealevel.jl:185-191,
    232-235)."""
    def __init__(self, t_vec, z_vec, z=None, flat_bc=False, ref=None):
        self.ref = ref or ReferenceBSL()
        self.t_vec = np.asarray(t_vec, dtype=np.float64)
        self.z_vec = np.asarray(z_vec, dtype=np.float64)
        self.flat_bc = bool(flat_bc)
        self.z = float(self.z_vec[0] if z is None else z)
        self.A = self.ref.A


def update_bsl(bsl, delta_V, t):
    """update_bsl!(bsl, delta_V, t) (src/barystatic_sealevel.jl:215-235) on the host objects; the
    engine applies the same rules on the device at every sparse update, with -delta_V of the
    volume contributions."""
    if isinstance(bsl, ConstantOceanSurfaceBSL):
        bsl.A = bsl.ref.A
        bsl.z += delta_V / bsl.A
    elif isinstance(bsl, PiecewiseConstantBSL):
        bsl.z += delta_V / bsl.A
        bsl.A = interpolate0d(bsl.ref.z_vec, bsl.ref.A_vec, bsl.z, bsl.ref.flat_bc)
    elif isinstance(bsl, ImposedBSL):
        bsl.z = interpolate0d(bsl.t_vec, bsl.z_vec, t, bsl.flat_bc)
    else:
        bsl.z = bsl.ref.z


def bsl_inputs(bsl, update_bsl):
    """(rule, update, A_ref, x, y, flat_bc) of fi_set_bsl for a BSL rule and update mechanism."""
    update = 1 if isinstance(update_bsl, ExternalBSLUpdate) else 0
    if isinstance(bsl, PiecewiseConstantBSL):
        return (2, update, bsl.ref.A, bsl.ref.z_vec, bsl.ref.A_vec, bsl.ref.flat_bc)
    if isinstance(bsl, ImposedBSL):
        return (3, update, bsl.ref.A, bsl.t_vec, bsl.z_vec, bsl.flat_bc)
    if isinstance(bsl, ConstantOceanSurfaceBSL):
        return (1, update, bsl.ref.A, None, None, False)
    return (0, update, bsl.ref.A, None, None, False)


class RegionalSeaLevel:
    def __init__(self, surface=None, load=None, bsl=None, update_bsl=None, volume_contribution=None,
                 density_contribution=None, adjustment_contribution=None):
        self.surface = surface or LaterallyConstantSeaSurface()
        self.load = load or NoSealevelLoad()
        self.bsl = bsl or ConstantBSL()
        self.update_bsl = update_bsl or InternalBSLUpdate()
        self.volume_contribution = volume_contribution or GoelzerVolumeContribution()
        self.density_contribution = density_contribution or GoelzerDensityContribution()
        self.adjustment_contribution = adjustment_contribution or NoAdjustmentContribution()
        if not isinstance(self.bsl, (ConstantBSL, ConstantOceanSurfaceBSL, PiecewiseConstantBSL, ImposedBSL)):
            raise NotImplementedError("BSL rule not supported by this engine (PiecewiseLinearOceanSurfaceBSL "
                                      "needs NLsolve; CombinedBSL has no update_bsl! in the reference)")
