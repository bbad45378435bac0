"""Boundary conditions and ice forcing (src/boundary_conditions.jl, src/interpolations.jl)."""
import numpy as np


class RegularBCSpace:  pass   # src/boundary_conditions.jl:93
class ExtendedBCSpace: pass   # src/boundary_conditions.jl:101


class CornerBC:
    """Offset so that the mean of the four corners equals x_border (src/boundary_conditions.jl:163-166)."""
    def __init__(self, space=None, x_border=0.0):
        self.space = space or RegularBCSpace()
        self.x_border = float(x_border)


class TimeInterpolatedIceThickness:
    """Piecewise-linear ice thickness between snapshots H_vec at times t_vec (yr)
    (src/boundary_conditions.jl:25-40, src/interpolations.jl:12-74)."""
    def __init__(self, t_vec, H_vec, domain=None, flat_bc=False):
        self.t_vec = np.asarray(t_vec, dtype=np.float64)
        self.H_vec = [np.asarray(H, dtype=np.float64) for H in H_vec]
        if len(self.t_vec) != len(self.H_vec):
            raise ValueError("t_vec and H_vec must have the same length")
        self.flat_bc = flat_bc

    def at(self, t):
        """interpolate! (src/interpolations.jl:51-74) on the host."""
        tv = self.t_vec
        if t < tv[0] or t > tv[-1]:
            if not self.flat_bc:
                raise ValueError("TimeInterpolation2D out of range.")
            return self.H_vec[0 if t < tv[0] else -1].copy()
        i = int(np.searchsorted(tv, t, side="left"))
        if i < len(tv) and tv[i] == t:
            return self.H_vec[i].copy()
        i -= 1
        return self.H_vec[i] + (self.H_vec[i + 1] - self.H_vec[i]) / (tv[i + 1] - tv[i]) * (t - tv[i])


class SeparableIceThickness:
    """Ice thickness a(t) * H_shape with a(t) piecewise linear over (t_vec, a_vec): the same forcing
    as TimeInterpolatedIceThickness on the snapshots a_k * H_shape, stored as one field."""
    def __init__(self, t_vec, a_vec, H_shape, flat_bc=False):
        self.t_vec = np.asarray(t_vec, dtype=np.float64)
        self.a_vec = np.asarray(a_vec, dtype=np.float64)
        self.H_shape = np.asarray(H_shape, dtype=np.float64)
        self.flat_bc = flat_bc

    def snapshots(self):
        return TimeInterpolatedIceThickness(self.t_vec, [a * self.H_shape for a in self.a_vec],
                                            flat_bc=self.flat_bc)

    def at(self, t):
        return self.snapshots().at(t)


class ExternallyUpdatedIceThickness:
    """Ice thickness written by an external model between steps (src/boundary_conditions.jl:48)."""
    pass


class BoundaryConditions:
    """src/boundary_conditions.jl:297-325."""
    def __init__(self, domain, ice_thickness=None, viscous_displacement=None, elastic_displacement=None,
                 sea_surface_perturbation=None):
        self.ice_thickness = ice_thickness or ExternallyUpdatedIceThickness()
        self.viscous_displacement = viscous_displacement or CornerBC(RegularBCSpace(), 0.0)
        self.elastic_displacement = elastic_displacement or CornerBC(ExtendedBCSpace(), 0.0)
        self.sea_surface_perturbation = sea_surface_perturbation or CornerBC(ExtendedBCSpace(), 0.0)
        if not isinstance(self.viscous_displacement.space, RegularBCSpace):
            raise ValueError("viscous_displacement must be defined on a RegularBCSpace")
        for bc in (self.viscous_displacement, self.elastic_displacement, self.sea_surface_perturbation):
            if not isinstance(bc, CornerBC):
                raise NotImplementedError("only CornerBC (the reference default) is supported")
        if not isinstance(self.elastic_displacement.space, ExtendedBCSpace) or \
                not isinstance(self.sea_surface_perturbation.space, ExtendedBCSpace):
            raise NotImplementedError("elastic / sea-surface CornerBC must be on the ExtendedBCSpace")
