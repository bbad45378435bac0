"""Example utilities (src/utils.jl:250-262)."""
import numpy as np


def mask_disc(X, Y, R, center=(0.0, 0.0)):
    return (np.sqrt((X - center[0]) ** 2 + (Y - center[1]) ** 2) < R).astype(np.float64)


def uniform_ice_cylinder(domain, R, H, center=(0.0, 0.0)):
    """Disc load of radius R and height H (src/utils.jl:258-262)."""
    return mask_disc(domain.X, domain.Y, R, center) * H
