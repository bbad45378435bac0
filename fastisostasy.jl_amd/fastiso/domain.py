"""RegionalDomain and PhysicalConstants: host-side setup, FP64 numpy."""
from dataclasses import dataclass

import numpy as np

SECONDS_PER_YEAR = 60.0 ** 2 * 24 * 365.25  # src/utils.jl:7


@dataclass
class PhysicalConstants:
    """src/constants.jl:16-28."""
    mE: float = 5.972e24
    r_equator: float = 6371e3
    r_pole: float = 6357e3
    A_ocean_pd: float = 3.625e14
    g: float = 9.8
    G: float = 6.674e-11
    seconds_per_year: float = SECONDS_PER_YEAR
    rho_ice: float = 0.910e3
    rho_water: float = 1e3
    rho_seawater: float = 1.023e3

    @property
    def rho_sw_ice(self):
        return self.rho_seawater / self.rho_ice


def meshgrid(x, y):
    """ij-indexed meshgrid (src/coordinates.jl:13-16): X[:, 0] == x, Y[0, :] == y."""
    x = np.asarray(x, dtype=np.float64)
    y = np.asarray(y, dtype=np.float64)
    return np.outer(x, np.ones_like(y)), np.outer(np.ones_like(x), y)


def fftint(N):
    """src/derivatives.jl:216-222."""
    N2 = N // 2
    if N % 2 == 0:
        return np.concatenate([np.arange(0, N2 + 1), np.arange(N2 - 1, 0, -1)]).astype(np.float64)
    return np.concatenate([np.arange(0, N2 + 1), np.arange(N2, 0, -1)]).astype(np.float64)


def get_differential_fourier(Wx, Wy, nx, ny):
    """Pseudodifferential, harmonic and biharmonic symbols (src/derivatives.jl:201-214)."""
    xc = (np.pi / Wx) * fftint(nx)
    yc = (np.pi / Wy) * fftint(ny)
    X, Y = meshgrid(xc, yc)
    harmonic = X ** 2 + Y ** 2
    return np.sqrt(harmonic), harmonic, harmonic ** 2


def samesize_conv_indices(N, M):
    """src/convolutions.jl:196-204 (1-based, as the reference)."""
    j1 = M if N % 2 == 0 else M + 1
    return j1, 2 * N - 1 - M


def nextfastfft(n):
    """nextprod((2,3,5,7), n) (src/convolutions.jl:17-18)."""
    m = max(int(n), 1)
    while True:
        r = m
        for p in (2, 3, 5, 7):
            while r % p == 0:
                r //= p
        if r == 1:
            return m
        m += 1


class RegionalDomain:
    """RegionalDomain(W, n) or RegionalDomain(Wx, Wy, nx, ny) (src/domain.jl:35-169).

    Cell-centred grid x_i = -Wx + i*dx (i = 1..nx), dx = 2Wx/nx.

    Map distortion: the reference computes the scale factor K from the latitudes of a Proj
    stereographic projection (src/domain.jl:123-147). Proj is not available here, so a distorted
    domain takes K directly (K=array of shape (nx, ny)); correct_distortion=True without K is
    refused. K gives Dx = K dx, Dy = K dy and the cell area dx dy K^2 (src/domain.jl:167)."""

    def __init__(self, *args, correct_distortion=False, K=None):
        if len(args) == 2:
            W, n = args
            Wx = Wy = float(W)
            nx = ny = 2 ** int(n)
        elif len(args) == 4:
            Wx, Wy, nx, ny = float(args[0]), float(args[1]), int(args[2]), int(args[3])
        else:
            raise TypeError("RegionalDomain(W, n) or RegionalDomain(Wx, Wy, nx, ny)")
        if correct_distortion and K is None:
            raise NotImplementedError("correct_distortion needs the Proj latitudes of the reference; "
                                      "pass the scale factor as K=...")
        self.Wx, self.Wy, self.nx, self.ny = Wx, Wy, nx, ny
        self.mx, self.my = nx // 2, ny // 2
        self.dx, self.dy = 2 * Wx / nx, 2 * Wy / ny
        self.x = -Wx + self.dx * np.arange(1, nx + 1)
        self.y = -Wy + self.dy * np.arange(1, ny + 1)
        self.X, self.Y = meshgrid(self.x, self.y)
        self.R = np.sqrt(self.X ** 2 + self.Y ** 2)
        self.K = np.ones((nx, ny)) if K is None else np.asarray(K, dtype=np.float64).reshape(nx, ny).copy()
        if not np.all(self.K > 0):
            raise ValueError("the scale factor K must be positive")
        self.Dx = self.K * self.dx
        self.Dy = self.K * self.dy
        self.A = (self.dx * self.dy) * self.K ** 2
        pd, _, _ = get_differential_fourier(Wx, Wy, nx, ny)
        pd[0, 0] = 1e-3 * np.mean([pd[0, 1], pd[1, 0]])  # src/domain.jl:155
        self.pseudodiff = pd
        self.i1, self.i2 = samesize_conv_indices(nx, self.mx)
        self.j1, self.j2 = samesize_conv_indices(ny, self.my)
        self.convo_offset = 0
        self.correct_distortion = bool(np.any(self.K != 1.0))

    def zeros(self):
        return np.zeros((self.nx, self.ny))

    def fill(self, v):
        return np.full((self.nx, self.ny), float(v))
