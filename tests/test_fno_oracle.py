"""Pins of the FP64 surrogate restatement (oracle/fno_ref.py) by mathematics: the truncated
spectral convolution two ways (DFT matrices vs numpy rfft2/irfft2) and the backpropagated
gradient against central finite differences."""
import numpy as np

from oracle import fno_ref as F

SMALL = dict(in_channels=2, width=3, modes_x=3, modes_y=2, layers=2, proj_hidden=4)


def test_spectral_conv_matches_numpy_fft():
    cfg = dict(F.default_config(), width=3, modes_x=5, modes_y=4)
    nx, ny = 40, 36
    rng = np.random.default_rng(0)
    v = rng.standard_normal((nx * ny, 3))
    R = rng.standard_normal((5, 8, 3, 3, 2))
    S, _ = F._spec_fwd(v, R, F.dft_mats(nx, ny, cfg), nx, ny, 3)
    S2 = F.spectral_conv_fft(v, R, nx, ny, cfg)
    np.testing.assert_allclose(S, S2, rtol=0, atol=1e-12 * np.abs(S2).max())


def test_spectral_conv_of_a_kept_mode_is_exact():
    """A single kept Fourier mode with identity weights passes through unchanged; a dropped one
    vanishes."""
    cfg = dict(F.default_config(), width=1, modes_x=4, modes_y=3)
    nx, ny = 24, 20
    i, j = np.meshgrid(np.arange(nx), np.arange(ny), indexing="ij")
    R = np.zeros((4, 6, 1, 1, 2))
    R[..., 0] = 1.0
    M = F.dft_mats(nx, ny, cfg)
    kept = np.cos(2 * np.pi * (2 * i / nx + 1 * j / ny)).ravel(order="F")[:, None]
    S, _ = F._spec_fwd(kept, R, M, nx, ny, 1)
    np.testing.assert_allclose(S, kept, atol=1e-12)
    dropped = np.cos(2 * np.pi * (6 * i / nx)).ravel(order="F")[:, None]
    S, _ = F._spec_fwd(dropped, R, M, nx, ny, 1)
    np.testing.assert_allclose(S, 0 * dropped, atol=1e-12)


def test_gradient_matches_finite_differences():
    nx, ny = 12, 10
    rng = np.random.default_rng(3)
    th = F.init_params(SMALL, 7)
    x = rng.standard_normal((nx * ny, SMALL["in_channels"]))
    t = rng.standard_normal(nx * ny)
    L0, g = F.loss_grad(th, x, t, nx, ny, SMALL)
    assert abs(L0 - F.loss(th, x, t, nx, ny, SMALL)) < 1e-14
    lay, n = F.param_layout(SMALL)
    # every block, a few entries each
    idx = []
    for k, (o, s) in lay.items():
        m = int(np.prod(s))
        idx += list(o + rng.choice(m, size=min(m, 4), replace=False))
    h = 1e-6
    for q in idx:
        e = np.zeros(n)
        e[q] = h
        fd = (F.loss(th + e, x, t, nx, ny, SMALL) - F.loss(th - e, x, t, nx, ny, SMALL)) / (2 * h)
        assert abs(fd - g[q]) <= 1e-6 * max(1.0, abs(fd)), (q, fd, g[q])
