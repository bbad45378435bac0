"""The device surrogate network (include/fastiso_surrogate.h) against the FP64 oracle
(oracle/fno_ref.py, itself pinned by numpy's FFT and finite differences).

Tolerances: every product runs on BF16 MFMA (8 significand bits, unit roundoff 2^-9 = 2e-3) with
FP32 accumulation, through about 30 chained products; forward outputs are compared at 2e-2 relative
L2 and parameter gradients at 5e-2 relative L2 per parameter block.
"""
import numpy as np
import pytest

from oracle import fno_ref as F

pytestmark = pytest.mark.gpu

FWD_TOL = 2e-2
GRAD_TOL = 5e-2


def make(nx=48, ny=40, seed=0, **kw):
    from fastiso.surrogate import Surrogate
    s = Surrogate(seed=seed, **kw)
    s.set_grid(nx, ny)
    return s


def sample(nx, ny, ci, seed):
    rng = np.random.default_rng(seed)
    x = rng.standard_normal((nx * ny, ci))
    t = rng.standard_normal(nx * ny)
    return x, t


def relerr(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def test_forward_matches_oracle():
    nx, ny = 48, 40
    s = make(nx, ny)
    cfg = s.config_dict()
    th = F.init_params(cfg, 5)
    s.set_params(th)
    x, _ = sample(nx, ny, cfg["in_channels"], 1)
    y = s.forward(x)
    yr = F.forward(th.astype(np.float32).astype(np.float64), x.astype(np.float32).astype(np.float64), nx, ny, cfg)
    assert relerr(y, yr) < FWD_TOL, relerr(y, yr)


def test_gradient_matches_oracle():
    nx, ny = 40, 36
    s = make(nx, ny)
    cfg = s.config_dict()
    th = F.init_params(cfg, 9)
    s.set_params(th)
    x, t = sample(nx, ny, cfg["in_channels"], 2)
    th64 = th.astype(np.float32).astype(np.float64)
    x64 = x.astype(np.float32).astype(np.float64)
    t64 = t.astype(np.float32).astype(np.float64)
    L, g = s.loss_grad(x, t)
    Lr, gr = F.loss_grad(th64, x64, t64, nx, ny, cfg)
    assert abs(L - Lr) <= FWD_TOL * Lr
    lay, _ = F.param_layout(cfg)
    for k, (o, shp) in lay.items():
        m = int(np.prod(shp))
        e = relerr(g[o:o + m], gr[o:o + m])
        assert e < GRAD_TOL, (k, e)


def test_training_fits_a_smoothing_operator():
    """Learn y = low-pass(x0) + 0.5 x1 on random fields: Adam on the device lowers the loss by
    more than 10x."""
    nx, ny = 40, 40
    s = make(nx, ny, width=16, modes_x=8, modes_y=8, layers=2, proj_hidden=32)
    rng = np.random.default_rng(4)
    X, Y = [], []
    kx = np.fft.fftfreq(ny)[:, None] ** 2 + np.fft.fftfreq(nx)[None, :] ** 2
    for _ in range(16):
        f = rng.standard_normal((ny, nx, 4))
        low = np.real(np.fft.ifft2(np.fft.fft2(f[..., 0]) * np.exp(-kx * 200)))
        X.append(f.reshape(nx * ny, 4))
        Y.append((low + 0.5 * f[..., 1]).reshape(nx * ny))
    s.set_data(np.array(X), np.array(Y))
    losses = s.train(300, batch=4, lr=3e-3, lr_min=1e-4, seed=3, wd=0.0)
    assert np.all(np.isfinite(losses))
    assert losses[-20:].mean() < 0.1 * losses[:5].mean(), (losses[:5].mean(), losses[-20:].mean())


def test_samples_side_by_side_match_one_at_a_time():
    """fi_surrogate_loss_grad_n runs the samples through the network together (the batch strides of
    every product, the per-sample spectral modes, the gradient sums over samples): its loss and
    gradient equal the sums of the single-sample calls up to FP32 summation order, and the oracle's
    sums within the BF16 bounds. 19 samples: one pass of 16 and one of 3."""
    nx, ny = 40, 36
    s = make(nx, ny)
    cfg = s.config_dict()
    th = F.init_params(cfg, 11)
    s.set_params(th)
    n = 19
    X = np.stack([sample(nx, ny, cfg["in_channels"], 20 + k)[0] for k in range(n)])
    T = np.stack([sample(nx, ny, cfg["in_channels"], 20 + k)[1] for k in range(n)])
    Ln, gn = s.loss_grad_n(X, T)
    L1, g1 = 0.0, np.zeros(s.n_params)
    for k in range(n):
        lk, gk = s.loss_grad(X[k], T[k])
        L1 += lk
        g1 += gk
    assert abs(Ln - L1) <= 1e-9 * abs(L1), (Ln, L1)
    lay, _ = F.param_layout(cfg)
    for key, (o, shp) in lay.items():
        m = int(np.prod(shp))
        e = relerr(gn[o:o + m], g1[o:o + m])
        assert e < 1e-5, (key, e)
    # the oracle's sums over three of the samples, through the batched call on those three
    th64 = th.astype(np.float32).astype(np.float64)
    L3, g3 = s.loss_grad_n(X[:3], T[:3])
    Lr, gr = 0.0, 0.0
    for k in range(3):
        lk, gk = F.loss_grad(th64, X[k].astype(np.float32).astype(np.float64),
                             T[k].astype(np.float32).astype(np.float64), nx, ny, cfg)
        Lr += lk
        gr = gr + gk
    assert abs(L3 - Lr) <= FWD_TOL * Lr
    for key, (o, shp) in lay.items():
        m = int(np.prod(shp))
        e = relerr(g3[o:o + m], gr[o:o + m])
        assert e < GRAD_TOL, (key, e)


def _splitmix64(state):
    """the training loop's sample picker (splitmix64), for replaying its picks"""
    m = (1 << 64) - 1
    state = (state + 0x9E3779B97F4A7C15) & m
    z = state
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & m
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & m
    return state, z ^ (z >> 31)


def test_training_step_loss_over_chunked_batch():
    """fi_surrogate_train gathers each step's samples on the device and runs them side by side in
    passes of at most 16: the reported loss of a 20-sample step (passes of 16 and 4) is the mean of
    the single-sample losses of the samples its seeded picker chose."""
    nx, ny = 40, 36
    s = make(nx, ny)
    cfg = s.config_dict()
    th = F.init_params(cfg, 13)
    s.set_params(th)
    n = 7
    X = np.stack([sample(nx, ny, cfg["in_channels"], 40 + k)[0] for k in range(n)])
    T = np.stack([sample(nx, ny, cfg["in_channels"], 40 + k)[1] for k in range(n)])
    s.set_data(X, T)
    seed, batch = 5, 20
    st, picks = seed, []
    for _ in range(batch):
        st, z = _splitmix64(st)
        picks.append(z % n)
    assert len(set(picks)) > 1
    single = [s.loss_grad(X[k], T[k])[0] for k in range(n)]
    want = float(np.mean([single[k] for k in picks]))
    losses = s.train(1, batch=batch, lr=0.0, lr_min=0.0, seed=seed, wd=0.0)
    assert abs(losses[0] - want) <= 1e-9 * want, (losses[0], want)
