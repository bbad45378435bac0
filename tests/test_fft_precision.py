"""FP32 twiddle precision of the in-LDS mixed-radix FFT.

The device FFT (csrc/fi_fft.hpp) runs Stockham passes in FP32 with a twiddle table rounded once
from FP64. A compensated (double-FP32) twiddle w = w_hi + w_lo would remove the table's rounding,
at two more complex products per twiddle and a second table in LDS. This emulation of the same
pass structure in FP32 (numpy complex64 arithmetic; radix-R butterflies as FP32 DFTs with
rounded constants) measures what that buys at the grid's lengths: the relative L2 error against
an FP64 FFT of the same FP32 input changes by about 1 %, because the butterflies' own FP32
rounding dominates. The plain table is kept; this test holds that measurement.
"""
import numpy as np


def _split(w):
    hi = w.astype(np.complex64)
    lo = (w - hi.astype(np.complex128)).astype(np.complex64)
    return hi, lo


def _twiddle(v, w, comp):
    hi, lo = _split(w)
    out = v * hi
    if comp:
        out = out + v * lo
    return out


def stockham_fp32(x, radices, comp):
    L = len(x)
    a = x.astype(np.complex64)
    Ns = 1
    for R in radices:
        m = L // R
        j = np.arange(m)
        k, jq = j % Ns, j // Ns
        v = [a[j + r * m] for r in range(R)]
        for r in range(1, R):
            v[r] = _twiddle(v[r], np.exp(-2j * np.pi * r * k / (Ns * R)), comp)
        b = np.zeros(L, np.complex64)
        for p in range(R):
            acc = np.zeros(m, np.complex64)
            for q in range(R):
                acc = acc + _twiddle(v[q], np.full(m, np.exp(-2j * np.pi * p * q / R)), comp)
            b[jq * Ns * R + k + p * Ns] = acc
        a = b
        Ns *= R
    return a


def rel_err(L, radices, comp, seeds=4):
    rng = np.random.default_rng(7)
    out = []
    for _ in range(seeds):
        x = (rng.standard_normal(L) + 1j * rng.standard_normal(L)).astype(np.complex64)
        ref = np.fft.fft(x.astype(np.complex128))
        y = stockham_fp32(x, radices, comp)
        out.append(np.linalg.norm(y - ref) / np.linalg.norm(ref))
    return float(np.mean(out))


def test_emulation_is_an_fft():
    x = np.random.default_rng(1).standard_normal(350) + 0j
    np.testing.assert_allclose(stockham_fp32(x, (2, 5, 5, 7), False), np.fft.fft(x), rtol=0, atol=1e-4)


def test_compensated_twiddles_change_the_fp32_fft_error_by_under_5_percent():
    # the engine's plans (largest radix first, then 3, 5, 7) for the N=350 grid, its padded
    # transform, and a power of two
    for L, radices in ((350, (2, 5, 5, 7)), (700, (4, 5, 5, 7)), (1024, (8, 8, 8, 2))):
        plain, comp = rel_err(L, radices, False), rel_err(L, radices, True)
        assert 5e-8 < plain < 5e-7, (L, plain)
        assert abs(plain - comp) < 0.05 * plain, (L, plain, comp)
