// ORACLE -- test infrastructure only (see fastiso_ref.cpp header).
//
// FP64 complex FFT of any length: iterative Stockham autosort passes over the prime
// factorisation (radix 4 first, then 2, 3, 5, 7, then any remaining prime by a direct DFT).
// It stands in for FFTW's plan_fft!/plan_rfft used by the reference (src/tools.jl:100-109,
// src/convolutions.jl:50-53); its output is checked against numpy.fft by the tests.
#pragma once
#include <cmath>
#include <complex>
#include <vector>

namespace fiorc {

typedef std::complex<double> cd;

struct FFT {
    int n = 0;
    std::vector<int> radix;
    std::vector<cd> w;  // w[k] = exp(-2 pi i k / n)

    FFT() {}
    explicit FFT(int n_) : n(n_) {
        int m = n;
        while (m % 4 == 0) { radix.push_back(4); m /= 4; }
        for (int p : {2, 3, 5, 7}) while (m % p == 0) { radix.push_back(p); m /= p; }
        for (int p = 11; m > 1; p += 2) while (m % p == 0) { radix.push_back(p); m /= p; }
        w.resize(n);
        for (int k = 0; k < n; ++k) {
            double a = -2.0 * M_PI * (double)k / (double)n;
            w[k] = cd(std::cos(a), std::sin(a));
        }
    }

    // Unnormalised transform of data[0..n) in place; sign -1 forward, +1 inverse.
    // work must hold n values.
    void exec(cd* data, cd* work, int sign) const {
        if (n <= 1) return;
        cd* in = data;
        cd* out = work;
        int Ns = 1;
        for (size_t s = 0; s < radix.size(); ++s) {
            pass(in, out, radix[s], Ns, sign);
            Ns *= radix[s];
            std::swap(in, out);
        }
        if (in != data)
            for (int k = 0; k < n; ++k) data[k] = in[k];
    }

  private:
    // One Stockham pass of radix R over span Ns: butterfly j reads in[j + r*n/R], applies the
    // twiddle exp(sign 2 pi i r (j mod Ns) / (Ns R)), does a length-R DFT and writes
    // out[(j / Ns) Ns R + (j mod Ns) + r Ns].
    void pass(const cd* in, cd* out, int R, int Ns, int sign) const {
        const int m = n / R;
        const int tstride = n / (Ns * R);
        cd v[64];
        std::vector<cd> big;
        cd* vv = v;
        if (R > 64) { big.resize(R); vv = big.data(); }
        for (int j = 0; j < m; ++j) {
            const int k = j % Ns;
            for (int r = 0; r < R; ++r) {
                cd x = in[j + r * m];
                if (r && k) {
                    cd tw = w[r * k * tstride];
                    if (sign > 0) tw = std::conj(tw);
                    x *= tw;
                }
                vv[r] = x;
            }
            dft(vv, R, sign);
            const int base = (j / Ns) * Ns * R + k;
            for (int r = 0; r < R; ++r) out[base + r * Ns] = vv[r];
        }
    }

    void dft(cd* v, int R, int sign) const {
        if (R == 2) {
            cd a = v[0], b = v[1];
            v[0] = a + b;
            v[1] = a - b;
            return;
        }
        if (R == 4) {
            cd a0 = v[0] + v[2], a1 = v[0] - v[2];
            cd b0 = v[1] + v[3], b1 = v[1] - v[3];
            // multiply b1 by (sign) i
            cd jb1 = (sign < 0) ? cd(b1.imag(), -b1.real()) : cd(-b1.imag(), b1.real());
            v[0] = a0 + b0;
            v[2] = a0 - b0;
            v[1] = a1 + jb1;
            v[3] = a1 - jb1;
            return;
        }
        // direct DFT with roots exp(sign 2 pi i q / R) taken from the length-n table
        cd t[64];
        std::vector<cd> big;
        cd* tt = t;
        if (R > 64) { big.resize(R); tt = big.data(); }
        const int step = n / R;
        for (int q = 0; q < R; ++q) {
            cd s = 0.0;
            for (int r = 0; r < R; ++r) {
                cd rt = w[((r * q) % R) * step];
                if (sign > 0) rt = std::conj(rt);
                s += v[r] * rt;
            }
            tt[q] = s;
        }
        for (int q = 0; q < R; ++q) v[q] = tt[q];
    }
};

}  // namespace fiorc
