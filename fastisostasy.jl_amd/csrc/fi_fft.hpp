// In-LDS mixed-radix complex FFT for one workgroup (gfx950, wave64).
//
// A workgroup transforms nb independent sequences of length L that sit in LDS, with Stockham
// autosort passes (no bit reversal) ping-ponging between two LDS buffers, one barrier per pass.
// Radices 2, 3, 4, 5, 7, 8 cover every length whose prime factors are 2, 3, 5, 7 (the sizes
// nextprod((2,3,5,7), .) produces and every grid the reference benchmarks). Twiddles come from a
// per-length table w[k] = exp(-2 pi i k / L) rounded once from FP64 on the host. The same code is
// instantiated for float2 (run time) and double2 (set-up of the FP64 Green spectra).
#pragma once
#include <hip/hip_runtime.h>

#ifndef FI_NS
#define FI_NS fi
#endif

namespace FI_NS {

struct FftPlan {
    int L;                   // transform length
    int nr;                  // number of passes
    unsigned long long code; // radix of pass s in bits [4s, 4s+4)
    int tw;                  // element offset of this length's twiddle table
    int pad;
};

template <typename C> struct Real;
template <> struct Real<float2> { typedef float T; };
template <> struct Real<double2> { typedef double T; };

template <typename C> __device__ __forceinline__ C cadd(C a, C b) { C r; r.x = a.x + b.x; r.y = a.y + b.y; return r; }
template <typename C> __device__ __forceinline__ C csub(C a, C b) { C r; r.x = a.x - b.x; r.y = a.y - b.y; return r; }
template <typename C> __device__ __forceinline__ C cmul(C a, C b) {
    C r;
    r.x = a.x * b.x - a.y * b.y;
    r.y = a.x * b.y + a.y * b.x;
    return r;
}
template <typename C, typename T> __device__ __forceinline__ C cscale(C a, T s) { C r; r.x = a.x * s; r.y = a.y * s; return r; }
// multiply by (sign * i)
template <typename C> __device__ __forceinline__ C cmuli(C a, int sign) {
    C r;
    if (sign < 0) { r.x = a.y; r.y = -a.x; }
    else { r.x = -a.y; r.y = a.x; }
    return r;
}

// cos / sin of 2 pi m / R for the odd radices
template <int R> struct OddRoots;
template <> struct OddRoots<3> {
    static __device__ __forceinline__ double c(int m) { return m == 0 ? 1.0 : -0.5; }
    static __device__ __forceinline__ double s(int m) {
        return m == 0 ? 0.0 : (m == 1 ? 0.86602540378443864676 : -0.86602540378443864676);
    }
};
template <> struct OddRoots<5> {
    static __device__ __forceinline__ double c(int m) {
        return m == 0 ? 1.0 : (m == 1 || m == 4) ? 0.30901699437494742410 : -0.80901699437494742410;
    }
    static __device__ __forceinline__ double s(int m) {
        return m == 0 ? 0.0 : m == 1 ? 0.95105651629515357212 : m == 2 ? 0.58778525229247312917
             : m == 3 ? -0.58778525229247312917 : -0.95105651629515357212;
    }
};
template <> struct OddRoots<7> {
    static __device__ __forceinline__ double c(int m) {
        return m == 0 ? 1.0 : (m == 1 || m == 6) ? 0.62348980185873353053
             : (m == 2 || m == 5) ? -0.22252093395631440429 : -0.90096886790241912624;
    }
    static __device__ __forceinline__ double s(int m) {
        return m == 0 ? 0.0 : m == 1 ? 0.78183148246802980871 : m == 2 ? 0.97492791218182360702
             : m == 3 ? 0.43388373911755812048 : m == 4 ? -0.43388373911755812048
             : m == 5 ? -0.97492791218182360702 : -0.78183148246802980871;
    }
};

// length-R DFT in registers, out[p] = sum_r v[r] exp(sign 2 pi i r p / R)
template <typename C, int R> struct Dft;

template <typename C> struct Dft<C, 2> {
    static __device__ __forceinline__ void run(C* v, int) {
        C a = v[0], b = v[1];
        v[0] = cadd(a, b);
        v[1] = csub(a, b);
    }
};
template <typename C> struct Dft<C, 4> {
    static __device__ __forceinline__ void run(C* v, int sign) {
        C a0 = cadd(v[0], v[2]), a1 = csub(v[0], v[2]);
        C b0 = cadd(v[1], v[3]), b1 = cmuli(csub(v[1], v[3]), sign);
        v[0] = cadd(a0, b0);
        v[2] = csub(a0, b0);
        v[1] = cadd(a1, b1);
        v[3] = csub(a1, b1);
    }
};
template <typename C> struct Dft<C, 8> {
    static __device__ __forceinline__ void run(C* v, int sign) {
        typedef typename Real<C>::T T;
        C e[4] = {v[0], v[2], v[4], v[6]};
        C o[4] = {v[1], v[3], v[5], v[7]};
        Dft<C, 4>::run(e, sign);
        Dft<C, 4>::run(o, sign);
        const T h = (T)0.70710678118654752440;
        // w^1 = (1 + sign i) h, w^2 = sign i, w^3 = (-1 + sign i) h
        C w1o, w2o, w3o;
        {
            C t = o[1];
            C it = cmuli(t, sign);
            w1o.x = (t.x + it.x) * h;
            w1o.y = (t.y + it.y) * h;
        }
        w2o = cmuli(o[2], sign);
        {
            C t = o[3];
            C it = cmuli(t, sign);
            w3o.x = (it.x - t.x) * h;
            w3o.y = (it.y - t.y) * h;
        }
        v[0] = cadd(e[0], o[0]);
        v[4] = csub(e[0], o[0]);
        v[1] = cadd(e[1], w1o);
        v[5] = csub(e[1], w1o);
        v[2] = cadd(e[2], w2o);
        v[6] = csub(e[2], w2o);
        v[3] = cadd(e[3], w3o);
        v[7] = csub(e[3], w3o);
    }
};
template <typename C, int R> struct DftOdd {
    static __device__ __forceinline__ void run(C* v, int sign) {
        typedef typename Real<C>::T T;
        constexpr int H = (R - 1) / 2;
        C a[H + 1], b[H + 1];
        C v0 = v[0];
        C s0 = v0;
#pragma unroll
        for (int q = 1; q <= H; ++q) {
            a[q] = cadd(v[q], v[R - q]);
            b[q] = csub(v[q], v[R - q]);
            s0 = cadd(s0, a[q]);
        }
#pragma unroll
        for (int p = 1; p <= H; ++p) {
            C A = v0, B;
            B.x = 0;
            B.y = 0;
#pragma unroll
            for (int q = 1; q <= H; ++q) {
                const int m = (p * q) % R;
                const T cm = (T)OddRoots<R>::c(m), sm = (T)OddRoots<R>::s(m);
                A.x += a[q].x * cm;
                A.y += a[q].y * cm;
                B.x += b[q].x * sm;
                B.y += b[q].y * sm;
            }
            C iB = cmuli(B, sign);
            v[p] = cadd(A, iB);
            v[R - p] = csub(A, iB);
        }
        v[0] = s0;
    }
};
template <typename C> struct Dft<C, 3> { static __device__ __forceinline__ void run(C* v, int s) { DftOdd<C, 3>::run(v, s); } };
template <typename C> struct Dft<C, 5> { static __device__ __forceinline__ void run(C* v, int s) { DftOdd<C, 5>::run(v, s); } };
template <typename C> struct Dft<C, 7> { static __device__ __forceinline__ void run(C* v, int s) { DftOdd<C, 7>::run(v, s); } };

// One Stockham pass of radix R over span Ns for nb sequences (stride seq apart):
// butterfly j reads in[j + r L/R], twiddles by w^(r (j mod Ns) L/(Ns R)), and writes
// out[(j / Ns) Ns R + (j mod Ns) + r Ns].
template <typename C, int R>
__device__ __forceinline__ void fft_pass(const C* in, C* out, int L, int nb, int seq, int Ns,
                                         const C* __restrict__ tw, int sign) {
    const int m = L / R;
    const int total = nb * m;
    const int tstride = L / (Ns * R);
    // exact quotients of small non-negative integers by float reciprocal: (x + 0.5) / d has a
    // fractional part >= 0.5/d, far above the float rounding error for x, d < 2^20
    const float inv_m = 1.0f / (float)m, inv_ns = 1.0f / (float)Ns;
    for (int b = threadIdx.x; b < total; b += blockDim.x) {
        const int s = (nb == 1) ? 0 : (int)(((float)b + 0.5f) * inv_m);
        const int j = b - s * m;
        const int jq = (int)(((float)j + 0.5f) * inv_ns);
        const int k = j - jq * Ns;
        const C* ip = in + s * seq;
        C v[R];
#pragma unroll
        for (int r = 0; r < R; ++r) v[r] = ip[j + r * m];
        if (k) {
#pragma unroll
            for (int r = 1; r < R; ++r) {
                C w = tw[r * k * tstride];
                if (sign > 0) w.y = -w.y;
                v[r] = cmul(v[r], w);
            }
        }
        Dft<C, R>::run(v, sign);
        C* op = out + s * seq + jq * Ns * R + k;
#pragma unroll
        for (int r = 0; r < R; ++r) op[r * Ns] = v[r];
    }
}

// Transform nb sequences held in buffer a (seq elements apart); b is scratch of the same size.
// tw is this length's twiddle table (in LDS for the run-time transforms). Unnormalised; sign -1
// forward, +1 inverse. Returns the buffer that holds the result. The caller synchronises the
// workgroup before the call; the result is synchronised.
template <typename C>
__device__ __forceinline__ C* fft_lds(C* a, C* b, int nb, int seq, const FftPlan& pl, const C* __restrict__ tw, int sign) {
    const int L = pl.L;
    C* in = a;
    C* out = b;
    int Ns = 1;
    for (int s = 0; s < pl.nr; ++s) {
        const int rad = (int)((pl.code >> (4 * s)) & 15ull);
        switch (rad) {
            case 2: fft_pass<C, 2>(in, out, L, nb, seq, Ns, tw, sign); break;
            case 3: fft_pass<C, 3>(in, out, L, nb, seq, Ns, tw, sign); break;
            case 4: fft_pass<C, 4>(in, out, L, nb, seq, Ns, tw, sign); break;
            case 5: fft_pass<C, 5>(in, out, L, nb, seq, Ns, tw, sign); break;
            case 7: fft_pass<C, 7>(in, out, L, nb, seq, Ns, tw, sign); break;
            default: fft_pass<C, 8>(in, out, L, nb, seq, Ns, tw, sign); break;
        }
        Ns *= rad;
        __syncthreads();
        C* t = in;
        in = out;
        out = t;
    }
    return in;
}

}  // namespace FI_NS
