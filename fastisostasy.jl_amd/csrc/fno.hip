// Fourier-neural-operator surrogate of the LV-ELVA macro step (include/fastiso_surrogate.h).
//
// Every product of the network -- the per-cell channel mixes and the truncated spectral
// transforms, written as partial DFTs -- is one batched strided GEMM on BF16 MFMA with FP32
// accumulation (k_gemm). Forward and backward passes are sequences of those GEMMs plus a few
// elementwise kernels; training (Adam) runs on the device, the host only sequences launches.
//
// Truncated real 2D transform of a channel v(i, j) (i fastest), modes kx < mx and the 2 my
// modes ky in {0..my-1, ny-my..ny-1}:
//   x-DFT   Fx(j, kx)   = sum_i v(i, j) e^{-2 pi i kx i/nx}          (rows of Ex)
//   y-DFT   G(ky, kx)   = sum_j Fx(j, kx) e^{-2 pi i ky j/ny}         (Ar on Re Fx, Ai on Im Fx)
//   mix     H(ky, kx)_o = sum_c G(ky, kx)_c R(kx, ky)_{c o}            (complex)
//   y-iDFT  F'(j, kx)   = sum_ky H(ky, kx) e^{+2 pi i ky j/ny}        (Ar^T, Ai^T)
//   x-iDFT  s(i, j)     = 1/(nx ny) sum_kx w_kx Re(F'(j, kx) e^{+2 pi i kx i/nx}), w = 1 (kx=0), 2
// which is irfft2 of the mode-mixed rfft2 restricted to the kept modes.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/fastiso_surrogate.h"

namespace fno {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int GEMM_THREADS = 256;   // 4 waves; block tile 64 (M) x 32 (N)

// C[b] = alpha A[b] B[b] + beta C[b] with element (m, k) of A[b] at A + b sAb + m sAm + k sAk
// (likewise B, C). With kchunk > 0, batch b is the K-slice [b kchunk, (b+1) kchunk) of one product
// and writes its partial into C + b sCb (split-K; reduced by k_splitk_reduce). With nslice > 0 as
// well, grid batch z is slice (z mod nslice) of product (z / nslice) (A and B batch strides apply
// to the product), its partial at C + z sCb (batched split-K; reduced by k_splitk_reduce_batched).
struct Gemm {
    const float* A;
    const float* B;
    float* C;
    long sAb, sAm, sAk, sBb, sBk, sBn, sCb, sCm, sCn;
    int M, N, K, kchunk;
    float alpha, beta;
    int nslice;
};

// The 8 K values k .. k+7 of one operand row or column: two 16-byte loads when the operand is
// contiguous in K, the row is in range, all 8 values are inside the K range and the address is
// 16-byte aligned; element loads (zero outside the ranges) otherwise.
__device__ __forceinline__ void load8(const float* p, long sk, int k, int kend, bool ok, float (&v)[8]) {
    if (ok && sk == 1 && k + 8 <= kend && ((reinterpret_cast<unsigned long long>(p + k) & 15ull) == 0)) {
        const float4 lo = *reinterpret_cast<const float4*>(p + k);
        const float4 hi = *reinterpret_cast<const float4*>(p + k + 4);
        v[0] = lo.x;
        v[1] = lo.y;
        v[2] = lo.z;
        v[3] = lo.w;
        v[4] = hi.x;
        v[5] = hi.y;
        v[6] = hi.z;
        v[7] = hi.w;
        return;
    }
#pragma unroll
    for (int e = 0; e < 8; ++e) v[e] = (ok && k + e < kend) ? p[(long)(k + e) * sk] : 0.f;
}

// Wave w of the block owns rows m0 + 16 w .. +15 and both 16-column halves of the block's 32
// columns: two 16x16x32 MFMAs per 32-wide K step. Lane l feeds row/column (l mod 16) with the 8
// K values 8 (l / 16) .. +7 (the same K assignment for A and B), and receives rows
// 4 (l / 16) .. +3 of column (l mod 16) of each 16x16 result.
__global__ void __launch_bounds__(GEMM_THREADS) k_gemm(const Gemm g) {
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int b = blockIdx.z;
    const int m0 = blockIdx.x * 64 + wave * 16;
    const int n0 = blockIdx.y * 32;
    const int r = lane & 15, kg = lane >> 4;
    int kbeg = 0, kend = g.K;
    const float* A = g.A + (long)b * g.sAb;
    const float* B = g.B + (long)b * g.sBb;
    if (g.nslice > 0) {
        const int pb = b / g.nslice, sl = b - pb * g.nslice;
        kbeg = sl * g.kchunk;
        kend = min(g.K, kbeg + g.kchunk);
        A = g.A + (long)pb * g.sAb;
        B = g.B + (long)pb * g.sBb;
    } else if (g.kchunk > 0) {
        kbeg = b * g.kchunk;
        kend = min(g.K, kbeg + g.kchunk);
        A = g.A;
        B = g.B;
    }
    float* C = g.C + (long)b * g.sCb;
    f32x4 acc0 = {0.f, 0.f, 0.f, 0.f}, acc1 = {0.f, 0.f, 0.f, 0.f};
    const int am = m0 + r, bn0 = n0 + r, bn1 = n0 + 16 + r;
    const bool aok = am < g.M, b0ok = bn0 < g.N, b1ok = bn1 < g.N;
    const float* Ar = A + (long)am * g.sAm;
    const float* B0 = B + (long)bn0 * g.sBn;
    const float* B1 = B + (long)bn1 * g.sBn;
    for (int k0 = kbeg; k0 < kend; k0 += 32) {
        bf16x8 a, x0, x1;
        float av[8], v0[8], v1[8];
        const int kk = k0 + 8 * kg;
        load8(Ar, g.sAk, kk, kend, aok, av);
        load8(B0, g.sBk, kk, kend, b0ok, v0);
        load8(B1, g.sBk, kk, kend, b1ok, v1);
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            a[e] = (__bf16)av[e];
            x0[e] = (__bf16)v0[e];
            x1[e] = (__bf16)v1[e];
        }
        acc0 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, x0, acc0, 0, 0, 0);
        acc1 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, x1, acc1, 0, 0, 0);
    }
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const int row = m0 + 4 * kg + i;
        if (row >= g.M) continue;
        float* cr = C + (long)row * g.sCm;
        if (b0ok) {
            float* c = cr + (long)bn0 * g.sCn;
            *c = g.alpha * acc0[i] + (g.beta != 0.f ? g.beta * *c : 0.f);
        }
        if (b1ok) {
            float* c = cr + (long)bn1 * g.sCn;
            *c = g.alpha * acc1[i] + (g.beta != 0.f ? g.beta * *c : 0.f);
        }
    }
}

// dst(m, n) = beta dst(m, n) + sum_b part[b][m][n]: one wave per output, lane l sums the slices
// b = l, l + 64, ... and the wave adds the lanes by a butterfly (fixed order: deterministic)
__global__ void __launch_bounds__(256) k_splitk_reduce(const float* __restrict__ part, int nsplit, int M, int N,
                                                       float* dst, long sm, long sn, float beta) {
    const int out = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
    if (out >= M * N) return;
    float s = 0.f;
    for (int b = lane; b < nsplit; b += 64) s += part[(long)b * M * N + out];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
    if (lane == 0) {
        const int m = out / N, n = out - m * N;
        float* d = dst + (long)m * sm + (long)n * sn;
        *d = beta * *d + s;
    }
}

// dst[p](m, n) = beta dst[p](m, n) + sum_s part[p][s][m][n] for products p: one thread per output
// (consecutive threads read consecutive partials), slices summed in order (deterministic)
__global__ void __launch_bounds__(256) k_splitk_reduce_batched(const float* __restrict__ part, int nslice, int M,
                                                               int N, int nprod, float* dst, long sp, long sm,
                                                               long sn, float beta) {
    const long MN = (long)M * N;
    for (long out = (long)blockIdx.x * blockDim.x + threadIdx.x; out < MN * nprod;
         out += (long)gridDim.x * blockDim.x) {
        const long pb = out / MN, mn = out - pb * MN;
        const float* pp = part + pb * nslice * MN + mn;
        float s = 0.f;
        for (int sl = 0; sl < nslice; ++sl) s += pp[(long)sl * MN];
        const int m = (int)(mn / N), n = (int)(mn - (long)m * N);
        float* d = dst + pb * sp + (long)m * sm + (long)n * sn;
        *d = (beta != 0.f) ? beta * *d + s : s;
    }
}

__device__ __forceinline__ float gelu(float z) { return 0.5f * z * (1.f + erff(z * 0.70710678118654752f)); }
__device__ __forceinline__ float gelu_d(float z) {
    return 0.5f * (1.f + erff(z * 0.70710678118654752f)) + z * 0.39894228040143268f * expf(-0.5f * z * z);
}

// z += bias (per channel); out = gelu(z) or z
__global__ void k_bias_act(float* z, const float* __restrict__ bias, long P, int C, int act, float* out) {
    const long n = P * C;
    for (long idx = (long)blockIdx.x * blockDim.x + threadIdx.x; idx < n; idx += (long)gridDim.x * blockDim.x) {
        const float v = z[idx] + bias[idx % C];
        z[idx] = v;
        out[idx] = act ? gelu(v) : v;
    }
}

// d *= gelu'(z)
__global__ void k_gelu_back(float* d, const float* __restrict__ z, long n) {
    for (long idx = (long)blockIdx.x * blockDim.x + threadIdx.x; idx < n; idx += (long)gridDim.x * blockDim.x)
        d[idx] *= gelu_d(z[idx]);
}

// H[kx][r][o] = complex sum_c G[kx][r][c] R[kx][m][c][o]   (r = 2 m + re/im)
__global__ void k_mix_fwd(const float* __restrict__ G, const float* __restrict__ R, float* H, int nmode, int C) {
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= nmode * C) return;
    const int mode = idx / C, o = idx - mode * C;
    const float* g = G + (long)mode * 2 * C;
    const float* w = R + (long)mode * C * C * 2;
    float hr = 0.f, hi = 0.f;
    for (int c = 0; c < C; ++c) {
        const float gr = g[c], gi = g[C + c];
        const float wr = w[(c * C + o) * 2], wi = w[(c * C + o) * 2 + 1];
        hr += gr * wr - gi * wi;
        hi += gr * wi + gi * wr;
    }
    H[(long)mode * 2 * C + o] = hr;
    H[(long)mode * 2 * C + C + o] = hi;
}

// dG[kx][r][c] = sum_o dH R^*-style terms (real form of the complex product's adjoint)
__global__ void k_mix_back_g(const float* __restrict__ dH, const float* __restrict__ R, float* dG, int nmode, int C) {
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= nmode * C) return;
    const int mode = idx / C, c = idx - mode * C;
    const float* d = dH + (long)mode * 2 * C;
    const float* w = R + (long)mode * C * C * 2 + (long)c * C * 2;
    float gr = 0.f, gi = 0.f;
    for (int o = 0; o < C; ++o) {
        const float dr = d[o], di = d[C + o];
        const float wr = w[2 * o], wi = w[2 * o + 1];
        gr += dr * wr + di * wi;
        gi += -dr * wi + di * wr;
    }
    dG[(long)mode * 2 * C + c] = gr;
    dG[(long)mode * 2 * C + C + c] = gi;
}

// dR[kx][m][c][o] += (dHr Gr + dHi Gi, -dHr Gi + dHi Gr)
__global__ void k_mix_back_r(const float* __restrict__ dH, const float* __restrict__ G, float* dR, int nmode, int C) {
    const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (long)nmode * C * C) return;
    const int mode = (int)(idx / (C * C));
    const int rem = (int)(idx - (long)mode * C * C);
    const int c = rem / C, o = rem - c * C;
    const float* d = dH + (long)mode * 2 * C;
    const float* g = G + (long)mode * 2 * C;
    const float dr = d[o], di = d[C + o], gr = g[c], gi = g[C + c];
    float* w = dR + idx * 2;
    w[0] += dr * gr + di * gi;
    w[1] += -dr * gi + di * gr;
}

// dy = 2/P (y - t); loss partial sums per block
__global__ void k_mse(const float* __restrict__ y, const float* __restrict__ t, long P, float* dy, double* part) {
    __shared__ double red[256];
    double s = 0.0;
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < P; i += (long)gridDim.x * blockDim.x) {
        const float d = y[i] - t[i];
        dy[i] = 2.f * d / (float)P;
        s += (double)d * d;
    }
    red[threadIdx.x] = s;
    __syncthreads();
    for (int w = blockDim.x / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = red[0] / (double)P;
}

__global__ void k_sum_parts(const double* __restrict__ part, int n, double* out) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        double s = 0.0;
        for (int i = 0; i < n; ++i) s += part[i];
        *out = s;
    }
}

__global__ void k_fill(float* p, long n, float v) {
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) p[i] = v;
}

// Adam with bias correction on the mean gradient g / nb, decoupled weight decay wd
__global__ void k_adam(float* p, float* m, float* v, const float* __restrict__ g, long n, float lr, float b1, float b2,
                       float eps, float c1, float c2, float inv_nb, float wd) {
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        const float gi = g[i] * inv_nb;
        const float mi = b1 * m[i] + (1.f - b1) * gi;
        const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
        m[i] = mi;
        v[i] = vi;
        p[i] -= lr * ((mi / c1) / (sqrtf(vi / c2) + eps) + wd * p[i]);
    }
}

}  // namespace fno

// ==========================================================================================
// host side
// ==========================================================================================
using namespace fno;

namespace {

template <typename T>
struct DB {
    T* p = nullptr;
    size_t n = 0;
    ~DB() { if (p) (void)hipFree(p); }
    hipError_t alloc(size_t cnt) {
        if (p) { (void)hipFree(p); p = nullptr; }
        n = cnt;
        return hipMalloc((void**)&p, std::max<size_t>(cnt, 1) * sizeof(T));
    }
};

uint64_t splitmix64(uint64_t& s) {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
double unif(uint64_t& s) { return (double)(splitmix64(s) >> 11) * (1.0 / 9007199254740992.0); }

}  // namespace

struct fi_surrogate {
    fi_fno_config cfg;
    std::string err;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    double device_ms = 0;
    // parameter offsets
    long oP, op, oQ1, oq1, oQ2, oq2, np;
    std::vector<long> oW, ob, oR;
    DB<float> par, grad, am, av, ones;
    // grid
    int nx = 0, ny = 0;
    long P = 0;
    DB<float> Ex, Ar, Ai, Bx;
    DB<float> x, y, dy, tgt;                // current sample
    std::vector<DB<float>> v, z;            // v[0..L], z[0..L-1]
    DB<float> Fx, G, H, Fy, dFx, dG, dH, dFy;
    std::vector<DB<float>> Gs;              // G per layer (kept for the backward pass)
    DB<float> q_pre, q, dq, dv, dz, part;
    DB<double> lpart, lsum;
    long part_cap = 0;
    // training set
    int nsamp = 0;
    DB<float> xs, ys;
    long long adam_t = 0;   // Adam steps taken (bias correction)
};

namespace {

int sfail(fi_surrogate* s, int code, const std::string& m) {
    if (s) s->err = m;
    return code;
}

#define SCHECK(s, x)                                                                            \
    do {                                                                                        \
        hipError_t e_ = (x);                                                                    \
        if (e_ != hipSuccess) return sfail((s), -5, std::string(#x " : ") + hipGetErrorString(e_)); \
    } while (0)

constexpr int S_ERR_INVALID = -1, S_ERR_STATE = -6;

// plain product. A batch of few output tiles over a long K (the y transforms per retained kx: 16
// products of one 64 x 32 tile over K = ny; the x transforms per row: ny products over K = nx) is
// split along K so that the launch has about 1024 workgroups (the K loop is load-latency bound),
// then reduced in a fixed order.
int gemm(fi_surrogate* s, const float* A, long sAm, long sAk, const float* B, long sBk, long sBn, float* C, long sCm,
         long sCn, int M, int N, int K, float beta, int batch = 1, long sAb = 0, long sBb = 0, long sCb = 0) {
    const long tiles = (long)((M + 63) / 64) * ((N + 31) / 32);
    if (batch > 1 && tiles * batch < 1024 && K >= 128) {
        const int steps = (K + 31) / 32;
        const int want = (int)std::max<long>(1, 1024 / (tiles * batch));
        const int nslice0 = std::min(steps, want);
        const int kchunk = 32 * ((steps + nslice0 - 1) / nslice0);
        const int nslice = (K + kchunk - 1) / kchunk;
        const long MN = (long)M * N, need = MN * nslice * batch;
        if (need > s->part_cap) {
            SCHECK(s, s->part.alloc(need));
            s->part_cap = need;
        }
        Gemm g{A, B, s->part.p, sAb, sAm, sAk, sBb, sBk, sBn, MN, N, 1, M, N, K, kchunk, 1.f, 0.f, nslice};
        dim3 grid((M + 63) / 64, (N + 31) / 32, batch * nslice);
        hipLaunchKernelGGL(k_gemm, grid, dim3(GEMM_THREADS), 0, s->stream, g);
        SCHECK(s, hipGetLastError());
        const long outs = MN * batch;
        hipLaunchKernelGGL(k_splitk_reduce_batched, dim3((unsigned)std::min<long>((outs + 255) / 256, 4096)), dim3(256),
                           0, s->stream, s->part.p, nslice, M, N, batch, C, sCb, sCm, sCn, beta);
        SCHECK(s, hipGetLastError());
        return 0;
    }
    Gemm g{A, B, C, sAb, sAm, sAk, sBb, sBk, sBn, sCb, sCm, sCn, M, N, K, 0, 1.f, beta, 0};
    dim3 grid((M + 63) / 64, (N + 31) / 32, batch);
    hipLaunchKernelGGL(k_gemm, grid, dim3(GEMM_THREADS), 0, s->stream, g);
    SCHECK(s, hipGetLastError());
    return 0;
}

// product with a long K (parameter gradients): split-K partials, then a fixed-order reduction
int gemm_splitk(fi_surrogate* s, const float* A, long sAm, long sAk, const float* B, long sBk, long sBn, float* C,
                long sCm, long sCn, int M, int N, int K, float beta) {
    const int kchunk = 128;   // ~960 slices over the 122 500 cells of an N=350 grid
    const int nsplit = (K + kchunk - 1) / kchunk;
    const long need = (long)nsplit * M * N;
    if (need > s->part_cap) {
        SCHECK(s, s->part.alloc(need));
        s->part_cap = need;
    }
    Gemm g{A, B, s->part.p, 0, sAm, sAk, 0, sBk, sBn, (long)M * N, N, 1, M, N, K, kchunk, 1.f, 0.f, 0};
    dim3 grid((M + 63) / 64, (N + 31) / 32, nsplit);
    hipLaunchKernelGGL(k_gemm, grid, dim3(GEMM_THREADS), 0, s->stream, g);
    SCHECK(s, hipGetLastError());
    hipLaunchKernelGGL(k_splitk_reduce, dim3((M * N + 3) / 4), dim3(256), 0, s->stream, s->part.p, nsplit, M, N, C,
                       sCm, sCn, beta);
    SCHECK(s, hipGetLastError());
    return 0;
}

int ew_grid(long n) { return (int)std::min<long>((n + 255) / 256, 8192); }

int forward(fi_surrogate* s, const float* x) {
    const fi_fno_config& c = s->cfg;
    const int C = c.width, L = c.layers, X2 = 2 * c.modes_x, R4 = 4 * c.modes_y, mx = c.modes_x, Hh = c.proj_hidden;
    const int nx = s->nx, ny = s->ny;
    const long P = s->P;
    float* prm = s->par.p;
    int st;
    // lift
    if ((st = gemm(s, x, c.in_channels, 1, prm + s->oP, C, 1, s->z[0].p, C, 1, (int)P, C, c.in_channels, 0.f))) return st;
    hipLaunchKernelGGL(k_bias_act, dim3(ew_grid(P * C)), dim3(256), 0, s->stream, s->z[0].p, prm + s->op, P, C, 0,
                       s->v[0].p);
    for (int l = 0; l < L; ++l) {
        const float* vl = s->v[l].p;
        float* zl = s->z[l + 1].p;
        float* G = s->Gs[l].p;
        // x-DFT per row j
        if ((st = gemm(s, s->Ex.p, nx, 1, vl, C, 1, s->Fx.p, C, 1, X2, C, nx, 0.f, ny, 0, (long)nx * C, (long)X2 * C)))
            return st;
        // y-DFT per kx: G = Ar Re(Fx) + Ai Im(Fx)
        if ((st = gemm(s, s->Ar.p, ny, 1, s->Fx.p, (long)X2 * C, 1, G, C, 1, R4, C, ny, 0.f, mx, 0, 2L * C, (long)R4 * C)))
            return st;
        if ((st = gemm(s, s->Ai.p, ny, 1, s->Fx.p + C, (long)X2 * C, 1, G, C, 1, R4, C, ny, 1.f, mx, 0, 2L * C,
                       (long)R4 * C)))
            return st;
        hipLaunchKernelGGL(k_mix_fwd, dim3((mx * 2 * c.modes_y * C + 255) / 256), dim3(256), 0, s->stream, G,
                           prm + s->oR[l], s->H.p, mx * 2 * c.modes_y, C);
        // y-iDFT per kx: Re F' = Ar^T H, Im F' = Ai^T H
        if ((st = gemm(s, s->Ar.p, 1, ny, s->H.p, C, 1, s->Fy.p, (long)X2 * C, 1, ny, C, R4, 0.f, mx, 0, (long)R4 * C,
                       2L * C)))
            return st;
        if ((st = gemm(s, s->Ai.p, 1, ny, s->H.p, C, 1, s->Fy.p + C, (long)X2 * C, 1, ny, C, R4, 0.f, mx, 0,
                       (long)R4 * C, 2L * C)))
            return st;
        // pointwise W, then + x-iDFT per row j, then bias and activation
        if ((st = gemm(s, vl, C, 1, prm + s->oW[l], C, 1, zl, C, 1, (int)P, C, C, 0.f))) return st;
        if ((st = gemm(s, s->Bx.p, X2, 1, s->Fy.p, C, 1, zl, C, 1, nx, C, X2, 1.f, ny, 0, (long)X2 * C, (long)nx * C)))
            return st;
        hipLaunchKernelGGL(k_bias_act, dim3(ew_grid(P * C)), dim3(256), 0, s->stream, zl, prm + s->ob[l], P, C,
                           l < L - 1 ? 1 : 0, s->v[l + 1].p);
    }
    // projection
    if ((st = gemm(s, s->v[L].p, C, 1, prm + s->oQ1, Hh, 1, s->q_pre.p, Hh, 1, (int)P, Hh, C, 0.f))) return st;
    hipLaunchKernelGGL(k_bias_act, dim3(ew_grid(P * Hh)), dim3(256), 0, s->stream, s->q_pre.p, prm + s->oq1, P, Hh, 1,
                       s->q.p);
    if ((st = gemm(s, s->q.p, Hh, 1, prm + s->oQ2, 1, 1, s->y.p, 1, 1, (int)P, 1, Hh, 0.f))) return st;
    hipLaunchKernelGGL(k_bias_act, dim3(ew_grid(P)), dim3(256), 0, s->stream, s->y.p, prm + s->oq2, P, 1, 0, s->y.p);
    SCHECK(s, hipGetLastError());
    return 0;
}

// gradient of the loss with respect to every parameter, accumulated into s->grad (after
// forward(x); s->dy holds dL/dy)
int backward(fi_surrogate* s, const float* x) {
    const fi_fno_config& c = s->cfg;
    const int C = c.width, L = c.layers, X2 = 2 * c.modes_x, R4 = 4 * c.modes_y, mx = c.modes_x, Hh = c.proj_hidden;
    const int nx = s->nx, ny = s->ny;
    const long P = s->P;
    float* prm = s->par.p;
    float* gr = s->grad.p;
    const float* one = s->ones.p;
    int st;
    // y = q Q2 + q2
    if ((st = gemm_splitk(s, s->q.p, 1, Hh, s->dy.p, 1, 1, gr + s->oQ2, 1, 1, Hh, 1, (int)P, 1.f))) return st;
    if ((st = gemm_splitk(s, one, 0, 1, s->dy.p, 1, 1, gr + s->oq2, 1, 1, 1, 1, (int)P, 1.f))) return st;
    if ((st = gemm(s, s->dy.p, 1, 1, prm + s->oQ2, 1, 1, s->dq.p, Hh, 1, (int)P, Hh, 1, 0.f))) return st;
    hipLaunchKernelGGL(k_gelu_back, dim3(ew_grid(P * Hh)), dim3(256), 0, s->stream, s->dq.p, s->q_pre.p, P * Hh);
    // q_pre = v_L Q1 + q1
    if ((st = gemm_splitk(s, s->v[L].p, 1, C, s->dq.p, Hh, 1, gr + s->oQ1, Hh, 1, C, Hh, (int)P, 1.f))) return st;
    if ((st = gemm_splitk(s, one, 0, 1, s->dq.p, Hh, 1, gr + s->oq1, 1, 1, 1, Hh, (int)P, 1.f))) return st;
    if ((st = gemm(s, s->dq.p, Hh, 1, prm + s->oQ1, 1, Hh, s->dv.p, C, 1, (int)P, C, Hh, 0.f))) return st;
    for (int l = L - 1; l >= 0; --l) {
        // dz = dv (last layer) or dv * gelu'(z)
        SCHECK(s, hipMemcpyAsync(s->dz.p, s->dv.p, P * C * sizeof(float), hipMemcpyDeviceToDevice, s->stream));
        if (l < L - 1)
            hipLaunchKernelGGL(k_gelu_back, dim3(ew_grid(P * C)), dim3(256), 0, s->stream, s->dz.p, s->z[l + 1].p, P * C);
        const float* vl = s->v[l].p;
        if ((st = gemm_splitk(s, one, 0, 1, s->dz.p, C, 1, gr + s->ob[l], 1, 1, 1, C, (int)P, 1.f))) return st;
        if ((st = gemm_splitk(s, vl, 1, C, s->dz.p, C, 1, gr + s->oW[l], C, 1, C, C, (int)P, 1.f))) return st;
        // dv_l = dz W^T
        if ((st = gemm(s, s->dz.p, C, 1, prm + s->oW[l], 1, C, s->dv.p, C, 1, (int)P, C, C, 0.f))) return st;
        // spectral path: dF' = Bx^T dz per row j
        if ((st = gemm(s, s->Bx.p, 1, X2, s->dz.p, C, 1, s->dFy.p, C, 1, X2, C, nx, 0.f, ny, 0, (long)nx * C,
                       (long)X2 * C)))
            return st;
        // dH = Ar dRe F' + Ai dIm F' per kx
        if ((st = gemm(s, s->Ar.p, ny, 1, s->dFy.p, (long)X2 * C, 1, s->dH.p, C, 1, R4, C, ny, 0.f, mx, 0, 2L * C,
                       (long)R4 * C)))
            return st;
        if ((st = gemm(s, s->Ai.p, ny, 1, s->dFy.p + C, (long)X2 * C, 1, s->dH.p, C, 1, R4, C, ny, 1.f, mx, 0, 2L * C,
                       (long)R4 * C)))
            return st;
        const int nmode = mx * 2 * c.modes_y;
        hipLaunchKernelGGL(k_mix_back_r, dim3((int)(((long)nmode * C * C + 255) / 256)), dim3(256), 0, s->stream,
                           s->dH.p, s->Gs[l].p, gr + s->oR[l], nmode, C);
        hipLaunchKernelGGL(k_mix_back_g, dim3((nmode * C + 255) / 256), dim3(256), 0, s->stream, s->dH.p,
                           prm + s->oR[l], s->dG.p, nmode, C);
        // dRe Fx = Ar^T dG, dIm Fx = Ai^T dG per kx
        if ((st = gemm(s, s->Ar.p, 1, ny, s->dG.p, C, 1, s->dFx.p, (long)X2 * C, 1, ny, C, R4, 0.f, mx, 0,
                       (long)R4 * C, 2L * C)))
            return st;
        if ((st = gemm(s, s->Ai.p, 1, ny, s->dG.p, C, 1, s->dFx.p + C, (long)X2 * C, 1, ny, C, R4, 0.f, mx, 0,
                       (long)R4 * C, 2L * C)))
            return st;
        // dv_l += Ex^T dFx per row j
        if ((st = gemm(s, s->Ex.p, 1, nx, s->dFx.p, C, 1, s->dv.p, C, 1, nx, C, X2, 1.f, ny, 0, (long)X2 * C,
                       (long)nx * C)))
            return st;
    }
    // dv now holds dL/dv_0; v_0 = x P + p
    if ((st = gemm_splitk(s, one, 0, 1, s->dv.p, C, 1, gr + s->op, 1, 1, 1, C, (int)P, 1.f))) return st;
    if ((st = gemm_splitk(s, x, 1, c.in_channels, s->dv.p, C, 1, gr + s->oP, C, 1, c.in_channels, C, (int)P, 1.f)))
        return st;
    SCHECK(s, hipGetLastError());
    return 0;
}

int loss_and_dy(fi_surrogate* s, const float* t, double* dev_out) {
    const int nb = 256;
    hipLaunchKernelGGL(k_mse, dim3(nb), dim3(256), 0, s->stream, s->y.p, t, s->P, s->dy.p, s->lpart.p);
    hipLaunchKernelGGL(k_sum_parts, dim3(1), dim3(64), 0, s->stream, s->lpart.p, nb, dev_out);
    SCHECK(s, hipGetLastError());
    return 0;
}

void timed_begin(fi_surrogate* s) { (void)hipEventRecord(s->ev0, s->stream); }
int timed_end(fi_surrogate* s) {
    SCHECK(s, hipEventRecord(s->ev1, s->stream));
    SCHECK(s, hipStreamSynchronize(s->stream));
    float ms = 0;
    SCHECK(s, hipEventElapsedTime(&ms, s->ev0, s->ev1));
    s->device_ms += ms;
    return 0;
}

}  // namespace

extern "C" {

void fi_fno_default_config(fi_fno_config* c) {
    std::memset(c, 0, sizeof(*c));
    c->in_channels = 4;
    c->width = 32;
    c->modes_x = 16;
    c->modes_y = 16;
    c->layers = 4;
    c->proj_hidden = 64;
    c->device = 0;
    c->seed = 12345;
}

int fi_surrogate_create(const fi_fno_config* cfg, fi_surrogate** out) {
    if (!cfg || !out) return S_ERR_INVALID;
    *out = nullptr;
    if (cfg->in_channels < 1 || cfg->width < 1 || cfg->width > 256 || cfg->modes_x < 1 || cfg->modes_y < 1 ||
        cfg->layers < 1 || cfg->proj_hidden < 1)
        return S_ERR_INVALID;
    fi_surrogate* s = new fi_surrogate();
    s->cfg = *cfg;
    const int C = cfg->width, Ci = cfg->in_channels, Hh = cfg->proj_hidden, L = cfg->layers;
    const long nR = (long)cfg->modes_x * 2 * cfg->modes_y * C * C * 2;
    long o = 0;
    s->oP = o; o += (long)Ci * C;
    s->op = o; o += C;
    for (int l = 0; l < L; ++l) {
        s->oW.push_back(o); o += (long)C * C;
        s->ob.push_back(o); o += C;
        s->oR.push_back(o); o += nR;
    }
    s->oQ1 = o; o += (long)C * Hh;
    s->oq1 = o; o += Hh;
    s->oQ2 = o; o += Hh;
    s->oq2 = o; o += 1;
    s->np = o;
    // initialisation: uniform(-1/sqrt(fan_in), 1/sqrt(fan_in)) for the channel maps and biases,
    // uniform(0, 1)/(C C) for the spectral weights
    std::vector<float> h(s->np, 0.f);
    uint64_t rs = cfg->seed;
    auto fillu = [&](long off, long n, double a) {
        for (long i = 0; i < n; ++i) h[off + i] = (float)((2.0 * unif(rs) - 1.0) * a);
    };
    fillu(s->oP, (long)Ci * C, 1.0 / std::sqrt((double)Ci));
    fillu(s->op, C, 1.0 / std::sqrt((double)Ci));
    for (int l = 0; l < L; ++l) {
        fillu(s->oW[l], (long)C * C, 1.0 / std::sqrt((double)C));
        fillu(s->ob[l], C, 1.0 / std::sqrt((double)C));
        for (long i = 0; i < nR; ++i) h[s->oR[l] + i] = (float)(unif(rs) / ((double)C * C));
    }
    fillu(s->oQ1, (long)C * Hh, 1.0 / std::sqrt((double)C));
    fillu(s->oq1, Hh, 1.0 / std::sqrt((double)C));
    fillu(s->oQ2, Hh, 1.0 / std::sqrt((double)Hh));
    fillu(s->oq2, 1, 1.0 / std::sqrt((double)Hh));
    if (hipSetDevice(cfg->device) != hipSuccess || hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreate(&s->ev0) != hipSuccess || hipEventCreate(&s->ev1) != hipSuccess ||
        s->par.alloc(s->np) != hipSuccess || s->grad.alloc(s->np) != hipSuccess || s->am.alloc(s->np) != hipSuccess ||
        s->av.alloc(s->np) != hipSuccess || s->lpart.alloc(256) != hipSuccess || s->lsum.alloc(1) != hipSuccess ||
        hipMemcpy(s->par.p, h.data(), s->np * sizeof(float), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemset(s->am.p, 0, s->np * sizeof(float)) != hipSuccess ||
        hipMemset(s->av.p, 0, s->np * sizeof(float)) != hipSuccess) {
        fi_surrogate_destroy(s);
        return -5;
    }
    *out = s;
    return 0;
}

void fi_surrogate_destroy(fi_surrogate* s) {
    if (!s) return;
    if (s->stream) (void)hipStreamSynchronize(s->stream);
    if (s->ev0) (void)hipEventDestroy(s->ev0);
    if (s->ev1) (void)hipEventDestroy(s->ev1);
    hipStream_t st = s->stream;
    delete s;
    if (st) (void)hipStreamDestroy(st);
}

const char* fi_surrogate_last_error(const fi_surrogate* s) { return s ? s->err.c_str() : "null surrogate"; }
int64_t fi_surrogate_num_params(const fi_surrogate* s) { return s ? s->np : 0; }

int fi_surrogate_get_params(fi_surrogate* s, float* h, int64_t n) {
    if (!s || !h || n != s->np) return sfail(s, S_ERR_INVALID, "get_params: size mismatch");
    SCHECK(s, hipMemcpy(h, s->par.p, n * sizeof(float), hipMemcpyDeviceToHost));
    return 0;
}

int fi_surrogate_set_params(fi_surrogate* s, const float* h, int64_t n) {
    if (!s || !h || n != s->np) return sfail(s, S_ERR_INVALID, "set_params: size mismatch");
    SCHECK(s, hipMemcpy(s->par.p, h, n * sizeof(float), hipMemcpyHostToDevice));
    SCHECK(s, hipMemset(s->am.p, 0, n * sizeof(float)));
    SCHECK(s, hipMemset(s->av.p, 0, n * sizeof(float)));
    s->adam_t = 0;
    return 0;
}

int fi_surrogate_set_grid(fi_surrogate* s, int32_t nx, int32_t ny) {
    if (!s) return S_ERR_INVALID;
    const fi_fno_config& c = s->cfg;
    if (nx <= 2 * (c.modes_x - 1) || ny < 2 * c.modes_y)
        return sfail(s, S_ERR_INVALID, "grid too small for the retained modes");
    if (nx == s->nx && ny == s->ny) return 0;
    SCHECK(s, hipSetDevice(c.device));
    s->nx = nx;
    s->ny = ny;
    s->P = (long)nx * ny;
    const long P = s->P;
    const int C = c.width, X2 = 2 * c.modes_x, R4 = 4 * c.modes_y, mx = c.modes_x, my = c.modes_y, Hh = c.proj_hidden;
    // DFT matrices (FP64 angles, FP32 storage)
    std::vector<float> Ex((size_t)X2 * nx), Ar((size_t)R4 * ny), Ai((size_t)R4 * ny), Bx((size_t)nx * X2);
    for (int k = 0; k < mx; ++k)
        for (int i = 0; i < nx; ++i) {
            const double th = 2.0 * M_PI * (double)((long)k * i % nx) / nx;
            Ex[(size_t)(2 * k) * nx + i] = (float)std::cos(th);
            Ex[(size_t)(2 * k + 1) * nx + i] = (float)-std::sin(th);
            const double w = (k == 0 ? 1.0 : 2.0) / ((double)nx * ny);
            Bx[(size_t)i * X2 + 2 * k] = (float)(w * std::cos(th));
            Bx[(size_t)i * X2 + 2 * k + 1] = (float)(-w * std::sin(th));
        }
    for (int m = 0; m < 2 * my; ++m) {
        const long ky = m < my ? m : ny - (2 * my - m);
        for (int j = 0; j < ny; ++j) {
            const double th = 2.0 * M_PI * (double)(ky * j % ny) / ny;
            Ar[(size_t)(2 * m) * ny + j] = (float)std::cos(th);
            Ar[(size_t)(2 * m + 1) * ny + j] = (float)-std::sin(th);
            Ai[(size_t)(2 * m) * ny + j] = (float)std::sin(th);
            Ai[(size_t)(2 * m + 1) * ny + j] = (float)std::cos(th);
        }
    }
    auto up = [&](DB<float>& d, const std::vector<float>& h) -> hipError_t {
        hipError_t e = d.alloc(h.size());
        if (e != hipSuccess) return e;
        return hipMemcpy(d.p, h.data(), h.size() * sizeof(float), hipMemcpyHostToDevice);
    };
    SCHECK(s, up(s->Ex, Ex));
    SCHECK(s, up(s->Ar, Ar));
    SCHECK(s, up(s->Ai, Ai));
    SCHECK(s, up(s->Bx, Bx));
    const int L = c.layers;
    s->v.clear();
    s->z.clear();
    s->Gs.clear();
    s->v.resize(L + 1);
    s->z.resize(L + 1);
    s->Gs.resize(L);
    for (int l = 0; l <= L; ++l) {
        SCHECK(s, s->v[l].alloc(P * C));
        SCHECK(s, s->z[l].alloc(P * C));
    }
    for (int l = 0; l < L; ++l) SCHECK(s, s->Gs[l].alloc((size_t)mx * R4 * C));
    SCHECK(s, s->x.alloc(P * c.in_channels));
    SCHECK(s, s->y.alloc(P));
    SCHECK(s, s->dy.alloc(P));
    SCHECK(s, s->tgt.alloc(P));
    SCHECK(s, s->Fx.alloc((size_t)ny * X2 * C));
    SCHECK(s, s->dFx.alloc((size_t)ny * X2 * C));
    SCHECK(s, s->Fy.alloc((size_t)ny * X2 * C));
    SCHECK(s, s->dFy.alloc((size_t)ny * X2 * C));
    SCHECK(s, s->G.alloc((size_t)mx * R4 * C));
    SCHECK(s, s->H.alloc((size_t)mx * R4 * C));
    SCHECK(s, s->dG.alloc((size_t)mx * R4 * C));
    SCHECK(s, s->dH.alloc((size_t)mx * R4 * C));
    SCHECK(s, s->q_pre.alloc(P * Hh));
    SCHECK(s, s->q.alloc(P * Hh));
    SCHECK(s, s->dq.alloc(P * Hh));
    SCHECK(s, s->dv.alloc(P * C));
    SCHECK(s, s->dz.alloc(P * C));
    SCHECK(s, s->ones.alloc(P));
    hipLaunchKernelGGL(k_fill, dim3(ew_grid(P)), dim3(256), 0, s->stream, s->ones.p, P, 1.f);
    SCHECK(s, hipGetLastError());
    SCHECK(s, hipStreamSynchronize(s->stream));
    s->nsamp = 0;   // a training set belongs to one grid
    return 0;
}

int fi_surrogate_forward(fi_surrogate* s, const float* x, float* y) {
    if (!s || !x || !y) return S_ERR_INVALID;
    if (!s->P) return sfail(s, S_ERR_STATE, "fi_surrogate_set_grid first");
    SCHECK(s, hipSetDevice(s->cfg.device));
    SCHECK(s, hipMemcpyAsync(s->x.p, x, s->P * s->cfg.in_channels * sizeof(float), hipMemcpyHostToDevice, s->stream));
    timed_begin(s);
    int st = forward(s, s->x.p);
    if (st) return st;
    if ((st = timed_end(s))) return st;
    SCHECK(s, hipMemcpy(y, s->y.p, s->P * sizeof(float), hipMemcpyDeviceToHost));
    return 0;
}

int fi_surrogate_loss_grad(fi_surrogate* s, const float* x, const float* t, double* loss, float* grad) {
    if (!s || !x || !t || !grad) return S_ERR_INVALID;
    if (!s->P) return sfail(s, S_ERR_STATE, "fi_surrogate_set_grid first");
    SCHECK(s, hipSetDevice(s->cfg.device));
    SCHECK(s, hipMemcpyAsync(s->x.p, x, s->P * s->cfg.in_channels * sizeof(float), hipMemcpyHostToDevice, s->stream));
    SCHECK(s, hipMemcpyAsync(s->tgt.p, t, s->P * sizeof(float), hipMemcpyHostToDevice, s->stream));
    hipLaunchKernelGGL(k_fill, dim3(ew_grid(s->np)), dim3(256), 0, s->stream, s->grad.p, s->np, 0.f);
    timed_begin(s);
    int st;
    if ((st = forward(s, s->x.p)) || (st = loss_and_dy(s, s->tgt.p, s->lsum.p)) || (st = backward(s, s->x.p))) return st;
    if ((st = timed_end(s))) return st;
    SCHECK(s, hipMemcpy(grad, s->grad.p, s->np * sizeof(float), hipMemcpyDeviceToHost));
    if (loss) SCHECK(s, hipMemcpy(loss, s->lsum.p, sizeof(double), hipMemcpyDeviceToHost));
    return 0;
}

int fi_surrogate_set_data(fi_surrogate* s, int32_t n, const float* x, const float* y) {
    if (!s || n < 1 || !x || !y) return S_ERR_INVALID;
    if (!s->P) return sfail(s, S_ERR_STATE, "fi_surrogate_set_grid first");
    SCHECK(s, hipSetDevice(s->cfg.device));
    SCHECK(s, s->xs.alloc((size_t)n * s->P * s->cfg.in_channels));
    SCHECK(s, s->ys.alloc((size_t)n * s->P));
    SCHECK(s, hipMemcpy(s->xs.p, x, (size_t)n * s->P * s->cfg.in_channels * sizeof(float), hipMemcpyHostToDevice));
    SCHECK(s, hipMemcpy(s->ys.p, y, (size_t)n * s->P * sizeof(float), hipMemcpyHostToDevice));
    s->nsamp = n;
    return 0;
}

int fi_surrogate_train(fi_surrogate* s, int32_t steps, int32_t batch, float lr, float lr_min, float wd, uint64_t seed,
                       double* loss_out) {
    if (!s || steps < 0 || batch < 1) return S_ERR_INVALID;
    if (s->nsamp < 1) return sfail(s, S_ERR_STATE, "no training set");
    SCHECK(s, hipSetDevice(s->cfg.device));
    DB<double> losses;
    SCHECK(s, losses.alloc((size_t)std::max(1, steps) * batch));
    uint64_t rs = seed;
    timed_begin(s);
    const long P = s->P;
    const int Ci = s->cfg.in_channels;
    for (int it = 0; it < steps; ++it) {
        hipLaunchKernelGGL(k_fill, dim3(ew_grid(s->np)), dim3(256), 0, s->stream, s->grad.p, s->np, 0.f);
        for (int b = 0; b < batch; ++b) {
            const int k = (int)(splitmix64(rs) % (uint64_t)s->nsamp);
            const float* xk = s->xs.p + (size_t)k * P * Ci;
            const float* yk = s->ys.p + (size_t)k * P;
            int st;
            if ((st = forward(s, xk)) || (st = loss_and_dy(s, yk, losses.p + (size_t)it * batch + b)) ||
                (st = backward(s, xk)))
                return st;
        }
        // Adam step with cosine learning-rate decay over this call's steps
        const double t = (double)(++s->adam_t);
        const double frac = steps > 1 ? (double)it / (double)(steps - 1) : 0.0;
        const float lr_t = (float)(lr_min + 0.5 * (lr - lr_min) * (1.0 + std::cos(M_PI * frac)));
        const float c1 = (float)(1.0 - std::pow(0.9, t)), c2 = (float)(1.0 - std::pow(0.999, t));
        hipLaunchKernelGGL(k_adam, dim3(ew_grid(s->np)), dim3(256), 0, s->stream, s->par.p, s->am.p, s->av.p, s->grad.p,
                           s->np, lr_t, 0.9f, 0.999f, 1e-8f, c1, c2, 1.f / (float)batch, wd);
        SCHECK(s, hipGetLastError());
    }
    int st;
    if ((st = timed_end(s))) return st;
    if (loss_out && steps > 0) {
        std::vector<double> h((size_t)steps * batch);
        SCHECK(s, hipMemcpy(h.data(), losses.p, h.size() * sizeof(double), hipMemcpyDeviceToHost));
        for (int it = 0; it < steps; ++it) {
            double m = 0;
            for (int b = 0; b < batch; ++b) m += h[(size_t)it * batch + b];
            loss_out[it] = m / batch;
        }
    }
    return 0;
}

double fi_surrogate_device_ms(const fi_surrogate* s) { return s ? s->device_ms : 0.0; }

}  // extern "C"
