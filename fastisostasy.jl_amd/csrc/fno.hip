// Fourier-neural-operator surrogate of the LV-ELVA macro step (include/fastiso_surrogate.h).
//
// Every product of the network -- the per-cell channel mixes and the truncated spectral
// transforms, written as partial DFTs -- runs on BF16 MFMA with FP32 accumulation: a batched
// strided GEMM (k_gemm), and for each layer's output one fused product (k_gemm2) that adds the
// channel mix and the inverse x-transform in the same accumulators and applies bias and GELU (or,
// backward, GELU's derivative) before it stores. Samples of a batch go through side by side, so a
// launch covers every row of every sample. Forward and backward passes are sequences of these plus
// a few small kernels (spectral mode mixing, loss, Adam); training runs on the device, the host
// only sequences launches.
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

// C[b] = alpha A[b] B[b] + beta C[b] with element (m, k) of A[b] at A + off_A(b) + m sAm + k sAk
// (likewise B, C). A and B hold FP32 or BF16 (abf, bbf); C is FP32. Batch b splits as
// b0 = b mod nb0, b1 = b / nb0 and off_X(b) = b0 sXb + b1 sXb1, so that a batch over samples
// and rows or spectral columns addresses buffers whose two strides do not compose. With
// kchunk > 0, batch b is the K-slice [b kchunk, (b+1) kchunk) of one product and writes its partial
// into C + b sCb (split-K; reduced by k_splitk_reduce). With nslice > 0 as well, grid batch z is
// slice (z mod nslice) of product (z / nslice) (the A and B batch offsets apply to the product),
// its partial at C + z sCb (batched split-K; reduced by k_splitk_reduce_batched).
struct Gemm {
    const void* A;
    const void* B;
    float* C;
    long sAb, sAm, sAk, sBb, sBk, sBn, sCb, sCm, sCn;
    int M, N, K, kchunk;
    float alpha, beta;
    int nslice;
    int nb0;
    long sAb1, sBb1, sCb1;
    int ones_row1;   // 1 + the row of A that reads 1.0 instead of memory (0: none; LDS-staged path)
    int batch;       // products in the launch (paired LDS-staged path)
    // second K segment (LDS-staged path): K values [K1, K) read A2, B2 at k - K1 with A's and B's
    // strides and batch offsets, so one product sums two (A B + A2 B2)
    const void* A2;
    const void* B2;
    int K1;
};

__device__ __forceinline__ long boff(int b, int nb0, long s0, long s1) {
    const int b1 = b / nb0;
    return (long)(b - b1 * nb0) * s0 + (long)b1 * s1;
}

// The 8 K values k .. k+7 of one operand row or column, as loaded (FP32 or BF16): one or two
// 16-byte loads when the operand is contiguous in K, the row is in range, all 8 values are inside
// the K range and the address is 16-byte aligned; element loads (zero outside the ranges)
// otherwise, none when k is past the K range.
template <typename T>
struct Raw;
template <>
struct Raw<float> {
    float v[8];
};
template <>
struct Raw<__bf16> {
    bf16x8 v;
};

__device__ __forceinline__ void load8(const float* p, long sk, int k, int kend, bool ok, Raw<float>& r) {
    if (ok && sk == 1 && k + 8 <= kend && ((reinterpret_cast<unsigned long long>(p + k) & 15ull) == 0)) {
        const float4 lo = *reinterpret_cast<const float4*>(p + k);
        const float4 hi = *reinterpret_cast<const float4*>(p + k + 4);
        r.v[0] = lo.x;
        r.v[1] = lo.y;
        r.v[2] = lo.z;
        r.v[3] = lo.w;
        r.v[4] = hi.x;
        r.v[5] = hi.y;
        r.v[6] = hi.z;
        r.v[7] = hi.w;
        return;
    }
#pragma unroll
    for (int e = 0; e < 8; ++e) r.v[e] = (ok && k + e < kend) ? p[(long)(k + e) * sk] : 0.f;
}

__device__ __forceinline__ void load8(const __bf16* p, long sk, int k, int kend, bool ok, Raw<__bf16>& r) {
    if (ok && sk == 1 && k + 8 <= kend && ((reinterpret_cast<unsigned long long>(p + k) & 15ull) == 0)) {
        r.v = *reinterpret_cast<const bf16x8*>(p + k);
        return;
    }
#pragma unroll
    for (int e = 0; e < 8; ++e) r.v[e] = (ok && k + e < kend) ? p[(long)(k + e) * sk] : (__bf16)0.f;
}

// MFMA operand: FP32 values round to BF16 (nearest even) here, exactly as a BF16 copy written by
// another kernel with the same conversion, so the product does not depend on the storage type
__device__ __forceinline__ bf16x8 operand(const Raw<float>& r) {
    bf16x8 a;
#pragma unroll
    for (int e = 0; e < 8; ++e) a[e] = (__bf16)r.v[e];
    return a;
}
__device__ __forceinline__ bf16x8 operand(const Raw<__bf16>& r) { return r.v; }

// acc0/acc1 += the K range [kbeg, kend) of row Ar of A times columns B0, B1 of B (one wave; lane
// group kg feeds K values 8 kg .. +7 of each 32-wide step). The loads of the next K step are issued
// before the MFMAs of the current one (two register sets), so the loop waits for one load round at
// its start instead of one per step.
template <typename TA, typename TB>
__device__ __forceinline__ void mma_loop(const TA* Ar, long sAk, const TB* B0, const TB* B1, long sBk, int kbeg,
                                         int kend, int kg, bool aok, bool b0ok, bool b1ok, f32x4& acc0, f32x4& acc1) {
    Raw<TA> ra;
    Raw<TB> r0, r1;
    load8(Ar, sAk, kbeg + 8 * kg, kend, aok, ra);
    load8(B0, sBk, kbeg + 8 * kg, kend, b0ok, r0);
    load8(B1, sBk, kbeg + 8 * kg, kend, b1ok, r1);
    for (int k0 = kbeg; k0 < kend; k0 += 32) {
        Raw<TA> na;
        Raw<TB> n0r, n1r;
        const int kn = k0 + 32 + 8 * kg;
        load8(Ar, sAk, kn, kend, aok, na);
        load8(B0, sBk, kn, kend, b0ok, n0r);
        load8(B1, sBk, kn, kend, b1ok, n1r);
        const bf16x8 a = operand(ra);
        acc0 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, operand(r0), acc0, 0, 0, 0);
        acc1 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, operand(r1), acc1, 0, 0, 0);
        ra = na;
        r0 = n0r;
        r1 = n1r;
    }
}

// Wave w of the block owns rows m0 + 16 w .. +15 and both 16-column halves of the block's 32
// columns: two 16x16x32 MFMAs per 32-wide K step. Lane l feeds row/column (l mod 16) with the 8
// K values 8 (l / 16) .. +7 (the same K assignment for A and B), and receives rows
// 4 (l / 16) .. +3 of column (l mod 16) of each 16x16 result.
template <typename TA, typename TB>
__global__ void __launch_bounds__(GEMM_THREADS) k_gemm(const Gemm g) {
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int b = blockIdx.z;
    const int m0 = blockIdx.x * 64 + wave * 16;
    const int n0 = blockIdx.y * 32;
    const int r = lane & 15, kg = lane >> 4;
    int kbeg = 0, kend = g.K;
    const TA* A;
    const TB* B;
    if (g.nslice > 0) {
        const int pb = b / g.nslice, sl = b - pb * g.nslice;
        kbeg = sl * g.kchunk;
        kend = min(g.K, kbeg + g.kchunk);
        A = (const TA*)g.A + boff(pb, g.nb0, g.sAb, g.sAb1);
        B = (const TB*)g.B + boff(pb, g.nb0, g.sBb, g.sBb1);
    } else if (g.kchunk > 0) {
        kbeg = b * g.kchunk;
        kend = min(g.K, kbeg + g.kchunk);
        A = (const TA*)g.A;
        B = (const TB*)g.B;
    } else {
        A = (const TA*)g.A + boff(b, g.nb0, g.sAb, g.sAb1);
        B = (const TB*)g.B + boff(b, g.nb0, g.sBb, g.sBb1);
    }
    float* C = g.C + ((g.nslice > 0 || g.kchunk > 0) ? (long)b * g.sCb : boff(b, g.nb0, g.sCb, g.sCb1));
    f32x4 acc0 = {0.f, 0.f, 0.f, 0.f}, acc1 = {0.f, 0.f, 0.f, 0.f};
    const int am = m0 + r, bn0 = n0 + r, bn1 = n0 + 16 + r;
    const bool aok = am < g.M, b0ok = bn0 < g.N, b1ok = bn1 < g.N;
    mma_loop(A + (long)am * g.sAm, g.sAk, B + (long)bn0 * g.sBn, B + (long)bn1 * g.sBn, g.sBk, kbeg, kend, kg, aok,
             b0ok, b1ok, acc0, acc1);
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

// k_gemm for operands strided in K (the per-row x-transform reads activations across cells; the
// parameter gradients read both operands across cells): the workgroup loads each 32-wide K step of
// its 64 x 32 A tile and 32 x 32 B tile cooperatively, consecutive threads on consecutive addresses
// of whichever index is contiguous, converts to BF16 into LDS ([row][k], two buffers), and the
// waves feed the MFMAs from LDS. The next step's loads are issued before the current step's MFMAs;
// one barrier per step. Same MFMA operands and order as k_gemm, so the same results bit for bit. A
// row of ones in A (ones_row1) makes a bias gradient, the sum over K of B, one more row of its
// weight gradient's product. PAIR (products of at most 32 rows, no split): the tile's upper and
// lower 32 rows are two consecutive products of the batch (waves 0-1 and 2-3), so no wave idles
// on rows past M (the x-transforms per row have 2 modes_x = 32 rows).
template <typename TA, typename TB, bool PAIR = false, bool SEG2 = false>
__global__ void __launch_bounds__(GEMM_THREADS) k_gemm_s(const Gemm g) {
    constexpr int PS = 40;   // LDS pitch (BF16) of a row of 32 k: 80 bytes, 16-byte aligned rows
    constexpr int NBL = PAIR ? 8 : 4;   // B elements per thread and step
    __shared__ __align__(16) __bf16 sa[2][64 * PS];
    __shared__ __align__(16) __bf16 sb[2][(PAIR ? 64 : 32) * PS];
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int b = PAIR ? 2 * (int)blockIdx.z : (int)blockIdx.z;
    const int mb = PAIR ? 0 : blockIdx.x * 64, n0 = blockIdx.y * 32;
    const int r = lane & 15, kg = lane >> 4;
    int kbeg = 0, kend = g.K;
    const TA* A;
    const TB* B;
    const TA* A1 = nullptr;   // PAIR: the second product's operands and output
    const TB* B1 = nullptr;
    float* C1 = nullptr;
    const bool have1 = PAIR && b + 1 < g.batch;
    if constexpr (PAIR) {
        A = (const TA*)g.A + boff(b, g.nb0, g.sAb, g.sAb1);
        B = (const TB*)g.B + boff(b, g.nb0, g.sBb, g.sBb1);
        if (have1) {
            A1 = (const TA*)g.A + boff(b + 1, g.nb0, g.sAb, g.sAb1);
            B1 = (const TB*)g.B + boff(b + 1, g.nb0, g.sBb, g.sBb1);
            C1 = g.C + boff(b + 1, g.nb0, g.sCb, g.sCb1);
        }
    } else {
        if (g.nslice > 0) {
            const int pb = b / g.nslice, sl = b - pb * g.nslice;
            kbeg = sl * g.kchunk;
            kend = min(g.K, kbeg + g.kchunk);
            A = (const TA*)g.A + boff(pb, g.nb0, g.sAb, g.sAb1);
            B = (const TB*)g.B + boff(pb, g.nb0, g.sBb, g.sBb1);
        } else if (g.kchunk > 0) {
            kbeg = b * g.kchunk;
            kend = min(g.K, kbeg + g.kchunk);
            A = (const TA*)g.A;
            B = (const TB*)g.B;
        } else {
            A = (const TA*)g.A + boff(b, g.nb0, g.sAb, g.sAb1);
            B = (const TB*)g.B + boff(b, g.nb0, g.sBb, g.sBb1);
        }
    }
    float* C = g.C + ((!PAIR && (g.nslice > 0 || g.kchunk > 0)) ? (long)b * g.sCb : boff(b, g.nb0, g.sCb, g.sCb1));
    const TA* As2 = nullptr;   // SEG2: the second segment's operands
    const TB* Bs2 = nullptr;
    if constexpr (SEG2) {
        As2 = (const TA*)g.A2 + (A - (const TA*)g.A);
        Bs2 = (const TB*)g.B2 + (B - (const TB*)g.B);
    }
    // this thread's elements of a K step: A e = tid + 256 j (j < 8), B e = tid + 256 j (j < NBL;
    // with PAIR, e >= 1024 is the second product's)
    const bool afk = g.sAk == 1, bfk = g.sBk == 1;
    int arow[8], ak[8], bcol[NBL], bk[NBL];
#pragma unroll
    for (int j = 0; j < 8; ++j) {
        const int e = tid + GEMM_THREADS * j;
        arow[j] = afk ? e >> 5 : e & 63;
        ak[j] = afk ? e & 31 : e >> 6;
    }
#pragma unroll
    for (int j = 0; j < NBL; ++j) {
        const int e = tid + GEMM_THREADS * j;
        if constexpr (PAIR) {
            const int ee = e & 1023, h = e >> 10;
            bcol[j] = (bfk ? ee >> 5 : ee & 31) + 32 * h;   // LDS row: column of product h
            bk[j] = bfk ? ee & 31 : ee >> 5;
        } else {
            bcol[j] = bfk ? e >> 5 : e & 31;
            bk[j] = bfk ? e & 31 : e >> 5;
        }
    }
    TA va[8];
    TB vb[NBL];
    auto gload = [&](int k0) {
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            if constexpr (PAIR) {
                const int h = arow[j] >> 5, m = arow[j] & 31, k = k0 + ak[j];
                const TA* Ah = h ? A1 : A;
                va[j] = (m < g.M && k < kend && (h == 0 || have1)) ? Ah[(long)m * g.sAm + (long)k * g.sAk] : (TA)0.f;
            } else if constexpr (SEG2) {
                const int m = mb + arow[j], k = k0 + ak[j];
                const bool s2 = k >= g.K1;
                const TA* Ap = s2 ? As2 : A;
                const int kk = s2 ? k - g.K1 : k;
                va[j] = (m < g.M && k < kend) ? Ap[(long)m * g.sAm + (long)kk * g.sAk] : (TA)0.f;
            } else {
                const int m = mb + arow[j], k = k0 + ak[j];
                va[j] = (m < g.M && k < kend)
                            ? (m + 1 == g.ones_row1 ? (TA)1.f : A[(long)m * g.sAm + (long)k * g.sAk])
                            : (TA)0.f;
            }
        }
#pragma unroll
        for (int j = 0; j < NBL; ++j) {
            if constexpr (PAIR) {
                const int h = bcol[j] >> 5, n = n0 + (bcol[j] & 31), k = k0 + bk[j];
                const TB* Bh = h ? B1 : B;
                vb[j] = (n < g.N && k < kend && (h == 0 || have1)) ? Bh[(long)k * g.sBk + (long)n * g.sBn] : (TB)0.f;
            } else if constexpr (SEG2) {
                const int n = n0 + bcol[j], k = k0 + bk[j];
                const bool s2 = k >= g.K1;
                const TB* Bp = s2 ? Bs2 : B;
                const int kk = s2 ? k - g.K1 : k;
                vb[j] = (n < g.N && k < kend) ? Bp[(long)kk * g.sBk + (long)n * g.sBn] : (TB)0.f;
            } else {
                const int n = n0 + bcol[j], k = k0 + bk[j];
                vb[j] = (n < g.N && k < kend) ? B[(long)k * g.sBk + (long)n * g.sBn] : (TB)0.f;
            }
        }
    };
    auto lstore = [&](int buf) {
#pragma unroll
        for (int j = 0; j < 8; ++j) sa[buf][arow[j] * PS + ak[j]] = (__bf16)va[j];
#pragma unroll
        for (int j = 0; j < NBL; ++j) sb[buf][bcol[j] * PS + bk[j]] = (__bf16)vb[j];
    };
    f32x4 acc0 = {0.f, 0.f, 0.f, 0.f}, acc1 = {0.f, 0.f, 0.f, 0.f};
    gload(kbeg);
    lstore(0);
    __syncthreads();
    int buf = 0;
    for (int k0 = kbeg; k0 < kend; k0 += 32) {
        const bool more = k0 + 32 < kend;
        if (more) gload(k0 + 32);
        const int bh = PAIR ? 32 * (wave >> 1) : 0;   // this wave's product's B rows in LDS
        const bf16x8 a = *reinterpret_cast<const bf16x8*>(&sa[buf][(wave * 16 + r) * PS + 8 * kg]);
        const bf16x8 x0 = *reinterpret_cast<const bf16x8*>(&sb[buf][(bh + r) * PS + 8 * kg]);
        const bf16x8 x1 = *reinterpret_cast<const bf16x8*>(&sb[buf][(bh + 16 + r) * PS + 8 * kg]);
        acc0 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, x0, acc0, 0, 0, 0);
        acc1 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, x1, acc1, 0, 0, 0);
        if (more) lstore(buf ^ 1);
        __syncthreads();
        buf ^= 1;
    }
    if constexpr (PAIR) {
        if ((wave >> 1) == 1) {
            if (!have1) return;
            C = C1;
        }
    }
    const int m0 = PAIR ? (wave & 1) * 16 : mb + wave * 16, bn0 = n0 + r, bn1 = n0 + 16 + r;
    const bool b0ok = bn0 < g.N, b1ok = bn1 < g.N;
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

// dst[p](m, n) = beta dst[p](m, n) + sum_s part[p][s][m][n] for products p, dst[p] at
// dst + (p mod nb0) sp + (p / nb0) sp1: one thread per output (consecutive threads read consecutive
// partials), slices summed in order (deterministic)
__global__ void __launch_bounds__(256) k_splitk_reduce_batched(const float* __restrict__ part, int nslice, int M,
                                                               int N, int nprod, float* dst, long sp, int nb0,
                                                               long sp1, long sm, long sn, float beta) {
    const long MN = (long)M * N;
    for (long out = (long)blockIdx.x * blockDim.x + threadIdx.x; out < MN * nprod;
         out += (long)gridDim.x * blockDim.x) {
        const long pb = out / MN, mn = out - pb * MN;
        const float* pp = part + pb * nslice * MN + mn;
        float s = 0.f;
        for (int sl = 0; sl < nslice; ++sl) s += pp[(long)sl * MN];
        const int m = (int)(mn / N), n = (int)(mn - (long)m * N);
        float* d = dst + boff((int)pb, nb0, sp, sp1) + (long)m * sm + (long)n * sn;
        *d = (beta != 0.f) ? beta * *d + s : s;
    }
}

__device__ __forceinline__ float gelu(float z) { return 0.5f * z * (1.f + erff(z * 0.70710678118654752f)); }
__device__ __forceinline__ float gelu_d(float z) {
    return 0.5f * (1.f + erff(z * 0.70710678118654752f)) + z * 0.39894228040143268f * expf(-0.5f * z * z);
}

// A layer's product and its elementwise step in one kernel. For product b (a row of a sample, or
// all cells at once):
//   acc = A[b] B[b] (+ A2[b] B2[b]: a second K segment accumulated in the same MFMA registers)
// then, at out + b sOb + m N + n,
//   EPI_BIAS_ACT: z = acc + bias[n]; zout = z (if zout); vout = BF16(act ? gelu(z) : z)
//   EPI_ACT_BACK: vout = BF16(zin ? acc gelu'(zin) : acc)
// A is FP32 or BF16 (TA); B, A2, B2 are FP32. Tile mapping as in k_gemm.
enum { EPI_BIAS_ACT = 0, EPI_ACT_BACK = 1 };
struct Gemm2 {
    const void* A;
    const float* B;
    long sAb, sAm, sAk, sBb, sBk, sBn;
    int K;
    const float* A2;
    const float* B2;
    long sA2b, sA2m, sA2k, sB2b, sB2k, sB2n;
    int K2;
    int M, N;
    long sOb;
    int epi, act;
    const float* bias;
    float* zout;
    __bf16* vout;
    const float* zin;
};

// B tile of one segment into LDS as BF16, [column][k] with k contiguous (pitch kr + 8, kr = K
// rounded up to 32, zeros past K), loaded by the whole workgroup; consecutive threads take
// consecutive addresses of whichever index is contiguous in memory.
__device__ __forceinline__ void stage_b(__bf16* sb, const float* B, long sBk, long sBn, int K, int kr, int n0, int N) {
    const int pitch = kr + 8;
    const bool nfast = sBn == 1;
    for (int e = threadIdx.x; e < 32 * kr; e += GEMM_THREADS) {
        const int n = nfast ? (e & 31) : e / kr;
        const int k = nfast ? (e >> 5) : e - (e / kr) * kr;
        const float v = (k < K && n0 + n < N) ? B[(long)k * sBk + (long)(n0 + n) * sBn] : 0.f;
        sb[n * pitch + k] = (__bf16)v;
    }
}

__device__ __forceinline__ bf16x8 lds8(const __bf16* sb, int pitch, int n, int k) {
    return *reinterpret_cast<const bf16x8*>(sb + n * pitch + k);
}

// Each workgroup stages its 32 columns of B (and of B2) in LDS once, so the four waves share them;
// the A rows of both segments and the epilogue's z values are loaded before the first MFMA, so a
// workgroup of the layer products (K = width, K2 = 2 modes_x: one step each) waits for one load
// round. Dynamic LDS: 32 (kr + 8) + 32 (kr2 + 8) BF16.
template <typename TA>
__global__ void __launch_bounds__(GEMM_THREADS) k_gemm2(const Gemm2 g) {
    extern __shared__ __align__(16) __bf16 sbuf[];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const long b = blockIdx.z;
    const int m0 = blockIdx.x * 64 + wave * 16;
    const int n0 = blockIdx.y * 32;
    const int r = lane & 15, kg = lane >> 4;
    const int kr = (g.K + 31) & ~31, kr2 = (g.K2 + 31) & ~31;
    __bf16* sb1 = sbuf;
    __bf16* sb2 = sbuf + 32 * (kr + 8);
    f32x4 acc0 = {0.f, 0.f, 0.f, 0.f}, acc1 = {0.f, 0.f, 0.f, 0.f};
    const int am = m0 + r;
    const bool aok = am < g.M, b0ok = n0 + r < g.N, b1ok = n0 + 16 + r < g.N;
    const TA* Ar = (const TA*)g.A + b * g.sAb + (long)am * g.sAm;
    const bool two = g.K2 > 0;
    const float* A2r = two ? g.A2 + b * g.sA2b + (long)am * g.sA2m : nullptr;
    Raw<TA> ra;
    Raw<float> qa;
    load8(Ar, g.sAk, 8 * kg, g.K, aok, ra);
    if (two) load8(A2r, g.sA2k, 8 * kg, g.K2, aok, qa);
    float zv[4][2];
    const bool zl = g.epi == EPI_ACT_BACK && g.zin;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const int row = m0 + 4 * kg + i;
        const long o = b * g.sOb + (long)row * g.N;
        zv[i][0] = (zl && row < g.M && b0ok) ? g.zin[o + n0 + r] : 0.f;
        zv[i][1] = (zl && row < g.M && b1ok) ? g.zin[o + n0 + 16 + r] : 0.f;
    }
    stage_b(sb1, g.B + b * g.sBb, g.sBk, g.sBn, g.K, kr, n0, g.N);
    if (two) stage_b(sb2, g.B2 + b * g.sB2b, g.sB2k, g.sB2n, g.K2, kr2, n0, g.N);
    __syncthreads();
    for (int k0 = 0; k0 < g.K; k0 += 32) {
        if (k0 > 0) load8(Ar, g.sAk, k0 + 8 * kg, g.K, aok, ra);
        const bf16x8 a = operand(ra);
        acc0 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, lds8(sb1, kr + 8, r, k0 + 8 * kg), acc0, 0, 0, 0);
        acc1 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, lds8(sb1, kr + 8, 16 + r, k0 + 8 * kg), acc1, 0, 0, 0);
    }
    if (two) {
        for (int k0 = 0; k0 < g.K2; k0 += 32) {
            if (k0 > 0) load8(A2r, g.sA2k, k0 + 8 * kg, g.K2, aok, qa);
            const bf16x8 a = operand(qa);
            acc0 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, lds8(sb2, kr2 + 8, r, k0 + 8 * kg), acc0, 0, 0, 0);
            acc1 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, lds8(sb2, kr2 + 8, 16 + r, k0 + 8 * kg), acc1, 0, 0, 0);
        }
    }
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const int row = m0 + 4 * kg + i;
        if (row >= g.M) continue;
        const long o = b * g.sOb + (long)row * g.N;
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const int n = n0 + 16 * h + r;
            if (!(h ? b1ok : b0ok)) continue;
            const float a = h ? acc1[i] : acc0[i];
            if (g.epi == EPI_BIAS_ACT) {
                const float z = a + g.bias[n];
                if (g.zout) g.zout[o + n] = z;
                g.vout[o + n] = (__bf16)(g.act ? gelu(z) : z);
            } else {
                g.vout[o + n] = (__bf16)(zl ? a * gelu_d(zv[i][h]) : a);
            }
        }
    }
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

// H[kx][r][o] = complex sum_c G[kx][r][c] R[kx][m][c][o]   (r = 2 m + re/im), for the nmode
// modes of each of nb samples (G, H: [nb][nmode][2][C])
__global__ void k_mix_fwd(const float* __restrict__ G, const float* __restrict__ R, float* H, int nmode, int C,
                          int nb) {
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= nb * nmode * C) return;
    const int mode = idx / C, o = idx - mode * C;
    const float* g = G + (long)mode * 2 * C;
    const float* w = R + (long)(mode % nmode) * C * C * 2;
    float hr = 0.f, hi = 0.f;
#pragma unroll 8
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
__global__ void k_mix_back_g(const float* __restrict__ dH, const float* __restrict__ R, float* dG, int nmode, int C,
                             int nb) {
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= nb * nmode * C) return;
    const int mode = idx / C, c = idx - mode * C;
    const float* d = dH + (long)mode * 2 * C;
    const float* w = R + (long)(mode % nmode) * C * C * 2 + (long)c * C * 2;
    float gr = 0.f, gi = 0.f;
#pragma unroll 8
    for (int o = 0; o < C; ++o) {
        const float dr = d[o], di = d[C + o];
        const float wr = w[2 * o], wi = w[2 * o + 1];
        gr += dr * wr + di * wi;
        gi += -dr * wi + di * wr;
    }
    dG[(long)mode * 2 * C + c] = gr;
    dG[(long)mode * 2 * C + C + c] = gi;
}

// dR[kx][m][c][o] += (dHr Gr + dHi Gi, -dHr Gi + dHi Gr), summed over the nb samples in order
__global__ void k_mix_back_r(const float* __restrict__ dH, const float* __restrict__ G, float* dR, int nmode, int C,
                             int nb) {
    const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (long)nmode * C * C) return;
    const int mode = (int)(idx / (C * C));
    const int rem = (int)(idx - (long)mode * C * C);
    const int c = rem / C, o = rem - c * C;
    float* w = dR + idx * 2;
    float w0 = w[0], w1 = w[1];
    for (int sm = 0; sm < nb; ++sm) {
        const float* d = dH + ((long)sm * nmode + mode) * 2 * C;
        const float* g = G + ((long)sm * nmode + mode) * 2 * C;
        const float dr = d[o], di = d[C + o], gr = g[c], gi = g[C + c];
        w0 += dr * gr + di * gi;
        w1 += -dr * gi + di * gr;
    }
    w[0] = w0;
    w[1] = w1;
}

// dy = 2/P (y - t) over n = nb P cells of nb samples; loss partial sums per block (the sum of the nb
// samples' mean squared errors)
__global__ void k_mse(const float* __restrict__ y, const float* __restrict__ t, long P, long n, float* dy,
                      double* part) {
    __shared__ double red[256];
    double s = 0.0;
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
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
#pragma unroll 16
        for (int i = 0; i < n; ++i) s += part[i];
        *out = s;
    }
}

// the inputs and targets of the training samples idx[0..nb-1], side by side
struct Picks {
    int k[64];
};
__global__ void k_gather(const float* __restrict__ xs, const float* __restrict__ ys, long nxs, long nys,
                         const Picks pk, int nb, float* x, float* y) {
    const long n = (long)nb * (nxs + nys);
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        if (i < nb * nxs) {
            const int sm = (int)(i / nxs);
            x[i] = xs[(long)pk.k[sm] * nxs + (i - sm * nxs)];
        } else {
            const long j = i - nb * nxs;
            const int sm = (int)(j / nys);
            y[j] = ys[(long)pk.k[sm] * nys + (j - sm * nys)];
        }
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
    DB<float> par, grad, am, av;
    // grid
    int nx = 0, ny = 0;
    long P = 0;
    DB<float> Ex, Ar, Ai, Bx, ExT;   // ExT: Ex transposed, [nx][2 modes_x]
    // activations of up to cap samples side by side ([sample][cell][channel]); the layer outputs v,
    // the hidden projection q and the backward products dq, dz feed only MFMA products and are
    // stored as BF16, the pre-activations z and q_pre as FP32
    int cap = 0;
    DB<float> x, y, dy, tgt;
    std::vector<DB<__bf16>> v;              // v[0..L]
    std::vector<DB<float>> z;               // z[0..L]
    DB<float> Fx, H, Fy, dFx, dG, dH, dFy;
    std::vector<DB<float>> Gs;              // G per layer (kept for the backward pass)
    DB<float> q_pre, part;
    DB<__bf16> q, dqb, dzb[2];
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

// an FP32 or BF16 operand
struct Src {
    const void* p;
    bool bf;
    Src(const float* q) : p(q), bf(false) {}
    Src(const __bf16* q) : p(q), bf(true) {}
};

void launch_gemm(fi_surrogate* s, const Gemm& g, dim3 grid, bool abf, bool bbf) {
    if (g.A2) {   // two K segments (FP32 operands: the y transforms of the real and imaginary parts)
        hipLaunchKernelGGL((k_gemm_s<float, float, false, true>), grid, dim3(GEMM_THREADS), 0, s->stream, g);
        return;
    }
    if (g.sAk != 1 || g.sBk != 1 || g.ones_row1 > 0) {   // an operand strided in K: stage the tiles in LDS
        if (abf && bbf)
            hipLaunchKernelGGL((k_gemm_s<__bf16, __bf16>), grid, dim3(GEMM_THREADS), 0, s->stream, g);
        else if (abf)
            hipLaunchKernelGGL((k_gemm_s<__bf16, float>), grid, dim3(GEMM_THREADS), 0, s->stream, g);
        else if (bbf)
            hipLaunchKernelGGL((k_gemm_s<float, __bf16>), grid, dim3(GEMM_THREADS), 0, s->stream, g);
        else
            hipLaunchKernelGGL((k_gemm_s<float, float>), grid, dim3(GEMM_THREADS), 0, s->stream, g);
        return;
    }
    if (abf && bbf)
        hipLaunchKernelGGL((k_gemm<__bf16, __bf16>), grid, dim3(GEMM_THREADS), 0, s->stream, g);
    else if (abf)
        hipLaunchKernelGGL((k_gemm<__bf16, float>), grid, dim3(GEMM_THREADS), 0, s->stream, g);
    else if (bbf)
        hipLaunchKernelGGL((k_gemm<float, __bf16>), grid, dim3(GEMM_THREADS), 0, s->stream, g);
    else
        hipLaunchKernelGGL((k_gemm<float, float>), grid, dim3(GEMM_THREADS), 0, s->stream, g);
}

constexpr int NB0_NONE = 1 << 30;

// plain product. A batch of few output tiles over a long K (the y transforms per retained kx: 16
// products of one 64 x 32 tile over K = ny per sample; the x transforms per row: ny products over
// K = nx per sample) is split along K so that the launch has about 1024 workgroups (the K loop is
// load-latency bound), then reduced in a fixed order. Product b's operands sit at
// X + (b mod nb0) sXb + (b / nb0) sXb1.
//
// With A2, B2 (FP32, the strides and batch offsets of A and B) the product is A B + A2 B2 over a K
// range of 2 K values, in one launch (and one reduction when split).
int gemm(fi_surrogate* s, Src A, long sAm, long sAk, Src B, long sBk, long sBn, float* C, long sCm, long sCn, int M,
         int N, int K, float beta, int batch = 1, long sAb = 0, long sBb = 0, long sCb = 0, int nb0 = NB0_NONE,
         long sAb1 = 0, long sBb1 = 0, long sCb1 = 0, const float* A2 = nullptr, const float* B2 = nullptr) {
    const int K1 = K;
    if (A2) {
        if (A.bf || B.bf || !B2) return sfail(s, S_ERR_INVALID, "two-segment product: FP32 operands only");
        K = 2 * K1;
    }
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
        Gemm g{A.p, B.p, s->part.p, sAb, sAm, sAk, sBb, sBk, sBn, MN, N, 1, M, N, K, kchunk, 1.f, 0.f, nslice,
               nb0, sAb1, sBb1, 0, 0, 0, A2, B2, K1};
        launch_gemm(s, g, dim3((M + 63) / 64, (N + 31) / 32, batch * nslice), A.bf, B.bf);
        SCHECK(s, hipGetLastError());
        const long outs = MN * batch;
        hipLaunchKernelGGL(k_splitk_reduce_batched, dim3((unsigned)std::min<long>((outs + 255) / 256, 4096)), dim3(256),
                           0, s->stream, s->part.p, nslice, M, N, batch, C, sCb, nb0, sCb1, sCm, sCn, beta);
        SCHECK(s, hipGetLastError());
        return 0;
    }
    Gemm g{A.p, B.p, C, sAb, sAm, sAk, sBb, sBk, sBn, sCb, sCm, sCn, M, N, K, 0, 1.f, beta, 0, nb0, sAb1, sBb1, sCb1,
           0, batch, A2, B2, K1};
    if (!A2 && M <= 32 && batch >= 2 && (sAk != 1 || sBk != 1)) {   // two products per workgroup
        const dim3 grid(1, (N + 31) / 32, (batch + 1) / 2);
        if (A.bf && B.bf)
            hipLaunchKernelGGL((k_gemm_s<__bf16, __bf16, true>), grid, dim3(GEMM_THREADS), 0, s->stream, g);
        else if (A.bf)
            hipLaunchKernelGGL((k_gemm_s<__bf16, float, true>), grid, dim3(GEMM_THREADS), 0, s->stream, g);
        else if (B.bf)
            hipLaunchKernelGGL((k_gemm_s<float, __bf16, true>), grid, dim3(GEMM_THREADS), 0, s->stream, g);
        else
            hipLaunchKernelGGL((k_gemm_s<float, float, true>), grid, dim3(GEMM_THREADS), 0, s->stream, g);
        SCHECK(s, hipGetLastError());
        return 0;
    }
    launch_gemm(s, g, dim3((M + 63) / 64, (N + 31) / 32, batch), A.bf, B.bf);
    SCHECK(s, hipGetLastError());
    return 0;
}

// product with a long K (parameter gradients, K = the cells of the samples): split-K partials over
// about 2048 slices of at least 128, then a fixed-order reduction. With bias, A gets one more row
// of ones (row M), whose product, the bias gradient, lands in row M of C: every bias follows its
// weight matrix in the parameter block.
int gemm_splitk(fi_surrogate* s, Src A, long sAm, long sAk, Src B, long sBk, long sBn, float* C, long sCm, long sCn,
                int M, int N, int K, float beta, bool bias = false) {
    const int ones_row1 = bias ? M + 1 : 0;
    if (bias) M += 1;
    const int kchunk = std::max(128, 32 * (int)(((long)K + 2048L * 32 - 1) / (2048L * 32)));
    const int nsplit = (K + kchunk - 1) / kchunk;
    const long need = (long)nsplit * M * N;
    if (need > s->part_cap) {
        SCHECK(s, s->part.alloc(need));
        s->part_cap = need;
    }
    Gemm g{A.p, B.p, s->part.p, 0, sAm, sAk, 0, sBk, sBn, (long)M * N, N, 1, M, N, K, kchunk, 1.f, 0.f, 0,
           NB0_NONE, 0, 0, 0, ones_row1};
    launch_gemm(s, g, dim3((M + 63) / 64, (N + 31) / 32, nsplit), A.bf, B.bf);
    SCHECK(s, hipGetLastError());
    hipLaunchKernelGGL(k_splitk_reduce, dim3((M * N + 3) / 4), dim3(256), 0, s->stream, s->part.p, nsplit, M, N, C,
                       sCm, sCn, beta);
    SCHECK(s, hipGetLastError());
    return 0;
}

// fused layer product (k_gemm2) over batch products
int gemm2(fi_surrogate* s, const Gemm2& g, bool abf, int batch) {
    const dim3 grid((g.M + 63) / 64, (g.N + 31) / 32, batch);
    const int kr = (g.K + 31) & ~31, kr2 = g.K2 > 0 ? (g.K2 + 31) & ~31 : 0;
    const size_t lds = (size_t)32 * ((kr + 8) + (g.K2 > 0 ? kr2 + 8 : 0)) * sizeof(__bf16);
    if (lds > 64 * 1024) return sfail(s, S_ERR_INVALID, "fused product: K too large");
    if (abf)
        hipLaunchKernelGGL(k_gemm2<__bf16>, grid, dim3(GEMM_THREADS), lds, s->stream, g);
    else
        hipLaunchKernelGGL(k_gemm2<float>, grid, dim3(GEMM_THREADS), lds, s->stream, g);
    SCHECK(s, hipGetLastError());
    return 0;
}

int ew_grid(long n) { return (int)std::min<long>((n + 255) / 256, 8192); }

// forward pass of nb samples side by side (x: [nb][cell][in_channels], nb <= cap)
int forward(fi_surrogate* s, const float* x, int nb) {
    const fi_fno_config& c = s->cfg;
    const int C = c.width, L = c.layers, X2 = 2 * c.modes_x, R4 = 4 * c.modes_y, mx = c.modes_x, Hh = c.proj_hidden;
    const int nx = s->nx, ny = s->ny, nmode = mx * 2 * c.modes_y;
    const long PB = s->P * nb;
    const long sFs = (long)ny * X2 * C, sGs = (long)mx * R4 * C;   // per-sample strides of Fx/Fy and G/H
    float* prm = s->par.p;
    int st;
    // lift: v_0 = x P + p
    {
        Gemm2 g{};
        g.A = x;
        g.sAm = c.in_channels;
        g.sAk = 1;
        g.B = prm + s->oP;
        g.sBk = C;
        g.sBn = 1;
        g.K = c.in_channels;
        g.M = (int)PB;
        g.N = C;
        g.epi = EPI_BIAS_ACT;
        g.bias = prm + s->op;
        g.vout = s->v[0].p;
        if ((st = gemm2(s, g, false, 1))) return st;
    }
    for (int l = 0; l < L; ++l) {
        const __bf16* vl = s->v[l].p;
        float* zl = s->z[l + 1].p;
        float* G = s->Gs[l].p;
        // x-DFT per row j of each sample
        if ((st = gemm(s, s->Ex.p, nx, 1, vl, C, 1, s->Fx.p, C, 1, X2, C, nx, 0.f, nb * ny, 0, (long)nx * C,
                       (long)X2 * C)))
            return st;
        // y-DFT per (sample, kx): G = Ar Re(Fx) + Ai Im(Fx)
        if ((st = gemm(s, s->Ar.p, ny, 1, s->Fx.p, (long)X2 * C, 1, G, C, 1, R4, C, ny, 0.f, nb * mx, 0, 2L * C,
                       (long)R4 * C, mx, 0, sFs, sGs, s->Ai.p, s->Fx.p + C)))
            return st;
        hipLaunchKernelGGL(k_mix_fwd, dim3((nb * nmode * C + 255) / 256), dim3(256), 0, s->stream, G, prm + s->oR[l],
                           s->H.p, nmode, C, nb);
        // y-iDFT per (sample, kx): Re F' = Ar^T H, Im F' = Ai^T H
        if ((st = gemm(s, s->Ar.p, 1, ny, s->H.p, C, 1, s->Fy.p, (long)X2 * C, 1, ny, C, R4, 0.f, nb * mx, 0,
                       (long)R4 * C, 2L * C, mx, 0, sGs, sFs)))
            return st;
        if ((st = gemm(s, s->Ai.p, 1, ny, s->H.p, C, 1, s->Fy.p + C, (long)X2 * C, 1, ny, C, R4, 0.f, nb * mx, 0,
                       (long)R4 * C, 2L * C, mx, 0, sGs, sFs)))
            return st;
        // per row j of each sample: z = v_l W + x-iDFT(F') + b, v_{l+1} = gelu(z) (z for the last layer);
        // z is kept where the backward pass needs gelu'(z)
        Gemm2 g{};
        g.A = vl;
        g.sAb = (long)nx * C;
        g.sAm = C;
        g.sAk = 1;
        g.B = prm + s->oW[l];
        g.sBk = C;
        g.sBn = 1;
        g.K = C;
        g.A2 = s->Bx.p;
        g.sA2m = X2;
        g.sA2k = 1;
        g.B2 = s->Fy.p;
        g.sB2b = (long)X2 * C;
        g.sB2k = C;
        g.sB2n = 1;
        g.K2 = X2;
        g.M = nx;
        g.N = C;
        g.sOb = (long)nx * C;
        g.epi = EPI_BIAS_ACT;
        g.act = l < L - 1;
        g.bias = prm + s->ob[l];
        g.zout = l < L - 1 ? zl : nullptr;
        g.vout = s->v[l + 1].p;
        if ((st = gemm2(s, g, true, nb * ny))) return st;
    }
    // projection: q = gelu(v_L Q1 + q1)
    {
        Gemm2 g{};
        g.A = s->v[L].p;
        g.sAm = C;
        g.sAk = 1;
        g.B = prm + s->oQ1;
        g.sBk = Hh;
        g.sBn = 1;
        g.K = C;
        g.M = (int)PB;
        g.N = Hh;
        g.epi = EPI_BIAS_ACT;
        g.act = 1;
        g.bias = prm + s->oq1;
        g.zout = s->q_pre.p;
        g.vout = s->q.p;
        if ((st = gemm2(s, g, true, 1))) return st;
    }
    if ((st = gemm(s, s->q.p, Hh, 1, prm + s->oQ2, 1, 1, s->y.p, 1, 1, (int)PB, 1, Hh, 0.f))) return st;
    hipLaunchKernelGGL(k_bias_act, dim3(ew_grid(PB)), dim3(256), 0, s->stream, s->y.p, prm + s->oq2, PB, 1, 0,
                       s->y.p);
    SCHECK(s, hipGetLastError());
    return 0;
}

// gradient of the loss summed over the nb samples with respect to every parameter, accumulated
// into s->grad (after forward(x, nb); s->dy holds dL/dy)
int backward(fi_surrogate* s, const float* x, int nb) {
    const fi_fno_config& c = s->cfg;
    const int C = c.width, L = c.layers, X2 = 2 * c.modes_x, R4 = 4 * c.modes_y, mx = c.modes_x, Hh = c.proj_hidden;
    const int nx = s->nx, ny = s->ny, nmode = mx * 2 * c.modes_y;
    const long PB = s->P * nb;
    const long sFs = (long)ny * X2 * C, sGs = (long)mx * R4 * C;
    float* prm = s->par.p;
    float* gr = s->grad.p;
    int st;
    // y = q Q2 + q2
    if ((st = gemm_splitk(s, s->q.p, 1, Hh, s->dy.p, 1, 1, gr + s->oQ2, 1, 1, Hh, 1, (int)PB, 1.f, true))) return st;
    {   // dq = (dy Q2^T) gelu'(q_pre)
        Gemm2 g{};
        g.A = s->dy.p;
        g.sAm = 1;
        g.sAk = 1;
        g.B = prm + s->oQ2;
        g.sBk = 1;
        g.sBn = 1;
        g.K = 1;
        g.M = (int)PB;
        g.N = Hh;
        g.epi = EPI_ACT_BACK;
        g.zin = s->q_pre.p;
        g.vout = s->dqb.p;
        if ((st = gemm2(s, g, false, 1))) return st;
    }
    // q_pre = v_L Q1 + q1
    if ((st = gemm_splitk(s, s->v[L].p, 1, C, s->dqb.p, Hh, 1, gr + s->oQ1, Hh, 1, C, Hh, (int)PB, 1.f, true)))
        return st;
    int cur = 0;
    {   // dz_L = dv_L = dq Q1^T (the last layer has no activation)
        Gemm2 g{};
        g.A = s->dqb.p;
        g.sAm = Hh;
        g.sAk = 1;
        g.B = prm + s->oQ1;
        g.sBk = 1;
        g.sBn = Hh;
        g.K = Hh;
        g.M = (int)PB;
        g.N = C;
        g.epi = EPI_ACT_BACK;
        g.vout = s->dzb[cur].p;
        if ((st = gemm2(s, g, true, 1))) return st;
    }
    for (int l = L - 1; l >= 0; --l) {
        // dz: gradient with respect to z_{l+1}
        const __bf16* vl = s->v[l].p;
        const __bf16* dz = s->dzb[cur].p;
        if ((st = gemm_splitk(s, vl, 1, C, dz, C, 1, gr + s->oW[l], C, 1, C, C, (int)PB, 1.f, true))) return st;
        // spectral path: dF' = Bx^T dz per row j of each sample
        if ((st = gemm(s, s->Bx.p, 1, X2, dz, C, 1, s->dFy.p, C, 1, X2, C, nx, 0.f, nb * ny, 0, (long)nx * C,
                       (long)X2 * C)))
            return st;
        // dH = Ar dRe F' + Ai dIm F' per (sample, kx)
        if ((st = gemm(s, s->Ar.p, ny, 1, s->dFy.p, (long)X2 * C, 1, s->dH.p, C, 1, R4, C, ny, 0.f, nb * mx, 0, 2L * C,
                       (long)R4 * C, mx, 0, sFs, sGs, s->Ai.p, s->dFy.p + C)))
            return st;
        hipLaunchKernelGGL(k_mix_back_r, dim3((int)(((long)nmode * C * C + 255) / 256)), dim3(256), 0, s->stream,
                           s->dH.p, s->Gs[l].p, gr + s->oR[l], nmode, C, nb);
        hipLaunchKernelGGL(k_mix_back_g, dim3((nb * nmode * C + 255) / 256), dim3(256), 0, s->stream, s->dH.p,
                           prm + s->oR[l], s->dG.p, nmode, C, nb);
        // dRe Fx = Ar^T dG, dIm Fx = Ai^T dG per (sample, kx)
        if ((st = gemm(s, s->Ar.p, 1, ny, s->dG.p, C, 1, s->dFx.p, (long)X2 * C, 1, ny, C, R4, 0.f, nb * mx, 0,
                       (long)R4 * C, 2L * C, mx, 0, sGs, sFs)))
            return st;
        if ((st = gemm(s, s->Ai.p, 1, ny, s->dG.p, C, 1, s->dFx.p + C, (long)X2 * C, 1, ny, C, R4, 0.f, nb * mx, 0,
                       (long)R4 * C, 2L * C, mx, 0, sGs, sFs)))
            return st;
        // per row j of each sample: dv_l = dz W^T + Ex^T dFx, then the gradient with respect to z_l,
        // dv_l gelu'(z_l) (layers 1 .. L-1 apply gelu; dv_0 as it is)
        Gemm2 g{};
        g.A = dz;
        g.sAb = (long)nx * C;
        g.sAm = C;
        g.sAk = 1;
        g.B = prm + s->oW[l];
        g.sBk = 1;
        g.sBn = C;
        g.K = C;
        g.A2 = s->ExT.p;
        g.sA2m = X2;
        g.sA2k = 1;
        g.B2 = s->dFx.p;
        g.sB2b = (long)X2 * C;
        g.sB2k = C;
        g.sB2n = 1;
        g.K2 = X2;
        g.M = nx;
        g.N = C;
        g.sOb = (long)nx * C;
        g.epi = EPI_ACT_BACK;
        g.zin = l >= 1 ? s->z[l].p : nullptr;
        g.vout = s->dzb[cur ^ 1].p;
        if ((st = gemm2(s, g, true, nb * ny))) return st;
        cur ^= 1;
    }
    // dzb[cur] now holds dL/dv_0; v_0 = x P + p
    const __bf16* dv0 = s->dzb[cur].p;
    if ((st = gemm_splitk(s, x, 1, c.in_channels, dv0, C, 1, gr + s->oP, C, 1, c.in_channels, C, (int)PB, 1.f, true)))
        return st;
    SCHECK(s, hipGetLastError());
    return 0;
}

// loss (sum of the nb samples' mean squared errors) into *dev_out and dL/dy
int loss_and_dy(fi_surrogate* s, const float* t, int nb, double* dev_out) {
    const int nblk = 256;
    hipLaunchKernelGGL(k_mse, dim3(nblk), dim3(256), 0, s->stream, s->y.p, t, s->P, s->P * nb, s->dy.p, s->lpart.p);
    hipLaunchKernelGGL(k_sum_parts, dim3(1), dim3(64), 0, s->stream, s->lpart.p, nblk, dev_out);
    SCHECK(s, hipGetLastError());
    return 0;
}

// samples that go through the network side by side: up to 16, within about 4 GB of activations
int chunk_for(const fi_surrogate* s, int batch) {
    const fi_fno_config& c = s->cfg;
    const double per_sample = (double)s->P * ((c.layers + 2) * c.width * 6.0 + c.proj_hidden * 12.0);
    return std::max(1, std::min({batch, 16, (int)std::min(64.0, 4e9 / per_sample)}));
}

// per-sample buffers for nb samples side by side
int ensure_cap(fi_surrogate* s, int nb) {
    if (nb <= s->cap) return 0;
    const fi_fno_config& c = s->cfg;
    const long P = s->P * nb;
    const int C = c.width, X2 = 2 * c.modes_x, R4 = 4 * c.modes_y, mx = c.modes_x, Hh = c.proj_hidden, L = c.layers;
    const size_t nF = (size_t)nb * s->ny * X2 * C, nG = (size_t)nb * mx * R4 * C;
    s->cap = 0;
    s->v.clear();
    s->z.clear();
    s->Gs.clear();
    s->v.resize(L + 1);
    s->z.resize(L + 1);
    s->Gs.resize(L);
    for (int l = 0; l <= L; ++l) {
        SCHECK(s, s->v[l].alloc(P * C));
        SCHECK(s, s->z[l].alloc(l >= 1 && l < L ? P * C : 1));   // z_1 .. z_{L-1}: the activated layers
    }
    for (int l = 0; l < L; ++l) SCHECK(s, s->Gs[l].alloc(nG));
    SCHECK(s, s->x.alloc(P * c.in_channels));
    SCHECK(s, s->y.alloc(P));
    SCHECK(s, s->dy.alloc(P));
    SCHECK(s, s->tgt.alloc(P));
    SCHECK(s, s->Fx.alloc(nF));
    SCHECK(s, s->dFx.alloc(nF));
    SCHECK(s, s->Fy.alloc(nF));
    SCHECK(s, s->dFy.alloc(nF));
    SCHECK(s, s->H.alloc(nG));
    SCHECK(s, s->dG.alloc(nG));
    SCHECK(s, s->dH.alloc(nG));
    SCHECK(s, s->q_pre.alloc(P * Hh));
    SCHECK(s, s->q.alloc(P * Hh));
    SCHECK(s, s->dqb.alloc(P * Hh));
    SCHECK(s, s->dzb[0].alloc(P * C));
    SCHECK(s, s->dzb[1].alloc(P * C));
    SCHECK(s, hipStreamSynchronize(s->stream));
    s->cap = nb;
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
    std::vector<float> Ex((size_t)X2 * nx), Ar((size_t)R4 * ny), Ai((size_t)R4 * ny), Bx((size_t)nx * X2),
        ExT((size_t)nx * X2);
    for (int k = 0; k < mx; ++k)
        for (int i = 0; i < nx; ++i) {
            const double th = 2.0 * M_PI * (double)((long)k * i % nx) / nx;
            Ex[(size_t)(2 * k) * nx + i] = (float)std::cos(th);
            Ex[(size_t)(2 * k + 1) * nx + i] = (float)-std::sin(th);
            ExT[(size_t)i * X2 + 2 * k] = Ex[(size_t)(2 * k) * nx + i];
            ExT[(size_t)i * X2 + 2 * k + 1] = Ex[(size_t)(2 * k + 1) * nx + i];
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
    SCHECK(s, up(s->ExT, ExT));
    s->cap = 0;
    int st = ensure_cap(s, 1);
    if (st) return st;
    s->nsamp = 0;   // a training set belongs to one grid
    return 0;
}

int fi_surrogate_forward(fi_surrogate* s, const float* x, float* y) {
    if (!s || !x || !y) return S_ERR_INVALID;
    if (!s->P) return sfail(s, S_ERR_STATE, "fi_surrogate_set_grid first");
    SCHECK(s, hipSetDevice(s->cfg.device));
    SCHECK(s, hipMemcpyAsync(s->x.p, x, s->P * s->cfg.in_channels * sizeof(float), hipMemcpyHostToDevice, s->stream));
    timed_begin(s);
    int st = forward(s, s->x.p, 1);
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
    if ((st = forward(s, s->x.p, 1)) || (st = loss_and_dy(s, s->tgt.p, 1, s->lsum.p)) || (st = backward(s, s->x.p, 1)))
        return st;
    if ((st = timed_end(s))) return st;
    SCHECK(s, hipMemcpy(grad, s->grad.p, s->np * sizeof(float), hipMemcpyDeviceToHost));
    if (loss) SCHECK(s, hipMemcpy(loss, s->lsum.p, sizeof(double), hipMemcpyDeviceToHost));
    return 0;
}

int fi_surrogate_loss_grad_n(fi_surrogate* s, int32_t n, const float* x, const float* t, double* loss, float* grad) {
    if (!s || n < 1 || !x || !t || !grad) return S_ERR_INVALID;
    if (!s->P) return sfail(s, S_ERR_STATE, "fi_surrogate_set_grid first");
    SCHECK(s, hipSetDevice(s->cfg.device));
    const long P = s->P;
    const int Ci = s->cfg.in_channels;
    const int chunk = chunk_for(s, n), nchunk = (n + chunk - 1) / chunk;
    int st = ensure_cap(s, chunk);
    if (st) return st;
    DB<double> ls;
    SCHECK(s, ls.alloc(nchunk));
    hipLaunchKernelGGL(k_fill, dim3(ew_grid(s->np)), dim3(256), 0, s->stream, s->grad.p, s->np, 0.f);
    for (int ci = 0; ci < nchunk; ++ci) {
        const int nb = std::min(chunk, n - ci * chunk);
        const size_t s0 = (size_t)ci * chunk;
        SCHECK(s, hipMemcpyAsync(s->x.p, x + s0 * P * Ci, (size_t)nb * P * Ci * sizeof(float), hipMemcpyHostToDevice,
                                 s->stream));
        SCHECK(s, hipMemcpyAsync(s->tgt.p, t + s0 * P, (size_t)nb * P * sizeof(float), hipMemcpyHostToDevice, s->stream));
        timed_begin(s);
        if ((st = forward(s, s->x.p, nb)) || (st = loss_and_dy(s, s->tgt.p, nb, ls.p + ci)) ||
            (st = backward(s, s->x.p, nb)))
            return st;
        if ((st = timed_end(s))) return st;
    }
    SCHECK(s, hipMemcpy(grad, s->grad.p, s->np * sizeof(float), hipMemcpyDeviceToHost));
    std::vector<double> h(nchunk);
    SCHECK(s, hipMemcpy(h.data(), ls.p, nchunk * sizeof(double), hipMemcpyDeviceToHost));
    if (loss) {
        double m = 0;
        for (double v : h) m += v;
        *loss = m;
    }
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
    const long P = s->P;
    const int Ci = s->cfg.in_channels;
    const int chunk = chunk_for(s, batch);
    const int nchunk = (batch + chunk - 1) / chunk;
    int st = ensure_cap(s, chunk);
    if (st) return st;
    DB<double> losses;
    SCHECK(s, losses.alloc((size_t)std::max(1, steps) * nchunk));
    uint64_t rs = seed;
    timed_begin(s);
    for (int it = 0; it < steps; ++it) {
        hipLaunchKernelGGL(k_fill, dim3(ew_grid(s->np)), dim3(256), 0, s->stream, s->grad.p, s->np, 0.f);
        for (int ci = 0; ci < nchunk; ++ci) {
            const int nb = std::min(chunk, batch - ci * chunk);
            Picks pk{};
            for (int b = 0; b < nb; ++b) pk.k[b] = (int)(splitmix64(rs) % (uint64_t)s->nsamp);
            hipLaunchKernelGGL(k_gather, dim3(ew_grid((long)nb * P * (Ci + 1))), dim3(256), 0, s->stream, s->xs.p,
                               s->ys.p, P * Ci, P, pk, nb, s->x.p, s->tgt.p);
            if ((st = forward(s, s->x.p, nb)) || (st = loss_and_dy(s, s->tgt.p, nb, losses.p + (size_t)it * nchunk + ci)) ||
                (st = backward(s, s->x.p, nb)))
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
    if ((st = timed_end(s))) return st;
    if (loss_out && steps > 0) {
        std::vector<double> h((size_t)steps * nchunk);
        SCHECK(s, hipMemcpy(h.data(), losses.p, h.size() * sizeof(double), hipMemcpyDeviceToHost));
        for (int it = 0; it < steps; ++it) {
            double m = 0;
            for (int ci = 0; ci < nchunk; ++ci) m += h[(size_t)it * nchunk + ci];
            loss_out[it] = m / batch;
        }
    }
    return 0;
}

double fi_surrogate_device_ms(const fi_surrogate* s) { return s ? s->device_ms : 0.0; }

}  // extern "C"
