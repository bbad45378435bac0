// Phases of the LV-ELVA right-hand side and the persistent integration kernel (gfx950).
//
// One RHS evaluation = update_diagnostics!(dudt, u, sim, t) (src/simulation.jl:335-393) runs as
// grid-wide phases separated by grid barriers:
//   rows  : [inverse real transform of the previous stage's spectrum for the row pairs covering
//           this block and its +-2-row halo] stage displacement (BS3 combination + viscous corner
//           BC, FP64), column loads, the two-sweep plate stencil (src/deformation.jl:110-131,
//           src/derivatives.jl CPU edge rules) on LDS rows, rhs * K/(2 eta), and a length-nx FFT
//           of the row pair packed as one complex sequence; the half spectrum is stored
//           transposed [kx][j];
//   cols  : length-ny FFT of each spectral column, multiplication by spy/(nx ny) / (R kappa),
//           inverse FFT (src/deformation.jl:95-98);
//   c2r   : standalone inverse real transform (last BS3 stage, with the error moments, and
//           single-RHS evaluations) -> dudt (m/yr) before the corner BC; the BC offset (mean of
//           the four raw corners) is published and applied by consumers.
// A BS3 step therefore costs 7 grid barriers: rows, cols, rows+c2r, cols, rows+c2r, cols, c2r.
// When the sparse trigger fires (src/simulation.jl:355-381) four more phases compute the elastic
// and sea-surface convolutions on the zero-padded nfx x nfy grid (both in one pass) and the
// pointwise sparse diagnostics.
#pragma once
#include <type_traits>

#include "fi_device.hpp"

namespace FI_NS {

// ------------------------------------------------------------------------------------------
// grid barrier (all workgroups co-resident: cooperative launch). Every workgroup publishes its
// arrival in its own 128-byte slot with an agent-scope release store; workgroup 0 polls all slots
// in parallel (one thread per slot) and then publishes the new generation, which the others poll.
// Depth is independent of the number of workgroups (no serialised atomics on one address).
// Waits are bounded: after spin_limit wall-clock ticks the abort flag is raised and every
// workgroup leaves the kernel.
// ------------------------------------------------------------------------------------------
__device__ __forceinline__ bool team_sync(const Team& tm, unsigned& gen, int* s_flag, unsigned long long spin_limit) {
    Bar* bar = tm.bar;
    unsigned* arrive = (unsigned*)(bar + 1);
    unsigned* release = arrive + (long)tm.size * BAR_SLOT;
    const unsigned g = gen + 1u;
    gen = g;
    // Every store that another workgroup reads after this barrier is a write-through agent-scope
    // store (st_shared). Waiting until each wave's stores have completed therefore publishes
    // them, and the arrival and release flags need no L2 writeback (release fence) of their own.
    // Every read of such data is a coherent load (G<T>, ldc), so the acquire side needs no L2
    // invalidate either: a workgroup-scope acquire keeps the compiler from moving those loads
    // above the flag wait.
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    if (tm.rank == 0) {
        for (int w = 1 + (int)threadIdx.x; w < tm.size; w += NT) {
            const unsigned long long t0 = wall_clock64();
            while (__hip_atomic_load(&arrive[w * BAR_SLOT], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != g) {
                if (__hip_atomic_load(&bar->abort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) break;
                if (wall_clock64() - t0 > spin_limit) {
                    __hip_atomic_store(&bar->abort, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    break;
                }
                __builtin_amdgcn_s_sleep(1);
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
        __syncthreads();
        // publish the generation into every workgroup's own release slot
        for (int w = 1 + (int)threadIdx.x; w < tm.size; w += NT)
            __hip_atomic_store(&release[w * BAR_SLOT], g, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (threadIdx.x == 0) *s_flag = __hip_atomic_load(&bar->abort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    } else if (threadIdx.x == 0) {
        __hip_atomic_store(&arrive[tm.rank * BAR_SLOT], g, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        const unsigned long long t0 = wall_clock64();
        unsigned* mine = &release[tm.rank * BAR_SLOT];
        while (__hip_atomic_load(mine, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != g) {
            if (__hip_atomic_load(&bar->abort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) break;
            if (wall_clock64() - t0 > spin_limit) {
                __hip_atomic_store(&bar->abort, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                break;
            }
            __builtin_amdgcn_s_sleep(1);
        }
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
        *s_flag = __hip_atomic_load(&bar->abort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    __syncthreads();
    return *s_flag == 0;
}

// team_sync with the optional per-phase timing of the team's rank 0 (p.ptime, 100 MHz ticks:
// [tag] work before the barrier, [8 + tag] barrier wait, [16 + tag] count)
__device__ __forceinline__ bool grid_sync(const KP& p, const Team& tm, unsigned& gen, int* s_flag, int tag,
                                          unsigned long long& tlast) {
    const bool timed = p.ptime && tm.rank == 0 && threadIdx.x == 0;
    unsigned long long tin = 0;
    if (timed) tin = wall_clock64();
    const bool ok = team_sync(tm, gen, s_flag, p.spin_limit);
    if (timed) {
        const unsigned long long tout = wall_clock64();
        p.ptime[tag] += tin - tlast;
        p.ptime[8 + tag] += tout - tin;
        p.ptime[16 + tag] += 1;
        tlast = tout;
    }
    return ok;
}

// Block-wide sums in the canonical order of a 512-thread workgroup (VNT lanes), whatever the
// build's width: thread t holds the partial sums of virtual lanes t + v NT (v < NV), each lane's
// elements in increasing order; each virtual wave reduces by the wave64 butterfly, then the
// virtual waves are added in fixed order. Identical rounding in every workgroup and in both
// builds. The result is in v[0]; red holds at least M * VNT/64 doubles.
template <int M>
__device__ __forceinline__ void block_sum(double* red, double (&v)[NV][M]) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
    for (int vv = 0; vv < NV; ++vv) {
#pragma unroll
        for (int q = 0; q < M; ++q) {
            double x = v[vv][q];
#pragma unroll
            for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o, 64);
            v[vv][q] = x;
        }
    }
    if (lane == 0) {
#pragma unroll
        for (int vv = 0; vv < NV; ++vv) {
#pragma unroll
            for (int q = 0; q < M; ++q) red[(w + vv * (NT / 64)) * M + q] = v[vv][q];
        }
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < M; ++q) {
        double s = 0.0;
        for (int ww = 0; ww < VNT / 64; ++ww) s += red[ww * M + q];
        v[0][q] = s;
    }
    __syncthreads();
}

// ------------------------------------------------------------------------------------------
// stage description and per-call scalars
// ------------------------------------------------------------------------------------------
struct StageDesc {
    int ubase;         // U buffer of the base state
    int ng;            // derivative terms read from global memory (0..3)
    int g0, g1, g2;
    double a0, a1, a2;
    int kin;           // K slot still held as a spectrum (inverse-transformed in the rows phase), or -1
    double ain;
    double dt;
    int write_u;       // U buffer receiving the corner-corrected stage displacement, or -1
    int kout;          // K slot of this call's derivative
    int do_c2r;        // this call ends with the standalone inverse transform
    double t;          // call time
    double t_next_max; // upper bound of the next call's time (for the stale-load store)
    int trig;          // set by rhs_call: this call ran the sparse update
    int final_err;     // accumulate the BS3 error moments in the c2r phase
    int e1, e2, e3;    // K slots of k1, k2, k3 for the error moments
    int u_prev;        // U buffer of u_n for the error moments
    int write_sigma;   // set by rhs_call: store this call's loads for a later sparse update
    int sb_in;         // spectrum buffer holding the pending derivative (kin)
    int sb_out;        // spectrum buffer this call writes
    int force_sigma;   // store this call's loads whatever the next call's time
};

struct Acc3 {
    double a, b, c;
};
struct KC4 {
    float v0, v1, v2, v3;  // (0,0), (nx-1,0), (0,ny-1), (nx-1,ny-1)
};

struct IceT {
    int i0, i1, exact, err;
    double tt, dtk, a;  // t - t_i0, t_i1 - t_i0, separable amplitude a(t)
};

__device__ __forceinline__ double ice_at(const KP& p, const IceT& it, long c) {
    if (p.ice_mode == ICE_EXTERNAL) return (double)p.H_ext[c];
    if (p.ice_mode == ICE_SEPARABLE) return it.a * (double)p.ice_H[c];
    const long N = (long)p.nx * p.ny;
    const double h0 = p.ice_H[it.i0 * N + c];
    if (it.exact) return h0;
    const double h1 = p.ice_H[it.i1 * N + c];
    return h0 + (h1 - h0) / it.dtk * it.tt;
}

__device__ __forceinline__ double hab(double H, double zd, double rsi) {
    return fmax(H + fmin(zd, 0.0) * rsi, 0.0);
}

__device__ __forceinline__ G<float> kslot(const KP& p, int k) { return p.Kb + (long)k * p.Nl; }

// t_computation! (src/simulation.jl:46-52), called by thread 0 at the end of RHS call k (0-based
// since fi_init): workgroup 0 of the team stores (t, wall clock) at index k; the host keeps the
// calls past t0. Stores only, no read back.
__device__ __forceinline__ void timer_push(const KP& p, const Team& tm, long long k, double t) {
    if (p.hist_t && tm.rank == 0 && k < p.hist_cap) {
        p.hist_t[k] = t;
        p.hist_clk[k] = wall_clock64();
    }
}
__device__ __forceinline__ G<float2> specbuf(const KP& p, int b) { return p.specb + (long)b * p.specN; }
// Element (kx, j) of a spectrum buffer, blocked [kx/4][j][4]: the four columns of a block are
// interleaved per row. A row pair's half spectrum is then 64 contiguous bytes per block (the rows
// phase's stores and its halo reads touch 16x fewer cache lines than in a [kx][j] layout) and a
// column is read and written with a 32-byte stride (columns phase). Every line of the buffer
// crosses the device's coherence point on each read (coherent loads), so lines, not bytes, are
// the traffic. The implicit step's two spectra use the same layout.
__device__ __forceinline__ long sidx(const KP& p, int kx, int j) {
    return (((long)(kx >> 2) * p.ny + j) << 2) + (kx & 3);
}
// the same blocked layout for a padded-grid half spectrum of `rows` rows (pA/pB: ny rows, the
// sparse update's oA/oB: ny + 2, the relaxed mantle's oA: ny)
__device__ __forceinline__ long bidx(int kx, int j, int rows) {
    return (((long)(kx >> 2) * rows + j) << 2) + (kx & 3);
}
__device__ __forceinline__ G<double> ubuf(const KP& p, int b) { return p.Ub + (long)b * p.Nl; }

// sum of the global derivative terms of a stage at cell c
__device__ __forceinline__ double kterms(const KP& p, const StageDesc& st, long c) {
    double v = 0.0;
    if (st.ng > 0) v = st.a0 * (double)kslot(p, st.g0)[c];
    if (st.ng > 1) v += st.a1 * (double)kslot(p, st.g1)[c];
    if (st.ng > 2) v += st.a2 * (double)kslot(p, st.g2)[c];
    return v;
}
// stage displacement before the BC when every term is in global memory
__device__ __forceinline__ double u_raw(const KP& p, const StageDesc& st, long c) {
    const double u0 = ubuf(p, st.ubase)[c];
    if (st.ng == 0) return u0;
    return u0 + st.dt * kterms(p, st, c);
}

// corner-BC offset of a raw derivative slot
__device__ __forceinline__ double k_offset(const KP& p, int slot) {
    const G<float> q = p.kcorner + 4 * slot;
    const double a = 0.25 * (double)q[0], b = 0.25 * (double)q[1], c = 0.25 * (double)q[2], d = 0.25 * (double)q[3];
    return ((a + b) + (c + d)) - p.vbc_x;
}

// The three inputs every RHS call needs before its rows phase, gathered in one round of global
// loads and one block reduction (every thread returns the same values):
//  * the TimeInterpolation2D bracket of st.t (src/interpolations.jl:51-74): i = first snapshot
//    with t_i >= t; exact hit at i, else linear between i-1 and i; outside [t_0, t_end] an error
//    unless the series is flat-extrapolated. Thread k tests snapshot k.
//  * the four corner values of the pending spectral derivative, by direct sums over spectral rows
//    0 and ny-1: x(0, j) = sum_k w_k Re X_k, x(nx-1, j) = sum_k w_k Re(X_k e^{-2 pi i k/nx})
//    (w = 1 for the DC and Nyquist terms, 2 otherwise);
//  * the apply_bc!(u, CornerBC) offset of the stage displacement (src/boundary_conditions.jl:
//    144-148): the mean of its four corners minus the border value.
__device__ __forceinline__ void stage_preamble(const KP& p, const StageDesc& st, const G<float2> spec, const float2* twx,
                                               double* red, IceT& it_out, KC4& kc_out, double& uoff_out) {
    __shared__ IceT sh_it;
    __shared__ double sh_uc[4], sh_u0[4];
    const int tid = threadIdx.x;
    const bool ice_search = p.ice_mode != ICE_EXTERNAL;
    const bool sep = p.ice_mode == ICE_SEPARABLE;
    const int nt = p.ice_nt;
    const double t = st.t;
    // ---- loads ----
    double tfirst = 0, tlast = 0, tk = 0, tp = 0, ak = 0, ap = 0;
    if (ice_search) {
        tfirst = p.ice_t[0];
        tlast = p.ice_t[nt - 1];
        if (tid < nt) {
            tk = p.ice_t[tid];
            tp = tid ? p.ice_t[tid - 1] : 0.0;
            if (sep) {
                ak = p.ice_a[tid];
                ap = tid ? p.ice_a[tid - 1] : 0.0;
            }
        }
    }
    const int nx = p.nx, ny = p.ny;
    const bool spec_on = st.kin >= 0;
    float2 sa = make_float2(0.f, 0.f), sb = make_float2(0.f, 0.f), tw = make_float2(1.f, 0.f);
    if (spec_on && tid < p.nhx) {
        sa = spec[sidx(p, tid, 0)];
        sb = spec[sidx(p, tid, ny - 1)];
        tw = twx[tid];
    }
    if (tid < 4) {
        // corner tid: base displacement and the sum of the global derivative terms
        const long n1 = nx - 1, n2 = (long)(ny - 1) * nx;
        const long c = tid == 0 ? 0 : tid == 1 ? n1 : tid == 2 ? n2 : n2 + n1;
        sh_u0[tid] = ubuf(p, st.ubase)[c];
        sh_uc[tid] = kterms(p, st, c);
    }
    // ---- ice bracket ----
    IceT it;
    it.i0 = it.i1 = 0;
    it.exact = 1;
    it.err = 0;
    it.tt = 0;
    it.dtk = 1;
    it.a = 0;
    bool in_range = true;
    if (ice_search && (t < tfirst || t > tlast)) {
        in_range = false;
        if (!p.ice_flat) it.err = 1;
        it.i0 = it.i1 = (t < tfirst) ? 0 : nt - 1;
    }
    if (ice_search && in_range) {
        for (int k = tid; k < nt; k += NT) {
            if (k >= NT) {
                tk = p.ice_t[k];
                tp = p.ice_t[k - 1];
                if (sep) {
                    ak = p.ice_a[k];
                    ap = p.ice_a[k - 1];
                }
            }
            if (tk >= t && (k == 0 || tp < t)) {
                IceT r = it;
                if (tk == t) {
                    r.i0 = r.i1 = k;
                    r.a = ak;
                } else {
                    r.i0 = k - 1;
                    r.i1 = k;
                    r.exact = 0;
                    r.tt = t - tp;
                    r.dtk = tk - tp;
                    r.a = ap + (ak - ap) / r.dtk * r.tt;
                }
                sh_it = r;
            }
        }
    }
    // ---- spectral corners: x(0, j) = sum_k w_k Re X_k, x(nx-1, j) = sum_k w_k Re(X_k e^{-2 pi i k/nx}) ----
    double vs[NV][4] = {};
    if (spec_on) {
        const bool even = (nx % 2) == 0;
        for (int k0 = tid; k0 < p.nhx; k0 += VNT) {
#pragma unroll
            for (int vv = 0; vv < NV; ++vv) {
                const int kx = k0 + vv * NT;
                if (kx >= p.nhx) break;
                float2 a = sa, b = sb, tw_k = tw;
                if (kx >= NT) {
                    a = spec[sidx(p, kx, 0)];
                    b = spec[sidx(p, kx, ny - 1)];
                    tw_k = twx[kx];
                }
                const bool nyq = even && kx == nx / 2;
                const double w = (kx == 0 || nyq) ? 1.0 : 2.0;
                double* v = vs[vv];
                v[0] += w * (double)a.x;
                v[2] += w * (double)b.x;
                if (nyq) {
                    v[1] -= (double)a.x;
                    v[3] -= (double)b.x;
                } else {
                    v[1] += w * ((double)a.x * tw_k.x - (double)a.y * tw_k.y);
                    v[3] += w * ((double)b.x * tw_k.x - (double)b.y * tw_k.y);
                }
            }
        }
        block_sum<4>(red, vs);
    } else {
        __syncthreads();
    }
    const double* v = vs[0];
    KC4 kc;
    kc.v0 = (float)v[0];
    kc.v1 = (float)v[1];
    kc.v2 = (float)v[2];
    kc.v3 = (float)v[3];
    if (ice_search && in_range) {
        it = sh_it;
        if (!sep) it.a = 0;
    } else if (ice_search && sep) {
        it.a = p.ice_a[it.i0];
    }
    // corner-BC offset: u0 + dt (sum of global terms + ain kc) at the four corners
    const float kcv[4] = {kc.v0, kc.v1, kc.v2, kc.v3};
    double q[4];
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        double vv = sh_uc[c];
        if (spec_on) vv += st.ain * (double)kcv[c];
        q[c] = 0.25 * ((st.ng == 0 && !spec_on) ? sh_u0[c] : sh_u0[c] + st.dt * vv);
    }
    uoff_out = ((q[0] + q[1]) + (q[2] + q[3])) - p.vbc_x;
    it_out = it;
    kc_out = kc;
    __syncthreads();
}

__device__ __forceinline__ int corner_index(const KP& p, int i, int j) {
    if (j == 0 && i == 0) return 0;
    if (j == 0 && i == p.nx - 1) return 1;
    if (j == p.ny - 1 && i == 0) return 2;
    if (j == p.ny - 1 && i == p.nx - 1) return 3;
    return -1;
}

// ------------------------------------------------------------------------------------------
// finite differences, CPU edge rules (src/derivatives.jl:66-76,171-191)
// ------------------------------------------------------------------------------------------
template <typename V>
__device__ __forceinline__ double d2x(const V* r, int i, int nx, double inv2) {
    if (i == 0) return ((double)r[2] - 2.0 * (double)r[1] + (double)r[0]) * inv2;
    if (i == nx - 1) return ((double)r[nx - 1] - 2.0 * (double)r[nx - 2] + (double)r[nx - 3]) * inv2;
    return ((double)r[i + 1] - 2.0 * (double)r[i] + (double)r[i - 1]) * inv2;
}
template <typename V>
__device__ __forceinline__ double d1x(const V* r, int i, int nx, double inv) {
    if (i == 0) return ((double)r[1] - (double)r[0]) * inv;
    if (i == nx - 1) return ((double)r[nx - 1] - (double)r[nx - 2]) * inv;
    return ((double)r[i + 1] - (double)r[i - 1]) * (0.5 * inv);
}
// second y-difference of field F at row j; rows are stored from row r0 with stride nx
template <typename V>
__device__ __forceinline__ double d2y(const V* F, int r0, int j, int i, int nx, int ny, double inv2) {
    auto at = [&](int jj) { return (double)F[(jj - r0) * nx + i]; };
    if (j == 0) return (at(2) - 2.0 * at(1) + at(0)) * inv2;
    if (j == ny - 1) return (at(ny - 1) - 2.0 * at(ny - 2) + at(ny - 3)) * inv2;
    return (at(j + 1) - 2.0 * at(j) + at(j - 1)) * inv2;
}
// dy(dx(F)) at row j (dxy! = dx! then dy!)
template <typename V>
__device__ __forceinline__ double dxy(const V* F, int r0, int j, int i, int nx, int ny, double invx, double invy) {
    auto dxr = [&](int jj) { return d1x(F + (jj - r0) * nx, i, nx, invx); };
    if (j == 0) return (dxr(1) - dxr(0)) * invy;
    if (j == ny - 1) return (dxr(ny - 1) - dxr(ny - 2)) * invy;
    return (dxr(j + 1) - dxr(j - 1)) * (0.5 * invy);
}

// Hermitian-extended pair value Z[k] = Aext[k] + i Bext[k] of spectral rows ja, jb (jb may be a
// phantom row) for the inverse real transform
__device__ __forceinline__ float2 c2r_pair_value(const KP& p, const G<float2> spec, int k, int ja, int jb) {
    const int nx = p.nx, ny = p.ny;
    float2 A, B;
    const int kk = (k < p.nhx) ? k : nx - k;
    A = spec[sidx(p, kk, ja)];
    if (jb < ny) B = spec[sidx(p, kk, jb)];
    else { B.x = 0; B.y = 0; }
    if (k >= p.nhx) { A.y = -A.y; B.y = -B.y; }
    else if (k == 0 || ((nx % 2) == 0 && k == nx / 2)) { A.y = 0; B.y = 0; }
    float2 Z;
    Z.x = A.x - B.y;
    Z.y = A.y + B.x;
    return Z;
}

// Sub-phase wall-clock marks of workgroup 0 inside the rows phase, compiled in only with
// -DFI_SUBPHASE_TIMING (each mark synchronises the workgroup, so timings there are upper bounds).
#ifdef FI_SUBPHASE_TIMING
#define FI_SUBMARK(k)                                                                   \
    do {                                                                                \
        __syncthreads();                                                                \
        if (p.ptime && tm.rank == 0 && threadIdx.x == 0) {                              \
            const unsigned long long t_ = wall_clock64();                               \
            p.ptime[24 + (k)] += t_ - tsub;                                             \
            tsub = t_;                                                                  \
        }                                                                               \
    } while (0)
#else
#define FI_SUBMARK(k) \
    do {              \
    } while (0)
#endif

// ------------------------------------------------------------------------------------------
// rows phase
// ------------------------------------------------------------------------------------------
// Stores of data that other workgroups read after the next grid barrier (spectra, derivative
// slots, the stage displacement, partial sums). Relaxed agent-scope atomic stores write through
// the L2 (sc1), so the barrier's release has no dirty lines of theirs to write back; the
// barrier's release/acquire pair orders them as before. Vector stores only.
// address of element i of a buffer, for the write-through stores below
template <typename T>
__device__ __forceinline__ T* addr(G<T> b, long i) { return b.p + i; }
template <typename T>
__device__ __forceinline__ T* addr(T* b, long i) { return b + i; }
__device__ __forceinline__ void st_shared(float* a, float v) {
    __hip_atomic_store(a, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_shared(double* a, double v) {
    __hip_atomic_store(a, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_shared(unsigned char* a, unsigned char v) {
    __hip_atomic_store(a, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_shared(float2* a, float2 v) {
    unsigned long long b;
    __builtin_memcpy(&b, &v, sizeof(b));
    __hip_atomic_store((unsigned long long*)a, b, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// exact x / d for small non-negative integers (fractional part of (x+0.5)/d >= 0.5/d)
__device__ __forceinline__ int fdivi(int x, float inv_d) { return (int)(((float)x + 0.5f) * inv_d); }

// Per-cell fields of the load update, prefetched into registers at the start of a row block.
struct CellF {
    float H, zo, hr, lt, lw, i2e, m;
};

// Rows phase for grids wider than 512 cells (FP32 moments in LDS). DIST: per-cell grid steps K dx,
// K dy (map distortion) as in phase_rows_lds; each thread keeps 1/K of its two columns' six window
// rows in registers (K = 1 where p.Kinv is null). The general passes (grids under 6 rows or 3
// columns) have no distortion; set-up refuses distorted grids of that shape.
template <typename MT, int QD, int QF, bool DIST>
__device__ __forceinline__ void phase_rows(const KP& p, const Team& tm, const StageDesc& st, double uoff, const IceT it,
                           const float2* twx, char* smem, const KC4 kc) {
    const int nx = p.nx, ny = p.ny;
    const float inv_nx = 1.0f / (float)nx;
    double* su = (double*)smem;                 // 6 rows of the stage displacement
    MT* sMxx = (MT*)(su + 6 * nx);              // 4 rows each of the plate moments
    MT* sMyy = sMxx + 4 * nx;
    MT* sMxy = sMyy + 4 * nx;
    float2* sf = (float2*)(sMxy + 4 * nx);      // R2C pair + scratch
    float2* sg = sf + nx;
    float* skin = (float*)(sg + nx);            // 6 rows of the spectral-stage derivative
    // the inverse transforms of the halo pairs use the moment region as scratch
    float2* ca = (float2*)sMxx;
    const G<float2> spin = specbuf(p, st.sb_in);
    G<float2> spout = specbuf(p, st.sb_out);
    const int maxpairs = (int)((12 * nx * sizeof(MT)) / (16 * nx));
    const double nu = p.nu;
    const bool wsig = st.write_sigma != 0;
    const G<double> U0 = ubuf(p, st.ubase);
    for (int item = tm.rank; item < p.nrowitems; item += tm.size) {
        const int j0 = 2 * item, je = min(j0 + 2, ny);
        const int jm0 = max(0, min(j0 - 1, ny - 3)), jm1 = min(ny, max(je + 1, 3));
        const int ju0 = max(0, min(jm0 - 1, ny - 3)), ju1 = min(ny, max(jm1 + 1, 3));
        const int totU = (ju1 - ju0) * nx, totM = (jm1 - jm0) * nx, totO = 2 * nx;
#ifdef FI_SUBPHASE_TIMING
        unsigned long long tsub = wall_clock64();
#endif
        // stage terms held in global memory (base state + derivative slots) -> su
        for (int base = 0; base < totU; base += 4 * NT) {
            double uv[4];
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int idx = min(base + (int)threadIdx.x + q * NT, totU - 1);
                const int r = fdivi(idx, inv_nx), i = idx - r * nx;
                const long c = i + (long)(ju0 + r) * nx;
                uv[q] = (st.ng == 0) ? U0[c] : U0[c] + st.dt * kterms(p, st, c);
            }
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int idx = base + (int)threadIdx.x + q * NT;
                if (idx < totU) su[idx] = uv[q];
            }
        }
        FI_SUBMARK(0);
        // previous stage's derivative on rows [ju0, ju1): inverse real transforms of row pairs
        if (st.kin >= 0) {
            G<float> Kin = p.Kb + (long)st.kin * p.Nl;
            const int p0 = ju0 & ~1;
            const int np = (ju1 - p0 + 1) / 2;
            for (int q0 = 0; q0 < np; q0 += maxpairs) {
                const int nb = min(maxpairs, np - q0);
                for (int b0 = 0; b0 < nb * nx; b0 += 4 * NT) {
                    float2 zv[4];
#pragma unroll
                    for (int q = 0; q < 4; ++q) {
                        const int idx = min(b0 + (int)threadIdx.x + q * NT, nb * nx - 1);
                        const int b = fdivi(idx, inv_nx), k = idx - b * nx;
                        const int ja = p0 + 2 * (q0 + b);
                        zv[q] = c2r_pair_value(p, spin, k, ja, ja + 1);
                    }
#pragma unroll
                    for (int q = 0; q < 4; ++q) {
                        const int idx = b0 + (int)threadIdx.x + q * NT;
                        if (idx < nb * nx) ca[idx] = zv[q];
                    }
                }
                __syncthreads();
                FI_SUBMARK(1);
                float2* res = fft_lds<float2>(ca, ca + nb * nx, nb, nx, p.px, twx, +1);
                FI_SUBMARK(2);
                for (int idx = threadIdx.x; idx < 2 * nb * nx; idx += NT) {
                    const int b = fdivi(idx, 0.5f * inv_nx), rem = idx - b * 2 * nx, r = rem >= nx ? 1 : 0,
                              i = rem - r * nx;
                    const int j = p0 + 2 * (q0 + b) + r;
                    if (j < ju0 || j >= ju1) continue;
                    float v = r ? res[b * nx + i].y : res[b * nx + i].x;
                    const int ci = corner_index(p, i, j);
                    if (ci >= 0) v = ci == 0 ? kc.v0 : ci == 1 ? kc.v1 : ci == 2 ? kc.v2 : kc.v3;
                    skin[(j - ju0) * nx + i] = v;
                    if (j >= j0 && j < je) {
                        st_shared(addr(Kin, i + (long)j * nx), v);
                        if (ci >= 0) st_shared(addr(p.kcorner, 4 * st.kin + ci), v);
                    }
                }
            }
        }
        FI_SUBMARK(3);
        // ---- plate rigidity and per-cell fields: loads in flight during the LDS-only work below.
        // Column-major per thread (nx <= 2 NT): Dr[2m + h] is the rigidity of moment row jm0 + m
        // and fr[2o + h] the fields of output row j0 + o, both at column threadIdx.x + h NT ----
        static_assert(QD == 8 && QF == 4, "two columns per thread: 4 moment rows, 2 output rows");
        float Dr[QD];
        CellF fr[QF];
#pragma unroll
        for (int q = 0; q < QD; ++q) {
            const int i = min((int)threadIdx.x + (q & 1) * NT, nx - 1), j = min(jm0 + (q >> 1), jm1 - 1);
            Dr[q] = p.D[i + (long)j * nx];
        }
#pragma unroll
        for (int q = 0; q < QF; ++q) {
            const int i = min((int)threadIdx.x + (q & 1) * NT, nx - 1), j = min(j0 + (q >> 1), ny - 1);
            const long c = i + (long)j * nx;
            fr[q].H = (float)ice_at(p, it, c);
            fr[q].zo = st.trig ? p.zoff_trig[c] : p.zoff[c];
            fr[q].hr = p.H_af_ref[c];
            fr[q].lt = p.LT[c];
            fr[q].lw = p.interactive ? p.LW[c] : 0.0f;
            fr[q].i2e = p.inv2eta[c];
            fr[q].m = p.mask[c] ? 1.0f : 0.0f;
        }
        // 1/K of window row r (u rows ju0 ..) at column threadIdx.x + h NT: kr[2 r + h]
        float kr[DIST ? 12 : 1];
        if constexpr (DIST) {
#pragma unroll
            for (int q = 0; q < 12; ++q) {
                const int i = min((int)threadIdx.x + (q & 1) * NT, nx - 1), j = min(ju0 + (q >> 1), ju1 - 1);
                kr[q] = p.Kinv ? p.Kinv[i + (long)j * nx] : 1.0f;
            }
        }
        __syncthreads();
        const bool spec_in = st.kin >= 0;
        // right-hand side of output cell (i, j0 + o) from the moment derivatives, its displacement
        // u and its fields; stores the loads for a later sparse update when asked
        auto rhs_cell = [&](int i, int j, double bxx, double bxy, double byy, double u, const CellF& F) {
            const long c = i + (long)j * nx;
            const double m = F.m;
            const double Haf = hab((double)F.H, (double)F.zo + u, p.rho_sw_ice);
            const double s_ice = p.rho_ice * (Haf - (double)F.hr);
            const double s_load = m * (s_ice + (double)F.lw);
            const double s_mant = p.rho_um * u;
            const double s_full = s_load + m * ((double)F.lt + s_mant);
            const double rhs = -p.g * s_full + (bxx + (2.0 * bxy + byy));
            if (wsig) {
                st_shared(addr(p.sig_load_prev, c), (float)s_load);
                st_shared(addr(p.sig_geo_prev, c), (float)(p.compressible_column
                                                ? s_load + m * (s_mant - p.rho_um * (double)p.u_ref[c])
                                                : s_full));
            }
            return (float)(rhs * (double)F.i2e);
        };
        // Register-window plate stencil for two columns per thread (as in phase_rows_lds, whose
        // comments give the row rules): column i = threadIdx.x + h NT finishes the stage
        // displacement on an NU-row x 3-column window in registers and writes its moment rows;
        // after one workgroup barrier each column evaluates its output rows.
        auto window2 = [&](auto JU_, auto NU_, auto JM_, auto NM_, auto NO_, auto TOP_, auto BOT_, auto LO_) {
            constexpr int JU = decltype(JU_)::value, NU = decltype(NU_)::value, JM = decltype(JM_)::value,
                          NM = decltype(NM_)::value, NO = decltype(NO_)::value, LO = decltype(LO_)::value;
            constexpr bool TOP = decltype(TOP_)::value, BOT = decltype(BOT_)::value;
            double uo[2][NO];
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                int i = (int)threadIdx.x + h * NT;
                asm volatile("" : "+v"(i));
                if (i < nx) {
                    const int a = min(max(i - 1, 0), nx - 3);
                    const int ci = i - a;
                    double cv[NU], dd[NU], xx[NU];
#pragma unroll
                    for (int r = 0; r < NU; ++r) {
                        const int b0 = r * nx + a;
                        double x0 = su[b0], x1 = su[b0 + 1], x2 = su[b0 + 2];
                        if (spec_in) {
                            x0 += st.dt * (st.ain * (double)skin[b0]);
                            x1 += st.dt * (st.ain * (double)skin[b0 + 1]);
                            x2 += st.dt * (st.ain * (double)skin[b0 + 2]);
                        }
                        x0 -= uoff;
                        x1 -= uoff;
                        x2 -= uoff;
                        cv[r] = ci == 0 ? x0 : (ci == 1 ? x1 : x2);
                        dd[r] = ci == 1 ? (x2 - x0) * (0.5 * p.inv_dx)
                                        : (ci == 0 ? (x1 - x0) * p.inv_dx : (x2 - x1) * p.inv_dx);
                        xx[r] = (x2 - 2.0 * x1 + x0) * p.inv_dx2;
                        if constexpr (DIST) {   // Dx = K dx at this row's cell
                            const double k = (double)kr[2 * r + h];
                            dd[r] *= k;
                            xx[r] *= k * k;
                        }
                    }
#pragma unroll
                    for (int o = 0; o < NO; ++o) uo[h][o] = cv[o - JU];
                    if (st.write_u >= 0) {
                        const G<double> un = ubuf(p, st.write_u);
#pragma unroll
                        for (int o = 0; o < NO; ++o) st_shared(addr(un, i + (long)(j0 + o) * nx), uo[h][o]);
                    }
#pragma unroll
                    for (int m = 0; m < NM; ++m) {
                        const int ur = JM + m - JU;   // window row of this moment row
                        const bool first = TOP && m == 0, last = BOT && m == NM - 1;
                        const int c = first ? ur + 1 : (last ? ur - 1 : ur);
                        const int lo = first ? ur : ur - 1, hi = last ? ur : ur + 1;
                        const double D = (double)(h ? Dr[2 * m + 1] : Dr[2 * m]);
                        const double uxx = xx[ur];
                        double uyy = (cv[c + 1] - 2.0 * cv[c] + cv[c - 1]) * p.inv_dy2;
                        double uxy = (first || last) ? (dd[hi] - dd[lo]) * p.inv_dy
                                                     : (dd[hi] - dd[lo]) * (0.5 * p.inv_dy);
                        if constexpr (DIST) {   // Dy = K dy at the moment's cell
                            const double k = (double)kr[2 * ur + h];
                            uyy *= k * k;
                            uxy *= k;
                        }
                        sMxx[m * nx + i] = (MT)(-D * (nu * uyy + uxx));
                        sMyy[m * nx + i] = (MT)(-D * (nu * uxx + uyy));
                        sMxy[m * nx + i] = (MT)(-D * (1.0 - nu) * uxy);
                    }
                }
            }
            __syncthreads();
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                const int i = (int)threadIdx.x + h * NT;
                if (i < nx) {
                    const int a = min(max(i - 1, 0), nx - 3);
                    const int ci = i - a;
                    auto mx = [&](const MT* M, int mm, int cc) { return (double)M[mm * nx + a + cc]; };
                    auto d1m = [&](int mm) {
                        const double d = ci == 1 ? (mx(sMxy, mm, 2) - mx(sMxy, mm, 0)) * (0.5 * p.inv_dx)
                                                 : (ci == 0 ? (mx(sMxy, mm, 1) - mx(sMxy, mm, 0)) * p.inv_dx
                                                            : (mx(sMxy, mm, 2) - mx(sMxy, mm, 1)) * p.inv_dx);
                        if constexpr (DIST) return d * (double)kr[2 * (JM + mm - JU) + h];
                        return d;
                    };
                    float f[2] = {0.0f, 0.0f};
#pragma unroll
                    for (int o = 0; o < NO; ++o) {
                        const int mr = o - JM;   // moment row of output row j0 + o
                        const bool first = TOP && o == 0, last = o == LO;
                        const int c = first ? mr + 1 : (last ? mr - 1 : mr);
                        const int lo = first ? mr : mr - 1, hi = last ? mr : mr + 1;
                        double bxx = (mx(sMxx, mr, 2) - 2.0 * mx(sMxx, mr, 1) + mx(sMxx, mr, 0)) * p.inv_dx2;
                        double byy = ((double)sMyy[(c + 1) * nx + i] - 2.0 * (double)sMyy[c * nx + i] +
                                      (double)sMyy[(c - 1) * nx + i]) * p.inv_dy2;
                        double bxy = (first || last) ? (d1m(hi) - d1m(lo)) * p.inv_dy
                                                     : (d1m(hi) - d1m(lo)) * (0.5 * p.inv_dy);
                        if constexpr (DIST) {   // Dx, Dy at the output cell
                            const double k = (double)kr[2 * (o - JU) + h];
                            bxx *= k * k;
                            byy *= k * k;
                            bxy *= k;
                        }
                        const CellF F = h ? fr[2 * o + 1] : fr[2 * o];
                        f[o] = rhs_cell(i, j0 + o, bxx, bxy, byy, uo[h][o], F);
                    }
                    sf[i] = make_float2(f[0], f[1]);
                }
            }
        };
        using std::integral_constant;
#define FI_WIN2(JU, NU, JM, NM, NO, TOP, BOT, LO)                                                          \
    window2(integral_constant<int, JU>{}, integral_constant<int, NU>{}, integral_constant<int, JM>{},     \
            integral_constant<int, NM>{}, integral_constant<int, NO>{}, integral_constant<bool, TOP>{},    \
            integral_constant<bool, BOT>{}, integral_constant<int, LO>{})
        // the kind of this row pair (as in phase_rows_lds); anything else takes the general passes
        const int rj = ju0 - j0, rm = jm0 - j0, nu_ = ju1 - ju0, nm_ = jm1 - jm0, no_ = je - j0;
        const int kind = (ny < 6 || nx < 3) ? -1
                         : (rj == -2 && nu_ == 6 && rm == -1 && nm_ == 4 && no_ == 2 && j0 >= 2 && j0 + 3 <= ny - 1) ? 0
                         : (j0 == 0 && rj == 0 && nu_ == 4 && rm == 0 && nm_ == 3 && no_ == 2)                     ? 1
                         : (j0 == ny - 2 && rj == -2 && nu_ == 4 && rm == -1 && nm_ == 3 && no_ == 2)              ? 2
                         : (j0 == ny - 3 && rj == -2 && nu_ == 5 && rm == -1 && nm_ == 4 && no_ == 2)              ? 3
                         : (j0 == ny - 1 && rj == -3 && nu_ == 4 && rm == -2 && nm_ == 3 && no_ == 1)              ? 4
                                                                                                                  : -1;
        if (kind == 0) {
            FI_WIN2(-2, 6, -1, 4, 2, false, false, -1);
        } else if (kind == 1) {
            FI_WIN2(0, 4, 0, 3, 2, true, false, -1);
        } else if (kind == 2) {
            FI_WIN2(-2, 4, -1, 3, 2, false, true, 1);
        } else if (kind == 3) {
            FI_WIN2(-2, 5, -1, 4, 2, false, true, -1);
        } else if (kind == 4) {
            FI_WIN2(-3, 4, -2, 3, 1, false, true, 0);
        } else {
            // general passes (grids with fewer than 6 rows): rigidity and fields read where used
            for (int idx = threadIdx.x; idx < totU; idx += NT) {
                double u = su[idx];
                if (spec_in) u += st.dt * (st.ain * (double)skin[idx]);
                u -= uoff;
                su[idx] = u;
                const int r = fdivi(idx, inv_nx);
                const int j = ju0 + r;
                if (st.write_u >= 0 && j >= j0 && j < je) st_shared(addr(ubuf(p, st.write_u), idx - r * nx + (long)j * nx), u);
            }
            __syncthreads();
            for (int idx = threadIdx.x; idx < totM; idx += NT) {
                const int r = fdivi(idx, inv_nx), i = idx - r * nx, j = jm0 + r;
                const double D = (double)p.D[i + (long)j * nx];
                const double uxx = d2x(su + (j - ju0) * nx, i, nx, p.inv_dx2);
                const double uyy = d2y(su, ju0, j, i, nx, ny, p.inv_dy2);
                const double uxy = dxy(su, ju0, j, i, nx, ny, p.inv_dx, p.inv_dy);
                sMxx[idx] = (MT)(-D * (nu * uyy + uxx));
                sMyy[idx] = (MT)(-D * (nu * uxx + uyy));
                sMxy[idx] = (MT)(-D * (1.0 - nu) * uxy);
            }
            __syncthreads();
            for (int idx = threadIdx.x; idx < totO; idx += NT) {
                const int r = idx >= nx ? 1 : 0, i = idx - r * nx, j = j0 + r;
                float f = 0.0f;
                if (j < ny) {
                    const long c = i + (long)j * nx;
                    CellF F;
                    F.H = (float)ice_at(p, it, c);
                    F.zo = st.trig ? p.zoff_trig[c] : p.zoff[c];
                    F.hr = p.H_af_ref[c];
                    F.lt = p.LT[c];
                    F.lw = p.interactive ? p.LW[c] : 0.0f;
                    F.i2e = p.inv2eta[c];
                    F.m = p.mask[c] ? 1.0f : 0.0f;
                    const double bxx = d2x(sMxx + (j - jm0) * nx, i, nx, p.inv_dx2);
                    const double byy = d2y(sMyy, jm0, j, i, nx, ny, p.inv_dy2);
                    const double bxy = dxy(sMxy, jm0, j, i, nx, ny, p.inv_dx, p.inv_dy);
                    f = rhs_cell(i, j, bxx, bxy, byy, su[(j - ju0) * nx + i], F);
                }
                if (r == 0) sf[i].x = f;
                else sf[i].y = f;
            }
        }
#undef FI_WIN2
        __syncthreads();
        FI_SUBMARK(6);
        float2* res = fft_lds<float2>(sf, sg, 1, nx, p.px, twx, -1);
        for (int kx = threadIdx.x; kx < p.nhx; kx += NT) {
            const float2 Z = res[kx], Zm = res[kx == 0 ? 0 : nx - kx];
            float2 X0, X1;
            X0.x = 0.5f * (Z.x + Zm.x);
            X0.y = 0.5f * (Z.y - Zm.y);
            X1.x = 0.5f * (Z.y + Zm.y);
            X1.y = -0.5f * (Z.x - Zm.x);
            st_shared(addr(spout, sidx(p, kx, j0)), X0);
            if (j0 + 1 < ny) st_shared(addr(spout, sidx(p, kx, j0 + 1)), X1);
        }
        __syncthreads();
        FI_SUBMARK(7);
    }
}

// Rows phase for grids up to 512 wide (FP64 moments): every global input of a row item -- the
// stage terms, the pending spectral derivative's halo pairs, the rigidity and the per-cell fields
// -- is loaded in one round at the start of the item and staged in LDS, so the rest of the item
// (inverse transforms, stencils, loads, forward transform) is LDS work. DIST: per-cell grid steps
// K dx, K dy (map distortion); 1/K is staged with the displacement rows (K = 1 where p.Kinv is
// null, so one ensemble launch can mix distorted and undistorted members).
template <bool DIST>
__device__ __forceinline__ void phase_rows_lds(const KP& p, const Team& tm, const StageDesc& st, double uoff,
                                               const IceT it, const float2* twx, char* smem, const KC4 kc) {
    constexpr int QU = 6, QC = 4, QD = 4, QF = 2;   // loads per thread: (6|4|4|2) x NT >= (6|4|4|2) x 512
    const int nx = p.nx, ny = p.ny;
    const float inv_nx = 1.0f / (float)nx;
    double* su = (double*)smem;                 // 6 rows of the stage displacement
    double* sMxx = su + 6 * nx;                 // 4 rows each of the plate moments
    double* sMyy = sMxx + 4 * nx;
    double* sMxy = sMyy + 4 * nx;
    float2* sf = (float2*)(sMxy + 4 * nx);      // R2C pair + scratch
    float2* sg = sf + nx;
    float* skin = (float*)(sg + nx);            // 6 rows of the spectral-stage derivative
    float* sD = skin + 6 * nx;                  // 4 rows of the rigidity
    float* sH = sD + 4 * nx;                    // 2 rows each of the per-cell fields
    float* sZo = sH + 2 * nx;
    float* sHr = sZo + 2 * nx;
    float* sLt = sHr + 2 * nx;
    float* sLw = sLt + 2 * nx;
    float* sI2e = sLw + 2 * nx;
    float* sM = sI2e + 2 * nx;
    float* sKi = sM + 2 * nx;                   // 6 rows of 1/K (map distortion only)
    float2* ca = (float2*)sMxx;                 // inverse-transform scratch (moment region)
    const G<float2> spin = specbuf(p, st.sb_in);
    G<float2> spout = specbuf(p, st.sb_out);
    const double nu = p.nu;
    const bool wsig = st.write_sigma != 0;
    const G<double> U0 = ubuf(p, st.ubase);
    const G<float> zoff = st.trig ? p.zoff_trig : p.zoff;
    for (int item = tm.rank; item < p.nrowitems; item += tm.size) {
        const int j0 = 2 * item, je = min(j0 + 2, ny);
        const int jm0 = max(0, min(j0 - 1, ny - 3)), jm1 = min(ny, max(je + 1, 3));
        const int ju0 = max(0, min(jm0 - 1, ny - 3)), ju1 = min(ny, max(jm1 + 1, 3));
        const int totU = (ju1 - ju0) * nx, totM = (jm1 - jm0) * nx, totO = 2 * nx;
        const int p0 = ju0 & ~1;
        const int np = st.kin >= 0 ? (ju1 - p0 + 1) / 2 : 0;   // halo pairs (<= 4)
        const int totC = np * nx;
#ifdef FI_SUBPHASE_TIMING
        unsigned long long tsub = wall_clock64();
#endif
        // ---- one round of global loads ----
        double uv[QU];
        float kv[QU];
        float2 zv[QC];
        float dv[QD];
        float fH[QF], fZo[QF], fHr[QF], fLt[QF], fLw[QF], fI2e[QF], fM[QF];
#pragma unroll
        for (int q = 0; q < QU; ++q) {
            const int idx = min((int)threadIdx.x + q * NT, totU - 1);
            const int r = fdivi(idx, inv_nx), i = idx - r * nx;
            const long c = i + (long)(ju0 + r) * nx;
            uv[q] = (st.ng == 0) ? U0[c] : U0[c] + st.dt * kterms(p, st, c);
        }
        if (DIST) {
#pragma unroll
            for (int q = 0; q < QU; ++q) {
                const int idx = min((int)threadIdx.x + q * NT, totU - 1);
                const int r = fdivi(idx, inv_nx), i = idx - r * nx;
                kv[q] = p.Kinv ? p.Kinv[i + (long)(ju0 + r) * nx] : 1.0f;
            }
        }
#pragma unroll
        for (int q = 0; q < QC; ++q) {
            const int idx = min((int)threadIdx.x + q * NT, max(totC - 1, 0));
            const int b = fdivi(idx, inv_nx), k = idx - b * nx;
            const int ja = p0 + 2 * b;
            zv[q] = (totC > 0) ? c2r_pair_value(p, spin, k, ja, ja + 1) : make_float2(0.f, 0.f);
        }
#pragma unroll
        for (int q = 0; q < QD; ++q) {
            const int idx = min((int)threadIdx.x + q * NT, totM - 1);
            const int r = fdivi(idx, inv_nx), i = idx - r * nx;
            dv[q] = p.D[i + (long)(jm0 + r) * nx];
        }
#pragma unroll
        for (int q = 0; q < QF; ++q) {
            const int idx = min((int)threadIdx.x + q * NT, totO - 1);
            const int r = fdivi(idx, inv_nx), i = idx - r * nx;
            const long c = i + (long)min(j0 + r, ny - 1) * nx;
            fH[q] = (float)ice_at(p, it, c);
            fZo[q] = zoff[c];
            fHr[q] = p.H_af_ref[c];
            fLt[q] = p.LT[c];
            fLw[q] = p.interactive ? p.LW[c] : 0.0f;
            fI2e[q] = p.inv2eta[c];
            fM[q] = p.mask[c] ? 1.0f : 0.0f;
        }
#pragma unroll
        for (int q = 0; q < QU; ++q) {
            const int idx = (int)threadIdx.x + q * NT;
            if (idx < totU) su[idx] = uv[q];
        }
        if (DIST) {
#pragma unroll
            for (int q = 0; q < QU; ++q) {
                const int idx = (int)threadIdx.x + q * NT;
                if (idx < totU) sKi[idx] = kv[q];
            }
        }
#pragma unroll
        for (int q = 0; q < QC; ++q) {
            const int idx = (int)threadIdx.x + q * NT;
            if (idx < totC) ca[idx] = zv[q];
        }
#pragma unroll
        for (int q = 0; q < QD; ++q) {
            const int idx = (int)threadIdx.x + q * NT;
            if (idx < totM) sD[idx] = dv[q];
        }
#pragma unroll
        for (int q = 0; q < QF; ++q) {
            const int idx = (int)threadIdx.x + q * NT;
            if (idx < totO) {
                sH[idx] = fH[q];
                sZo[idx] = fZo[q];
                sHr[idx] = fHr[q];
                sLt[idx] = fLt[q];
                sLw[idx] = fLw[q];
                sI2e[idx] = fI2e[q];
                sM[idx] = fM[q];
            }
        }
        __syncthreads();
        FI_SUBMARK(0);
        // previous stage's derivative on rows [ju0, ju1): inverse real transforms of the halo pairs
        if (np > 0) {
            G<float> Kin = p.Kb + (long)st.kin * p.Nl;
            float2* res = fft_lds<float2>(ca, ca + totC, np, nx, p.px, twx, +1);
            FI_SUBMARK(2);
            for (int idx = threadIdx.x; idx < 2 * totC; idx += NT) {
                const int b = fdivi(idx, 0.5f * inv_nx), rem = idx - b * 2 * nx, r = rem >= nx ? 1 : 0,
                          i = rem - r * nx;
                const int j = p0 + 2 * b + r;
                if (j < ju0 || j >= ju1) continue;
                float v = r ? res[b * nx + i].y : res[b * nx + i].x;
                const int ci = corner_index(p, i, j);
                if (ci >= 0) v = ci == 0 ? kc.v0 : ci == 1 ? kc.v1 : ci == 2 ? kc.v2 : kc.v3;
                skin[(j - ju0) * nx + i] = v;
                if (j >= j0 && j < je) {
                    st_shared(addr(Kin, i + (long)j * nx), v);
                    if (ci >= 0) st_shared(addr(p.kcorner, 4 * st.kin + ci), v);
                }
            }
            __syncthreads();
        }
        FI_SUBMARK(3);
        // Row-pair kinds with fixed row offsets (relative to j0: first u row JU, count NU; first
        // moment row JM, count NM; NO output rows; TOP: row 0 is the first u and moment row; BOT:
        // row ny-1 is the last one, and output row LO when LO >= 0). Thread i owns column i: it
        // finishes u on an NU-row x 3-column window in registers, writes its moments to LDS, and
        // after one workgroup barrier evaluates its output rows. The row rules are the CPU edge
        // rules of d2y/dxy: one-sided second difference = central one at the neighbouring row,
        // one-sided first difference at the grid edge.
        auto window = [&](auto JU_, auto NU_, auto JM_, auto NM_, auto NO_, auto TOP_, auto BOT_, auto LO_) {
            constexpr int JU = decltype(JU_)::value, NU = decltype(NU_)::value, JM = decltype(JM_)::value,
                          NM = decltype(NM_)::value, NO = decltype(NO_)::value, LO = decltype(LO_)::value;
            constexpr bool TOP = decltype(TOP_)::value, BOT = decltype(BOT_)::value;
            constexpr bool dist = DIST;
            int i = threadIdx.x;
            asm volatile("" : "+v"(i));   // keep column-derived addresses inside the item loop
            double uo[NO];
            const int a = min(max(i - 1, 0), nx - 3);   // window columns a..a+2 (d2x/d1x edge rules)
            const int ci = i - a;                       // 0 at i = 0, 2 at i = nx-1, else 1
            if (i < nx) {
                double cv[NU], dd[NU], xx[NU];   // centre value, x-first-derivative, x-second-difference
#pragma unroll
                for (int r = 0; r < NU; ++r) {
                    const int b0 = r * nx + a;
                    double x0 = su[b0], x1 = su[b0 + 1], x2 = su[b0 + 2];
                    if (np > 0) {
                        x0 += st.dt * (st.ain * (double)skin[b0]);
                        x1 += st.dt * (st.ain * (double)skin[b0 + 1]);
                        x2 += st.dt * (st.ain * (double)skin[b0 + 2]);
                    }
                    x0 -= uoff;
                    x1 -= uoff;
                    x2 -= uoff;
                    cv[r] = ci == 0 ? x0 : (ci == 1 ? x1 : x2);
                    dd[r] = ci == 1 ? (x2 - x0) * (0.5 * p.inv_dx)
                                    : (ci == 0 ? (x1 - x0) * p.inv_dx : (x2 - x1) * p.inv_dx);
                    xx[r] = (x2 - 2.0 * x1 + x0) * p.inv_dx2;
                    if (dist) {   // Dx = K dx at this row's cell
                        const double k = sKi[r * nx + i];
                        dd[r] *= k;
                        xx[r] *= k * k;
                    }
                }
#pragma unroll
                for (int o = 0; o < NO; ++o) uo[o] = cv[o - JU];
                if (st.write_u >= 0) {
                    G<double> un = ubuf(p, st.write_u);
#pragma unroll
                    for (int o = 0; o < NO; ++o) st_shared(addr(un, i + (long)(j0 + o) * nx), uo[o]);
                }
#pragma unroll
                for (int m = 0; m < NM; ++m) {
                    const int ur = JM + m - JU;   // window row of this moment row
                    const bool first = TOP && m == 0, last = BOT && m == NM - 1;
                    const int c = first ? ur + 1 : (last ? ur - 1 : ur);
                    const int lo = first ? ur : ur - 1, hi = last ? ur : ur + 1;
                    const double D = (double)sD[m * nx + i];
                    const double uxx = xx[ur];
                    double uyy = (cv[c + 1] - 2.0 * cv[c] + cv[c - 1]) * p.inv_dy2;
                    double uxy = (first || last) ? (dd[hi] - dd[lo]) * p.inv_dy
                                                 : (dd[hi] - dd[lo]) * (0.5 * p.inv_dy);
                    if (dist) {   // Dy = K dy at the moment's cell
                        const double k = sKi[ur * nx + i];
                        uyy *= k * k;
                        uxy *= k;
                    }
                    sMxx[m * nx + i] = -D * (nu * uyy + uxx);
                    sMyy[m * nx + i] = -D * (nu * uxx + uyy);
                    sMxy[m * nx + i] = -D * (1.0 - nu) * uxy;
                }
            }
            __syncthreads();
            FI_SUBMARK(5);
            if (i < nx) {
                auto mx = [&](const double* M, int mm, int cc) { return M[mm * nx + a + cc]; };
                // 1/K of window row wr at this column (1 without map distortion)
                auto kw = [&](int wr) { return dist ? (double)sKi[wr * nx + i] : 1.0; };
                auto d1m = [&](int mm) {
                    const double d = ci == 1 ? (mx(sMxy, mm, 2) - mx(sMxy, mm, 0)) * (0.5 * p.inv_dx)
                                             : (ci == 0 ? (mx(sMxy, mm, 1) - mx(sMxy, mm, 0)) * p.inv_dx
                                                        : (mx(sMxy, mm, 2) - mx(sMxy, mm, 1)) * p.inv_dx);
                    return dist ? d * kw(JM + mm - JU) : d;
                };
                float f[2] = {0.0f, 0.0f};
#pragma unroll
                for (int o = 0; o < NO; ++o) {
                    const int mr = o - JM;   // moment row of output row j0 + o
                    const bool first = TOP && o == 0, last = o == LO;
                    const int c = first ? mr + 1 : (last ? mr - 1 : mr);
                    const int lo = first ? mr : mr - 1, hi = last ? mr : mr + 1;
                    double bxx = (mx(sMxx, mr, 2) - 2.0 * mx(sMxx, mr, 1) + mx(sMxx, mr, 0)) * p.inv_dx2;
                    double byy =
                        (sMyy[(c + 1) * nx + i] - 2.0 * sMyy[c * nx + i] + sMyy[(c - 1) * nx + i]) * p.inv_dy2;
                    double bxy = (first || last) ? (d1m(hi) - d1m(lo)) * p.inv_dy
                                                 : (d1m(hi) - d1m(lo)) * (0.5 * p.inv_dy);
                    if (dist) {   // Dx, Dy at the output cell
                        const double k = kw(o - JU);
                        bxx *= k * k;
                        byy *= k * k;
                        bxy *= k;
                    }
                    const int idx = o * nx + i;
                    const long cc = i + (long)(j0 + o) * nx;
                    const double u = uo[o];
                    const double msk = sM[idx];
                    const double Haf = hab((double)sH[idx], (double)sZo[idx] + u, p.rho_sw_ice);
                    const double s_ice = p.rho_ice * (Haf - (double)sHr[idx]);
                    const double s_load = msk * (s_ice + (double)sLw[idx]);
                    const double s_mant = p.rho_um * u;
                    const double s_full = s_load + msk * ((double)sLt[idx] + s_mant);
                    const double rhs = -p.g * s_full + (bxx + (2.0 * bxy + byy));
                    f[o] = (float)(rhs * (double)sI2e[idx]);
                    if (wsig) {
                        st_shared(addr(p.sig_load_prev, cc), (float)s_load);
                        st_shared(addr(p.sig_geo_prev, cc), (float)(p.compressible_column
                                                         ? s_load + msk * (s_mant - p.rho_um * (double)p.u_ref[cc])
                                                         : s_full));
                    }
                }
                sf[i] = make_float2(f[0], f[1]);
            }
        };
        using std::integral_constant;
#define FI_WIN(JU, NU, JM, NM, NO, TOP, BOT, LO)                                                          \
    window(integral_constant<int, JU>{}, integral_constant<int, NU>{}, integral_constant<int, JM>{},      \
           integral_constant<int, NM>{}, integral_constant<int, NO>{}, integral_constant<bool, TOP>{},     \
           integral_constant<bool, BOT>{}, integral_constant<int, LO>{})
        // the kind of this row pair; each is used only where the row ranges computed above are
        // exactly its own (ny >= 6), anything else takes the general path
        const int rj = ju0 - j0, rm = jm0 - j0, nu_ = ju1 - ju0, nm_ = jm1 - jm0, no_ = je - j0;
        const int kind = (ny < 6 || nx < 3) ? -1
                         : (rj == -2 && nu_ == 6 && rm == -1 && nm_ == 4 && no_ == 2 && j0 >= 2 && j0 + 3 <= ny - 1) ? 0
                         : (j0 == 0 && rj == 0 && nu_ == 4 && rm == 0 && nm_ == 3 && no_ == 2)                     ? 1
                         : (j0 == ny - 2 && rj == -2 && nu_ == 4 && rm == -1 && nm_ == 3 && no_ == 2)              ? 2
                         : (j0 == ny - 3 && rj == -2 && nu_ == 5 && rm == -1 && nm_ == 4 && no_ == 2)              ? 3
                         : (j0 == ny - 1 && rj == -3 && nu_ == 4 && rm == -2 && nm_ == 3 && no_ == 1)              ? 4
                                                                                                                  : -1;
        if (kind == 0) {
            FI_WIN(-2, 6, -1, 4, 2, false, false, -1);
        } else if (kind == 1) {
            FI_WIN(0, 4, 0, 3, 2, true, false, -1);
        } else if (kind == 2) {
            FI_WIN(-2, 4, -1, 3, 2, false, true, 1);
        } else if (kind == 3) {
            FI_WIN(-2, 5, -1, 4, 2, false, true, -1);
        } else if (kind == 4) {
            FI_WIN(-3, 4, -2, 3, 1, false, true, 0);
        } else {
            // finish the stage displacement: spectral term, corner BC; keep u_{n+1} when asked
            for (int idx = threadIdx.x; idx < totU; idx += NT) {
                double u = su[idx];
                if (np > 0) u += st.dt * (st.ain * (double)skin[idx]);
                u -= uoff;
                su[idx] = u;
                const int r = fdivi(idx, inv_nx);
                const int j = ju0 + r;
                if (st.write_u >= 0 && j >= j0 && j < je) st_shared(addr(ubuf(p, st.write_u), idx - r * nx + (long)j * nx), u);
            }
            __syncthreads();
            FI_SUBMARK(4);
            // plate moments M = -D (...) on rows [jm0, jm1)
            for (int idx = threadIdx.x; idx < totM; idx += NT) {
                const int r = fdivi(idx, inv_nx), i = idx - r * nx, j = jm0 + r;
                const double D = (double)sD[idx];
                const double uxx = d2x(su + (j - ju0) * nx, i, nx, p.inv_dx2);
                const double uyy = d2y(su, ju0, j, i, nx, ny, p.inv_dy2);
                const double uxy = dxy(su, ju0, j, i, nx, ny, p.inv_dx, p.inv_dy);
                sMxx[idx] = -D * (nu * uyy + uxx);
                sMyy[idx] = -D * (nu * uxx + uyy);
                sMxy[idx] = -D * (1.0 - nu) * uxy;
            }
            __syncthreads();
            FI_SUBMARK(5);
            // rhs = -g sigma_full + dxx Mxx + 2 dxy Mxy + dyy Myy, times K/(2 eta)
            for (int idx = threadIdx.x; idx < totO; idx += NT) {
                const int r = idx >= nx ? 1 : 0, i = idx - r * nx, j = j0 + r;
                float f = 0.0f;
                if (j < ny) {
                    const long c = i + (long)j * nx;
                    const double bxx = d2x(sMxx + (j - jm0) * nx, i, nx, p.inv_dx2);
                    const double byy = d2y(sMyy, jm0, j, i, nx, ny, p.inv_dy2);
                    const double bxy = dxy(sMxy, jm0, j, i, nx, ny, p.inv_dx, p.inv_dy);
                    const double u = su[(j - ju0) * nx + i];
                    const double m = sM[idx];
                    const double Haf = hab((double)sH[idx], (double)sZo[idx] + u, p.rho_sw_ice);
                    const double s_ice = p.rho_ice * (Haf - (double)sHr[idx]);
                    const double s_load = m * (s_ice + (double)sLw[idx]);
                    const double s_mant = p.rho_um * u;
                    const double s_full = s_load + m * ((double)sLt[idx] + s_mant);
                    const double rhs = -p.g * s_full + (bxx + (2.0 * bxy + byy));
                    f = (float)(rhs * (double)sI2e[idx]);
                    if (wsig) {
                        st_shared(addr(p.sig_load_prev, c), (float)s_load);
                        st_shared(addr(p.sig_geo_prev, c), (float)(p.compressible_column
                                                        ? s_load + m * (s_mant - p.rho_um * (double)p.u_ref[c])
                                                        : s_full));
                    }
                }
                if (r == 0) sf[i].x = f;
                else sf[i].y = f;
            }
        }
#undef FI_WIN
        __syncthreads();
        FI_SUBMARK(6);
        float2* res = fft_lds<float2>(sf, sg, 1, nx, p.px, twx, -1);
        for (int kx = threadIdx.x; kx < p.nhx; kx += NT) {
            const float2 Z = res[kx], Zm = res[kx == 0 ? 0 : nx - kx];
            float2 X0, X1;
            X0.x = 0.5f * (Z.x + Zm.x);
            X0.y = 0.5f * (Z.y - Zm.y);
            X1.x = 0.5f * (Z.y + Zm.y);
            X1.y = -0.5f * (Z.x - Zm.x);
            st_shared(addr(spout, sidx(p, kx, j0)), X0);
            if (j0 + 1 < ny) st_shared(addr(spout, sidx(p, kx, j0 + 1)), X1);
        }
        __syncthreads();
        FI_SUBMARK(7);
    }
}

// ------------------------------------------------------------------------------------------
// columns phase: FFT_y, * spy/(nx ny)/(R kappa), inverse FFT_y
// ------------------------------------------------------------------------------------------
__device__ __forceinline__ void phase_cols(const KP& p, const Team& tm, G<float2> spec, const float2* twy, char* smem) {
    const int ny = p.ny, nb = p.colnb;
    float2* sa = (float2*)smem;
    float2* sb = sa + nb * ny;
    for (int item = tm.rank; item < p.ncolitems; item += tm.size) {
        const int kx0 = item * nb;
        const int nbb = min(nb, p.nhx - kx0);
        const long base = (long)kx0 * ny;   // the multiplier table is [kx][ky]
        const int tot = nbb * ny;
        const float inv_ny = 1.0f / (float)ny;
        // spectrum element of column-phase index idx (column kx0 + idx / ny, row idx % ny)
        auto se = [&](int idx) {
            const int b = fdivi(idx, inv_ny);
            return sidx(p, kx0 + b, idx - b * ny);
        };
        if (tot <= 2 * NT) {
            // the multiplier is loaded with the column, in the same round
            float mv[2];
#pragma unroll
            for (int q = 0; q < 2; ++q) {
                const int idx = min((int)threadIdx.x + q * NT, tot - 1);
                sa[idx] = spec[se(idx)];
                mv[q] = p.mult[base + idx];
            }
            __syncthreads();
            float2* r = fft_lds<float2>(sa, sb, nbb, ny, p.py, twy, -1);
#pragma unroll
            for (int q = 0; q < 2; ++q) {
                const int idx = (int)threadIdx.x + q * NT;
                if (idx < tot) {
                    r[idx].x *= mv[q];
                    r[idx].y *= mv[q];
                }
            }
            __syncthreads();
            float2* o = (r == sa) ? sb : sa;
            float2* r2 = fft_lds<float2>(r, o, nbb, ny, p.py, twy, +1);
            for (int idx = threadIdx.x; idx < tot; idx += NT) st_shared(addr(spec, se(idx)), r2[idx]);
            __syncthreads();
            continue;
        }
        for (int idx = threadIdx.x; idx < tot; idx += NT) sa[idx] = spec[se(idx)];
        __syncthreads();
        float2* r = fft_lds<float2>(sa, sb, nbb, ny, p.py, twy, -1);
        for (int idx = threadIdx.x; idx < tot; idx += NT) {
            const float m = p.mult[base + idx];
            r[idx].x *= m;
            r[idx].y *= m;
        }
        __syncthreads();
        float2* o = (r == sa) ? sb : sa;
        float2* r2 = fft_lds<float2>(r, o, nbb, ny, p.py, twy, +1);
        for (int idx = threadIdx.x; idx < nbb * ny; idx += NT) st_shared(addr(spec, se(idx)), r2[idx]);
        __syncthreads();
    }
}

// ------------------------------------------------------------------------------------------
// standalone c2r phase: inverse real transform of row pairs into K[slot]; optional BS3 error
// moments, one partial per row pair in p.part (so the step-size decision does not depend on how
// many workgroups share the work)
// ------------------------------------------------------------------------------------------
__device__ __forceinline__ void phase_c2r(const KP& p, const Team& tm, const StageDesc& st, const G<float2> spec, int slot,
                                          const float2* twx, char* smem, double* red) {
    const int nx = p.nx, ny = p.ny, nbp = p.c2rnb;
    float2* sf = (float2*)smem;
    float2* sg = sf + nbp * nx;
    const float inv_nx = 1.0f / (float)nx;
    G<float> K = p.Kb + (long)slot * p.Nl;
    // an item is c2rnb consecutive row pairs, inverse-transformed as one batch
    for (int item = tm.rank; item < p.nc2ritems; item += tm.size) {
        const int pr0 = item * nbp, npr = min(nbp, p.nrowitems - pr0);
        const bool pre = st.final_err && npr == 1 && 2 * nx <= 2 * NT;
        // error-norm inputs of the pair's cells, loaded with the spectrum (one round)
        float e1v[2] = {0.f, 0.f}, e2v[2] = {0.f, 0.f}, e3v[2] = {0.f, 0.f};
        double unv[2] = {0.0, 0.0}, upv[2] = {0.0, 0.0};
        if (pre) {
            const int j0 = 2 * pr0;
#pragma unroll
            for (int q = 0; q < 2; ++q) {
                const int idx = min((int)threadIdx.x + q * NT, 2 * nx - 1);
                const int r = idx >= nx ? 1 : 0, i = idx - r * nx, j = min(j0 + r, ny - 1);
                const long c = i + (long)j * nx;
                e1v[q] = kslot(p, st.e1)[c];
                e2v[q] = kslot(p, st.e2)[c];
                e3v[q] = kslot(p, st.e3)[c];
                unv[q] = ubuf(p, st.u_prev)[c];
                upv[q] = ubuf(p, st.write_u)[c];
            }
        }
        for (int idx = threadIdx.x; idx < npr * nx; idx += NT) {
            const int b = fdivi(idx, inv_nx), k = idx - b * nx, j0 = 2 * (pr0 + b);
            sf[idx] = c2r_pair_value(p, spec, k, j0, j0 + 1);
        }
        __syncthreads();
        float2* resb = fft_lds<float2>(sf, sg, npr, nx, p.px, twx, +1);
        for (int pb = 0; pb < npr; ++pb) {
            const int j0 = 2 * (pr0 + pb), pair = pr0 + pb;
            const float2* res = resb + pb * nx;
            double accv[NV][3] = {};
            for (int i0 = threadIdx.x, q0 = 0; i0 < 2 * nx; i0 += VNT, q0 += NV)
    #pragma unroll
            for (int vv = 0; vv < NV; ++vv) {
                const int idx = i0 + vv * NT, q = q0 + vv;
                if (idx >= 2 * nx) break;
                double* acc = accv[vv];
                const int r = idx >= nx ? 1 : 0, i = idx - r * nx, j = j0 + r;
                if (j >= ny) continue;
                const long c = i + (long)j * nx;
                const float v = r ? res[i].y : res[i].x;
                st_shared(addr(K, c), v);
                const int ci = corner_index(p, i, j);
                if (ci >= 0) st_shared(addr(p.kcorner, 4 * slot + ci), v);
                if (st.final_err) {
                    double k1, k2, k3, un, up;
                    if (pre) {
                        const int qq = q & 1;   // q <= 1 when pre
                        k1 = qq ? e1v[1] : e1v[0];
                        k2 = qq ? e2v[1] : e2v[0];
                        k3 = qq ? e3v[1] : e3v[0];
                        un = qq ? unv[1] : unv[0];
                        up = qq ? upv[1] : upv[0];
                    } else {
                        k1 = kslot(p, st.e1)[c];
                        k2 = kslot(p, st.e2)[c];
                        k3 = kslot(p, st.e3)[c];
                        un = ubuf(p, st.u_prev)[c];
                        up = ubuf(p, st.write_u)[c];
                    }
                    const double a = st.dt * (BS3_BT1 * k1 + BS3_BT2 * k2 + BS3_BT3 * k3 + BS3_BT4 * (double)v);
                    const double w = 1.0 / (p.abstol + fmax(fabs(un), fabs(up)) * p.reltol);
                    acc[0] += (a * w) * (a * w);
                    acc[1] += a * w * w;
                    acc[2] += w * w;
                }
            }
            if (st.final_err) {
                block_sum<3>(red, accv);
                if (threadIdx.x == 0) {
                    st_shared(addr(p.part, 4 * pair + 0), accv[0][0]);
                    st_shared(addr(p.part, 4 * pair + 1), accv[0][1]);
                    st_shared(addr(p.part, 4 * pair + 2), accv[0][2]);
                }
            }
        }
        __syncthreads();
    }
}

// ------------------------------------------------------------------------------------------
// sparse update phases (src/simulation.jl:359-381)
// ------------------------------------------------------------------------------------------
__device__ __forceinline__ void stage_twiddles(const KP& p, const FftPlan& pl, float2* dst) {
    for (int k = threadIdx.x; k < pl.L; k += NT) dst[k] = p.tw32[pl.tw + k];
    __syncthreads();
}

// padded rows: the two convolution inputs of row j packed as one complex sequence
__device__ __forceinline__ void phase_sp_rows(const KP& p, const Team& tm, char* smem) {
    const int nx = p.nx, ny = p.ny, nfx = p.nfx;
    float2* tw = (float2*)smem;
    float2* sf = tw + nfx;
    float2* sg = sf + nfx;
    stage_twiddles(p, p.pfx, tw);
    for (int j = tm.rank; j < ny; j += tm.size) {
        for (int i = threadIdx.x; i < nfx; i += NT) {
            float2 z;
            z.x = 0;
            z.y = 0;
            if (i < nx) {
                const long c = i + (long)j * nx;
                // sigma_load K^2 (src/deformation.jl:177); mass anomaly / (dx dy) = sigma K^2
                // (src/sealevel.jl:193-201; the constant dx dy is folded in the Green spectrum)
                const float k2 = p.K2 ? p.K2[c] : 1.0f;
                if (!p.rigid_litho) z.x = p.sig_load_prev[c] * k2;
                if (p.lv_surface) z.y = p.sig_geo_prev[c] * k2;
            }
            sf[i] = z;
        }
        __syncthreads();
        float2* res = fft_lds<float2>(sf, sg, 1, nfx, p.pfx, tw, -1);
        for (int kx = threadIdx.x; kx < p.nfhx; kx += NT) {
            const float2 Z = res[kx], Zm = res[kx == 0 ? 0 : nfx - kx];
            float2 A, B;
            A.x = 0.5f * (Z.x + Zm.x);
            A.y = 0.5f * (Z.y - Zm.y);
            B.x = 0.5f * (Z.y + Zm.y);
            B.y = -0.5f * (Z.x - Zm.x);
            st_shared(addr(p.pA, bidx(kx, j, ny)), A);
            st_shared(addr(p.pB, bidx(kx, j, ny)), B);
        }
        __syncthreads();
    }
}

__device__ __forceinline__ int sp_row_of(const KP& p, int jj) {
    if (jj == 0) return 0;
    if (jj == p.ny + 1) return 2 * p.ny - 2;
    return (p.j1 - 1) + (jj - 1);
}

// padded columns: FFT_y of both half spectra, multiply by the FP64 Green spectra, inverse FFT_y,
// keep the rows the crop and the extended-space corner BC need
__device__ __forceinline__ void phase_sp_cols(const KP& p, const Team& tm, char* smem) {
    const int ny = p.ny, nfy = p.nfy, no = ny + 2;
    float2* tw = (float2*)smem;
    float2* sa = tw + nfy;
    float2* sb = sa + 2 * nfy;
    stage_twiddles(p, p.pfy, tw);
    for (int kx = tm.rank; kx < p.nfhx; kx += tm.size) {
        for (int idx = threadIdx.x; idx < 2 * nfy; idx += NT) {
            const int s = idx / nfy, j = idx - s * nfy;
            float2 v;
            v.x = 0;
            v.y = 0;
            if (j < ny) v = (s ? p.pB : p.pA)[bidx(kx, j, ny)];
            sa[idx] = v;
        }
        __syncthreads();
        float2* r = fft_lds<float2>(sa, sb, 2, nfy, p.pfy, tw, -1);
        for (int idx = threadIdx.x; idx < 2 * nfy; idx += NT) {
            const int s = idx / nfy, ky = idx - s * nfy;
            const double2 G = (s ? p.gg_hat : p.ge_hat)[(long)kx * nfy + ky];
            const double xr = r[idx].x, xi = r[idx].y;
            r[idx].x = (float)(xr * G.x - xi * G.y);
            r[idx].y = (float)(xr * G.y + xi * G.x);
        }
        __syncthreads();
        float2* o = (r == sa) ? sb : sa;
        float2* r2 = fft_lds<float2>(r, o, 2, nfy, p.pfy, tw, +1);
        for (int idx = threadIdx.x; idx < 2 * no; idx += NT) {
            const int s = idx / no, jj = idx - s * no;
            st_shared(addr((s ? p.oB : p.oA), bidx(kx, jj, no)), r2[s * nfy + sp_row_of(p, jj)]);
        }
        __syncthreads();
    }
}

// needed padded rows: inverse real transforms, crop [i1:i2] and the extended corners
__device__ __forceinline__ void phase_sp_c2r(const KP& p, const Team& tm, char* smem) {
    const int nx = p.nx, nfx = p.nfx, nfhx = p.nfhx, no = p.ny + 2;
    const bool even = (nfx % 2) == 0;
    float2* tw = (float2*)smem;
    float2* sf = tw + nfx;
    float2* sg = sf + nfx;
    stage_twiddles(p, p.pfx, tw);
    for (int jj = tm.rank; jj < no; jj += tm.size) {
        for (int k = threadIdx.x; k < nfx; k += NT) {
            float2 A, B;
            if (k < nfhx) {
                A = p.oA[bidx(k, jj, no)];
                B = p.oB[bidx(k, jj, no)];
                if (k == 0 || (even && k == nfx / 2)) { A.y = 0; B.y = 0; }
            } else {
                const int kk = nfx - k;
                A = p.oA[bidx(kk, jj, no)];
                B = p.oB[bidx(kk, jj, no)];
                A.y = -A.y;
                B.y = -B.y;
            }
            float2 Z;
            Z.x = A.x - B.y;
            Z.y = A.y + B.x;
            sf[k] = Z;
        }
        __syncthreads();
        float2* res = fft_lds<float2>(sf, sg, 1, nfx, p.pfx, tw, +1);
        if (jj == 0 || jj == no - 1) {
            if (threadIdx.x == 0) {
                const int q = (jj == 0) ? 0 : 2;
                st_shared(addr(p.ext_corners, q), res[0].x);
                st_shared(addr(p.ext_corners, q + 1), res[2 * nx - 2].x);
                st_shared(addr(p.ext_corners, 4 + q), res[0].y);
                st_shared(addr(p.ext_corners, 4 + q + 1), res[2 * nx - 2].y);
            }
        } else {
            const int j = jj - 1;
            for (int ii = threadIdx.x; ii < nx; ii += NT) {
                const float2 v = res[p.i1 - 1 + ii];
                st_shared(addr(p.ue_raw, ii + (long)j * nx), v.x);
                st_shared(addr(p.dz_raw, ii + (long)j * nx), v.y);
            }
        }
        __syncthreads();
    }
}

// volume contributions of the state before a sparse update, for update_delta_V!
// (src/sealevel.jl:254-331): H_af with the current bedrock and the old sea surface, H_ice, and the
// ocean column max(z_ss - z_b, 0), each times the cell area; one partial per row pair in p.vpart
__device__ __forceinline__ void phase_sp_volumes(const KP& p, const Team& tm, const StageDesc& st, double uoff,
                                                 const IceT it, double* red) {
    const int nx = p.nx;
    for (int item = tm.rank; item < p.nrowitems; item += tm.size) {
        double accv[NV][3] = {};
        const int ncell = (min(2 * item + 2, p.ny) - 2 * item) * nx;
        for (int q0 = threadIdx.x; q0 < ncell; q0 += VNT)
#pragma unroll
        for (int vv = 0; vv < NV; ++vv) {
            const int q = q0 + vv * NT;
            if (q >= ncell) break;
            double* acc = accv[vv];
            const long c = (long)2 * item * nx + q;
            const double u = u_raw(p, st, c) - uoff;
            const double H = ice_at(p, it, c);
            const double zss_old = p.z_ss[c];
            const double zb_old = (double)p.z_b_ref[c] + (double)p.ue[c] + u;
            const double a = p.K2 ? p.A * (double)p.K2[c] : p.A;   // cell area dx dy K^2
            acc[0] += hab(H, zb_old - zss_old, p.rho_sw_ice) * a;
            acc[1] += H * a;
            acc[2] += fmax(zss_old - zb_old, 0.0) * a;
        }
        block_sum<3>(red, accv);
        if (threadIdx.x == 0) {
            st_shared(addr(p.vpart, 4 * item + 0), accv[0][0]);
            st_shared(addr(p.vpart, 4 * item + 1), accv[0][1]);
            st_shared(addr(p.vpart, 4 * item + 2), accv[0][2]);
        }
    }
}

// TimeInterpolation0D (src/interpolations.jl:28-49): linear between ascending knots, the knot
// value on an exact hit; outside the knots the end value, flagged in *err unless flat
__device__ __forceinline__ double interp0d(const double* x, const double* y, int n, int flat, double t, int* err) {
    if (t < x[0]) {
        if (!flat) *err = 1;
        return y[0];
    }
    if (t > x[n - 1]) {
        if (!flat) *err = 1;
        return y[n - 1];
    }
    int lo = 0, hi = n - 1;   // first knot with x >= t (searchsortedfirst)
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (x[mid] >= t) hi = mid;
        else lo = mid + 1;
    }
    if (x[lo] == t) return y[lo];
    const int i = lo - 1;
    return y[i] + (y[i + 1] - y[i]) / (x[i + 1] - x[i]) * (t - x[i]);
}

// Sum the volume partials (fixed order, identical in every workgroup), update V_af, V_den, V_pov
// and delta_V (update_delta_V!, src/sealevel.jl:254-262) and, with an internal update, the
// barystatic sea level and ocean surface (update_bsl!(bsl, -delta_V, t),
// src/barystatic_sealevel.jl:215-234); count the sparse update.
__device__ __forceinline__ void volumes_and_bsl(const KP& p, Ctl& s, double t, double* red) {
    double vsv[NV][3] = {};
    for (int w0 = threadIdx.x; w0 < p.nrowitems; w0 += VNT)
#pragma unroll
    for (int vv = 0; vv < NV; ++vv) {
        const int w = w0 + vv * NT;
        if (w >= p.nrowitems) break;
        vsv[vv][0] += p.vpart[4 * w];
        vsv[vv][1] += p.vpart[4 * w + 1];
        vsv[vv][2] += p.vpart[4 * w + 2];
    }
    block_sum<3>(red, vsv);
    const double* vs = vsv[0];
    if (threadIdx.x == 0) {
        const double Vold = s.V_af + s.V_den + s.V_pov;
        s.V_af = p.vol_goelzer ? vs[0] * p.rho_ice_sw : 0.0;
        s.V_den = p.den_goelzer ? vs[1] * (p.rho_ice / p.rho_water - p.rho_ice / p.rho_seawater) : 0.0;
        s.V_pov = p.pov_goelzer ? vs[2] : 0.0;
        {
            s.delta_V = (s.V_af + s.V_den + s.V_pov) - Vold;
            if (p.bsl_internal) {
                const double dV = -s.delta_V;
                int e = 0;
                if (p.bsl_rule == 1) {          // ConstantOceanSurfaceBSL
                    s.A_bsl = p.bsl_A_ref;
                    s.z_bsl += dV / s.A_bsl;
                } else if (p.bsl_rule == 2) {   // PiecewiseConstantBSL
                    s.z_bsl += dV / s.A_bsl;
                    s.A_bsl = interp0d(p.bsl_x, p.bsl_y, p.bsl_n, p.bsl_flat, s.z_bsl, &e);
                } else if (p.bsl_rule == 3) {   // ImposedBSL
                    s.z_bsl = interp0d(p.bsl_x, p.bsl_y, p.bsl_n, p.bsl_flat, t, &e);
                } else {                        // ConstantBSL
                    s.z_bsl = p.bsl_z;
                }
                if (e) s.err = ERR_BSL_RANGE;
            }
            // update_dz_ss! of an ImposedSeaSurface runs after the BSL update and shares its
            // out-of-range error (src/sealevel.jl:188-191)
            if (p.imposed_ss && !p.ss_flat && !s.err && (t < p.ss_t[0] || t > p.ss_t[p.ss_nt - 1]))
                s.err = ERR_TIME_RANGE;
            s.count += 1;
        }
    }
    __syncthreads();
}

// TimeInterpolation2D bracket of the ImposedSeaSurface snapshots at t (src/interpolations.jl:
// 51-74): dz_ss = X[i0] + (X[i1] - X[i0]) / dtk * tt, or X[i0] when exact. Every thread runs the
// same binary search (uniform loads); out of range clamps to the end snapshot (the range error
// is raised in volumes_and_bsl).
struct SsBracket {
    int i0, i1, exact;
    double tt, dtk;
};
__device__ __forceinline__ SsBracket ss_bracket(const KP& p, double t) {
    SsBracket b{0, 0, 1, 0.0, 1.0};
    const int nt = p.ss_nt;
    if (t < p.ss_t[0]) return b;
    if (t > p.ss_t[nt - 1]) {
        b.i0 = b.i1 = nt - 1;
        return b;
    }
    int lo = 0, hi = nt - 1;   // first k with ss_t[k] >= t (exists: t <= ss_t[nt-1])
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (p.ss_t[mid] >= t) hi = mid; else lo = mid + 1;
    }
    const double tk = p.ss_t[lo];
    if (tk == t) {
        b.i0 = b.i1 = lo;
        return b;
    }
    const double tp = p.ss_t[lo - 1];
    b.i0 = lo - 1;
    b.i1 = lo;
    b.exact = 0;
    b.tt = t - tp;
    b.dtk = tk - tp;
    return b;
}

// pointwise sparse diagnostics with the updated barystatic sea level z_bsl. Row pairs go to the
// workgroups as in the relaxed-mantle rows phase (items of p.erb consecutive pairs; one pair per
// item elsewhere), which reads these fields after a workgroup barrier only.
__device__ __forceinline__ void phase_sp_point(const KP& p, const Team& tm, const StageDesc& st, double uoff,
                                               const IceT it, double z_bsl) {
    const G<float> ec = p.ext_corners;
    SsBracket sb{0, 0, 1, 0.0, 1.0};
    if (p.imposed_ss) sb = ss_bracket(p, st.t);
    const double oe = ((0.25 * ec[0] + 0.25 * ec[1]) + (0.25 * ec[2] + 0.25 * ec[3])) - p.ebc_x;
    const double og = ((0.25 * ec[4] + 0.25 * ec[5]) + (0.25 * ec[6] + 0.25 * ec[7])) - p.sbc_x;
    const int nx = p.nx;
    for (int item = tm.rank; item < p.nerowitems; item += tm.size) {
      const int pr0 = item * p.erb, pr1 = min(pr0 + p.erb, p.nrowitems);
      const int ncell = (min(2 * pr1, p.ny) - 2 * pr0) * nx;
      for (int q = threadIdx.x; q < ncell; q += NT) {
        const long c = (long)2 * pr0 * nx + q;
        const double u = u_raw(p, st, c) - uoff;
        const double H = ice_at(p, it, c);
        const double ue_old = p.ue[c];
        const double zb_old = (double)p.z_b_ref[c] + ue_old + u;
        const double ue_new = p.rigid_litho ? 0.0 : (double)p.ue_raw[c] - oe;
        double dz_new;
        if (p.lv_surface) {
            dz_new = (double)p.dz_raw[c] - og;
        } else if (p.imposed_ss) {
            const double x0 = p.ss_X[sb.i0 * p.Nl + c];
            dz_new = sb.exact ? x0 : x0 + (p.ss_X[sb.i1 * p.Nl + c] - x0) / sb.dtk * sb.tt;
        } else {
            dz_new = (double)p.dz_ss[c];
        }
        const double zss_new = (double)p.z_ss_ref[c] + dz_new + z_bsl;
        const double haf = hab(H, zb_old - zss_new, p.rho_sw_ice);
        const bool mgc = haf > 0;
        const bool moc = (zss_new - zb_old > 0) && !mgc;
        const double buf = fmax(zss_new - zb_old, 0.0);
        const double hw = (H <= 1 ? buf : 0.0) + ((!mgc && H > 1) ? (buf - H * p.rho_ice_sw) : 0.0);
        const double m = p.mask[c] ? 1.0 : 0.0;
        const double sw = p.interactive ? p.rho_seawater * (zss_new - (double)p.z_ss_ref[c]) * (moc ? 1.0 : 0.0) * m : 0.0;
        const double ca_litho = p.rho_litho * (ue_new - (double)p.ue_ref[c]);
        st_shared(addr(p.ue, c), (float)ue_new);
        st_shared(addr(p.dz_ss, c), (float)dz_new);
        st_shared(addr(p.z_ss, c), (float)zss_new);
        st_shared(addr(p.mg, c), (unsigned char)mgc);
        st_shared(addr(p.mo, c), (unsigned char)moc);
        st_shared(addr(p.H_water, c), (float)hw);
        st_shared(addr(p.LW, c), (float)sw);
        st_shared(addr(p.LT, c), (float)(ca_litho - p.rho_um * (double)p.u_ref[c]));
        st_shared(addr(p.zoff, c), (float)((double)p.z_b_ref[c] + ue_new - zss_new));
        st_shared(addr(p.zoff_trig, c), (float)((double)p.z_b_ref[c] + ue_old - zss_new));
      }
    }
}

// the sparse update of an RHS call (src/simulation.jl:355-381): elastic and sea-surface
// convolutions, volumes and barystatic sea level, pointwise diagnostics; false when the call stops.
// With SKIP_EMPTY a rigid lithosphere under a laterally constant sea surface (both convolution
// inputs unused) skips the convolution phases. With SAME_CELL_NEXT the phase that follows reads
// the pointwise diagnostics only at the cells of this workgroup's own row pairs (the ELRA and
// implicit-step rows phases): the last grid barrier becomes a workgroup barrier after this
// workgroup's stores have completed.
template <bool SKIP_EMPTY, bool SAME_CELL_NEXT = false>
__device__ __forceinline__ bool sparse_phases(const KP& p, const Team& tm, Ctl& s, const StageDesc& st, unsigned& gen,
                                              char* smem, double* red, int* flag, unsigned long long& tl,
                                              double uoff, const IceT it) {
    const bool conv = !SKIP_EMPTY || !p.rigid_litho || p.lv_surface;
    if (conv) {
        phase_sp_rows(p, tm, smem);
        if (!grid_sync(p, tm, gen, flag, TAG_SPARSE, tl)) return false;
        phase_sp_cols(p, tm, smem);
        if (!grid_sync(p, tm, gen, flag, TAG_SPARSE, tl)) return false;
        phase_sp_c2r(p, tm, smem);
    }
    // the volume partials read only the state before the update: they follow the last
    // convolution phase without a barrier of their own (the BSL update needs all of them)
    phase_sp_volumes(p, tm, st, uoff, it, red);
    if (!grid_sync(p, tm, gen, flag, TAG_SPARSE, tl)) return false;
    volumes_and_bsl(p, s, st.t, red);
    if (s.err) return false;   // every workgroup sees the same error
    phase_sp_point(p, tm, st, uoff, it, s.z_bsl);
    if constexpr (SAME_CELL_NEXT) {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
        __builtin_amdgcn_s_waitcnt(0);
        __syncthreads();
        return true;
    }
    return grid_sync(p, tm, gen, flag, TAG_SPARSE, tl);
}

// ------------------------------------------------------------------------------------------
// relaxed mantle (ELRA, src/deformation.jl:29-45): u_eq = samesize_conv(-(sigma_load +
// sigma_litho) g K^2, green_viscous) in the regular BC space (crop to the domain, then the corner
// BC; src/convolutions.jl:145-154) and dudt = (u_eq - u) / tau, on the zero-padded nfx x nfy grid:
//   elra_rows : stage displacement (BS3 combination, corner BC) and loads of a row pair, the two
//               load rows packed as one complex sequence, length-nfx FFT, half spectra stored
//               transposed [kx][j] in pA;
//   elra_cols : length-nfy FFTs of two spectral columns, multiplication by the FP64 Green
//               spectrum, inverse FFTs; the ny rows of the crop [j1:j2] are kept in oA [kx][j];
//   elra_c2r  : the corner-BC offset of the cropped u_eq by four direct sums over the spectral
//               rows 0 and ny-1 (each workgroup computes it), inverse real transforms of row
//               pairs, crop [i1:i2], u_eq and dudt into the K slot (no BC on dudt, so no corner
//               offset is left for consumers); BS3 error moments on the last stage.
// update_deformation_rhs!, which the reference calls first in this method, only fills scratch
// buffers that the method does not read afterwards.
// ------------------------------------------------------------------------------------------
// An item of elra_rows and elra_c2r is p.erb consecutive row pairs, transformed as one batch (an
// ensemble sets erb so that a team covers a member's rows in one round); an item of elra_cols is
// p.ecb columns. Partial sums stay per row pair, so results do not depend on erb or ecb.
__device__ __forceinline__ void phase_elra_rows(const KP& p, const Team& tm, const StageDesc& st, double uoff,
                                                const IceT it, char* smem) {
    const int nx = p.nx, ny = p.ny, nfx = p.nfx, erb = p.erb;
    const float inv_nfx = 1.0f / (float)nfx, inv_nfhx = 1.0f / (float)p.nfhx;
    float2* tw = (float2*)smem;
    float2* sf = tw + nfx;
    float2* sg = sf + erb * nfx;
    stage_twiddles(p, p.pfx, tw);
    const bool wsig = st.write_sigma != 0;
    for (int item = tm.rank; item < p.nerowitems; item += tm.size) {
        const int pr0 = item * erb, npair = min(erb, p.nrowitems - pr0);
        for (int idx = threadIdx.x; idx < npair * nfx; idx += NT) {
            const int b = fdivi(idx, inv_nfx), i = idx - b * nfx;
            const int j0 = 2 * (pr0 + b);
            float2 z = make_float2(0.f, 0.f);
            if (i < nx) {
#pragma unroll
                for (int r = 0; r < 2; ++r) {
                    const int j = j0 + r;
                    if (j >= ny) continue;
                    const long c = i + (long)j * nx;
                    const double u = u_raw(p, st, c) - uoff;
                    if (st.write_u >= 0) st_shared(addr(ubuf(p, st.write_u), c), u);
                    const double m = p.mask[c] ? 1.0 : 0.0;
                    const double zo = st.trig ? p.zoff_trig[c] : p.zoff[c];
                    const double Haf = hab(ice_at(p, it, c), zo + u, p.rho_sw_ice);
                    const double s_ice = p.rho_ice * (Haf - (double)p.H_af_ref[c]);
                    const double s_load = m * (s_ice + (p.interactive ? (double)p.LW[c] : 0.0));
                    const double lt = p.LT[c];   // sigma_litho - rho_um u_ref
                    const double uref = p.u_ref[c];
                    const double k2 = p.K2 ? (double)p.K2[c] : 1.0;
                    const float x = (float)(-(s_load + (lt + p.rho_um * uref)) * p.g * k2);
                    if (r == 0) z.x = x;
                    else z.y = x;
                    if (wsig) {
                        const double s_mant = p.rho_um * u;
                        const double s_full = s_load + m * (lt + s_mant);
                        st_shared(addr(p.sig_load_prev, c), (float)s_load);
                        st_shared(addr(p.sig_geo_prev, c), (float)(p.compressible_column
                                                        ? s_load + m * (s_mant - p.rho_um * uref)
                                                        : s_full));
                    }
                }
            }
            sf[idx] = z;
        }
        __syncthreads();
        float2* res = fft_lds<float2>(sf, sg, npair, nfx, p.pfx, tw, -1);
        for (int idx = threadIdx.x; idx < npair * p.nfhx; idx += NT) {
            const int b = fdivi(idx, inv_nfhx), kx = idx - b * p.nfhx;
            const int j0 = 2 * (pr0 + b);
            const float2* rb = res + b * nfx;
            const float2 Z = rb[kx], Zm = rb[kx == 0 ? 0 : nfx - kx];
            float2 A, B;
            A.x = 0.5f * (Z.x + Zm.x);
            A.y = 0.5f * (Z.y - Zm.y);
            B.x = 0.5f * (Z.y + Zm.y);
            B.y = -0.5f * (Z.x - Zm.x);
            st_shared(addr(p.pA, bidx(kx, j0, ny)), A);
            if (j0 + 1 < ny) st_shared(addr(p.pA, bidx(kx, j0 + 1, ny)), B);
        }
        __syncthreads();
    }
}

__device__ __forceinline__ void phase_elra_cols(const KP& p, const Team& tm, char* smem) {
    const int ny = p.ny, nfy = p.nfy, nfhx = p.nfhx, ecb = p.ecb;
    const float inv_nfy = 1.0f / (float)nfy, inv_ny = 1.0f / (float)ny;
    float2* tw = (float2*)smem;
    float2* sa = tw + nfy;
    float2* sb = sa + ecb * nfy;
    stage_twiddles(p, p.pfy, tw);
    for (int item = tm.rank; item < p.necolitems; item += tm.size) {
        const int kx0 = ecb * item, nb = min(ecb, nfhx - kx0);
        for (int idx = threadIdx.x; idx < nb * nfy; idx += NT) {
            const int s = fdivi(idx, inv_nfy), j = idx - s * nfy;
            float2 v = make_float2(0.f, 0.f);
            if (j < ny) v = p.pA[bidx(kx0 + s, j, ny)];
            sa[idx] = v;
        }
        __syncthreads();
        float2* r = fft_lds<float2>(sa, sb, nb, nfy, p.pfy, tw, -1);
        for (int idx = threadIdx.x; idx < nb * nfy; idx += NT) {
            const int s = fdivi(idx, inv_nfy), ky = idx - s * nfy;
            const double2 G = p.gv_hat[(long)(kx0 + s) * nfy + ky];
            const double xr = r[idx].x, xi = r[idx].y;
            r[idx].x = (float)(xr * G.x - xi * G.y);
            r[idx].y = (float)(xr * G.y + xi * G.x);
        }
        __syncthreads();
        float2* o = (r == sa) ? sb : sa;
        float2* r2 = fft_lds<float2>(r, o, nb, nfy, p.pfy, tw, +1);
        for (int idx = threadIdx.x; idx < nb * ny; idx += NT) {
            const int s = fdivi(idx, inv_ny), j = idx - s * ny;
            st_shared(addr(p.oA, bidx(kx0 + s, j, ny)), r2[s * nfy + (p.j1 - 1) + j]);
        }
        __syncthreads();
    }
}

__device__ __forceinline__ void phase_elra_c2r(const KP& p, const Team& tm, const StageDesc& st, double uoff,
                                               char* smem, double* red) {
    const int nx = p.nx, ny = p.ny, nfx = p.nfx, nfhx = p.nfhx, erb = p.erb;
    const float inv_nfx = 1.0f / (float)nfx;
    const bool even = (nfx % 2) == 0;
    float2* tw = (float2*)smem;
    float2* sf = tw + nfx;
    float2* sg = sf + erb * nfx;
    stage_twiddles(p, p.pfx, tw);
    // corners of the cropped u_eq before its BC: x(m, j) = sum_k w_k Re(X_k(j) e^{2 pi i k m / nfx})
    // at m = i1-1 and i1-1+nx-1 of the kept rows j = 0 and ny-1 (w = 1 for the DC and Nyquist
    // terms, whose imaginary parts the real transform ignores; 2 otherwise)
    const int m0 = p.i1 - 1, m1 = p.i1 - 1 + nx - 1;
    double vs[NV][4] = {};
    for (int k0 = threadIdx.x; k0 < nfhx; k0 += VNT)
#pragma unroll
    for (int vv = 0; vv < NV; ++vv) {
        const int k = k0 + vv * NT;
        if (k >= nfhx) break;
        double* v = vs[vv];
        float2 a = p.oA[bidx(k, 0, ny)], b = p.oA[bidx(k, ny - 1, ny)];
        const bool dcn = k == 0 || (even && k == nfx / 2);
        if (dcn) {
            a.y = 0.f;
            b.y = 0.f;
        }
        const double w = dcn ? 1.0 : 2.0;
        double s0, c0, s1, c1;
        sincospi(2.0 * (double)(((long)k * m0) % nfx) / (double)nfx, &s0, &c0);
        sincospi(2.0 * (double)(((long)k * m1) % nfx) / (double)nfx, &s1, &c1);
        v[0] += w * ((double)a.x * c0 - (double)a.y * s0);
        v[1] += w * ((double)a.x * c1 - (double)a.y * s1);
        v[2] += w * ((double)b.x * c0 - (double)b.y * s0);
        v[3] += w * ((double)b.x * c1 - (double)b.y * s1);
    }
    block_sum<4>(red, vs);
    const double* v = vs[0];
    const double off = ((0.25 * v[0] + 0.25 * v[1]) + (0.25 * v[2] + 0.25 * v[3])) - p.vbc_x;
    G<float> K = p.Kb + (long)st.kout * p.Nl;
    for (int item = tm.rank; item < p.nerowitems; item += tm.size) {
        const int pr0 = item * erb, npair = min(erb, p.nrowitems - pr0);
        for (int idx = threadIdx.x; idx < npair * nfx; idx += NT) {
            const int b = fdivi(idx, inv_nfx), k = idx - b * nfx;
            const int j0 = 2 * (pr0 + b);
            const bool two = j0 + 1 < ny;
            const bool cj = k >= nfhx;
            const int kk = cj ? nfx - k : k;
            float2 A = p.oA[bidx(kk, j0, ny)];
            float2 B = two ? p.oA[bidx(kk, j0 + 1, ny)] : make_float2(0.f, 0.f);
            if (k == 0 || (even && k == nfx / 2)) {
                A.y = 0.f;
                B.y = 0.f;
            }
            if (cj) {
                A.y = -A.y;
                B.y = -B.y;
            }
            sf[idx] = make_float2(A.x - B.y, A.y + B.x);
        }
        __syncthreads();
        float2* resb = fft_lds<float2>(sf, sg, npair, nfx, p.pfx, tw, +1);
        for (int b = 0; b < npair; ++b) {
            const int pair = pr0 + b, j0 = 2 * pair;
            const float2* res = resb + b * nfx;
            double accv[NV][3] = {};
            for (int i0 = threadIdx.x; i0 < 2 * nx; i0 += VNT)
    #pragma unroll
            for (int vv = 0; vv < NV; ++vv) {
                const int idx = i0 + vv * NT;
                if (idx >= 2 * nx) break;
                double* acc = accv[vv];
                const int r = idx >= nx ? 1 : 0, i = idx - r * nx, j = j0 + r;
                if (j >= ny) continue;
                const long c = i + (long)j * nx;
                const float2 z = res[m0 + i];
                const double ueq = (double)(r ? z.y : z.x) - off;
                // the stage displacement: stored by the rows phase when this call keeps it
                const double u = st.write_u >= 0 ? ubuf(p, st.write_u)[c] : u_raw(p, st, c) - uoff;
                const float dv = (float)((ueq - u) * (double)p.inv_tau[c]);
                st_shared(addr(K, c), dv);
                st_shared(addr(p.ueq, c), (float)ueq);
                if (st.final_err) {
                    const double k1 = kslot(p, st.e1)[c], k2 = kslot(p, st.e2)[c], k3 = kslot(p, st.e3)[c];
                    const double un = ubuf(p, st.u_prev)[c];
                    const double a = st.dt * (BS3_BT1 * k1 + BS3_BT2 * k2 + BS3_BT3 * k3 + BS3_BT4 * (double)dv);
                    const double wgt = 1.0 / (p.abstol + fmax(fabs(un), fabs(u)) * p.reltol);
                    acc[0] += (a * wgt) * (a * wgt);
                    acc[1] += a * wgt * wgt;
                    acc[2] += wgt * wgt;
                }
            }
            if (st.final_err) {
                block_sum<3>(red, accv);
                if (threadIdx.x == 0) {
                    st_shared(addr(p.part, 4 * pair + 0), accv[0][0]);
                    st_shared(addr(p.part, 4 * pair + 1), accv[0][1]);
                    st_shared(addr(p.part, 4 * pair + 2), accv[0][2]);
                }
            }
        }
        __syncthreads();
    }
}

// one ELRA RHS call: preamble, the sparse update when due, then the three convolution phases
__device__ __forceinline__ bool elra_call(const KP& p, const Team& tm, Ctl& s, StageDesc& st, unsigned& gen,
                                          const float2* twx, char* smem, double* red, int* flag, bool sparse_only,
                                          unsigned long long& tl) {
    KC4 kc;
    const bool trig = sparse_only || (((st.t - s.t0) / p.dt_sparse) >= (double)s.count);
    st.trig = trig ? 1 : 0;
    st.kin = -1;   // every derivative term is held in global memory
    IceT it;
    double uoff;
    stage_preamble(p, st, specbuf(p, 0), twx, red, it, kc, uoff);
    if (it.err) {
        if (threadIdx.x == 0) s.err = ERR_TIME_RANGE;
        __syncthreads();
        return false;
    }
    if (trig) {
        if (!sparse_phases<true, true>(p, tm, s, st, gen, smem, red, flag, tl, uoff, it)) return false;
        if (sparse_only) return true;
    }
    st.write_sigma = (st.force_sigma || p.write_sigma_always ||
                      (((st.t_next_max - s.t0) / p.dt_sparse) >= (double)s.count)) ? 1 : 0;
    phase_elra_rows(p, tm, st, uoff, it, smem);
    if (!grid_sync(p, tm, gen, flag, TAG_ROWS, tl)) return false;
    phase_elra_cols(p, tm, smem);
    if (!grid_sync(p, tm, gen, flag, TAG_COLS, tl)) return false;
    phase_elra_c2r(p, tm, st, uoff, smem, red);
    if (!grid_sync(p, tm, gen, flag, TAG_C2R, tl)) return false;
    if (threadIdx.x == 0) {
        s.nrhs += 1;
        timer_push(p, tm, s.nrhs - 1, st.t);
        s.t_last = st.t;
        s.last_k = st.kout;
        s.last_trig = trig ? 1 : 0;
    }
    __syncthreads();
    return true;
}

// ------------------------------------------------------------------------------------------
// implicit step of a Maxwell mantle under a rigid or laterally constant lithosphere
// (src/deformation.jl:52-89): u <- Re IFFT[a FFT(u) + b FFT(F)] on the n x n grid with
// a = (nabla - dt/2 beta) / (nabla + dt/2 beta), b = dt / (nabla + dt/2 beta) and
// F = -(sigma_load + sigma_litho) g K^2, then the corner BC. It is evaluated as the increment
// u + Re IFFT[(a - 1) FFT(u) + b FFT(F)], so the FP32 transforms only carry the change of u:
//   imp_rows : stage displacement and loads of a row pair; u and F rows packed as two complex
//              sequences, length-nx FFTs; half spectra of u in spectrum buffer 0, of F in 1;
//   imp_cols : length-ny FFTs of both columns, (a-1) U + b F, inverse FFT into buffer 0;
//   imp_c2r  : corner offset of the new u from its FP64 corners and direct sums over spectral
//              rows 0 and ny-1; inverse real transforms of row pairs; u + increment - offset.
// ------------------------------------------------------------------------------------------
__device__ __forceinline__ void phase_imp_rows(const KP& p, const Team& tm, const StageDesc& st, double uoff,
                                               const IceT it, const float2* twx, char* smem) {
    const int nx = p.nx, ny = p.ny;
    float2* sa = (float2*)smem;   // [2][nx]: u rows, F rows
    float2* sb = sa + 2 * nx;
    const bool wsig = st.write_sigma != 0;
    const G<double> U0 = ubuf(p, st.ubase);
    G<float2> specU = specbuf(p, 0);
    G<float2> specF = specbuf(p, 1);
    for (int item = tm.rank; item < p.nrowitems; item += tm.size) {
        const int j0 = 2 * item;
        for (int i = threadIdx.x; i < nx; i += NT) {
            float2 zu = make_float2(0.f, 0.f), zf = make_float2(0.f, 0.f);
#pragma unroll
            for (int r = 0; r < 2; ++r) {
                const int j = j0 + r;
                if (j >= ny) continue;
                const long c = i + (long)j * nx;
                const double u = U0[c] - uoff;
                const double m = p.mask[c] ? 1.0 : 0.0;
                const double zo = st.trig ? p.zoff_trig[c] : p.zoff[c];
                const double Haf = hab(ice_at(p, it, c), zo + u, p.rho_sw_ice);
                const double s_ice = p.rho_ice * (Haf - (double)p.H_af_ref[c]);
                const double s_load = m * (s_ice + (p.interactive ? (double)p.LW[c] : 0.0));
                const double lt = p.LT[c];   // sigma_litho - rho_um u_ref
                const double uref = p.u_ref[c];
                const double k2 = p.K2 ? (double)p.K2[c] : 1.0;
                const float f = (float)(-(s_load + (lt + p.rho_um * uref)) * p.g * k2);
                if (r == 0) {
                    zu.x = (float)u;
                    zf.x = f;
                } else {
                    zu.y = (float)u;
                    zf.y = f;
                }
                if (wsig) {
                    const double s_mant = p.rho_um * u;
                    const double s_full = s_load + m * (lt + s_mant);
                    st_shared(addr(p.sig_load_prev, c), (float)s_load);
                    st_shared(addr(p.sig_geo_prev, c), (float)(p.compressible_column
                                                    ? s_load + m * (s_mant - p.rho_um * uref)
                                                    : s_full));
                }
            }
            sa[i] = zu;
            sa[nx + i] = zf;
        }
        __syncthreads();
        float2* res = fft_lds<float2>(sa, sb, 2, nx, p.px, twx, -1);
        for (int idx = threadIdx.x; idx < 2 * p.nhx; idx += NT) {
            const int s = idx >= p.nhx ? 1 : 0, kx = idx - s * p.nhx;
            const float2 Z = res[s * nx + kx], Zm = res[s * nx + (kx == 0 ? 0 : nx - kx)];
            float2 A, B;
            A.x = 0.5f * (Z.x + Zm.x);
            A.y = 0.5f * (Z.y - Zm.y);
            B.x = 0.5f * (Z.y + Zm.y);
            B.y = -0.5f * (Z.x - Zm.x);
            G<float2> spec = s ? specF : specU;
            st_shared(addr(spec, sidx(p, kx, j0)), A);
            if (j0 + 1 < ny) st_shared(addr(spec, sidx(p, kx, j0 + 1)), B);
        }
        __syncthreads();
    }
}

__device__ __forceinline__ void phase_imp_cols(const KP& p, const Team& tm, const float2* twy, char* smem) {
    const int ny = p.ny;
    float2* sa = (float2*)smem;   // [2][ny]: U column, F column
    float2* sb = sa + 2 * ny;
    G<float2> specU = specbuf(p, 0);
    const G<float2> specF = specbuf(p, 1);
    for (int kx = tm.rank; kx < p.nhx; kx += tm.size) {
        for (int idx = threadIdx.x; idx < 2 * ny; idx += NT) {
            const int s = idx >= ny ? 1 : 0, j = idx - s * ny;
            sa[idx] = (s ? specF : specU)[sidx(p, kx, j)];
        }
        __syncthreads();
        float2* r = fft_lds<float2>(sa, sb, 2, ny, p.py, twy, -1);
        float2* o = (r == sa) ? sb : sa;
        for (int ky = threadIdx.x; ky < ny; ky += NT) {
            const float am1 = p.imp_am1[(long)kx * ny + ky], b = p.imp_b[(long)kx * ny + ky];
            const float2 U = r[ky], F = r[ny + ky];
            o[ky] = make_float2(am1 * U.x + b * F.x, am1 * U.y + b * F.y);
        }
        __syncthreads();
        float2* r2 = fft_lds<float2>(o, r, 1, ny, p.py, twy, +1);
        for (int j = threadIdx.x; j < ny; j += NT) st_shared(addr(specU, sidx(p, kx, j)), r2[j]);
        __syncthreads();
    }
}

__device__ __forceinline__ void phase_imp_c2r(const KP& p, const Team& tm, const StageDesc& st, double uoff,
                                              const double* ucorner, const float2* twx, char* smem, double* red) {
    const int nx = p.nx, ny = p.ny;
    const bool even = (nx % 2) == 0;
    const G<float2> spec = specbuf(p, 0);
    const G<double> U0 = ubuf(p, st.ubase);
    G<double> Uw = ubuf(p, st.write_u);
    float2* sf = (float2*)smem;
    float2* sg = sf + nx;
    // increment at the four corners: x(0, j) = sum_k w_k Re X_k, x(nx-1, j) = sum_k w_k Re(X_k e^{-2 pi i k/nx})
    double vs[NV][4] = {};
    for (int k0 = threadIdx.x; k0 < p.nhx; k0 += VNT)
#pragma unroll
    for (int vv = 0; vv < NV; ++vv) {
        const int kx = k0 + vv * NT;
        if (kx >= p.nhx) break;
        double* v = vs[vv];
        const float2 a = spec[sidx(p, kx, 0)], b = spec[sidx(p, kx, ny - 1)], tw = twx[kx];
        const bool nyq = even && kx == nx / 2;
        const double w = (kx == 0 || nyq) ? 1.0 : 2.0;
        v[0] += w * (double)a.x;
        v[2] += w * (double)b.x;
        if (nyq) {
            v[1] -= (double)a.x;
            v[3] -= (double)b.x;
        } else {
            v[1] += w * ((double)a.x * tw.x - (double)a.y * tw.y);
            v[3] += w * ((double)b.x * tw.x - (double)b.y * tw.y);
        }
    }
    block_sum<4>(red, vs);
    const double* v = vs[0];
    // corners of the new u: the stage's corners (read before this phase overwrites u) + increment
    const double off = ((0.25 * (ucorner[0] + v[0]) + 0.25 * (ucorner[1] + v[1])) +
                        (0.25 * (ucorner[2] + v[2]) + 0.25 * (ucorner[3] + v[3]))) - p.vbc_x;
    for (int item = tm.rank; item < p.nrowitems; item += tm.size) {
        const int j0 = 2 * item;
        for (int k = threadIdx.x; k < nx; k += NT) sf[k] = c2r_pair_value(p, spec, k, j0, j0 + 1);
        __syncthreads();
        float2* res = fft_lds<float2>(sf, sg, 1, nx, p.px, twx, +1);
        // each cell is read and rewritten by the same thread; no other workgroup reads this row
        // pair in this phase
        for (int idx = threadIdx.x; idx < 2 * nx; idx += NT) {
            const int r = idx >= nx ? 1 : 0, i = idx - r * nx, j = j0 + r;
            if (j >= ny) continue;
            const long c = i + (long)j * nx;
            st_shared(addr(Uw, c), (U0[c] - uoff) + (double)(r ? res[i].y : res[i].x) - off);
        }
        __syncthreads();
    }
}

// one implicit-step call: preamble, the sparse update when due, the three spectral phases; u is
// replaced in U[st.write_u] (st.write_u == st.ubase: in place)
__device__ __forceinline__ bool implicit_call(const KP& p, const Team& tm, Ctl& s, StageDesc& st, unsigned& gen,
                                              const float2* twx, const float2* twy, char* smem, double* red,
                                              int* flag, bool sparse_only, unsigned long long& tl) {
    KC4 kc;
    const bool trig = sparse_only || (((st.t - s.t0) / p.dt_sparse) >= (double)s.count);
    st.trig = trig ? 1 : 0;
    st.kin = -1;
    st.ng = 0;
    IceT it;
    double uoff;
    stage_preamble(p, st, specbuf(p, 0), twx, red, it, kc, uoff);
    if (it.err) {
        if (threadIdx.x == 0) s.err = ERR_TIME_RANGE;
        __syncthreads();
        return false;
    }
    if (trig) {
        if (!sparse_phases<true, true>(p, tm, s, st, gen, smem, red, flag, tl, uoff, it)) return false;
        if (sparse_only) return true;
    }
    // the stage's corners (with their BC), before the last phase overwrites u in place
    const G<double> U0 = ubuf(p, st.ubase);
    const long n1 = p.nx - 1, n2 = (long)(p.ny - 1) * p.nx;
    const double ucorner[4] = {U0[0] - uoff, U0[n1] - uoff, U0[n2] - uoff, U0[n2 + n1] - uoff};
    st.write_sigma = (st.force_sigma || p.write_sigma_always ||
                      (((st.t_next_max - s.t0) / p.dt_sparse) >= (double)s.count)) ? 1 : 0;
    phase_imp_rows(p, tm, st, uoff, it, twx, smem);
    if (!grid_sync(p, tm, gen, flag, TAG_ROWS, tl)) return false;
    phase_imp_cols(p, tm, twy, smem);
    if (!grid_sync(p, tm, gen, flag, TAG_COLS, tl)) return false;
    phase_imp_c2r(p, tm, st, uoff, ucorner, twx, smem, red);
    if (!grid_sync(p, tm, gen, flag, TAG_C2R, tl)) return false;
    if (threadIdx.x == 0) {
        s.nrhs += 1;
        timer_push(p, tm, s.nrhs - 1, st.t);
        s.t_last = st.t;
        s.last_trig = trig ? 1 : 0;
    }
    __syncthreads();
    return true;
}

// ------------------------------------------------------------------------------------------
// one RHS call
// ------------------------------------------------------------------------------------------
template <typename MT, bool DIST>
__device__ __forceinline__ bool rhs_call(const KP& p, const Team& tm, Ctl& s, StageDesc& st, unsigned& gen,
                                         const float2* twx, const float2* twy, char* smem, double* red, int* flag,
                                         bool sparse_only, unsigned long long& tl) {
    st.sb_out = (int)(s.nrhs & 1);
    st.sb_in = 1 - st.sb_out;   // the previous call's output
    KC4 kc;
    kc.v0 = kc.v1 = kc.v2 = kc.v3 = 0.0f;
    const bool trig = sparse_only || (((st.t - s.t0) / p.dt_sparse) >= (double)s.count);
    st.trig = trig ? 1 : 0;
    if (trig && st.kin >= 0) {
        // the sparse update needs this call's displacement everywhere: materialise the pending
        // derivative first and treat it as a global term
        StageDesc sk = st;
        sk.final_err = 0;
        phase_c2r(p, tm, sk, specbuf(p, st.sb_in), st.kin, twx, smem, red);
        if (!grid_sync(p, tm, gen, flag, TAG_KIN, tl)) return false;
        if (st.ng == 0) { st.g0 = st.kin; st.a0 = st.ain; }
        else if (st.ng == 1) { st.g1 = st.kin; st.a1 = st.ain; }
        else { st.g2 = st.kin; st.a2 = st.ain; }
        st.ng += 1;
        st.kin = -1;
    }
    IceT it;
    double uoff;
    stage_preamble(p, st, specbuf(p, st.sb_in), twx, red, it, kc, uoff);
    if (it.err) {
        if (threadIdx.x == 0) s.err = ERR_TIME_RANGE;
        __syncthreads();
        return false;
    }
    if (trig) {
        if (!sparse_phases<false>(p, tm, s, st, gen, smem, red, flag, tl, uoff, it)) return false;
        if (sparse_only) return true;
    }
    // store this call's loads if the next call may trigger (the sparse update reads the
    // column anomalies of the previous call, src/simulation.jl:362-385)
    st.write_sigma = (st.force_sigma || p.write_sigma_always ||
                      (((st.t_next_max - s.t0) / p.dt_sparse) >= (double)s.count)) ? 1 : 0;
#ifdef FI_SUBPHASE_TIMING
    __syncthreads();
    if (p.ptime && tm.rank == 0 && threadIdx.x == 0) p.ptime[23] += wall_clock64() - tl;
#endif
    if constexpr (sizeof(MT) == 8) phase_rows_lds<DIST>(p, tm, st, uoff, it, twx, smem, kc);
    else phase_rows<MT, 8, 4, DIST>(p, tm, st, uoff, it, twx, smem, kc);
    if (!grid_sync(p, tm, gen, flag, TAG_ROWS, tl)) return false;
    phase_cols(p, tm, specbuf(p, st.sb_out), twy, smem);
    if (!grid_sync(p, tm, gen, flag, TAG_COLS, tl)) return false;
    if (st.do_c2r) {
        phase_c2r(p, tm, st, specbuf(p, st.sb_out), st.kout, twx, smem, red);
        if (!grid_sync(p, tm, gen, flag, TAG_C2R, tl)) return false;
    }
    if (threadIdx.x == 0) {
        s.nrhs += 1;
        timer_push(p, tm, s.nrhs - 1, st.t);
        s.t_last = st.t;
        s.last_k = st.kout;
        s.last_trig = trig ? 1 : 0;
    }
    __syncthreads();
    return true;
}

}  // namespace FI_NS
