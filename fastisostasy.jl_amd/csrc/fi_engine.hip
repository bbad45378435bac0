// libfastiso_hip.so -- the MI355X (gfx950) LV-ELVA engine behind include/fastiso.h.
//
// Device-resident state; one cooperative persistent kernel runs either a single RHS
// evaluation (fi_rhs) or the adaptive BS3 loop (fi_integrate) until the next output time,
// with every step-control decision taken on the device (no host round trip per step).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/fastiso.h"
#include "fi_loop.hpp"

namespace fi256 {
// the ensemble kernels' 256-thread build (fi_batch256.hip)
const void* batch_kernel_nt256(bool dist, bool elra);
}  // namespace fi256

namespace fi {

// ------------------------------------------------------------------------------------------
// persistent kernel of one context: the whole grid is one team
// ------------------------------------------------------------------------------------------

template <typename MT, bool DIST>
__global__ void __launch_bounds__(NT) fi_persistent(const KP* __restrict__ pp) {
    const KP& p = *(const KP*)(const ConstKP*)(unsigned long long)pp;
    extern __shared__ __align__(16) char smem_all[];
    __shared__ Ctl s;
    __shared__ double red[64];
    __shared__ int flag;
    const Team tm{(int)blockIdx.x, (int)gridDim.x, p.bar};
    unsigned gen = 0;
    unsigned long long tl = wall_clock64();
    float2* twx = (float2*)smem_all;
    float2* twy = twx + p.nx;
    char* smem = (char*)(twy + p.ny);
    load_twiddles(p, twx, twy);
    if (threadIdx.x == 0) s = *p.ctl;
    __syncthreads();
    if (p.mode == MODE_RHS || p.mode == MODE_SPARSE_ONLY)
        (void)single_rhs<MT, DIST, false>(p, tm, s, gen, twx, twy, smem, red, &flag, tl, p.rhs_ubuf, p.rhs_kslot,
                                         s.t_rhs, p.mode == MODE_SPARSE_ONLY, false);
    else
        integrate_loop<MT, DIST, false>(p, tm, s, gen, twx, twy, smem, red, &flag, tl);
    if (tm.rank == 0 && threadIdx.x == 0) *p.ctl = s;
}

// the relaxed mantle (ELRA): per-cell K^2 is read at run time, so one instance serves both grids.
// (The body repeats fi_persistent's: sharing it through a common inline function changed the
// register allocation of the LV-ELVA instances.)
__global__ void __launch_bounds__(NT) fi_persistent_elra(const KP* __restrict__ pp) {
    const KP& p = *(const KP*)(const ConstKP*)(unsigned long long)pp;
    extern __shared__ __align__(16) char smem_all[];
    __shared__ Ctl s;
    __shared__ double red[64];
    __shared__ int flag;
    const Team tm{(int)blockIdx.x, (int)gridDim.x, p.bar};
    unsigned gen = 0;
    unsigned long long tl = wall_clock64();
    float2* twx = (float2*)smem_all;
    float2* twy = twx + p.nx;
    char* smem = (char*)(twy + p.ny);
    load_twiddles(p, twx, twy);
    if (threadIdx.x == 0) s = *p.ctl;
    __syncthreads();
    if (p.mode == MODE_RHS || p.mode == MODE_SPARSE_ONLY)
        (void)single_rhs<double, false, true>(p, tm, s, gen, twx, twy, smem, red, &flag, tl, p.rhs_ubuf, p.rhs_kslot,
                                              s.t_rhs, p.mode == MODE_SPARSE_ONLY, false);
    else
        integrate_loop<double, false, true>(p, tm, s, gen, twx, twy, smem, red, &flag, tl);
    if (tm.rank == 0 && threadIdx.x == 0) *p.ctl = s;
}

// Euler with dt = dt_min for the implicit step (src/simulation.jl:244-250): each step replaces
// U[s.cur] by one implicit step at the new time (update_diagnostics! overwrites u and leaves the
// derivative at zero, so u = uprev + dt * 0 first). Steps land on s.stop; the loop returns after
// the step that reaches s.next_out (the host then fires the output and the FSAL call).
__device__ __forceinline__ void implicit_loop(const KP& p, const Team& tm, Ctl& s, unsigned& gen, const float2* twx,
                                              const float2* twy, char* smem, double* red, int* flag,
                                              unsigned long long& tl) {
    long long attempts = 0;
    for (;;) {
        if (s.err || s.t >= s.stop || attempts >= p.max_steps) break;
        const bool clip = s.t + p.dt_min >= s.stop;
        const double tn = clip ? s.stop : s.t + p.dt_min;
        StageDesc st{};
        st.ubase = s.cur;
        st.write_u = s.cur;
        st.kin = -1;
        st.kout = s.k1;
        st.t = tn;
        st.t_next_max = tn + p.dt_min;
        if (!implicit_call(p, tm, s, st, gen, twx, twy, smem, red, flag, false, tl)) break;
        ++attempts;
        if (threadIdx.x == 0) {
            s.t = tn;
            s.nacc += 1;
            s.last_u = s.cur;
            s.steps = 1;
        }
        __syncthreads();
        if (s.t >= s.next_out) break;
    }
}

__global__ void __launch_bounds__(NT) fi_persistent_implicit(const KP* __restrict__ pp) {
    const KP& p = *(const KP*)(const ConstKP*)(unsigned long long)pp;
    extern __shared__ __align__(16) char smem_all[];
    __shared__ Ctl s;
    __shared__ double red[64];
    __shared__ int flag;
    const Team tm{(int)blockIdx.x, (int)gridDim.x, p.bar};
    unsigned gen = 0;
    unsigned long long tl = wall_clock64();
    float2* twx = (float2*)smem_all;
    float2* twy = twx + p.nx;
    char* smem = (char*)(twy + p.ny);
    load_twiddles(p, twx, twy);
    if (threadIdx.x == 0) s = *p.ctl;
    __syncthreads();
    if (p.mode == MODE_RHS || p.mode == MODE_SPARSE_ONLY) {
        StageDesc st{};
        st.ubase = p.rhs_ubuf;
        st.write_u = p.rhs_ubuf;
        st.kin = -1;
        st.kout = p.rhs_kslot;
        st.t = s.t_rhs;
        st.t_next_max = s.t_rhs;
        (void)implicit_call(p, tm, s, st, gen, twx, twy, smem, red, &flag, p.mode == MODE_SPARSE_ONLY, tl);
    } else {
        implicit_loop(p, tm, s, gen, twx, twy, smem, red, &flag, tl);
    }
    if (tm.rank == 0 && threadIdx.x == 0) *p.ctl = s;
}


// kernel instances: FP64 or FP32 moments; with or without per-cell grid steps (map distortion)
inline const void* persistent_kernel(bool mt_double, bool dist, bool elra, bool implicit = false) {
    if (implicit) return (const void*)fi_persistent_implicit;
    if (elra) return (const void*)fi_persistent_elra;
    if (!mt_double) return dist ? (const void*)fi_persistent<float, true> : (const void*)fi_persistent<float, false>;
    return dist ? (const void*)fi_persistent<double, true> : (const void*)fi_persistent<double, false>;
}
inline const void* batch_kernel(bool mt_double, bool dist, bool elra = false) {
    if (elra) return (const void*)fi_batch_persistent_elra;
    if (!mt_double)
        return dist ? (const void*)fi_batch_persistent<float, true> : (const void*)fi_batch_persistent<float, false>;
    return dist ? (const void*)fi_batch_persistent<double, true> : (const void*)fi_batch_persistent<double, false>;
}

// ------------------------------------------------------------------------------------------
// set-up: FP64 half spectra of the zero-padded Green kernels (ConvolutionPlan,
// src/convolutions.jl:41-56,89-92), stored [kx][ky] and pre-scaled by 1/(nfx nfy) * scale
// ------------------------------------------------------------------------------------------
__global__ void __launch_bounds__(NT) k_green_rows(const double* __restrict__ G, int nx, int ny, int nfx, int nfhx,
                                                    FftPlan pl, const double2* __restrict__ tw, double2* out) {
    extern __shared__ __align__(16) char smem[];
    double2* a = (double2*)smem;
    double2* b = a + nfx;
    const int j = blockIdx.x;
    for (int i = threadIdx.x; i < nfx; i += NT) {
        double2 z;
        z.x = (i < nx) ? G[i + (long)j * nx] : 0.0;
        z.y = 0.0;
        a[i] = z;
    }
    __syncthreads();
    double2* r = fft_lds<double2>(a, b, 1, nfx, pl, tw + pl.tw, -1);
    for (int kx = threadIdx.x; kx < nfhx; kx += NT) out[(long)kx * ny + j] = r[kx];
}

__global__ void __launch_bounds__(NT) k_green_cols(const double2* __restrict__ in, int ny, int nfy, FftPlan pl,
                                                    const double2* __restrict__ tw, double scale, double2* out) {
    extern __shared__ __align__(16) char smem[];
    double2* a = (double2*)smem;
    double2* b = a + nfy;
    const int kx = blockIdx.x;
    for (int j = threadIdx.x; j < nfy; j += NT) {
        double2 z;
        z.x = 0;
        z.y = 0;
        if (j < ny) z = in[(long)kx * ny + j];
        a[j] = z;
    }
    __syncthreads();
    double2* r = fft_lds<double2>(a, b, 1, nfy, pl, tw + pl.tw, -1);
    for (int ky = threadIdx.x; ky < nfy; ky += NT) {
        double2 v = r[ky];
        v.x *= scale;
        v.y *= scale;
        out[(long)kx * nfy + ky] = v;
    }
}

// dudt = k - offset (FP64), for device-array callers
__global__ void k_dudt_out(const float* __restrict__ k, double off, long n, double* out) {
    for (long c = (long)blockIdx.x * blockDim.x + threadIdx.x; c < n; c += (long)gridDim.x * blockDim.x)
        out[c] = (double)k[c] - off;
}
__global__ void k_sub_offset(double* u, double off, long n) {
    for (long c = (long)blockIdx.x * blockDim.x + threadIdx.x; c < n; c += (long)gridDim.x * blockDim.x) u[c] -= off;
}
// One RKC2 stage on the FP64 state: y = a0 y0 + a1 ym1 + a2 ym2 + b1 f1 + b0 f0, the derivatives
// f read from their FP32 K slots minus their corner-BC offsets (zero for ELRA, whose derivatives
// are stored whole). y may be ym2 (the stage recursion overwrites Y_{j-2}).
__device__ __forceinline__ double kslot_off_dev(const float* kc, double vbc) {
    const double a = 0.25 * (double)kc[0], b = 0.25 * (double)kc[1], c = 0.25 * (double)kc[2], d = 0.25 * (double)kc[3];
    return ((a + b) + (c + d)) - vbc;
}
// The single-RHS launches of an RKC chain follow each other without host work: the kernel that
// runs between two of them (a stage, or the sub-step's corner offset) prepares the next one in
// its workgroup 0: it keeps the previous launch's barrier abort word in *abort_acc, clears the
// barrier region (fresh generations for the next launch) and writes the next right-hand side's
// time into the control block from the chain's time table (idx < 0: no next right-hand side).
struct RkcPrep {
    Bar* bar;          // null: nothing to prepare
    int words;         // 32-bit words of the barrier region
    int* abort_acc;
    Ctl* ctl;
    const double* times;
    int idx;
    int first;         // the chain's first kernel: clears *abort_acc
};
__device__ __forceinline__ void rkc_prep(const RkcPrep& q) {
    if (!q.bar || blockIdx.x != 0) return;
    if (threadIdx.x == 0) {
        // the chain's first kernel follows launches of other calls: their abort words are theirs
        *q.abort_acc = q.first ? 0 : (*q.abort_acc | q.bar->abort);
        if (q.idx >= 0) q.ctl->t_rhs = q.times[q.idx];
    }
    __syncthreads();
    unsigned* w = (unsigned*)q.bar;
    for (int k = threadIdx.x; k < q.words; k += blockDim.x) w[k] = 0u;
}
__global__ void k_rkc_prep(RkcPrep q) { rkc_prep(q); }

__global__ void __launch_bounds__(256) k_rkc_stage(double* y, const double* __restrict__ y0, const double* __restrict__ ym1,
                                                   const double* ym2, const float* __restrict__ f1,
                                                   const float* __restrict__ f0, const float* __restrict__ kc1,
                                                   const float* __restrict__ kc0, int offsets, double vbc, double a0,
                                                   double a1, double a2, double b1, double b0, long n, RkcPrep q) {
    rkc_prep(q);
    const double off1 = (offsets && b1 != 0.0) ? kslot_off_dev(kc1, vbc) : 0.0;
    const double off0 = (offsets && b0 != 0.0) ? kslot_off_dev(kc0, vbc) : 0.0;
    for (long c = (long)blockIdx.x * blockDim.x + threadIdx.x; c < n; c += (long)gridDim.x * blockDim.x) {
        double v = a0 * y0[c];
        if (a1 != 0.0) v += a1 * ym1[c];
        if (a2 != 0.0) v += a2 * ym2[c];
        if (b1 != 0.0) v += b1 * ((double)f1[c] - off1);
        if (b0 != 0.0) v += b0 * ((double)f0[c] - off0);
        y[c] = v;
    }
}
// the corner-BC offset of u (src/boundary_conditions.jl:144-148) into *off, then u -= *off
__global__ void k_corner_offset(const double* __restrict__ u, int nx, int ny, double xb, double* off) {
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        const long n2 = (long)(ny - 1) * nx;
        const double a = 0.25 * u[0], b = 0.25 * u[nx - 1], c = 0.25 * u[n2], d = 0.25 * u[n2 + nx - 1];
        *off = ((a + b) + (c + d)) - xb;
    }
}
__global__ void k_f64_to_f32(const double* __restrict__ in, float* __restrict__ out, long n) {
    for (long c = (long)blockIdx.x * blockDim.x + threadIdx.x; c < n; c += (long)gridDim.x * blockDim.x)
        out[c] = (float)in[c];
}
__global__ void k_sub_offset_dev(double* u, const double* __restrict__ off, long n, RkcPrep q) {
    rkc_prep(q);
    const double o = *off;
    for (long c = (long)blockIdx.x * blockDim.x + threadIdx.x; c < n; c += (long)gridDim.x * blockDim.x) u[c] -= o;
}

// thinplate_horizontal_displacement!(u_x, u_y, u + ue, h, domain) (src/deformation.jl:141-164):
// -h/2 times the first derivative along x (axis 0) or y (axis 1) with the CPU edge rules
// (src/derivatives.jl:171-191) and the per-cell step K dx or K dy. A constant offset of u (its
// pending corner BC) does not change the derivative.
__global__ void k_thinplate(const double* __restrict__ u, const float* __restrict__ ue, const float* __restrict__ h,
                            const float* __restrict__ kinv, int nx, int ny, double step, int axis, double* out) {
    const long N = (long)nx * ny;
    for (long c = (long)blockIdx.x * blockDim.x + threadIdx.x; c < N; c += (long)gridDim.x * blockDim.x) {
        const long j = c / nx, i = c - j * nx;
        const long stride = axis == 0 ? 1 : nx;
        const long pos = axis == 0 ? i : j, len = axis == 0 ? nx : ny;
        const long a = pos == 0 ? c : c - stride, b = pos == len - 1 ? c : c + stride;
        const double span = (pos == 0 || pos == len - 1) ? 1.0 : 2.0;
        const double d = ((u[b] + (double)ue[b]) - (u[a] + (double)ue[a])) / (span * step);
        out[c] = -(double)h[c] / 2 * (kinv ? d * (double)kinv[c] : d);
    }
}

}  // namespace fi

// ==========================================================================================
// host side
// ==========================================================================================
using namespace fi;

namespace {

template <typename T>
struct DBuf {
    T* p = nullptr;
    size_t n = 0;
    ~DBuf() { if (p) (void)hipFree(p); }
    void reset() {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    hipError_t alloc(size_t cnt) {
        if (p) { (void)hipFree(p); p = nullptr; }
        n = cnt;
        return hipMalloc((void**)&p, std::max<size_t>(cnt, 1) * sizeof(T));
    }
};

bool plan_fft(int L, int twoff, FftPlan& pl) {
    std::memset(&pl, 0, sizeof(pl));
    pl.L = L;
    pl.tw = twoff;
    int m = L;
    auto push = [&](int r) {
        if (pl.nr >= 16) return false;
        pl.code |= (unsigned long long)r << (4 * pl.nr);
        pl.nr++;
        m /= r;
        return true;
    };
    while (m % 8 == 0) if (!push(8)) return false;
    while (m % 4 == 0) if (!push(4)) return false;
    while (m % 2 == 0) if (!push(2)) return false;
    for (int r : {3, 5, 7}) while (m % r == 0) if (!push(r)) return false;
    return m == 1;
}

int nextfastfft(int n) {
    for (int m = std::max(n, 1);; ++m) {
        int r = m;
        for (int q : {2, 3, 5, 7}) while (r % q == 0) r /= q;
        if (r == 1) return m;
    }
}

}  // namespace

struct fi_ctx {
    fi_config cfg;
    std::string err;
    int dev = 0;
    hipStream_t stream = nullptr;
    int nx = 0, ny = 0, nhx = 0, nfx = 0, nfy = 0, nfhx = 0;
    size_t N = 0;
    double dx = 0, dy = 0, A = 0;
    int i1 = 0, i2 = 0, j1 = 0, j2 = 0;
    // host copies of the static inputs
    std::vector<double> D, eta, spinv, mask, ge, gg, gv, tau, litho_h, u_ref, ue_ref, dz_ss_ref, z_b_ref, H_ref_in;
    bool has_D = false, has_eta = false, has_spinv = false, has_ge = false, has_gg = false, has_gv = false,
         has_tau = false, has_href = false;
    bool elra = false;   // RelaxedMantle (ELRA) instead of the LV-ELVA Maxwell mantle
    bool implicit = false;   // Maxwell mantle under a rigid or laterally constant lithosphere
    int ice_mode = ICE_EXTERNAL, ice_flat = 0;
    std::vector<double> ice_t, ice_a, ice_H;
    // ImposedSeaSurface snapshots of dz_ss (FP64 on the device: read once per sparse update)
    std::vector<double> ss_t, ss_X;
    int ss_flat = 0;
    // barystatic sea-level rule (fi_set_bsl); A_OCEAN_PD (src/barystatic_sealevel.jl:1) by default
    int bsl_rule = FI_BSL_CONSTANT, bsl_update = FI_BSL_INTERNAL_UPDATE, bsl_flat = 0;
    double bsl_A_ref = 3.625e14;
    std::vector<double> bsl_x, bsl_y;
    // map scale factor K (FI_DISTORTION); has_dist when any K != 1
    std::vector<double> dist;
    bool has_dist = false;
    std::vector<double> tout;
    std::vector<double> H_ice_ref, H_af_ref, H_water_ref, z_ss_ref;
    std::vector<double> H_ext_host;
    // device memory
    DBuf<double> U, ice_t_d, ice_a_d, part, vpart, bsl_x_d, bsl_y_d, ss_t_d, ss_X_d;
    DBuf<float> kinv_d, k2_d;   // 1/K and K^2 per cell (distortion only)
    DBuf<float> K, kcorner, mult, Df, inv2eta, H_af_ref_d, u_ref_d, z_b_ref_d, z_ss_ref_d, ue_ref_d, H_ext, ue, dz_ss,
        z_ss, zoff, zoff_trig, LW, LT, H_water, sig_load_prev, sig_geo_prev, ue_raw, dz_raw, ext_corners, ice_H_d;
    DBuf<unsigned char> maskd, mg, mo;
    DBuf<float2> spec, pA, pB, oA, oB, tw32;
    DBuf<double2> ge_hat, gg_hat, gv_hat, tw64, gtmp;
    DBuf<float> inv_tau_d, ueq, litho_h_d, imp_am1, imp_b;
    DBuf<double> uxy_d;   // horizontal displacement output (FI_U_X / FI_U_Y)
    DBuf<double> rkc_off;   // corner offset scratch of fi_integrate_rkc
    double rkc_rho = 0;     // spectral radius estimate of fi_integrate_rkc (1/yr)
    DBuf<KP> rkc_kp;        // single-RHS parameter blocks of fi_integrate_rkc: [U buffer][K slot 0/1]
    double* rkc_times = nullptr;   // pinned stage times of one fi_integrate_rkc call, then its abort word
    size_t rkc_times_n = 0;
    DBuf<double> rkc_tdev;         // the stage times on the device (read by the chain's prep)
    DBuf<int> rkc_abort;           // barrier abort words of the chain, or-ed on the device
    // captured chains, one per (sub-steps, stages, sub-step, starting buffer, grid, buffers) key
    // (at most four, the oldest replaced), with the buffer each chain's state ends in
    struct RkcGraph {
        std::vector<unsigned char> key;
        hipGraphExec_t exec = nullptr;
        int y_end = 0;
    };
    std::vector<RkcGraph> rkc_graphs;
    bool rkc_graph_off = false;    // capture failed once, or FASTISO_RKC_GRAPH=0
    DBuf<Ctl> ctl_d;
    DBuf<Bar> bar_d;
    DBuf<KP> kp_d;
    DBuf<unsigned long long> ptime;
    // Timer history: capacity (0: off) and the device buffers
    long long hist_cap = 0;
    DBuf<double> hist_t;
    DBuf<unsigned long long> hist_clk;
    KP kp;
    Ctl ctl;
    int G = 0, maxg = 0;   // persistent grid size in use; co-resident maximum (bar_d holds maxg)
    size_t lds = 0, lds_d = 0, lds_f = 0;   // LDS bytes of this context's launch; with FP64 / FP32 moments
    size_t lds_dist = 0;                     // FP64 moments with per-cell grid steps
    bool mt_double = true;
    FftPlan px64, py64, pfx64, pfy64;
    bool initialized = false, integ = false, k1_stale = false;
    long long kout = 0;
    double t_end = 0;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    double device_ms = 0;
    long long n_launch = 0, n_accept_timed = 0;
};

namespace {

int fail(fi_ctx* c, int code, const std::string& m) {
    if (c) c->err = m;
    return code;
}

#define HIPCHECK(c, x)                                                                         \
    do {                                                                                       \
        hipError_t e_ = (x);                                                                   \
        if (e_ != hipSuccess) return fail((c), FI_ERR_DEVICE, std::string(#x " : ") + hipGetErrorString(e_)); \
    } while (0)

std::vector<float> to_f(const std::vector<double>& v) { return std::vector<float>(v.begin(), v.end()); }

template <typename T>
int upload(fi_ctx* c, DBuf<T>& b, const std::vector<T>& h) {
    HIPCHECK(c, b.alloc(h.size()));
    HIPCHECK(c, hipMemcpy(b.p, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice));
    return FI_OK;
}

// TimeInterpolation2D on the host (src/interpolations.jl:51-74), for H_ice_ref and diagnostics
int host_ice(fi_ctx* c, double t, std::vector<double>& out) {
    out.assign(c->N, 0.0);
    if (c->ice_mode == ICE_EXTERNAL) {
        if (!c->H_ext_host.empty()) out = c->H_ext_host;
        return FI_OK;
    }
    const std::vector<double>& tv = c->ice_t;
    int nt = (int)tv.size();
    auto snap = [&](int k, size_t p) { return c->ice_mode == ICE_SNAPSHOTS ? c->ice_H[(size_t)k * c->N + p] : c->ice_a[k] * c->ice_H[p]; };
    if (t < tv.front() || t > tv.back()) {
        if (!c->ice_flat) return fail(c, FI_ERR_TIME_RANGE, "TimeInterpolation2D out of range.");
        int k = t < tv.front() ? 0 : nt - 1;
        for (size_t p = 0; p < c->N; ++p) out[p] = snap(k, p);
        return FI_OK;
    }
    int i = (int)(std::lower_bound(tv.begin(), tv.end(), t) - tv.begin());
    if (tv[i] == t) {
        for (size_t p = 0; p < c->N; ++p) out[p] = snap(i, p);
    } else {
        --i;
        for (size_t p = 0; p < c->N; ++p) {
            double a = snap(i, p), b = snap(i + 1, p);
            out[p] = a + (b - a) / (tv[i + 1] - tv[i]) * (t - tv[i]);
        }
    }
    return FI_OK;
}

double hab_h(double H, double zb, double zss, double rsi) { return std::max(H + std::min(zb - zss, 0.0) * rsi, 0.0); }

int launch(fi_ctx* c, int mode, int ubuf, int kslot, int max_steps, int write_sigma_always) {
    // the integration loop writes u_{n+1} into U[1 - cur]: cur must be 0 or 1
    if (mode == MODE_INTEGRATE && c->ctl.cur != 0 && c->ctl.cur != 1)
        return fail(c, FI_ERR_STATE, "state buffer index out of range");
    if (ubuf < 0 || ubuf > 2 || kslot < 0 || kslot > 3) return fail(c, FI_ERR_STATE, "buffer index out of range");
    KP kp = c->kp;
    kp.mode = mode;
    kp.rhs_ubuf = ubuf;
    kp.rhs_kslot = kslot;
    kp.max_steps = max_steps;
    kp.write_sigma_always = write_sigma_always;
    HIPCHECK(c, hipMemsetAsync(c->bar_d.p, 0, sizeof(Bar) * (1 + 2 * (size_t)c->G), c->stream));
    HIPCHECK(c, hipMemcpyAsync(c->ctl_d.p, &c->ctl, sizeof(Ctl), hipMemcpyHostToDevice, c->stream));
    HIPCHECK(c, hipMemcpyAsync(c->kp_d.p, &kp, sizeof(KP), hipMemcpyHostToDevice, c->stream));
    const KP* kptr = c->kp_d.p;
    void* args[] = {&kptr};
    const void* f = persistent_kernel(c->mt_double, c->has_dist, c->elra, c->implicit);
    const bool timed = mode == MODE_INTEGRATE;
    const long long nacc0 = c->ctl.nacc;
    if (timed) HIPCHECK(c, hipEventRecord(c->ev0, c->stream));
    HIPCHECK(c, hipLaunchCooperativeKernel(f, dim3(c->G), dim3(NT), args, (unsigned)c->lds, c->stream));
    if (timed) HIPCHECK(c, hipEventRecord(c->ev1, c->stream));
    HIPCHECK(c, hipMemcpyAsync(&c->ctl, c->ctl_d.p, sizeof(Ctl), hipMemcpyDeviceToHost, c->stream));
    Bar bar;
    HIPCHECK(c, hipMemcpyAsync(&bar, c->bar_d.p, sizeof(Bar), hipMemcpyDeviceToHost, c->stream));
    HIPCHECK(c, hipStreamSynchronize(c->stream));
    if (timed) {
        float ms = 0;
        HIPCHECK(c, hipEventElapsedTime(&ms, c->ev0, c->ev1));
        c->device_ms += ms;
        c->n_launch += 1;
        c->n_accept_timed += c->ctl.nacc - nacc0;
    }
    if (bar.abort) return fail(c, FI_ERR_DEVICE, "device barrier timed out; kernel aborted");
    if (c->ctl.err == ERR_TIME_RANGE) return fail(c, FI_ERR_TIME_RANGE, "TimeInterpolation2D out of range.");
    if (c->ctl.err == ERR_BSL_RANGE) return fail(c, FI_ERR_TIME_RANGE, "TimeInterpolation0D out of range.");
    if (c->ctl.err == ERR_DT_MIN) return fail(c, FI_ERR_DT_MIN, "dt below minimum");
    if (c->ctl.err == ERR_NONFINITE)
        return fail(c, FI_ERR_NONFINITE, "NaN dt detected: NaN in the state or its derivative");
    if (c->ctl.err) return fail(c, FI_ERR_DEVICE, "device error");
    return FI_OK;
}

// the kappa table with its DC fix (get_differential_fourier, src/derivatives.jl:201-222;
// src/domain.jl:151-155): sqrt((pi/Wx f_i)^2 + (pi/Wy f_j)^2) with f the FFT integer frequencies
// 0..N/2 then back down; the DC entry is 1e-3 times the mean of its two neighbours
std::vector<double> pseudodiff_table(int nx, int ny, double Wx, double Wy) {
    auto fint = [](int n, int k) { return (double)std::min(k, n - k); };
    std::vector<double> out((size_t)nx * ny);
    for (int j = 0; j < ny; ++j)
        for (int i = 0; i < nx; ++i) {
            const double X = M_PI / Wx * fint(nx, i), Y = M_PI / Wy * fint(ny, j);
            out[i + (size_t)j * nx] = std::sqrt(X * X + Y * Y);
        }
    out[0] = 1e-3 * ((out[(size_t)nx] + out[1]) / 2.0);
    return out;
}

// corner BC offset of a host array (src/boundary_conditions.jl:144-148)
double corner_offset(const double* X, int nx, int ny, double xb) {
    double a = 0.25 * X[0], b = 0.25 * X[nx - 1], cc = 0.25 * X[(size_t)(ny - 1) * nx],
           d = 0.25 * X[(size_t)(ny - 1) * nx + nx - 1];
    return ((a + b) + (cc + d)) - xb;
}

double kslot_offset(fi_ctx* c, int slot, float kc[4]) {
    if (c->elra) return 0.0;   // ELRA derivatives are stored whole (no BC on dudt)
    double a = 0.25 * (double)kc[0], b = 0.25 * (double)kc[1], cc = 0.25 * (double)kc[2], d = 0.25 * (double)kc[3];
    (void)slot;
    return ((a + b) + (cc + d)) - c->cfg.viscous_x_border;
}

int read_k(fi_ctx* c, int slot, std::vector<double>& out) {
    std::vector<float> k(c->N);
    float kc[4];
    HIPCHECK(c, hipMemcpy(k.data(), c->K.p + (size_t)slot * c->N, c->N * sizeof(float), hipMemcpyDeviceToHost));
    HIPCHECK(c, hipMemcpy(kc, c->kcorner.p + 4 * slot, 4 * sizeof(float), hipMemcpyDeviceToHost));
    double off = kslot_offset(c, slot, kc);
    out.resize(c->N);
    for (size_t p = 0; p < c->N; ++p) out[p] = (double)k[p] - off;
    return FI_OK;
}

int read_f(fi_ctx* c, const float* d, std::vector<double>& out) {
    std::vector<float> h(c->N);
    HIPCHECK(c, hipMemcpy(h.data(), d, c->N * sizeof(float), hipMemcpyDeviceToHost));
    out.assign(h.begin(), h.end());
    return FI_OK;
}

int read_u(fi_ctx* c, int buf, std::vector<double>& out) {
    out.resize(c->N);
    HIPCHECK(c, hipMemcpy(out.data(), c->U.p + (size_t)buf * c->N, c->N * sizeof(double), hipMemcpyDeviceToHost));
    return FI_OK;
}

// one RHS evaluation on U[ubuf] at time t into K[slot] (the input is not modified; the
// corner-corrected displacement is u - offset, computed identically on the host)
int rhs_on_device(fi_ctx* c, double t, int ubuf, int slot) {
    c->ctl.t_rhs = t;
    int st = launch(c, MODE_RHS, ubuf, slot, 0, 1);
    if (st) return st;
    return FI_OK;
}

}  // namespace

// ==========================================================================================
// C ABI
// ==========================================================================================
extern "C" {

int32_t fi_abi_version(void) { return FI_ABI_VERSION; }
const char* fi_engine_name(void) { return "hip-gfx950"; }

void fi_default_config(fi_config* c, int32_t nx, int32_t ny, double Wx, double Wy) {
    std::memset(c, 0, sizeof(*c));
    c->nx = nx;
    c->ny = ny;
    c->Wx = Wx;
    c->Wy = Wy;
    c->g = 9.8;
    c->seconds_per_year = 60.0 * 60.0 * 24.0 * 365.25;
    c->rho_ice = 0.910e3;
    c->rho_water = 1e3;
    c->rho_seawater = 1.023e3;
    c->r_pole = 6357e3;
    c->mE = 5.972e24;
    c->litho_poissonratio = 0.28;
    c->rho_uppermantle = 3.4e3;
    c->rho_litho = 3.2e3;
    c->lithosphere = FI_LATERALLY_VARIABLE_LITHOSPHERE;
    c->mantle = FI_MAXWELL_MANTLE;
    c->lithosphere_column = FI_INCOMPRESSIBLE_LITHOSPHERE_COLUMN;
    c->sea_surface = FI_LATERALLY_CONSTANT_SEA_SURFACE;
    c->sealevel_load = FI_NO_SEALEVEL_LOAD;
    c->volume_contribution = FI_GOELZER_CONTRIBUTION;
    c->density_contribution = FI_GOELZER_CONTRIBUTION;
    c->adjustment_contribution = FI_NO_CONTRIBUTION;
    c->dt_sparse = 10.0;
    c->reltol = 1e-5;
    c->abstol = 1e-6;
}

int fi_create(const fi_config* cfg, fi_ctx** out) {
    if (!cfg || !out) return FI_ERR_INVALID;
    *out = nullptr;
    if (cfg->nx < 4 || cfg->ny < 4 || !(cfg->Wx > 0) || !(cfg->Wy > 0)) return FI_ERR_INVALID;
    if (cfg->sea_surface < FI_LATERALLY_CONSTANT_SEA_SURFACE || cfg->sea_surface > FI_IMPOSED_SEA_SURFACE)
        return FI_ERR_INVALID;
    // RelaxedMantle with the laterally variable lithosphere errors in the reference
    // (src/deformation.jl:47-50)
    if (cfg->mantle == FI_RELAXED_MANTLE && cfg->lithosphere == FI_LATERALLY_VARIABLE_LITHOSPHERE)
        return FI_ERR_UNSUPPORTED;
    if (cfg->mantle == FI_RIGID_MANTLE) return FI_ERR_UNSUPPORTED;  // dudt = 0: nothing to accelerate
    fi_ctx* c = new fi_ctx();
    c->cfg = *cfg;
    c->elra = cfg->mantle == FI_RELAXED_MANTLE;
    c->implicit = cfg->mantle == FI_MAXWELL_MANTLE && cfg->lithosphere != FI_LATERALLY_VARIABLE_LITHOSPHERE;
    c->nx = cfg->nx;
    c->ny = cfg->ny;
    c->N = (size_t)c->nx * c->ny;
    c->nhx = c->nx / 2 + 1;
    c->nfx = nextfastfft(2 * c->nx - 1);
    c->nfy = nextfastfft(2 * c->ny - 1);
    c->nfhx = c->nfx / 2 + 1;
    c->dx = 2 * cfg->Wx / c->nx;
    c->dy = 2 * cfg->Wy / c->ny;
    c->A = c->dx * c->dy;
    int mx = c->nx / 2, my = c->ny / 2;
    c->i1 = (c->nx % 2 == 0) ? mx : mx + 1;
    c->i2 = 2 * c->nx - 1 - mx;
    c->j1 = (c->ny % 2 == 0) ? my : my + 1;
    c->j2 = 2 * c->ny - 1 - my;
    FftPlan tmp;
    if (!plan_fft(c->nx, 0, tmp) || !plan_fft(c->ny, 0, tmp)) {
        delete c;
        return FI_ERR_UNSUPPORTED;  // grid sides must factor into 2, 3, 5, 7
    }
    c->dev = cfg->device;
    if (hipSetDevice(c->dev) != hipSuccess || hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreate(&c->ev0) != hipSuccess || hipEventCreate(&c->ev1) != hipSuccess) {
        delete c;
        return FI_ERR_DEVICE;
    }
    size_t N = c->N;
    c->mask.assign(N, 1.0);
    c->u_ref.assign(N, 0.0);
    c->ue_ref.assign(N, 0.0);
    c->dz_ss_ref.assign(N, 0.0);
    c->z_b_ref.assign(N, 1e6);
    *out = c;
    return FI_OK;
}

void fi_destroy(fi_ctx* c) {
    if (!c) return;
    (void)hipSetDevice(c->dev);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    if (c->ptime.p) {
        // per-phase wall-clock totals of workgroup 0 (100 MHz ticks): work before the barrier and
        // the barrier wait
        unsigned long long h[32];
        if (hipMemcpy(h, c->ptime.p, sizeof(h), hipMemcpyDeviceToHost) == hipSuccess) {
            const char* names[5] = {"rows", "cols", "c2r", "sparse", "kin-c2r"};
            for (int t = 0; t < 5; ++t)
                if (h[16 + t])
                    fprintf(stderr, "fastiso phase %-8s n=%llu work %.3f us wait %.3f us (per phase)\n", names[t],
                            h[16 + t], h[t] / 100.0 / h[16 + t], h[8 + t] / 100.0 / h[16 + t]);
            if (h[23] && h[16]) fprintf(stderr, "fastiso rows preamble %.3f us (per rows phase)\n", h[23] / 100.0 / h[16]);
            if (h[22] && h[16]) fprintf(stderr, "fastiso step control %.3f us (per rows phase)\n", h[22] / 100.0 / h[16]);
            for (int k = 0; k < 8; ++k)
                if (h[24 + k] && h[16])
                    fprintf(stderr, "fastiso rows mark %d %.3f us (per rows phase)\n", k, h[24 + k] / 100.0 / h[16]);
        }
    }
    hipStream_t s = c->stream;
    if (c->ev0) (void)hipEventDestroy(c->ev0);
    if (c->ev1) (void)hipEventDestroy(c->ev1);
    if (c->rkc_times) (void)hipHostFree(c->rkc_times);
    for (auto& g : c->rkc_graphs) (void)hipGraphExecDestroy(g.exec);
    delete c;
    if (s) (void)hipStreamDestroy(s);
}

const char* fi_last_error(const fi_ctx* c) { return c ? c->err.c_str() : "null context"; }

int fi_set_field(fi_ctx* c, int32_t f, const double* h, size_t n) {
    if (!c || !h || n != c->N) return fail(c, FI_ERR_INVALID, "fi_set_field: size mismatch");
    std::vector<double> v(h, h + n);
    switch (f) {
        case FI_LITHO_RIGIDITY: c->D = v; c->has_D = true; return FI_OK;
        case FI_EFFECTIVE_VISCOSITY: c->eta = v; c->has_eta = true; return FI_OK;
        case FI_SCALED_PSEUDODIFF_INV: c->spinv = v; c->has_spinv = true; return FI_OK;
        case FI_MASK_ACTIVE: c->mask = v; return FI_OK;
        case FI_GREEN_ELASTIC: c->ge = v; c->has_ge = true; return FI_OK;
        case FI_GREEN_GEOID: c->gg = v; c->has_gg = true; return FI_OK;
        case FI_GREEN_VISCOUS: c->gv = v; c->has_gv = true; return FI_OK;
        case FI_LITHO_THICKNESS: c->litho_h = v; c->litho_h_d.n = 0; return FI_OK;
        case FI_TAU:
            for (double x : v)
                if (!(x > 0) || !std::isfinite(x)) return fail(c, FI_ERR_INVALID, "tau must be positive");
            c->tau = v;
            c->has_tau = true;
            return FI_OK;
        case FI_REF_U: c->u_ref = v; return FI_OK;
        case FI_REF_UE: c->ue_ref = v; return FI_OK;
        case FI_REF_DZ_SS: c->dz_ss_ref = v; return FI_OK;
        case FI_REF_Z_B: c->z_b_ref = v; return FI_OK;
        case FI_REF_H_ICE: c->H_ref_in = v; c->has_href = true; return FI_OK;
        case FI_DISTORTION:
            for (double k : v)
                if (!(k > 0) || !std::isfinite(k)) return fail(c, FI_ERR_INVALID, "distortion K must be positive");
            c->dist = v;
            c->has_dist = std::any_of(v.begin(), v.end(), [](double k) { return k != 1.0; });
            return FI_OK;
        default: break;
    }
    if (!c->initialized) return fail(c, FI_ERR_STATE, "state field written before fi_init");
    (void)hipSetDevice(c->dev);
    if (f == FI_U) {
        int buf = c->integ ? c->ctl.cur : 0;
        HIPCHECK(c, hipMemcpy(c->U.p + (size_t)buf * c->N, h, n * sizeof(double), hipMemcpyHostToDevice));
        c->ctl.last_u = buf;
        if (c->integ) c->k1_stale = true;
        return FI_OK;
    }
    if (f == FI_H_ICE) {
        if (c->ice_mode != ICE_EXTERNAL) return fail(c, FI_ERR_INVALID, "H_ice is interpolated in time; not writable");
        c->H_ext_host = v;
        std::vector<float> fv = to_f(v);
        HIPCHECK(c, hipMemcpy(c->H_ext.p, fv.data(), n * sizeof(float), hipMemcpyHostToDevice));
        return FI_OK;
    }
    return fail(c, FI_ERR_INVALID, "fi_set_field: field not writable");
}

int fi_set_field_dev(fi_ctx* c, int32_t f, const void* d, size_t n) {
    if (!c || !d || n != c->N) return fail(c, FI_ERR_INVALID, "fi_set_field_dev: size mismatch");
    HIPCHECK(c, hipSetDevice(c->dev));
    if (c->initialized && f == FI_U) {
        const int buf = c->integ ? c->ctl.cur : 0;
        HIPCHECK(c, hipMemcpyAsync(c->U.p + (size_t)buf * c->N, d, n * sizeof(double), hipMemcpyDeviceToDevice,
                                   c->stream));
        HIPCHECK(c, hipStreamSynchronize(c->stream));
        c->ctl.last_u = buf;
        if (c->integ) c->k1_stale = true;
        return FI_OK;
    }
    if (c->initialized && f == FI_H_ICE && c->ice_mode == ICE_EXTERNAL) {
        const int nb = (int)std::min<size_t>((n + 255) / 256, 4096);
        hipLaunchKernelGGL(k_f64_to_f32, dim3(nb), dim3(256), 0, c->stream, (const double*)d, c->H_ext.p, (long)n);
        HIPCHECK(c, hipGetLastError());
        // the host copy serves the host-side diagnostics (fi_get_field(FI_H_ICE), volumes)
        c->H_ext_host.resize(n);
        HIPCHECK(c, hipMemcpyAsync(c->H_ext_host.data(), d, n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
        HIPCHECK(c, hipStreamSynchronize(c->stream));
        return FI_OK;
    }
    std::vector<double> h(n);
    HIPCHECK(c, hipMemcpy(h.data(), d, n * sizeof(double), hipMemcpyDeviceToHost));
    return fi_set_field(c, f, h.data(), n);
}

int fi_get_field_dev(fi_ctx* c, int32_t f, void* d, size_t n) {
    if (!c || !d || n != c->N) return fail(c, FI_ERR_INVALID, "fi_get_field_dev: size mismatch");
    HIPCHECK(c, hipSetDevice(c->dev));
    if (c->initialized && f == FI_U) {
        HIPCHECK(c, hipMemcpyAsync(d, c->U.p + (size_t)c->ctl.last_u * c->N, n * sizeof(double),
                                   hipMemcpyDeviceToDevice, c->stream));
        if (c->ctl.last_u == 2) {   // fi_rhs input buffer kept unmodified: its corner BC applied here
            if (!c->rkc_off.p) HIPCHECK(c, c->rkc_off.alloc(1));
            const int nb = (int)std::min<size_t>((n + 255) / 256, 4096);
            hipLaunchKernelGGL(k_corner_offset, dim3(1), dim3(64), 0, c->stream, (const double*)d, c->nx, c->ny,
                               c->cfg.viscous_x_border, c->rkc_off.p);
            hipLaunchKernelGGL(k_sub_offset_dev, dim3(nb), dim3(256), 0, c->stream, (double*)d, c->rkc_off.p, (long)n,
                               RkcPrep{});
            HIPCHECK(c, hipGetLastError());
        }
        HIPCHECK(c, hipStreamSynchronize(c->stream));
        return FI_OK;
    }
    std::vector<double> h(n);
    const int st = fi_get_field(c, f, h.data(), n);
    if (st) return st;
    HIPCHECK(c, hipMemcpy(d, h.data(), n * sizeof(double), hipMemcpyHostToDevice));
    return FI_OK;
}

int fi_set_ice_snapshots(fi_ctx* c, const double* t, int32_t nt, const double* H, int32_t flat) {
    if (!c || !t || !H || nt < 1) return fail(c, FI_ERR_INVALID, "fi_set_ice_snapshots: bad args");
    for (int k = 1; k < nt; ++k)
        if (!(t[k] > t[k - 1])) return fail(c, FI_ERR_INVALID, "ice times must increase");
    c->ice_mode = ICE_SNAPSHOTS;
    c->ice_t.assign(t, t + nt);
    c->ice_H.assign(H, H + (size_t)nt * c->N);
    c->ice_a.assign(nt, 1.0);
    c->ice_flat = flat;
    return FI_OK;
}

int fi_set_timer_history(fi_ctx* c, int64_t capacity) {
    if (!c || capacity < 0) return fail(c, FI_ERR_INVALID, "fi_set_timer_history: bad args");
    if (c->initialized) return fail(c, FI_ERR_STATE, "fi_set_timer_history: call it before fi_init");
    c->hist_cap = capacity;
    return FI_OK;
}

int fi_get_timer_history(fi_ctx* c, double* t, double* seconds, int64_t capacity, int64_t* n_total) {
    if (!c || !n_total || capacity < 0 || (capacity > 0 && (!t || !seconds)))
        return fail(c, FI_ERR_INVALID, "fi_get_timer_history: bad args");
    *n_total = 0;
    if (!c->initialized || !c->hist_t.p) return FI_OK;
    HIPCHECK(c, hipSetDevice(c->dev));
    HIPCHECK(c, hipStreamSynchronize(c->stream));
    const long long n = c->ctl.nrhs;
    const size_t m = (size_t)std::min(n, c->hist_cap);
    std::vector<double> tv(m);
    std::vector<unsigned long long> clk(m);
    if (m) {
        HIPCHECK(c, hipMemcpy(tv.data(), c->hist_t.p, m * sizeof(double), hipMemcpyDeviceToHost));
        HIPCHECK(c, hipMemcpy(clk.data(), c->hist_clk.p, m * sizeof(unsigned long long), hipMemcpyDeviceToHost));
    }
    int khz = 0;
    HIPCHECK(c, hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, c->dev));
    if (khz <= 0) return fail(c, FI_ERR_DEVICE, "fi_get_timer_history: no wall-clock rate");
    // t_computation! keeps the calls past t_span[0]; run! subtracts the first one's time
    int64_t j = 0;
    unsigned long long clk0 = 0;
    for (size_t k = 0; k < m; ++k) {
        if (!(tv[k] > c->ctl.t0)) continue;
        if (j == 0) clk0 = clk[k];
        if (j < capacity) {
            t[j] = tv[k];
            seconds[j] = (double)(clk[k] - clk0) / (1e3 * (double)khz);
        }
        ++j;
    }
    *n_total = j;
    return FI_OK;
}

int fi_set_sea_surface_snapshots(fi_ctx* c, const double* t, int32_t nt, const double* X, int32_t flat) {
    if (!c || !t || !X || nt < 1) return fail(c, FI_ERR_INVALID, "fi_set_sea_surface_snapshots: bad args");
    for (int k = 1; k < nt; ++k)
        if (!(t[k] > t[k - 1])) return fail(c, FI_ERR_INVALID, "sea-surface times must increase");
    c->ss_t.assign(t, t + nt);
    c->ss_X.assign(X, X + (size_t)nt * c->N);
    c->ss_flat = flat;
    return FI_OK;
}

int fi_set_ice_separable(fi_ctx* c, const double* t, const double* a, int32_t nt, const double* H, int32_t flat) {
    if (!c || !t || !a || !H || nt < 1) return fail(c, FI_ERR_INVALID, "fi_set_ice_separable: bad args");
    for (int k = 1; k < nt; ++k)
        if (!(t[k] > t[k - 1])) return fail(c, FI_ERR_INVALID, "ice times must increase");
    c->ice_mode = ICE_SEPARABLE;
    c->ice_t.assign(t, t + nt);
    c->ice_a.assign(a, a + nt);
    c->ice_H.assign(H, H + c->N);
    c->ice_flat = flat;
    return FI_OK;
}

int fi_set_bsl(fi_ctx* c, int32_t rule, int32_t update, double A_ref, const double* x, const double* y,
               int32_t n, int32_t flat) {
    if (!c) return FI_ERR_INVALID;
    if (rule < FI_BSL_CONSTANT || rule > FI_BSL_IMPOSED) return fail(c, FI_ERR_INVALID, "fi_set_bsl: unknown rule");
    if (update != FI_BSL_INTERNAL_UPDATE && update != FI_BSL_EXTERNAL_UPDATE)
        return fail(c, FI_ERR_INVALID, "fi_set_bsl: unknown update");
    if ((rule == FI_BSL_CONSTANT_OCEAN_SURFACE || rule == FI_BSL_PIECEWISE_CONSTANT) && !(A_ref > 0))
        return fail(c, FI_ERR_INVALID, "fi_set_bsl: the reference ocean surface must be positive");
    const bool table = rule == FI_BSL_PIECEWISE_CONSTANT || rule == FI_BSL_IMPOSED;
    if (table) {
        if (!x || !y || n < 1) return fail(c, FI_ERR_INVALID, "fi_set_bsl: the rule needs a table");
        for (int k = 1; k < n; ++k)
            if (!(x[k] > x[k - 1])) return fail(c, FI_ERR_INVALID, "fi_set_bsl: table knots must increase");
    }
    c->bsl_rule = rule;
    c->bsl_update = update;
    c->bsl_A_ref = A_ref > 0 ? A_ref : 3.625e14;
    c->bsl_flat = flat;
    c->bsl_x.assign(table ? x : nullptr, table ? x + n : nullptr);
    c->bsl_y.assign(table ? y : nullptr, table ? y + n : nullptr);
    return FI_OK;
}

int fi_set_output_times(fi_ctx* c, const double* t, int32_t n) {
    if (!c || n < 0 || (n > 0 && !t)) return fail(c, FI_ERR_INVALID, "fi_set_output_times: bad args");
    c->tout.assign(t, t + n);
    std::sort(c->tout.begin(), c->tout.end());
    c->kout = 0;
    return FI_OK;
}

int fi_init(fi_ctx* c, double t0, double t_end) {
    if (!c) return FI_ERR_INVALID;
    const fi_config& g = c->cfg;
    const size_t N = c->N;
    (void)hipSetDevice(c->dev);
    if (!c->elra && (!c->has_D || !c->has_eta || !c->has_spinv))
        return fail(c, FI_ERR_INVALID, "litho rigidity, effective viscosity and scaled_pseudodiff_inv required");
    if (c->elra && (!c->has_gv || !c->has_tau)) return fail(c, FI_ERR_INVALID, "viscous Green kernel and tau required");
    if (c->implicit && !(g.dt_min > 0))
        return fail(c, FI_ERR_INVALID, "the implicit step needs dt_min (DiffEqOptions.dt_min with Euler)");
    if (g.lithosphere != FI_RIGID_LITHOSPHERE && !c->has_ge) return fail(c, FI_ERR_INVALID, "elastic Green kernel required");
    if (g.sea_surface == FI_LATERALLY_VARIABLE_SEA_SURFACE && !c->has_gg)
        return fail(c, FI_ERR_INVALID, "sea-surface Green kernel required");
    if (g.sea_surface == FI_IMPOSED_SEA_SURFACE && c->ss_t.empty())
        return fail(c, FI_ERR_INVALID, "ImposedSeaSurface needs dz_ss snapshots (fi_set_sea_surface_snapshots)");
    if (g.sealevel_load == FI_INTERACTIVE_SEALEVEL_LOAD) {
        double s = 0;
        for (double m : c->mask) s += m;
        if (s > 0.6 * c->nx * c->ny)
            return fail(c, FI_ERR_INVALID, "Mask defining regions of active load must not cover more than 60% "
                                           "of the cells when using an interactive sea level.");
    }
    const int nx = c->nx, ny = c->ny;
    // FFT plans and twiddle tables
    std::vector<int> lens = {nx, ny, c->nfx, c->nfy};
    std::vector<int> offs;
    std::vector<float2> tw32;
    for (int L : lens) {
        offs.push_back((int)tw32.size());
        for (int k = 0; k < L; ++k) {
            double a = -2.0 * M_PI * (double)k / (double)L;
            tw32.push_back(make_float2((float)std::cos(a), (float)std::sin(a)));
        }
    }
    std::vector<double2> tw64;
    int off64x = 0, off64y = 0;
    for (int q = 0; q < 2; ++q) {
        int L = q ? c->nfy : c->nfx;
        (q ? off64y : off64x) = (int)tw64.size();
        for (int k = 0; k < L; ++k) {
            double a = -2.0 * M_PI * (double)k / (double)L;
            tw64.push_back(make_double2(std::cos(a), std::sin(a)));
        }
    }
    FftPlan px, py, pfx, pfy, pfx64, pfy64;
    if (!plan_fft(nx, offs[0], px) || !plan_fft(ny, offs[1], py) || !plan_fft(c->nfx, offs[2], pfx) ||
        !plan_fft(c->nfy, offs[3], pfy) || !plan_fft(c->nfx, off64x, pfx64) || !plan_fft(c->nfy, off64y, pfy64))
        return fail(c, FI_ERR_UNSUPPORTED, "FFT length not 2,3,5,7-smooth");
    if (upload(c, c->tw32, tw32) || upload(c, c->tw64, tw64)) return FI_ERR_DEVICE;

    // spectral multiplier spy/(nx ny) * even part of scaled_pseudodiff_inv, transposed [kx][ky];
    // the DC term is zeroed: the corner BC on dudt removes any constant exactly
    if (c->implicit) {
        // u <- a u + b F per Fourier index (src/deformation.jl:58-79): nabla = 2 eta kappa R =
        // 2 eta / scaled_pseudodiff_inv, beta = rho_um g + D kappa^4, dt = dt_min in seconds;
        // stored as the even parts of a - 1 and b (the real part of the inverse transform of a
        // real field), transposed [kx][ky] and scaled by 1/(nx ny)
        const std::vector<double> kap = pseudodiff_table(nx, ny, c->cfg.Wx, c->cfg.Wy);
        const double dt = g.dt_min * g.seconds_per_year, sc = 1.0 / ((double)nx * ny);
        auto coef = [&](size_t p, double& am1, double& b) {
            const double k2 = kap[p] * kap[p];
            const double nabla = 2.0 * c->eta[p] / c->spinv[p];
            const double beta = g.rho_uppermantle * g.g + c->D[p] * k2 * k2;
            am1 = -dt * beta / (nabla + 0.5 * dt * beta);
            b = dt / (nabla + 0.5 * dt * beta);
        };
        std::vector<float> A((size_t)c->nhx * ny), B((size_t)c->nhx * ny);
        for (int kx = 0; kx < c->nhx; ++kx)
            for (int ky = 0; ky < ny; ++ky) {
                const int mkx = (nx - kx) % nx, mky = (ny - ky) % ny;
                double a0, b0, a1, b1;
                coef(kx + (size_t)ky * nx, a0, b0);
                coef(mkx + (size_t)mky * nx, a1, b1);
                A[(size_t)kx * ny + ky] = (float)(0.5 * (a0 + a1) * sc);
                B[(size_t)kx * ny + ky] = (float)(0.5 * (b0 + b1) * sc);
            }
        if (upload(c, c->imp_am1, A) || upload(c, c->imp_b, B)) return FI_ERR_DEVICE;
    } else if (!c->elra) {
        std::vector<float> mult((size_t)c->nhx * ny);
        const double sc = g.seconds_per_year / ((double)nx * ny);
        for (int kx = 0; kx < c->nhx; ++kx)
            for (int ky = 0; ky < ny; ++ky) {
                int mkx = (nx - kx) % nx, mky = (ny - ky) % ny;
                double m = 0.5 * (c->spinv[kx + (size_t)ky * nx] + c->spinv[mkx + (size_t)mky * nx]);
                mult[(size_t)kx * ny + ky] = (kx == 0 && ky == 0) ? 0.0f : (float)(m * sc);
            }
        if (upload(c, c->mult, mult)) return FI_ERR_DEVICE;
        // rhs K / (2 eta) (src/deformation.jl:95)
        std::vector<float> inv2eta(N);
        for (size_t p = 0; p < N; ++p)
            inv2eta[p] = (float)((c->has_dist ? c->dist[p] : 1.0) * (1.0 / (2.0 * c->eta[p])));
        if (upload(c, c->inv2eta, inv2eta) || upload(c, c->Df, to_f(c->D))) return FI_ERR_DEVICE;
    } else {
        // ELRA: 1/tau per cell and the equilibrium displacement, u_ref until the first RHS
        // (CurrentState, src/state.jl:87)
        std::vector<float> itau(N);
        for (size_t p = 0; p < N; ++p) itau[p] = (float)(1.0 / c->tau[p]);
        if (upload(c, c->inv_tau_d, itau) || upload(c, c->ueq, to_f(c->u_ref))) return FI_ERR_DEVICE;
    }
    if (c->has_dist) {
        std::vector<float> kinv(N), k2(N);
        for (size_t p = 0; p < N; ++p) {
            kinv[p] = (float)(1.0 / c->dist[p]);
            k2[p] = (float)(c->dist[p] * c->dist[p]);
        }
        if (upload(c, c->kinv_d, kinv) || upload(c, c->k2_d, k2)) return FI_ERR_DEVICE;
    }
    std::vector<unsigned char> m8(N);
    for (size_t p = 0; p < N; ++p) m8[p] = c->mask[p] != 0;
    if (upload(c, c->maskd, m8)) return FI_ERR_DEVICE;

    // reference state (src/simulation.jl:132-152)
    c->t_end = t_end;
    c->H_ice_ref.assign(N, 0.0);
    if (c->has_href) c->H_ice_ref = c->H_ref_in;
    else if (c->ice_mode != ICE_EXTERNAL) {
        int st = host_ice(c, t0, c->H_ice_ref);
        if (st) return st;
    }
    const double rsi = g.rho_seawater / g.rho_ice;
    c->z_ss_ref.resize(N);
    c->H_af_ref.resize(N);
    c->H_water_ref.resize(N);
    std::vector<unsigned char> mg(N), mo(N);
    for (size_t p = 0; p < N; ++p) {
        c->z_ss_ref[p] = g.bsl_z + c->dz_ss_ref[p];
        double h = hab_h(c->H_ice_ref[p], c->z_b_ref[p], c->z_ss_ref[p], rsi);
        c->H_af_ref[p] = h;
        mg[p] = h > 0;
        mo[p] = (c->z_ss_ref[p] - c->z_b_ref[p] > 0) && !mg[p];
        double b = std::max(c->z_ss_ref[p] - c->z_b_ref[p], 0.0), H = c->H_ice_ref[p];
        c->H_water_ref[p] = (H <= 1 ? b : 0.0) + ((!mg[p] && H > 1) ? (b - H * (g.rho_ice / g.rho_seawater)) : 0.0);
    }
    if (c->ice_mode == ICE_EXTERNAL) c->H_ext_host = c->H_ice_ref;
    // device state: CurrentState(domain, ref, z_bsl) (src/state.jl:80-103)
    HIPCHECK(c, c->U.alloc(3 * N));
    for (int b = 0; b < 3; ++b)
        HIPCHECK(c, hipMemcpy(c->U.p + b * N, c->u_ref.data(), N * sizeof(double), hipMemcpyHostToDevice));
    HIPCHECK(c, c->K.alloc(4 * N));
    HIPCHECK(c, hipMemset(c->K.p, 0, 4 * N * sizeof(float)));
    HIPCHECK(c, c->kcorner.alloc(16));
    HIPCHECK(c, hipMemset(c->kcorner.p, 0, 16 * sizeof(float)));
    // two spectrum buffers, each with its columns padded to a multiple of 4 (blocked [kx/4][j][4])
    HIPCHECK(c, c->spec.alloc(2 * (size_t)((c->nhx + 3) & ~3) * ny));
    if (upload(c, c->H_af_ref_d, to_f(c->H_af_ref)) || upload(c, c->u_ref_d, to_f(c->u_ref)) ||
        upload(c, c->z_b_ref_d, to_f(c->z_b_ref)) || upload(c, c->z_ss_ref_d, to_f(c->z_ss_ref)) ||
        upload(c, c->ue_ref_d, to_f(c->ue_ref)) || upload(c, c->H_ext, to_f(c->H_ext_host.empty() ? c->H_ice_ref : c->H_ext_host)) ||
        upload(c, c->ue, to_f(c->ue_ref)) || upload(c, c->dz_ss, std::vector<float>(N, 0.0f)) ||
        upload(c, c->z_ss, to_f(c->z_ss_ref)) || upload(c, c->H_water, to_f(c->H_water_ref)) ||
        upload(c, c->mg, mg) || upload(c, c->mo, mo) || upload(c, c->LW, std::vector<float>(N, 0.0f)) ||
        upload(c, c->sig_load_prev, std::vector<float>(N, 0.0f)) ||
        upload(c, c->sig_geo_prev, std::vector<float>(N, 0.0f)))
        return FI_ERR_DEVICE;
    {
        // column anomalies start at zero (ColumnAnomalies(domain), src/state.jl:11-14):
        // LT = sigma_litho - rho_um u_ref with sigma_litho = 0
        std::vector<float> LT(N), zoff(N);
        for (size_t p = 0; p < N; ++p) {
            LT[p] = (float)(-g.rho_uppermantle * c->u_ref[p]);
            zoff[p] = (float)(c->z_b_ref[p] + c->ue_ref[p] - c->z_ss_ref[p]);
        }
        if (upload(c, c->LT, LT) || upload(c, c->zoff, zoff) || upload(c, c->zoff_trig, zoff)) return FI_ERR_DEVICE;
    }
    HIPCHECK(c, c->ue_raw.alloc(N));
    HIPCHECK(c, c->dz_raw.alloc(N));
    HIPCHECK(c, c->ext_corners.alloc(8));
    HIPCHECK(c, hipMemset(c->ext_corners.p, 0, 8 * sizeof(float)));
    // padded half spectra, blocked [kx/4][row][4] like the n x n spectrum: columns padded to a
    // multiple of 4
    const size_t nfhx4 = (size_t)((c->nfhx + 3) & ~3);
    HIPCHECK(c, c->pA.alloc(nfhx4 * ny));
    HIPCHECK(c, c->pB.alloc(nfhx4 * ny));
    HIPCHECK(c, c->oA.alloc(nfhx4 * (ny + 2)));
    HIPCHECK(c, c->oB.alloc(nfhx4 * (ny + 2)));
    if (!c->bsl_x.empty()) {
        if (upload(c, c->bsl_x_d, c->bsl_x) || upload(c, c->bsl_y_d, c->bsl_y)) return FI_ERR_DEVICE;
    } else {
        HIPCHECK(c, c->bsl_x_d.alloc(1));
        HIPCHECK(c, c->bsl_y_d.alloc(1));
    }
    if (g.sea_surface == FI_IMPOSED_SEA_SURFACE) {
        if (upload(c, c->ss_t_d, c->ss_t) || upload(c, c->ss_X_d, c->ss_X)) return FI_ERR_DEVICE;
    }
    if (c->ice_mode != ICE_EXTERNAL) {
        if (upload(c, c->ice_t_d, c->ice_t) || upload(c, c->ice_a_d, c->ice_a) || upload(c, c->ice_H_d, to_f(c->ice_H)))
            return FI_ERR_DEVICE;
    } else {
        HIPCHECK(c, c->ice_t_d.alloc(1));
        HIPCHECK(c, c->ice_a_d.alloc(1));
        HIPCHECK(c, c->ice_H_d.alloc(1));
    }

    // Green spectra in FP64
    HIPCHECK(c, c->ge_hat.alloc((size_t)c->nfhx * c->nfy));
    HIPCHECK(c, c->gg_hat.alloc((size_t)c->nfhx * c->nfy));
    HIPCHECK(c, c->gv_hat.alloc(c->elra ? (size_t)c->nfhx * c->nfy : 1));
    HIPCHECK(c, c->gtmp.alloc((size_t)c->nfhx * ny));
    {
        DBuf<double> gd;
        const double nrm = 1.0 / ((double)c->nfx * c->nfy);
        for (int q = 0; q < 3; ++q) {
            if (q == 2 && !c->elra) break;
            const std::vector<double>* G = q == 2 ? &c->gv : q ? &c->gg : &c->ge;
            bool have = q == 2 ? c->has_gv : q ? c->has_gg : c->has_ge;
            DBuf<double2>& out = q == 2 ? c->gv_hat : q ? c->gg_hat : c->ge_hat;
            if (!have) {
                HIPCHECK(c, hipMemset(out.p, 0, out.n * sizeof(double2)));
                continue;
            }
            if (upload(c, gd, *G)) return FI_ERR_DEVICE;
            size_t l1 = 2 * (size_t)c->nfx * sizeof(double2), l2 = 2 * (size_t)c->nfy * sizeof(double2);
            HIPCHECK(c, hipFuncSetAttribute((const void*)k_green_rows, hipFuncAttributeMaxDynamicSharedMemorySize, (int)l1));
            HIPCHECK(c, hipFuncSetAttribute((const void*)k_green_cols, hipFuncAttributeMaxDynamicSharedMemorySize, (int)l2));
            hipLaunchKernelGGL(k_green_rows, dim3(ny), dim3(NT), l1, c->stream, gd.p, nx, ny, c->nfx, c->nfhx, pfx64,
                               c->tw64.p, c->gtmp.p);
            HIPCHECK(c, hipGetLastError());
            // the sea-surface input is the mass anomaly divided by the cell area: A goes in the kernel
            double scale = nrm * (q == 1 ? c->A : 1.0);
            hipLaunchKernelGGL(k_green_cols, dim3(c->nfhx), dim3(NT), l2, c->stream, c->gtmp.p, ny, c->nfy, pfy64,
                               c->tw64.p, scale, out.p);
            HIPCHECK(c, hipGetLastError());
            HIPCHECK(c, hipStreamSynchronize(c->stream));
        }
    }

    // launch geometry
    KP& kp = c->kp;
    std::memset(&kp, 0, sizeof(kp));
    kp.nx = nx;
    kp.ny = ny;
    kp.nhx = c->nhx;
    kp.nfx = c->nfx;
    kp.nfy = c->nfy;
    kp.nfhx = c->nfhx;
    kp.nrowitems = (ny + 1) / 2;
    // spectral columns per columns-phase item: one, unless there are more columns than CUs (one
    // workgroup per CU), where a workgroup transforms its columns as one batch instead of one
    // after another (N=1024: 513 columns, 3 per item); bitwise neutral
    {
        int ncu = 0;
        HIPCHECK(c, hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, c->dev));
        kp.colnb = (ncu > 0 && c->nhx > ncu) ? (c->nhx + ncu - 1) / ncu : 1;
        // likewise row pairs per item of the standalone inverse transform (N=1024: 2)
        kp.c2rnb = (ncu > 0 && kp.nrowitems > ncu) ? (kp.nrowitems + ncu - 1) / ncu : 1;
    }
    kp.ncolitems = (c->nhx + kp.colnb - 1) / kp.colnb;
    kp.nc2ritems = (kp.nrowitems + kp.c2rnb - 1) / kp.c2rnb;
    // relaxed mantle: one row pair per rows/inverse item, two columns per columns item (an
    // ensemble batches more per item for its team size)
    kp.erb = 1;
    kp.nerowitems = kp.nrowitems;
    kp.ecb = 2;
    kp.necolitems = (c->nfhx + 1) / 2;
    kp.i1 = c->i1;
    kp.i2 = c->i2;
    kp.j1 = c->j1;
    kp.j2 = c->j2;
    kp.dx = c->dx;
    kp.dy = c->dy;
    kp.A = c->A;
    kp.inv_dx = 1.0 / c->dx;
    kp.inv_dy = 1.0 / c->dy;
    kp.inv_dx2 = 1.0 / (c->dx * c->dx);
    kp.inv_dy2 = 1.0 / (c->dy * c->dy);
    kp.g = g.g;
    kp.spy = g.seconds_per_year;
    kp.rho_ice = g.rho_ice;
    kp.rho_sw_ice = g.rho_seawater / g.rho_ice;
    kp.rho_ice_sw = g.rho_ice / g.rho_seawater;
    kp.rho_seawater = g.rho_seawater;
    kp.rho_water = g.rho_water;
    kp.rho_um = g.rho_uppermantle;
    kp.rho_litho = g.rho_litho;
    kp.nu = g.litho_poissonratio;
    kp.dt_sparse = g.dt_sparse;
    kp.reltol = g.reltol;
    kp.abstol = g.abstol;
    kp.vbc_x = g.viscous_x_border;
    kp.ebc_x = g.elastic_x_border;
    kp.sbc_x = g.ssh_x_border;
    kp.bsl_z = g.bsl_z;
    kp.N = (double)N;
    kp.rigid_litho = g.lithosphere == FI_RIGID_LITHOSPHERE;
    kp.rigid_mantle = g.mantle == FI_RIGID_MANTLE;
    kp.compressible_column = g.lithosphere_column == FI_COMPRESSIBLE_LITHOSPHERE_COLUMN;
    kp.lv_surface = g.sea_surface == FI_LATERALLY_VARIABLE_SEA_SURFACE;
    kp.interactive = g.sealevel_load == FI_INTERACTIVE_SEALEVEL_LOAD;
    kp.vol_goelzer = g.volume_contribution == FI_GOELZER_CONTRIBUTION;
    kp.den_goelzer = g.density_contribution == FI_GOELZER_CONTRIBUTION;
    kp.pov_goelzer = g.adjustment_contribution == FI_GOELZER_CONTRIBUTION;
    kp.bsl_rule = c->bsl_rule;
    kp.bsl_internal = c->bsl_update == FI_BSL_INTERNAL_UPDATE;
    kp.bsl_n = (int)c->bsl_x.size();
    kp.bsl_flat = c->bsl_flat;
    kp.bsl_A_ref = c->bsl_A_ref;
    kp.bsl_x = c->bsl_x_d.p;
    kp.bsl_y = c->bsl_y_d.p;
    kp.ice_mode = c->ice_mode;
    kp.ice_nt = (int)c->ice_t.size();
    kp.ice_flat = c->ice_flat;
    kp.ice_t = c->ice_t_d.p;
    kp.ice_a = c->ice_a_d.p;
    kp.ice_H = c->ice_H_d.p;
    kp.imposed_ss = g.sea_surface == FI_IMPOSED_SEA_SURFACE;
    kp.ss_nt = (int)c->ss_t.size();
    kp.ss_flat = c->ss_flat;
    kp.ss_t = c->ss_t_d.p;
    kp.ss_X = c->ss_X_d.p;
    kp.Nl = (long)N;
    kp.Ub = c->U.p;
    kp.Kb = c->K.p;
    kp.kcorner = c->kcorner.p;
    kp.specb = c->spec.p;
    kp.specN = (long)((c->nhx + 3) & ~3) * ny;
    kp.mult = c->mult.p;
    kp.D = c->Df.p;
    kp.inv2eta = c->inv2eta.p;
    kp.Kinv = c->has_dist ? c->kinv_d.p : nullptr;
    kp.K2 = c->has_dist ? c->k2_d.p : nullptr;
    kp.mask = c->maskd.p;
    kp.H_af_ref = c->H_af_ref_d.p;
    kp.u_ref = c->u_ref_d.p;
    kp.z_b_ref = c->z_b_ref_d.p;
    kp.z_ss_ref = c->z_ss_ref_d.p;
    kp.ue_ref = c->ue_ref_d.p;
    kp.H_ext = c->H_ext.p;
    kp.ue = c->ue.p;
    kp.dz_ss = c->dz_ss.p;
    kp.z_ss = c->z_ss.p;
    kp.zoff = c->zoff.p;
    kp.zoff_trig = c->zoff_trig.p;
    kp.LW = c->LW.p;
    kp.LT = c->LT.p;
    kp.H_water = c->H_water.p;
    kp.mg = c->mg.p;
    kp.mo = c->mo.p;
    kp.sig_load_prev = c->sig_load_prev.p;
    kp.sig_geo_prev = c->sig_geo_prev.p;
    kp.pA = c->pA.p;
    kp.pB = c->pB.p;
    kp.oA = c->oA.p;
    kp.oB = c->oB.p;
    kp.ue_raw = c->ue_raw.p;
    kp.dz_raw = c->dz_raw.p;
    kp.ext_corners = c->ext_corners.p;
    kp.ge_hat = c->ge_hat.p;
    kp.gg_hat = c->gg_hat.p;
    kp.elra = c->elra ? 1 : 0;
    kp.implicit = c->implicit ? 1 : 0;
    kp.imp_am1 = c->imp_am1.p;
    kp.imp_b = c->imp_b.p;
    kp.dt_min = g.dt_min;
    kp.gv_hat = c->gv_hat.p;
    kp.inv_tau = c->inv_tau_d.p;
    kp.ueq = c->ueq.p;
    kp.px = px;
    kp.py = py;
    kp.pfx = pfx;
    kp.pfy = pfy;
    kp.tw32 = c->tw32.p;
    kp.spin_limit = 2000000000ull;  // 20 s of the 100 MHz wall clock

    // LDS: twiddles of nx and ny for the whole kernel, then the largest phase scratch: rows phase
    // (u FP64 6 rows, 3 moment fields 4 rows, R2C pair, 6 rows of the spectral-stage derivative)
    // or the padded-grid phases (their twiddles staged in the scratch)
    const size_t twb = (size_t)(nx + ny) * 8;
    // FP64-moment rows phase also stages the rigidity (4 rows) and 7 per-cell fields (2 rows)
    size_t rows_d = (size_t)nx * (6 * 8 + 12 * 8 + 16 + 24 + 16 + 56), rows_f = (size_t)nx * (6 * 8 + 12 * 4 + 16 + 24);
    size_t other = std::max({(size_t)2 * kp.colnb * ny * 8, (size_t)2 * kp.c2rnb * nx * 8, (size_t)3 * c->nfx * 8,
                             (size_t)5 * c->nfy * 8});
    c->mt_double = nx <= 512 && twb + rows_d + (c->has_dist ? (size_t)nx * 24 : 0) <= 160 * 1024;   // FP64 moments, register prefetch sized for nx <= 512
    c->lds_dist = twb + std::max(rows_d + (size_t)nx * 24, other);   // also 6 rows of 1/K
    c->lds_d = c->has_dist ? c->lds_dist : twb + std::max(rows_d, other);
    c->lds_f = twb + std::max(rows_f, other);
    c->lds = c->mt_double ? c->lds_d : c->lds_f;
    if (c->elra) c->lds = twb + other;   // ELRA runs only the padded-grid phases
    if (c->implicit) c->lds = twb + std::max({other, (size_t)32 * nx, (size_t)32 * ny});
    if (c->lds > 160 * 1024 || nx > 1024) return fail(c, FI_ERR_UNSUPPORTED, "grid too large for one workgroup's LDS");
    if (!c->elra && !c->implicit && c->has_dist && (ny < 6 || nx < 3))
        return fail(c, FI_ERR_UNSUPPORTED, "map distortion needs a grid of at least 3 columns and 6 rows");
    const void* f = persistent_kernel(c->mt_double, c->has_dist, c->elra, c->implicit);
    HIPCHECK(c, hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)c->lds));
    int per_cu = 0, ncu = 0;
    HIPCHECK(c, hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, f, NT, c->lds));
    HIPCHECK(c, hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, c->dev));
    int maxg = per_cu * ncu;
    int want = std::max(kp.nrowitems, c->elra ? (c->nfhx + 1) / 2 : c->implicit ? c->nhx : kp.ncolitems);
    c->G = std::max(1, std::min(maxg, want));
    c->maxg = maxg;
    if (maxg < 1) return fail(c, FI_ERR_UNSUPPORTED, "persistent kernel does not fit on the device");
    // one error-norm and one volume partial per row pair, whatever the number of workgroups
    HIPCHECK(c, c->part.alloc(4 * (size_t)kp.nrowitems));
    HIPCHECK(c, c->vpart.alloc(4 * (size_t)kp.nrowitems));
    kp.part = c->part.p;
    kp.vpart = c->vpart.p;
    HIPCHECK(c, c->ctl_d.alloc(1));
    // words + 128-byte arrival and release slots per workgroup, for any grid size up to maxg
    HIPCHECK(c, c->bar_d.alloc(1 + 2 * (size_t)maxg));
    HIPCHECK(c, c->kp_d.alloc(1));
    kp.ctl = c->ctl_d.p;
    kp.bar = c->bar_d.p;
    kp.hist_t = nullptr;
    kp.hist_clk = nullptr;
    kp.hist_cap = 0;
    if (c->hist_cap > 0) {
        HIPCHECK(c, c->hist_t.alloc((size_t)c->hist_cap));
        HIPCHECK(c, c->hist_clk.alloc((size_t)c->hist_cap));
        kp.hist_t = c->hist_t.p;
        kp.hist_clk = c->hist_clk.p;
        kp.hist_cap = c->hist_cap;
    }
    kp.ptime = nullptr;
    if (getenv("FASTISO_PHASE_TIMING")) {
        HIPCHECK(c, c->ptime.alloc(32));
        HIPCHECK(c, hipMemset(c->ptime.p, 0, 32 * sizeof(unsigned long long)));
        kp.ptime = c->ptime.p;
    }

    std::memset(&c->ctl, 0, sizeof(Ctl));
    c->ctl.t0 = t0;
    c->ctl.t = t0;
    c->ctl.t_last = t0;
    c->ctl.z_bsl = g.bsl_z;
    c->ctl.A_bsl = c->bsl_A_ref;
    c->ctl.qold = PI_QOLDINIT;
    c->ctl.cur = 0;
    c->ctl.k1 = 0;
    c->ctl.last_k = 0;
    c->ctl.last_u = 0;
    c->initialized = true;
    c->integ = false;
    c->kout = 0;
    return FI_OK;
}

// diagnostic fields that the engine does not store are assembled from stored ones with the
// reference formulas (src/topography.jl:30-38, src/loads.jl:31-106)
int fi_get_field(fi_ctx* c, int32_t f, double* h, size_t n) {
    if (!c || !h || n != c->N) return fail(c, FI_ERR_INVALID, "fi_get_field: size mismatch");
    const size_t N = c->N;
    auto put = [&](const std::vector<double>& v) -> int {
        if (v.size() != N) return fail(c, FI_ERR_STATE, "fi_get_field: field not set");
        std::memcpy(h, v.data(), N * sizeof(double));
        return FI_OK;
    };
    switch (f) {
        case FI_LITHO_RIGIDITY: return put(c->D);
        case FI_EFFECTIVE_VISCOSITY: return put(c->eta);
        case FI_SCALED_PSEUDODIFF_INV: return put(c->spinv);
        case FI_MASK_ACTIVE: return put(c->mask);
        case FI_GREEN_ELASTIC: return put(c->ge);
        case FI_GREEN_GEOID: return put(c->gg);
        case FI_GREEN_VISCOUS: return put(c->gv);
        case FI_TAU: return put(c->tau);
        case FI_LITHO_THICKNESS: return put(c->litho_h);
        case FI_REF_U: return put(c->u_ref);
        case FI_REF_UE: return put(c->ue_ref);
        case FI_REF_DZ_SS: return put(c->dz_ss_ref);
        case FI_REF_Z_B: return put(c->z_b_ref);
        case FI_REF_H_ICE: return put(c->H_ice_ref);
        case FI_DISTORTION: return put(c->has_dist ? c->dist : std::vector<double>(c->N, 1.0));
        default: break;
    }
    if (!c->initialized) return fail(c, FI_ERR_STATE, "fi_get_field before fi_init");
    (void)hipSetDevice(c->dev);
    const fi_config& g = c->cfg;
    std::vector<double> v, u, ue, zss, H;
    int st = 0;
    auto umain = [&](std::vector<double>& out) -> int {
        int s = read_u(c, c->ctl.last_u, out);
        if (s) return s;
        if (c->ctl.last_u == 2) {  // fi_rhs input buffer is kept unmodified: apply the corner BC here
            double off = corner_offset(out.data(), c->nx, c->ny, g.viscous_x_border);
            for (double& x : out) x -= off;
        }
        return FI_OK;
    };
    switch (f) {
        case FI_U: if ((st = umain(v))) return st; return put(v);
        case FI_DUDT: if ((st = read_k(c, c->ctl.last_k, v))) return st; return put(v);
        case FI_UE: if ((st = read_f(c, c->ue.p, v))) return st; return put(v);
        case FI_DZ_SS: if ((st = read_f(c, c->dz_ss.p, v))) return st; return put(v);
        case FI_Z_SS: if ((st = read_f(c, c->z_ss.p, v))) return st; return put(v);
        case FI_H_WATER: if ((st = read_f(c, c->H_water.p, v))) return st; return put(v);
        case FI_U_X:
        case FI_U_Y: {
            if (c->litho_h.size() != N) return fail(c, FI_ERR_STATE, "fi_get_field: litho thickness not set");
            if (c->litho_h_d.n != N && upload(c, c->litho_h_d, to_f(c->litho_h))) return FI_ERR_DEVICE;
            if (c->uxy_d.n != N) HIPCHECK(c, c->uxy_d.alloc(N));
            const int nb = (int)std::min<size_t>((N + 255) / 256, 4096);
            const bool ax = f == FI_U_X;
            hipLaunchKernelGGL(k_thinplate, dim3(nb), dim3(256), 0, c->stream, c->U.p + (size_t)c->ctl.last_u * N,
                               c->ue.p, c->litho_h_d.p, c->has_dist ? c->kinv_d.p : nullptr, c->nx, c->ny,
                               ax ? c->dx : c->dy, ax ? 0 : 1, c->uxy_d.p);
            HIPCHECK(c, hipGetLastError());
            HIPCHECK(c, hipMemcpyAsync(h, c->uxy_d.p, N * sizeof(double), hipMemcpyDeviceToHost, c->stream));
            HIPCHECK(c, hipStreamSynchronize(c->stream));
            return FI_OK;
        }
        case FI_U_EQ:
            if (!c->elra) return put(c->u_ref);   // written only by the relaxed mantle's RHS
            if ((st = read_f(c, c->ueq.p, v))) return st;
            return put(v);
        case FI_CANOM_SEAWATER: if ((st = read_f(c, c->LW.p, v))) return st; return put(v);
        case FI_MASK_GROUNDED:
        case FI_MASK_OCEAN: {
            std::vector<unsigned char> m(N);
            HIPCHECK(c, hipMemcpy(m.data(), f == FI_MASK_GROUNDED ? c->mg.p : c->mo.p, N, hipMemcpyDeviceToHost));
            for (size_t p = 0; p < N; ++p) h[p] = m[p] ? 1.0 : 0.0;
            return FI_OK;
        }
        default: break;
    }
    if ((st = umain(u)) || (st = read_f(c, c->ue.p, ue)) || (st = read_f(c, c->z_ss.p, zss))) return st;
    if ((st = host_ice(c, c->ctl.t_last, H))) return st;
    v.resize(N);
    const double rsi = g.rho_seawater / g.rho_ice;
    for (size_t p = 0; p < N; ++p) {
        double zb = c->z_b_ref[p] + ue[p] + u[p];
        double haf = hab_h(H[p], zb, zss[p], rsi);
        double s_ice = g.rho_ice * (haf - c->H_af_ref[p]);
        double s_litho = g.rho_litho * (ue[p] - c->ue_ref[p]);
        double s_mant = g.rho_uppermantle * (u[p] - c->u_ref[p]);
        double m = c->mask[p];
        switch (f) {
            case FI_Z_B: v[p] = zb; break;
            case FI_H_ICE: v[p] = H[p]; break;
            case FI_H_AF: v[p] = haf; break;
            case FI_CANOM_ICE: v[p] = s_ice; break;
            case FI_CANOM_LITHO: v[p] = s_litho; break;
            case FI_CANOM_MANTLE: v[p] = s_mant; break;
            case FI_CANOM_LOAD:
            case FI_CANOM_FULL: v[p] = 0; break;
            default: return fail(c, FI_ERR_INVALID, "fi_get_field: unknown field");
        }
    }
    if (f == FI_CANOM_LOAD || f == FI_CANOM_FULL) {
        std::vector<double> sw;
        if ((st = read_f(c, c->LW.p, sw))) return st;
        for (size_t p = 0; p < N; ++p) {
            double zb = c->z_b_ref[p] + ue[p] + u[p];
            double haf = hab_h(H[p], zb, zss[p], rsi);
            double s_load = c->mask[p] * (g.rho_ice * (haf - c->H_af_ref[p]) + sw[p]);
            v[p] = f == FI_CANOM_LOAD ? s_load
                                      : s_load + c->mask[p] * (g.rho_litho * (ue[p] - c->ue_ref[p]) +
                                                               g.rho_uppermantle * (u[p] - c->u_ref[p]));
        }
    }
    return put(v);
}

int fi_rhs(fi_ctx* c, double t, double* u, double* dudt) {
    if (!c || !u || !dudt) return FI_ERR_INVALID;
    if (!c->initialized) return fail(c, FI_ERR_STATE, "fi_rhs before fi_init");
    (void)hipSetDevice(c->dev);
    const int ubuf = 2, slot = (c->ctl.k1 + 1) % 4;
    HIPCHECK(c, hipMemcpy(c->U.p + 2 * c->N, u, c->N * sizeof(double), hipMemcpyHostToDevice));
    int st = rhs_on_device(c, t, ubuf, slot);
    if (st) return st;
    if (c->implicit) {   // u is replaced by the implicit step; dudt is not written
        HIPCHECK(c, hipMemcpy(u, c->U.p + 2 * c->N, c->N * sizeof(double), hipMemcpyDeviceToHost));
        c->ctl.last_u = ubuf;
        return FI_OK;
    }
    std::vector<double> k;
    if ((st = read_k(c, slot, k))) return st;
    std::memcpy(dudt, k.data(), c->N * sizeof(double));
    double off = corner_offset(u, c->nx, c->ny, c->cfg.viscous_x_border);
    for (size_t p = 0; p < c->N; ++p) u[p] -= off;
    c->ctl.last_u = ubuf;
    return FI_OK;
}

int fi_rhs_device(fi_ctx* c, double t, void* u, void* dudt) {
    if (!c || !u || !dudt) return FI_ERR_INVALID;
    if (!c->initialized) return fail(c, FI_ERR_STATE, "fi_rhs_device before fi_init");
    (void)hipSetDevice(c->dev);
    const int ubuf = 2, slot = (c->ctl.k1 + 1) % 4;
    HIPCHECK(c, hipMemcpyAsync(c->U.p + 2 * c->N, u, c->N * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
    int st = rhs_on_device(c, t, ubuf, slot);
    if (st) return st;
    if (c->implicit) {   // u is replaced by the implicit step; dudt is not written
        HIPCHECK(c, hipMemcpyAsync(u, c->U.p + 2 * c->N, c->N * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
        HIPCHECK(c, hipStreamSynchronize(c->stream));
        c->ctl.last_u = ubuf;
        return FI_OK;
    }
    float kc[4];
    double uc[4];
    HIPCHECK(c, hipMemcpy(kc, c->kcorner.p + 4 * slot, sizeof(kc), hipMemcpyDeviceToHost));
    const size_t N = c->N;
    const size_t idx[4] = {0, (size_t)c->nx - 1, (size_t)(c->ny - 1) * c->nx, (size_t)(c->ny - 1) * c->nx + c->nx - 1};
    for (int q = 0; q < 4; ++q) HIPCHECK(c, hipMemcpy(&uc[q], (double*)u + idx[q], sizeof(double), hipMemcpyDeviceToHost));
    double uoff = ((0.25 * uc[0] + 0.25 * uc[1]) + (0.25 * uc[2] + 0.25 * uc[3])) - c->cfg.viscous_x_border;
    double koff = kslot_offset(c, slot, kc);
    int nb = (int)std::min<size_t>((N + 255) / 256, 4096);
    hipLaunchKernelGGL(k_dudt_out, dim3(nb), dim3(256), 0, c->stream, c->K.p + slot * N, koff, (long)N, (double*)dudt);
    hipLaunchKernelGGL(k_sub_offset, dim3(nb), dim3(256), 0, c->stream, (double*)u, uoff, (long)N);
    HIPCHECK(c, hipGetLastError());
    HIPCHECK(c, hipStreamSynchronize(c->stream));
    c->ctl.last_u = ubuf;
    return FI_OK;
}

int fi_sparse_update(fi_ctx* c, double t) {
    if (!c) return FI_ERR_INVALID;
    if (!c->initialized) return fail(c, FI_ERR_STATE, "fi_sparse_update before fi_init");
    (void)hipSetDevice(c->dev);
    c->ctl.t_rhs = t;
    return launch(c, MODE_SPARSE_ONLY, c->ctl.last_u, (c->ctl.k1 + 1) % 4, 0, 1);
}

// init_problem! + FSAL + initial step (src/simulation.jl:293-311)
int fi_integrator_init(fi_ctx* c) {
    if (!c) return FI_ERR_INVALID;
    if (!c->initialized) return fail(c, FI_ERR_STATE, "fi_integrator_init before fi_init");
    (void)hipSetDevice(c->dev);
    const fi_config& g = c->cfg;
    const size_t N = c->N;
    int st;
    // update_V_af!, update_V_den!, update_V_pov! on the current state
    {
        std::vector<double> haf(N), H(N), zb(N), zss(N);
        if ((st = fi_get_field(c, FI_H_AF, haf.data(), N)) || (st = fi_get_field(c, FI_H_ICE, H.data(), N)) ||
            (st = fi_get_field(c, FI_Z_B, zb.data(), N)) || (st = fi_get_field(c, FI_Z_SS, zss.data(), N)))
            return st;
        double s0 = 0, s1 = 0, s2 = 0;
        for (size_t p = 0; p < N; ++p) {
            const double a = c->has_dist ? c->A * (c->dist[p] * c->dist[p]) : c->A;   // dx dy K^2
            s0 += haf[p] * a;
            s1 += H[p] * a;
            s2 += std::max(zss[p] - zb[p], 0.0) * a;
        }
        c->ctl.V_af = g.volume_contribution ? s0 * g.rho_ice / g.rho_seawater : 0.0;
        c->ctl.V_den = g.density_contribution ? s1 * (g.rho_ice / g.rho_water - g.rho_ice / g.rho_seawater) : 0.0;
        c->ctl.V_pov = g.adjustment_contribution ? s2 : 0.0;
    }
    // u0 = now.u with the corner BC (update_diagnostics! mutates it)
    std::vector<double> u0;
    if ((st = read_u(c, c->ctl.last_u, u0))) return st;
    double off = corner_offset(u0.data(), c->nx, c->ny, g.viscous_x_border);
    for (double& x : u0) x -= off;
    HIPCHECK(c, hipMemcpy(c->U.p, u0.data(), N * sizeof(double), hipMemcpyHostToDevice));
    const double t0 = c->ctl.t0;
    c->ctl.cur = 0;
    c->ctl.last_u = 0;
    if ((st = rhs_on_device(c, t0, 0, 1))) return st;   // init_problem!
    if ((st = rhs_on_device(c, t0, 0, 0))) return st;   // FSAL: k1 = f(u0, t0)
    if (c->implicit) {
        // Euler with dt = dt_min (src/simulation.jl:244-250): both calls above replaced U[0] by an
        // implicit step; the derivative slots stay zero
        c->ctl.dt = g.dt_min;
        c->ctl.t = t0;
        c->ctl.k1 = 0;
        c->ctl.nacc = c->ctl.nrej = 0;
        c->integ = true;
        c->k1_stale = false;
        c->kout = 0;
        return FI_OK;
    }
    // initial step (Hairer, Norsett & Wanner, Solving ODEs I, II.4)
    std::vector<double> f0, f1;
    if ((st = read_k(c, 0, f0))) return st;
    std::vector<double> sk(N), tq(N);
    auto rms = [&](const std::vector<double>& a) {
        double s = 0;
        for (double x : a) s += x * x;
        return std::sqrt(s / (double)a.size());
    };
    for (size_t p = 0; p < N; ++p) sk[p] = g.abstol + std::fabs(u0[p]) * g.reltol;
    for (size_t p = 0; p < N; ++p) tq[p] = u0[p] / sk[p];
    double d0 = rms(tq);
    for (size_t p = 0; p < N; ++p) tq[p] = f0[p] / sk[p];
    double d1 = rms(tq);
    double dtmax = c->t_end - t0;
    double dt0 = (d0 < 1e-5 || d1 < 1e-5) ? 1e-6 : 0.01 * (d0 / d1);
    dt0 = std::min(dt0, dtmax);
    std::vector<double> u1(N);
    for (size_t p = 0; p < N; ++p) u1[p] = u0[p] + dt0 * f0[p];
    HIPCHECK(c, hipMemcpy(c->U.p + 2 * N, u1.data(), N * sizeof(double), hipMemcpyHostToDevice));
    if ((st = rhs_on_device(c, t0 + dt0, 2, 1))) return st;
    if ((st = read_k(c, 1, f1))) return st;
    for (size_t p = 0; p < N; ++p) tq[p] = (f1[p] - f0[p]) / sk[p];
    double d2 = rms(tq) / dt0;
    double mx = std::max(d1, d2);
    double dt1 = (mx <= 1e-15) ? std::max(1e-6, dt0 * 1e-3) : std::pow(10.0, -(2.0 + std::log10(mx)) / 3.0);
    c->ctl.dt = std::min(std::min(100 * dt0, dt1), dtmax);
    c->ctl.t = t0;
    c->ctl.qold = PI_QOLDINIT;
    c->ctl.k1 = 0;
    c->ctl.cur = 0;
    c->ctl.last_u = 0;
    c->ctl.nacc = c->ctl.nrej = 0;
    c->integ = true;
    c->k1_stale = false;
    c->kout = 0;
    return FI_OK;
}

int fi_integrate(fi_ctx* c, double t_target, fi_output_cb cb, void* user) {
    if (!c) return FI_ERR_INVALID;
    if (!c->integ) return fail(c, FI_ERR_STATE, "fi_integrate before fi_integrator_init");
    if (t_target < c->ctl.t) return fail(c, FI_ERR_INVALID, "fi_integrate: target before current time");
    (void)hipSetDevice(c->dev);
    int st;
    while (c->ctl.t < t_target) {
        if (c->k1_stale) {
            if ((st = rhs_on_device(c, c->ctl.t, c->ctl.cur, c->ctl.k1))) return st;
            c->ctl.last_u = c->ctl.cur;
            c->k1_stale = false;
        }
        double stop = t_target;
        for (double to : c->tout)
            if (to > c->ctl.t) { stop = std::min(stop, to); break; }
        c->ctl.stop = stop;
        c->ctl.next_out = (c->kout < (long long)c->tout.size()) ? c->tout[c->kout] : INFINITY;
        if ((st = launch(c, MODE_INTEGRATE, 0, 0, 20000, 0))) return st;
        c->ctl.last_u = c->ctl.cur;
        if (c->ctl.steps == 1 && c->kout < (long long)c->tout.size() && c->ctl.t >= c->tout[c->kout]) {
            if (cb) cb(user, c->ctl.t, (int32_t)c->kout);
            c->kout += 1;
            // a DiscreteCallback marks u as modified: the FSAL derivative is re-evaluated
            if ((st = rhs_on_device(c, c->ctl.t, c->ctl.cur, c->ctl.k1))) return st;
            c->ctl.last_u = c->ctl.cur;
        }
    }
    return FI_OK;
}

int fi_set_time(fi_ctx* c, double t) {
    if (!c) return FI_ERR_INVALID;
    if (!c->integ) return fail(c, FI_ERR_STATE, "fi_set_time before fi_integrator_init");
    // its load-storing call would advance u under the implicit step
    if (c->implicit) return fail(c, FI_ERR_UNSUPPORTED, "fi_set_time: not with the implicit step");
    if (!(t >= c->ctl.t0)) return fail(c, FI_ERR_INVALID, "fi_set_time: t before the start time");
    (void)hipSetDevice(c->dev);
    const long long due = (long long)std::floor((t - c->ctl.t0) / c->cfg.dt_sparse);
    // store the loads of (t, u) with a call that cannot trigger the sparse update
    c->ctl.count = due + 1;
    int st = rhs_on_device(c, t, c->ctl.cur, (c->ctl.k1 + 1) % 4);
    if (st) return st;
    // the sparse update at t with those loads; the next one is due one interval later
    c->ctl.t_rhs = t;
    if ((st = launch(c, MODE_SPARSE_ONLY, c->ctl.cur, (c->ctl.k1 + 1) % 4, 0, 1))) return st;
    c->ctl.count = due + 1;
    c->ctl.t = t;
    c->ctl.last_u = c->ctl.cur;
    c->k1_stale = true;
    return FI_OK;
}

}  // extern "C"

namespace {

// RKC2 (Sommeijer, Shampine & Verwer 1997): with w0 = 1 + eps/s^2 (eps = 2/13), Chebyshev values
// T_j(w0) and their first two derivatives, w1 = T_s'/T_s'', b_j = T_j''/T_j'^2 (b_0 = b_1 = b_2):
//   Y_1 = Y_0 + mut_1 h F_0,  Y_j = (1 - mu_j - nu_j) Y_0 + mu_j Y_{j-1} + nu_j Y_{j-2}
//                                   + mut_j h F_{j-1} + gt_j h F_0,
//   mu_j = 2 b_j w0/b_{j-1}, nu_j = -b_j/b_{j-2}, mut_j = 2 b_j w1/b_{j-1},
//   gt_j = -(1 - b_{j-1} T_{j-1}) mut_j, stage times c_j = mu_j c_{j-1} + nu_j c_{j-2} + mut_j + gt_j;
// real stability interval beta(s) = (w0 + 1) T_s''/T_s'.
struct Rkc {
    int s = 0;
    std::vector<double> mu, nu, mut, gt, c;
    double beta = 0;
};
Rkc rkc_make(int s) {
    Rkc k;
    k.s = s;
    const double w0 = 1.0 + (2.0 / 13.0) / ((double)s * (double)s);
    std::vector<double> T(s + 1), d1(s + 1), d2(s + 1), b(s + 1);
    T[0] = 1.0, T[1] = w0, d1[0] = 0.0, d1[1] = 1.0, d2[0] = 0.0, d2[1] = 0.0;
    for (int j = 2; j <= s; ++j) {
        T[j] = 2.0 * w0 * T[j - 1] - T[j - 2];
        d1[j] = 2.0 * T[j - 1] + 2.0 * w0 * d1[j - 1] - d1[j - 2];
        d2[j] = 4.0 * d1[j - 1] + 2.0 * w0 * d2[j - 1] - d2[j - 2];
    }
    const double w1 = d1[s] / d2[s];
    for (int j = 2; j <= s; ++j) b[j] = d2[j] / (d1[j] * d1[j]);
    b[0] = b[1] = b[2];
    k.mu.assign(s + 1, 0.0), k.nu.assign(s + 1, 0.0), k.mut.assign(s + 1, 0.0), k.gt.assign(s + 1, 0.0);
    k.c.assign(s + 1, 0.0);
    k.mut[1] = b[1] * w1;
    k.c[1] = k.mut[1];
    for (int j = 2; j <= s; ++j) {
        k.mu[j] = 2.0 * b[j] * w0 / b[j - 1];
        k.nu[j] = -b[j] / b[j - 2];
        k.mut[j] = 2.0 * b[j] * w1 / b[j - 1];
        k.gt[j] = -(1.0 - b[j - 1] * T[j - 1]) * k.mut[j];
        k.c[j] = k.mu[j] * k.c[j - 1] + k.nu[j] * k.c[j - 2] + k.mut[j] + k.gt[j];
    }
    k.beta = (w0 + 1.0) * d2[s] / d1[s];
    return k;
}
// the fewest stages (>= 2) with beta(s) >= x
int rkc_stage_count(double x) {
    int s = 2;
    while (rkc_make(s).beta < x && s < 100000) s = std::max(s + 1, (int)(s * 1.05));
    while (s > 2 && rkc_make(s - 1).beta >= x) --s;
    return s;
}
// start vector of the power iteration: uniform in [-1, 1) per cell (splitmix64 of the cell index)
double rkc_start(uint64_t p) {
    uint64_t z = 0x9E3779B97F4A7C15ull * (p + 1) + 0x5EEDull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    return (double)(z >> 11) * (2.0 / 9007199254740992.0) - 1.0;
}

// spectral radius of the right-hand side's Jacobian at (U[cur], t) by 40 power iterations on
// J v = f(u + v) - f(u) (the right-hand side is affine in u); U[scratch] and K slots 0/1 are
// overwritten
int rkc_spectral_radius(fi_ctx* c, double t, int cur, double* rho) {
    const size_t N = c->N;
    std::vector<double> u, f0, f1, v(N), w(N);
    int st;
    if ((st = read_u(c, cur, u))) return st;
    if ((st = rhs_on_device(c, t, cur, 0)) || (st = read_k(c, 0, f0))) return st;
    auto rms = [&](const std::vector<double>& a) {
        double q = 0;
        for (double x : a) q += x * x;
        return std::sqrt(q / (double)N);
    };
    for (size_t p = 0; p < N; ++p) v[p] = rkc_start(p);
    const double n0 = rms(v);
    for (double& x : v) x /= n0;
    const int sb = (cur + 1) % 3;
    double lam = 0;
    for (int it = 0; it < 40; ++it) {
        for (size_t p = 0; p < N; ++p) w[p] = u[p] + v[p];
        HIPCHECK(c, hipMemcpy(c->U.p + (size_t)sb * N, w.data(), N * sizeof(double), hipMemcpyHostToDevice));
        if ((st = rhs_on_device(c, t, sb, 1)) || (st = read_k(c, 1, f1))) return st;
        for (size_t p = 0; p < N; ++p) v[p] = f1[p] - f0[p];
        lam = rms(v);
        if (!(lam > 0) || !std::isfinite(lam)) return fail(c, FI_ERR_NONFINITE, "fi_integrate_rkc: power iteration");
        for (double& x : v) x /= lam;
    }
    *rho = lam;
    return FI_OK;
}

}  // namespace

extern "C" {

int fi_integrate_rkc(fi_ctx* c, double t_target, double dt_max, double rho, fi_rkc_info* info) {
    if (!c || !(dt_max > 0)) return FI_ERR_INVALID;
    if (!c->integ) return fail(c, FI_ERR_STATE, "fi_integrate_rkc before fi_integrator_init");
    if (c->implicit) return fail(c, FI_ERR_UNSUPPORTED, "fi_integrate_rkc: not with the implicit step");
    if (!(t_target >= c->ctl.t)) return fail(c, FI_ERR_INVALID, "fi_integrate_rkc: target before current time");
    HIPCHECK(c, hipSetDevice(c->dev));
    const long long r0 = c->ctl.nrhs;
    const size_t N = c->N;
    int st;
    if (!(rho > 0)) {
        if (!(c->rkc_rho > 0) && (st = rkc_spectral_radius(c, c->ctl.t, c->ctl.cur, &c->rkc_rho))) return st;
        rho = c->rkc_rho;
    }
    const double t_start = c->ctl.t, span = t_target - t_start;
    const int nsub = span > 0 ? (int)std::ceil(span / dt_max - 1e-12) : 0;
    const double h = nsub > 0 ? span / nsub : 0.0;
    const int s = nsub > 0 ? rkc_stage_count(1.2 * rho * h) : 0;
    int graph_state = 0;
    if (nsub > 0) {
        if (!c->rkc_off.p) HIPCHECK(c, c->rkc_off.alloc(1));
        // Every right-hand side of the call is queued on the stream without a host round trip:
        // the control block stays on the device between the launches (its sparse-update count
        // evolves there), each launch's stage time is written into it on the device from a table
        // copied once per call, and its U buffer and K slot come from one of six prepared
        // parameter blocks.
        if (!c->rkc_kp.p) {
            HIPCHECK(c, c->rkc_kp.alloc(6));
            KP kps[6];
            for (int q = 0; q < 6; ++q) {
                kps[q] = c->kp;
                kps[q].mode = MODE_RHS;
                kps[q].rhs_ubuf = q / 2;
                kps[q].rhs_kslot = q % 2;
                kps[q].max_steps = 0;
                kps[q].write_sigma_always = 1;
            }
            HIPCHECK(c, hipMemcpy(c->rkc_kp.p, kps, sizeof(kps), hipMemcpyHostToDevice));
        }
        // pinned: the stage times in launch order, then the chain's abort word
        const size_t ntimes = (size_t)nsub * (size_t)s;
        if (c->rkc_times_n < ntimes) {
            if (c->rkc_times) (void)hipHostFree(c->rkc_times);
            c->rkc_times = nullptr;
            c->rkc_times_n = 0;
            HIPCHECK(c, hipHostMalloc((void**)&c->rkc_times, (ntimes + 1) * sizeof(double)));
            c->rkc_times_n = ntimes;
        }
        if (c->rkc_tdev.n < ntimes) HIPCHECK(c, c->rkc_tdev.alloc(ntimes));
        if (!c->rkc_abort.p) HIPCHECK(c, c->rkc_abort.alloc(1));
        int* abort_h = (int*)(c->rkc_times + c->rkc_times_n);
        const Rkc k = rkc_make(s);
        {
            size_t q = 0;
            for (int sub = 0; sub < nsub; ++sub) {
                const double t0 = t_start + sub * h;
                c->rkc_times[q++] = t0;   // F_0
                for (int j = 2; j <= s; ++j) c->rkc_times[q++] = t0 + k.c[j - 1] * h;   // F_{j-1}
            }
        }
        HIPCHECK(c, hipMemcpy(c->ctl_d.p, &c->ctl, sizeof(Ctl), hipMemcpyHostToDevice));
        HIPCHECK(c, hipMemcpyAsync(c->rkc_tdev.p, c->rkc_times, ntimes * sizeof(double), hipMemcpyHostToDevice,
                                   c->stream));
        const void* fk = persistent_kernel(c->mt_double, c->has_dist, c->elra, c->implicit);
        const int words = (int)((1 + 2 * (size_t)c->G) * sizeof(Bar) / sizeof(unsigned));
        auto prep = [&](int idx, int first) {
            return RkcPrep{c->bar_d.p, words, c->rkc_abort.p, c->ctl_d.p, c->rkc_tdev.p, idx, first};
        };
        const int offs = c->elra ? 0 : 1;
        const double vbc = c->cfg.viscous_x_border;
        const int nb = (int)std::min<size_t>((N + 255) / 256, 4096);
        double* U = c->U.p;
        const float* K0 = c->K.p;
        const float* K1 = c->K.p + N;
        const float* kc0 = c->kcorner.p;
        const float* kc1 = c->kcorner.p + 4;
        const int y_start = c->ctl.cur;
        int y_end = y_start;
        // the chain: the first prep; per sub-step F_0, stage 1, then F_{j-1} and stage j for
        // j = 2..s, and the sub-step's corner offset. Each kernel between two right-hand sides
        // prepares the next one; the last one of the chain keeps the last launch's abort word.
        auto enqueue = [&]() -> hipError_t {
            hipError_t e;
            int idx = 0;   // the next right-hand side's entry in the time table
            hipLaunchKernelGGL(k_rkc_prep, dim3(1), dim3(256), 0, c->stream, prep(idx, 1));
            if ((e = hipGetLastError()) != hipSuccess) return e;
            int y0 = y_start;
            for (int sub = 0; sub < nsub; ++sub) {
                const int a = (y0 + 1) % 3, b = (y0 + 2) % 3;
                {
                    const KP* kptr = c->rkc_kp.p + 2 * y0 + 0;   // F_0 into K slot 0
                    void* args[] = {&kptr};
                    if ((e = hipLaunchCooperativeKernel(fk, dim3(c->G), dim3(NT), args, (unsigned)c->lds, c->stream)))
                        return e;
                    ++idx;
                }
                hipLaunchKernelGGL(k_rkc_stage, dim3(nb), dim3(256), 0, c->stream, U + a * N, U + y0 * N, U + y0 * N,
                                   (const double*)nullptr, K1, K0, kc1, kc0, offs, vbc, 1.0, 0.0, 0.0, 0.0,
                                   k.mut[1] * h, (long)N, prep(s >= 2 ? idx : -1, 0));
                if ((e = hipGetLastError()) != hipSuccess) return e;
                int pm1 = a, pm2 = y0;
                for (int j = 2; j <= s; ++j) {
                    const KP* kptr = c->rkc_kp.p + 2 * pm1 + 1;   // F_{j-1} into K slot 1
                    void* args[] = {&kptr};
                    if ((e = hipLaunchCooperativeKernel(fk, dim3(c->G), dim3(NT), args, (unsigned)c->lds, c->stream)))
                        return e;
                    ++idx;
                    const int out = (j == 2) ? b : pm2;   // never Y_0
                    hipLaunchKernelGGL(k_rkc_stage, dim3(nb), dim3(256), 0, c->stream, U + out * N, U + y0 * N,
                                       U + pm1 * N, U + pm2 * N, K1, K0, kc1, kc0, offs, vbc,
                                       1.0 - k.mu[j] - k.nu[j], k.mu[j], k.nu[j], k.mut[j] * h, k.gt[j] * h, (long)N,
                                       prep(j < s ? idx : -1, 0));
                    if ((e = hipGetLastError()) != hipSuccess) return e;
                    pm2 = pm1;
                    pm1 = out;
                }
                hipLaunchKernelGGL(k_corner_offset, dim3(1), dim3(64), 0, c->stream, U + pm1 * N, c->nx, c->ny, vbc,
                                   c->rkc_off.p);
                hipLaunchKernelGGL(k_sub_offset_dev, dim3(nb), dim3(256), 0, c->stream, U + pm1 * N, c->rkc_off.p,
                                   (long)N, prep(sub + 1 < nsub ? idx : -1, 0));
                if ((e = hipGetLastError()) != hipSuccess) return e;
                y0 = pm1;
            }
            // the BS3 loop ping-pongs u_n between U[0] and U[1] (U[2] is scratch): a state left in
            // U[2] moves to U[0]
            if (y0 == 2) {
                if ((e = hipMemcpyAsync(U, U + 2 * N, N * sizeof(double), hipMemcpyDeviceToDevice, c->stream)))
                    return e;
                y0 = 0;
            }
            y_end = y0;
            return hipSuccess;
        };
        // every value the captured chain holds: counts, sub-step, buffers and launch shape
        struct Key {
            int nsub, s, y0, G;
            size_t lds;
            double h, vbc;
            const void *kp, *tdev, *abrt, *U, *K, *fk;
        };
        Key key;
        std::memset(&key, 0, sizeof(key));
        key.nsub = nsub;
        key.s = s;
        key.y0 = y_start;
        key.G = c->G;
        key.lds = c->lds;
        key.h = h;
        key.vbc = vbc;
        key.kp = c->rkc_kp.p;
        key.tdev = c->rkc_tdev.p;
        key.abrt = c->rkc_abort.p;
        key.U = c->U.p;
        key.K = c->K.p;
        key.fk = fk;
        std::vector<unsigned char> kb((const unsigned char*)&key, (const unsigned char*)&key + sizeof(key));
        const char* genv = getenv("FASTISO_RKC_GRAPH");
        if (genv && genv[0] == '0') c->rkc_graph_off = true;
        hipGraphExec_t exec = nullptr;
        int y0 = y_start;
        if (!c->rkc_graph_off) {
            for (auto& g : c->rkc_graphs)
                if (g.key == kb) {
                    exec = g.exec;
                    y0 = g.y_end;
                    graph_state = 1;
                }
            if (!exec) {
                hipGraph_t gr = nullptr;
                hipError_t e = hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal);
                if (e == hipSuccess) {
                    const hipError_t e1 = enqueue();
                    const hipError_t e2 = hipStreamEndCapture(c->stream, &gr);
                    e = e1 != hipSuccess ? e1 : e2;
                }
                if (e == hipSuccess && gr) e = hipGraphInstantiate(&exec, gr, nullptr, nullptr, 0);
                if (gr) (void)hipGraphDestroy(gr);
                if (e != hipSuccess || !exec) {
                    // the chain cannot be captured here: it is launched directly from now on
                    (void)hipGetLastError();
                    if (exec) (void)hipGraphExecDestroy(exec);
                    exec = nullptr;
                    c->rkc_graph_off = true;
                } else {
                    if (c->rkc_graphs.size() >= 4) {
                        (void)hipGraphExecDestroy(c->rkc_graphs.front().exec);
                        c->rkc_graphs.erase(c->rkc_graphs.begin());
                    }
                    c->rkc_graphs.push_back({kb, exec, y_end});
                    y0 = y_end;
                    graph_state = 2;
                }
            }
        }
        if (exec) {
            HIPCHECK(c, hipGraphLaunch(exec, c->stream));
        } else {
            HIPCHECK(c, enqueue());
            y0 = y_end;
        }
        HIPCHECK(c, hipMemcpyAsync(abort_h, c->rkc_abort.p, sizeof(int), hipMemcpyDeviceToHost, c->stream));
        HIPCHECK(c, hipMemcpyAsync(&c->ctl, c->ctl_d.p, sizeof(Ctl), hipMemcpyDeviceToHost, c->stream));
        HIPCHECK(c, hipStreamSynchronize(c->stream));
        c->ctl.cur = y0;
        c->ctl.last_u = y0;
        if (*abort_h) return fail(c, FI_ERR_DEVICE, "device barrier timed out; kernel aborted");
        if (c->ctl.err == ERR_TIME_RANGE) return fail(c, FI_ERR_TIME_RANGE, "TimeInterpolation2D out of range.");
        if (c->ctl.err == ERR_BSL_RANGE) return fail(c, FI_ERR_TIME_RANGE, "TimeInterpolation0D out of range.");
        if (c->ctl.err) return fail(c, FI_ERR_DEVICE, "device error");
    }
    c->ctl.t = t_target;
    c->k1_stale = true;
    if (info) {
        info->rho = rho;
        info->h = h;
        info->substeps = nsub;
        info->stages = s;
        info->n_rhs = c->ctl.nrhs - r0;
        info->graph = graph_state;
        info->reserved = 0;
    }
    return FI_OK;
}

}  // extern "C"

namespace {

// Checkpoint blob: header, host control block and flags, then the dynamic device buffers in a
// fixed order. The same routine measures, writes or reads it (mode 0/1/2).
struct CkHeader {
    uint32_t magic, version;
    int32_t nx, ny, ice_mode, reserved;
};
constexpr uint32_t CK_MAGIC = 0x46494b31u;   // "FIK1"

template <typename T>
int ck_dev(fi_ctx* c, int mode, DBuf<T>& b, size_t n, unsigned char* blob, size_t& off) {
    const size_t bytes = n * sizeof(T);
    if (mode == 1) HIPCHECK(c, hipMemcpy(blob + off, b.p, bytes, hipMemcpyDeviceToHost));
    if (mode == 2) HIPCHECK(c, hipMemcpy(b.p, blob + off, bytes, hipMemcpyHostToDevice));
    off += bytes;
    return FI_OK;
}

template <typename T>
void ck_host(int mode, T& v, unsigned char* blob, size_t& off) {
    if (mode == 1) std::memcpy(blob + off, &v, sizeof(T));
    if (mode == 2) std::memcpy(&v, blob + off, sizeof(T));
    off += sizeof(T);
}

int ck_walk(fi_ctx* c, int mode, unsigned char* blob, size_t& off) {
    const size_t N = c->N;
    CkHeader h{CK_MAGIC, 1, c->nx, c->ny, c->ice_mode, 0};
    CkHeader hr = h;
    ck_host(mode, hr, blob, off);
    if (mode == 2 && (hr.magic != h.magic || hr.version != h.version || hr.nx != h.nx || hr.ny != h.ny ||
                      hr.ice_mode != h.ice_mode))
        return fail(c, FI_ERR_INVALID, "checkpoint of another grid, forcing or library");
    ck_host(mode, c->ctl, blob, off);
    int32_t flags[2] = {c->integ ? 1 : 0, c->k1_stale ? 1 : 0};
    ck_host(mode, flags, blob, off);
    ck_host(mode, c->kout, blob, off);
    if (mode == 2) {
        c->integ = flags[0] != 0;
        c->k1_stale = flags[1] != 0;
    }
    int st;
    if ((st = ck_dev(c, mode, c->U, 3 * N, blob, off)) || (st = ck_dev(c, mode, c->K, 4 * N, blob, off)) ||
        (st = ck_dev(c, mode, c->kcorner, 16, blob, off)) || (st = ck_dev(c, mode, c->ue, N, blob, off)) ||
        (st = ck_dev(c, mode, c->dz_ss, N, blob, off)) || (st = ck_dev(c, mode, c->z_ss, N, blob, off)) ||
        (st = ck_dev(c, mode, c->zoff, N, blob, off)) || (st = ck_dev(c, mode, c->zoff_trig, N, blob, off)) ||
        (st = ck_dev(c, mode, c->LW, N, blob, off)) || (st = ck_dev(c, mode, c->LT, N, blob, off)) ||
        (st = ck_dev(c, mode, c->H_water, N, blob, off)) || (st = ck_dev(c, mode, c->mg, N, blob, off)) ||
        (st = ck_dev(c, mode, c->mo, N, blob, off)) || (st = ck_dev(c, mode, c->sig_load_prev, N, blob, off)) ||
        (st = ck_dev(c, mode, c->sig_geo_prev, N, blob, off)) || (st = ck_dev(c, mode, c->H_ext, N, blob, off)))
        return st;
    if (c->elra && (st = ck_dev(c, mode, c->ueq, N, blob, off))) return st;
    if (c->ice_mode == ICE_EXTERNAL) {
        if (mode == 2) c->H_ext_host.assign(N, 0.0);
        for (size_t p = 0; p < N; ++p) ck_host(mode, c->H_ext_host[p], blob, off);
    }
    return FI_OK;
}

}  // namespace

extern "C" {

int fi_checkpoint_size(fi_ctx* c, size_t* bytes) {
    if (!c || !bytes) return FI_ERR_INVALID;
    if (!c->initialized) return fail(c, FI_ERR_STATE, "fi_checkpoint_size before fi_init");
    size_t off = 0;
    int st = ck_walk(c, 0, nullptr, off);
    *bytes = off;
    return st;
}

int fi_checkpoint_save(fi_ctx* c, void* buf, size_t bytes) {
    size_t need = 0;
    int st = fi_checkpoint_size(c, &need);
    if (st) return st;
    if (!buf || bytes != need) return fail(c, FI_ERR_INVALID, "fi_checkpoint_save: buffer size");
    (void)hipSetDevice(c->dev);
    HIPCHECK(c, hipStreamSynchronize(c->stream));
    size_t off = 0;
    return ck_walk(c, 1, (unsigned char*)buf, off);
}

int fi_checkpoint_load(fi_ctx* c, const void* buf, size_t bytes) {
    size_t need = 0;
    int st = fi_checkpoint_size(c, &need);
    if (st) return st;
    if (!buf || bytes != need) return fail(c, FI_ERR_INVALID, "fi_checkpoint_load: buffer size");
    (void)hipSetDevice(c->dev);
    HIPCHECK(c, hipStreamSynchronize(c->stream));
    size_t off = 0;
    if ((st = ck_walk(c, 2, (unsigned char*)const_cast<void*>(buf), off))) return st;
    const Ctl& g = c->ctl;
    if ((g.cur != 0 && g.cur != 1) || g.k1 < 0 || g.k1 > 3 || g.last_k < 0 || g.last_k > 3 || g.last_u < 0 || g.last_u > 2)
        return fail(c, FI_ERR_INVALID, "fi_checkpoint_load: buffer indices out of range");
    return FI_OK;
}

int fi_get_stats(fi_ctx* c, fi_stats* o) {
    if (!c || !o) return FI_ERR_INVALID;
    std::memset(o, 0, sizeof(*o));
    o->t = c->integ ? c->ctl.t : c->ctl.t_last;
    o->dt = c->ctl.dt;
    o->n_rhs = c->ctl.nrhs;
    o->n_accept = c->ctl.nacc;
    o->n_reject = c->ctl.nrej;
    o->count_sparse_updates = c->ctl.count;
    o->n_output = c->kout;
    o->V_af = c->ctl.V_af;
    o->V_den = c->ctl.V_den;
    o->V_pov = c->ctl.V_pov;
    o->delta_V = c->ctl.delta_V;
    o->z_bsl = c->ctl.z_bsl;
    o->A_ocean = c->ctl.A_bsl;
    o->last_EEst = c->ctl.EEst;
    o->device_ms = c->device_ms;
    o->n_launch = c->n_launch;
    o->n_accept_timed = c->n_accept_timed;
    return FI_OK;
}

}  // extern "C"

namespace {

// ---- autotuner: persistent grid size and FFT radix orders, timed on the context's own run ----

// every distinct radix order of length L (powers of two grouped into radices 8, 4, 2 in every
// way, fewest passes first, and each multiset in every order), the default plan first, at most cap
std::vector<FftPlan> plan_candidates(int L, int twoff, size_t cap) {
    std::vector<FftPlan> out;
    FftPlan def;
    if (!plan_fft(L, twoff, def)) return out;
    out.push_back(def);
    int m = L, a2 = 0;
    while (m % 2 == 0) {
        m /= 2;
        ++a2;
    }
    std::vector<int> odd;
    for (int r : {3, 5, 7})
        while (m % r == 0) {
            odd.push_back(r);
            m /= r;
        }
    for (int n8 = a2 / 3; n8 >= 0; --n8)
        for (int n4 = (a2 - 3 * n8) / 2; n4 >= 0; --n4) {
            std::vector<int> f(n8, 8);
            f.insert(f.end(), n4, 4);
            f.insert(f.end(), a2 - 3 * n8 - 2 * n4, 2);
            f.insert(f.end(), odd.begin(), odd.end());
            if (f.size() > 16) continue;
            std::sort(f.begin(), f.end());
            do {
                FftPlan pl;
                std::memset(&pl, 0, sizeof(pl));
                pl.L = L;
                pl.tw = twoff;
                pl.nr = (int)f.size();
                for (size_t q = 0; q < f.size(); ++q) pl.code |= (unsigned long long)f[q] << (4 * q);
                if (pl.code != def.code) out.push_back(pl);
                if (out.size() >= cap) return out;
            } while (std::next_permutation(f.begin(), f.end()));
        }
    return out;
}

bool plan_valid(int L, unsigned long long code, int& nr) {
    long long prod = 1;
    nr = 0;
    for (int q = 0; q < 16; ++q) {
        const int r = (int)((code >> (4 * q)) & 15ull);
        if (r == 0) break;
        if (r != 2 && r != 3 && r != 4 && r != 5 && r != 7 && r != 8) return false;
        prod *= r;
        nr = q + 1;
    }
    return nr > 0 && (code >> (4 * nr)) == 0 && prod == L;
}

// µs per step attempt of up to `attempts` integration attempts from the checkpointed state, the
// fastest of `reps` runs; the state is restored after each run
int tune_time(fi_ctx* c, const std::vector<unsigned char>& ck, int attempts, int reps, double* us) {
    *us = INFINITY;
    for (int r = 0; r < reps; ++r) {
        int st = fi_checkpoint_load(c, ck.data(), ck.size());
        if (st) return st;
        if (c->k1_stale) {
            if ((st = rhs_on_device(c, c->ctl.t, c->ctl.cur, c->ctl.k1))) return st;
            c->ctl.last_u = c->ctl.cur;
            c->k1_stale = false;
        }
        c->ctl.stop = c->t_end;
        c->ctl.next_out = INFINITY;
        const long long a0 = c->ctl.nacc + c->ctl.nrej;
        const double ms0 = c->device_ms;
        if ((st = launch(c, MODE_INTEGRATE, 0, 0, attempts, 0))) return st;
        const long long done = c->ctl.nacc + c->ctl.nrej - a0;
        if (done > 0) *us = std::min(*us, (c->device_ms - ms0) * 1e3 / (double)done);
    }
    return fi_checkpoint_load(c, ck.data(), ck.size());
}

}  // namespace

extern "C" {

int fi_set_fft_plans(fi_ctx* c, const uint64_t plans[4]) {
    if (!c || !plans) return FI_ERR_INVALID;
    if (!c->initialized) return fail(c, FI_ERR_STATE, "fi_set_fft_plans before fi_init");
    FftPlan* pl[4] = {&c->kp.px, &c->kp.py, &c->kp.pfx, &c->kp.pfy};
    for (int q = 0; q < 4; ++q) {
        int nr = 0;
        if (plans[q] && !plan_valid(pl[q]->L, plans[q], nr))
            return fail(c, FI_ERR_INVALID, "fi_set_fft_plans: radices (2, 3, 4, 5, 7, 8) must multiply to the length");
    }
    for (int q = 0; q < 4; ++q) {
        int nr = 0;
        if (plans[q] && plan_valid(pl[q]->L, plans[q], nr)) {
            pl[q]->code = plans[q];
            pl[q]->nr = nr;
        }
    }
    c->rkc_kp.reset();   // rebuilt with the new plans by the next fi_integrate_rkc
    return FI_OK;
}

int fi_get_fft_plans(fi_ctx* c, uint64_t plans[4]) {
    if (!c || !plans) return FI_ERR_INVALID;
    if (!c->initialized) return fail(c, FI_ERR_STATE, "fi_get_fft_plans before fi_init");
    plans[0] = c->kp.px.code;
    plans[1] = c->kp.py.code;
    plans[2] = c->kp.pfx.code;
    plans[3] = c->kp.pfy.code;
    return FI_OK;
}

int fi_set_workgroups(fi_ctx* c, int32_t g) {
    if (!c) return FI_ERR_INVALID;
    if (!c->initialized) return fail(c, FI_ERR_STATE, "fi_set_workgroups before fi_init");
    if (g < 1 || g > c->maxg) return fail(c, FI_ERR_INVALID, "fi_set_workgroups: 1 <= g <= the co-resident maximum");
    c->G = g;
    return FI_OK;
}

int fi_autotune(fi_ctx* c, int32_t what, int32_t attempts, fi_tune_info* info) {
    if (!c || !info || attempts < 1 || (what & ~(FI_TUNE_WORKGROUPS | FI_TUNE_FFT_PLANS)))
        return FI_ERR_INVALID;
    std::memset(info, 0, sizeof(*info));
    if (!c->integ) return fail(c, FI_ERR_STATE, "fi_autotune before fi_integrator_init");
    if (!(c->ctl.t < c->t_end)) return fail(c, FI_ERR_STATE, "fi_autotune: nothing left to integrate before t_end");
    (void)hipSetDevice(c->dev);
    size_t bytes = 0;
    int st = fi_checkpoint_size(c, &bytes);
    if (st) return st;
    std::vector<unsigned char> ck(bytes);
    if ((st = fi_checkpoint_save(c, ck.data(), bytes))) return st;
    const double dms = c->device_ms;
    const long long nl = c->n_launch, nat = c->n_accept_timed;
    const int reps = 2;
    double us = 0;
    if (what & FI_TUNE_WORKGROUPS) {
        const int nrow = c->kp.nrowitems, ncol = c->kp.ncolitems;
        std::vector<int> cand = {c->G, nrow, ncol, (c->G + 1) / 2, (3 * c->G + 3) / 4, c->maxg};
        std::vector<int> uniq;
        for (int g : cand)
            if (g >= 1 && g <= c->maxg && std::find(uniq.begin(), uniq.end(), g) == uniq.end()) uniq.push_back(g);
        const int g0 = c->G;
        int best = g0;
        double best_us = INFINITY;
        for (size_t q = 0; q < uniq.size() && q < 8; ++q) {
            c->G = uniq[q];
            if ((st = tune_time(c, ck, attempts, reps, &us))) {
                c->G = g0;
                break;
            }
            info->wg_candidates[q] = uniq[q];
            info->wg_us[q] = us;
            info->n_wg_candidates = (int32_t)q + 1;
            if (us < best_us) {
                best_us = us;
                best = uniq[q];
            }
        }
        c->G = st ? g0 : best;
    }
    // the transforms of the hot path: nx, ny for the Maxwell mantle, the padded lengths under ELRA
    const int first = c->elra ? 2 : 0;
    FftPlan* pl[4] = {&c->kp.px, &c->kp.py, &c->kp.pfx, &c->kp.pfy};
    if (!st && (what & FI_TUNE_FFT_PLANS)) {
        for (int q = first; q < first + 2 && !st; ++q) {
            const FftPlan def = *pl[q];
            const std::vector<FftPlan> cand = plan_candidates(def.L, def.tw, 24);
            FftPlan best = def;
            double best_us = INFINITY;
            for (size_t k = 0; k < cand.size(); ++k) {
                *pl[q] = cand[k];
                if ((st = tune_time(c, ck, attempts, reps, &us))) break;
                if (k == 0) info->plan_default_us[q] = us;
                if (us < best_us) {
                    best_us = us;
                    best = cand[k];
                }
            }
            *pl[q] = st ? def : best;
            info->n_plan_candidates[q] = (int32_t)cand.size();
            info->plan_best_us[q] = best_us;
        }
        c->rkc_kp.reset();
    }
    info->workgroups = c->G;
    for (int q = 0; q < 4; ++q) info->plans[q] = pl[q]->code;
    // the run continues from the state it had; the tuning launches are not counted
    const int st2 = fi_checkpoint_load(c, ck.data(), ck.size());
    c->device_ms = dms;
    c->n_launch = nl;
    c->n_accept_timed = nat;
    return st ? st : st2;
}

}  // extern "C"

// ==========================================================================================
// ensembles
// ==========================================================================================
struct fi_batch {
    std::vector<fi_ctx*> mem;
    std::string err;
    int dev = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    int tsize_fixed = 0;         // 0: choose by timing
    int tsize_best = 0;
    size_t cand_next = 0;        // next candidate team size under trial
    std::vector<int> cands;
    std::vector<double> cand_rate;
    // trials time whole windows: set once a window's first (capped) trial launch ended with most
    // members at their window end, i.e. the windows are shorter than TUNE_ATTEMPTS
    bool tune_window = false;
    int last_tsize = 0, last_nteams = 0;
    bool dist = false;           // the current launch uses the kernel with per-cell grid steps
    bool elra = false;           // the members are relaxed-mantle (ELRA) contexts
    int wg_threads = NT;         // threads per workgroup of the ensemble kernel in use (512 or 256)
    double device_ms = 0, cell_attempts = 0;
    long long n_launch = 0;
    DBuf<const KP*> kps;
    DBuf<int> order, queue, team_member;
    DBuf<Bar> bars;
    DBuf<BatchArgs> args;
    std::vector<int> order_h;
    // the members' parameter blocks and step controls of a launch, contiguous so that one copy
    // each way moves them (a member's own kp_d / ctl_d serve its single-context launches)
    DBuf<KP> kp_all;
    DBuf<Ctl> ctl_all;
    std::vector<Ctl> ctl_h;
    // An ensemble that mixes member kinds is a composite: one sub-ensemble per ensemble kernel
    // (LV-ELVA, ELRA), with the members' indices in the composite (gid), plus the members no
    // ensemble kernel runs (the implicit step), advanced one after another by fi_integrate.
    // mem then lists every member in the caller's order; the subs and solo contexts own them.
    std::vector<fi_batch*> subs;
    std::vector<int> solo;   // composite indices of the single-context members
    std::vector<int> gid;    // a sub-ensemble's member index -> the composite's
};

namespace {

int bfail(fi_batch* b, int code, const std::string& m) {
    if (b) b->err = m;
    return code;
}

// a member as the caller numbers it (the composite's index for a sub-ensemble's member)
std::string mname(const fi_batch* b, int m) {
    return "member " + std::to_string(b->gid.empty() ? m : b->gid[m]);
}

#define BHIPCHECK(b, x)                                                                        \
    do {                                                                                       \
        hipError_t e_ = (x);                                                                   \
        if (e_ != hipSuccess) return bfail((b), FI_ERR_DEVICE, std::string(#x " : ") + hipGetErrorString(e_)); \
    } while (0)

// attempts per member per trial launch while the team size is being chosen; pausing a member on
// an attempt count leaves its step sequence unchanged
constexpr int TUNE_ATTEMPTS = 64;
// share of a trial's capped work (cells x attempts) that must run for the trial to be timed
constexpr double TUNE_CAPPED_SHARE = 0.9;
// a candidate timed below this share of the best rate so far ends the search
constexpr double TUNE_STOP = 0.8;
// attempts per member per launch otherwise (bounds one launch's duration)
constexpr int LAUNCH_ATTEMPTS = 1 << 20;

// Device memory of one ensemble launch beyond the members' own buffers: the members' parameter
// blocks and step controls, the pointer table, the queue order and head, two member slots per team,
// the teams' barrier regions and the launch arguments.
struct BatchScratch {
    KP* kp_all;
    Ctl* ctl_all;
    const KP** kps;
    int* order;
    int* queue;
    int* team_member;
    Bar* bars;
    BatchArgs* args;
};

// the scratch of fi_batch_integrate_each, in buffers the batch allocates and grows as needed
int batch_scratch_alloc(fi_batch* b, int nm, int nteams, int ts, BatchScratch& sc) {
    if (b->kp_all.n < (size_t)nm) BHIPCHECK(b, b->kp_all.alloc(nm));
    if (b->ctl_all.n < (size_t)nm) BHIPCHECK(b, b->ctl_all.alloc(nm));
    const size_t nbar = (size_t)nteams * (1 + 2 * (size_t)ts);
    if (b->bars.n < nbar) BHIPCHECK(b, b->bars.alloc(nbar));
    if (b->team_member.n < 2 * (size_t)nteams) BHIPCHECK(b, b->team_member.alloc(2 * (size_t)nteams));
    sc.kp_all = b->kp_all.p;
    sc.ctl_all = b->ctl_all.p;
    sc.kps = b->kps.p;
    sc.order = b->order.p;
    sc.queue = b->queue.p;
    sc.team_member = b->team_member.p;
    sc.bars = b->bars.p;
    sc.args = b->args.p;
    return FI_OK;
}

// one persistent ensemble launch with teams of ts workgroups; returns the device time. Only the
// members still short of their targets are queued (largest first).
int batch_launch(fi_batch* b, const BatchScratch& sc, const void* f, size_t lds, int gmax, int ts, int max_steps,
                 int fsal, const std::vector<double>& t_target, float* ms_out, double* work_out,
                 double* cap_share_out) {
    const int nm = (int)b->mem.size();
    const int nteams = std::max(1, gmax / ts);
    b->ctl_h.resize(nm);
    std::vector<KP> kph(nm);
    std::vector<const KP*> ptrs(nm);
    std::vector<long long> att0(nm);
    double cap_work = 0;  // cells x attempts if every member still short of its target used the cap
    for (int m = 0; m < nm; ++m) {
        fi_ctx* c = b->mem[m];
        // the integration loop writes u_{n+1} into U[1 - cur]: cur must be 0 or 1
        if (c->ctl.cur != 0 && c->ctl.cur != 1)
            return bfail(b, FI_ERR_STATE, mname(b, m) + ": state buffer index out of range");
        if (c->ctl.t < t_target[m]) cap_work += (double)max_steps * (double)c->N;
        c->ctl.stop = t_target[m];
        c->ctl.next_out = t_target[m];
        att0[m] = c->ctl.nacc + c->ctl.nrej;
        kph[m] = c->kp;
        kph[m].hist_t = nullptr;   // the Timer history is kept for single contexts only
        kph[m].mode = MODE_INTEGRATE;
        kph[m].max_steps = max_steps;
        kph[m].write_sigma_always = 0;
        // columns per columns-phase item so that a team of ts workgroups transforms the member's
        // columns in one round (a batch per workgroup), as far as the launch's LDS holds the batch
        // (two buffers of nb columns after the twiddles); bitwise neutral
        if (!c->elra && !c->implicit) {
            const size_t twb = (size_t)(c->nx + c->ny) * 8;
            int nb = std::max(1, (c->nhx + ts - 1) / ts);
            while (nb > 1 && twb + (size_t)16 * nb * c->ny > lds) --nb;
            nb = std::max(nb, c->kp.colnb);
            kph[m].colnb = nb;
            kph[m].ncolitems = (c->nhx + nb - 1) / nb;
            // and row pairs per item of the standalone inverse transform
            int np2 = std::max(1, (c->kp.nrowitems + ts - 1) / ts);
            while (np2 > 1 && twb + (size_t)16 * np2 * c->nx > lds) --np2;
            np2 = std::max(np2, c->kp.c2rnb);
            kph[m].c2rnb = np2;
            kph[m].nc2ritems = (c->kp.nrowitems + np2 - 1) / np2;
        }
        // relaxed-mantle members likewise: row pairs per rows/inverse item and columns per
        // columns item so that a team covers them in one round where the LDS after the member's
        // twiddles holds the batch (twiddles and two buffers of the batch); bitwise neutral
        if (c->elra) {
            const size_t avail = lds - (size_t)(c->nx + c->ny) * 8;
            int nb = std::max(1, (c->kp.nrowitems + ts - 1) / ts);
            while (nb > 1 && (size_t)(1 + 2 * nb) * c->nfx * 8 > avail) --nb;
            kph[m].erb = nb;
            kph[m].nerowitems = (c->kp.nrowitems + nb - 1) / nb;
            int cb = std::max(2, (c->nfhx + ts - 1) / ts);
            while (cb > 2 && (size_t)(1 + 2 * cb) * c->nfy * 8 > avail) --cb;
            kph[m].ecb = cb;
            kph[m].necolitems = (c->nfhx + cb - 1) / cb;
        }
        kph[m].ctl = sc.ctl_all + m;
        ptrs[m] = sc.kp_all + m;
        b->ctl_h[m] = c->ctl;
    }
    // the queue: members short of their targets, in the batch's size order
    std::vector<int> act;
    act.reserve(nm);
    for (int m : b->order_h)
        if (b->mem[m]->ctl.t < t_target[m] && !b->mem[m]->ctl.err) act.push_back(m);
    BHIPCHECK(b, hipMemcpyAsync(sc.kp_all, kph.data(), nm * sizeof(KP), hipMemcpyHostToDevice, b->stream));
    BHIPCHECK(b, hipMemcpyAsync(sc.ctl_all, b->ctl_h.data(), nm * sizeof(Ctl), hipMemcpyHostToDevice, b->stream));
    const size_t nbar = (size_t)nteams * (1 + 2 * (size_t)ts);
    BHIPCHECK(b, hipMemsetAsync(sc.bars, 0, nbar * sizeof(Bar), b->stream));
    BHIPCHECK(b, hipMemsetAsync(sc.queue, 0, sizeof(int), b->stream));
    BHIPCHECK(b, hipMemcpyAsync(sc.kps, ptrs.data(), nm * sizeof(const KP*), hipMemcpyHostToDevice, b->stream));
    if (!act.empty())
        BHIPCHECK(b, hipMemcpyAsync(sc.order, act.data(), act.size() * sizeof(int), hipMemcpyHostToDevice, b->stream));
    BatchArgs a;
    a.kps = sc.kps;
    a.order = sc.order;
    a.nmem = (int)act.size();
    a.nteams = nteams;
    a.tsize = ts;
    a.fsal = fsal;
    a.bars = sc.bars;
    a.queue = sc.queue;
    a.team_member = sc.team_member;
    a.spin_limit = 2000000000ull;
    BHIPCHECK(b, hipMemcpyAsync(sc.args, &a, sizeof(a), hipMemcpyHostToDevice, b->stream));
    const BatchArgs* ap = sc.args;
    void* kargs[] = {&ap};
    BHIPCHECK(b, hipEventRecord(b->ev0, b->stream));
    BHIPCHECK(b, hipLaunchCooperativeKernel(f, dim3(nteams * ts), dim3(b->wg_threads), kargs, (unsigned)lds, b->stream));
    BHIPCHECK(b, hipEventRecord(b->ev1, b->stream));
    BHIPCHECK(b, hipMemcpyAsync(b->ctl_h.data(), sc.ctl_all, nm * sizeof(Ctl), hipMemcpyDeviceToHost, b->stream));
    std::vector<Bar> hb(nbar);
    BHIPCHECK(b, hipMemcpyAsync(hb.data(), sc.bars, nbar * sizeof(Bar), hipMemcpyDeviceToHost, b->stream));
    BHIPCHECK(b, hipStreamSynchronize(b->stream));
    float ms = 0;
    BHIPCHECK(b, hipEventElapsedTime(&ms, b->ev0, b->ev1));
    double work = 0;
    for (int m = 0; m < nm; ++m) {
        fi_ctx* c = b->mem[m];
        c->ctl = b->ctl_h[m];
        work += (double)(c->ctl.nacc + c->ctl.nrej - att0[m]) * (double)c->N;
        c->ctl.last_u = c->ctl.cur;
    }
    b->device_ms += ms;
    b->cell_attempts += work;
    b->n_launch += 1;
    b->last_tsize = ts;
    b->last_nteams = nteams;
    *ms_out = ms;
    *work_out = work;
    *cap_share_out = cap_work > 0 ? work / cap_work : 0.0;
    for (int t = 0; t < nteams; ++t)
        if (hb[(size_t)t * (1 + 2 * (size_t)ts)].abort) return bfail(b, FI_ERR_DEVICE, "device barrier timed out; kernel aborted");
    for (int m = 0; m < nm; ++m) {
        const int e = b->mem[m]->ctl.err;
        if (e == ERR_TIME_RANGE) return bfail(b, FI_ERR_TIME_RANGE, mname(b, m) + ": TimeInterpolation2D out of range.");
        if (e == ERR_BSL_RANGE) return bfail(b, FI_ERR_TIME_RANGE, mname(b, m) + ": TimeInterpolation0D out of range.");
        if (e == ERR_DT_MIN) return bfail(b, FI_ERR_DT_MIN, mname(b, m) + ": dt below minimum");
        if (e == ERR_NONFINITE)
            return bfail(b, FI_ERR_NONFINITE, mname(b, m) + ": NaN in the state or its derivative");
        if (e) return bfail(b, FI_ERR_DEVICE, mname(b, m) + ": device error");
    }
    return FI_OK;
}

}  // namespace

extern "C" {

int fi_create_batch(const fi_config* cfgs, int32_t nmem, fi_batch** out) {
    if (!cfgs || !out || nmem < 1) return FI_ERR_INVALID;
    *out = nullptr;
    // member kinds: 0 LV-ELVA (fi_batch_persistent), 1 ELRA (fi_batch_persistent_elra), 2 the
    // implicit step (no ensemble kernel: advanced by fi_integrate)
    std::vector<int> kind(nmem);
    for (int m = 0; m < nmem; ++m) {
        if (cfgs[m].device != cfgs[0].device) return FI_ERR_INVALID;  // one device per ensemble
        const fi_config& g = cfgs[m];
        kind[m] = g.mantle == FI_RELAXED_MANTLE ? 1 : g.lithosphere == FI_LATERALLY_VARIABLE_LITHOSPHERE ? 0 : 2;
    }
    const bool mixed = std::any_of(kind.begin(), kind.end(), [&](int k) { return k != kind[0]; }) || kind[0] == 2;
    fi_batch* b = new fi_batch();
    b->dev = cfgs[0].device;
    if (mixed) {
        b->mem.assign(nmem, nullptr);
        for (int k = 0; k < 2; ++k) {
            std::vector<fi_config> sub_cfg;
            std::vector<int> idx;
            for (int m = 0; m < nmem; ++m)
                if (kind[m] == k) {
                    sub_cfg.push_back(cfgs[m]);
                    idx.push_back(m);
                }
            if (idx.empty()) continue;
            fi_batch* sb = nullptr;
            const int st = fi_create_batch(sub_cfg.data(), (int32_t)idx.size(), &sb);
            if (st) {
                fi_batch_destroy(b);
                return st;
            }
            sb->gid = idx;
            for (size_t q = 0; q < idx.size(); ++q) b->mem[idx[q]] = sb->mem[q];
            b->subs.push_back(sb);
        }
        for (int m = 0; m < nmem; ++m)
            if (kind[m] == 2) {
                fi_ctx* c = nullptr;
                const int st = fi_create(&cfgs[m], &c);
                if (st) {
                    fi_batch_destroy(b);
                    return st;
                }
                b->mem[m] = c;
                b->solo.push_back(m);
            }
        *out = b;
        return FI_OK;
    }
    for (int m = 0; m < nmem; ++m) {
        fi_ctx* c = nullptr;
        const int st = fi_create(&cfgs[m], &c);
        if (st) {
            fi_batch_destroy(b);
            return st;
        }
        b->mem.push_back(c);
    }
    if (hipSetDevice(b->dev) != hipSuccess || hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreate(&b->ev0) != hipSuccess || hipEventCreate(&b->ev1) != hipSuccess ||
        b->kps.alloc(nmem) != hipSuccess || b->order.alloc(nmem) != hipSuccess || b->queue.alloc(1) != hipSuccess ||
        b->args.alloc(1) != hipSuccess) {
        fi_batch_destroy(b);
        return FI_ERR_DEVICE;
    }
    *out = b;
    return FI_OK;
}

fi_ctx* fi_batch_member(fi_batch* b, int32_t m) {
    if (!b || m < 0 || m >= (int)b->mem.size()) return nullptr;
    return b->mem[m];
}

void fi_batch_destroy(fi_batch* b) {
    if (!b) return;
    (void)hipSetDevice(b->dev);
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    if (!b->subs.empty() || !b->solo.empty()) {   // a composite owns its subs and solo members only
        for (fi_batch* sb : b->subs) fi_batch_destroy(sb);
        for (int m : b->solo) fi_destroy(b->mem[m]);
        b->mem.clear();
    }
    for (fi_ctx* c : b->mem) fi_destroy(c);
    if (b->ev0) (void)hipEventDestroy(b->ev0);
    if (b->ev1) (void)hipEventDestroy(b->ev1);
    hipStream_t s = b->stream;
    delete b;
    if (s) (void)hipStreamDestroy(s);
}

const char* fi_batch_last_error(const fi_batch* b) { return b ? b->err.c_str() : "null batch"; }

int fi_batch_set_team_size(fi_batch* b, int32_t ts) {
    if (!b || ts < 0) return FI_ERR_INVALID;
    for (fi_batch* sb : b->subs) {
        const int st = fi_batch_set_team_size(sb, ts);
        if (st) return st;
    }
    b->tsize_fixed = ts;
    b->tsize_best = 0;
    b->cand_next = 0;
    b->tune_window = false;
    b->cands.clear();
    b->cand_rate.clear();
    return FI_OK;
}

int fi_batch_integrate(fi_batch* b, double t_target, int32_t fsal) {
    if (!b) return FI_ERR_INVALID;
    std::vector<double> tt(b->mem.size(), t_target);
    return fi_batch_integrate_each(b, tt.data(), fsal);
}

}  // extern "C"

namespace {

// Every device buffer an ensemble launch needs beyond the members' own, for the entry point that
// allocates nothing (fi_batch_integrate_each_noalloc): one file-scope array per device, carved
// per launch. Its host-side name is not a device address; host code uses the address
// hipGetSymbolAddress returns for the current device (noalloc_scratch()).
constexpr int NOALLOC_MAX_MEMBERS = FI_NOALLOC_SCRATCH_MAX_MEMBERS;
constexpr int NOALLOC_MAX_WORKGROUPS = FI_NOALLOC_SCRATCH_MAX_WORKGROUPS;
constexpr size_t scratch_al(size_t x) { return (x + 255) & ~(size_t)255; }
constexpr size_t NOALLOC_SCRATCH_BYTES =
    scratch_al(NOALLOC_MAX_MEMBERS * sizeof(KP)) + scratch_al(NOALLOC_MAX_MEMBERS * sizeof(Ctl)) +
    scratch_al(NOALLOC_MAX_MEMBERS * sizeof(const KP*)) + scratch_al(NOALLOC_MAX_MEMBERS * sizeof(int)) +
    scratch_al(sizeof(int)) + scratch_al(2 * NOALLOC_MAX_WORKGROUPS * sizeof(int)) +
    scratch_al(3 * (size_t)NOALLOC_MAX_WORKGROUPS * sizeof(Bar)) + scratch_al(sizeof(BatchArgs));

}  // namespace

static __device__ __align__(256) unsigned char scratch[NOALLOC_SCRATCH_BYTES];

namespace {

constexpr int NOALLOC_MAX_DEVICES = 64;
std::mutex noalloc_mu;
unsigned char* noalloc_ptr[NOALLOC_MAX_DEVICES];   // device address of scratch, per device

// the device address of this device's scratch array (looked up once per device, with the device
// current)
int noalloc_scratch(fi_batch* b, unsigned char** out) {
    if (b->dev < 0 || b->dev >= NOALLOC_MAX_DEVICES) return bfail(b, FI_ERR_INVALID, "device ordinal out of range");
    std::lock_guard<std::mutex> lk(noalloc_mu);
    if (!noalloc_ptr[b->dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return bfail(b, FI_ERR_DEVICE, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        noalloc_ptr[b->dev] = (unsigned char*)p;
    }
    *out = noalloc_ptr[b->dev];
    return FI_OK;
}

// the scratch of one launch carved from the device's static array; sizes beyond its maxima are
// refused
int batch_scratch_static(fi_batch* b, unsigned char* base, int nm, int gmax, BatchScratch& sc) {
    if (nm > NOALLOC_MAX_MEMBERS)
        return bfail(b, FI_ERR_INVALID, "fi_batch_integrate_each_noalloc: " + std::to_string(nm) +
                                            " members exceed the static scratch maximum of " +
                                            std::to_string(NOALLOC_MAX_MEMBERS));
    if (gmax > NOALLOC_MAX_WORKGROUPS)
        return bfail(b, FI_ERR_INVALID, "fi_batch_integrate_each_noalloc: " + std::to_string(gmax) +
                                            " co-resident workgroups exceed the static scratch maximum of " +
                                            std::to_string(NOALLOC_MAX_WORKGROUPS));
    size_t off = 0;
    auto take = [&](size_t bytes) {
        unsigned char* p = base + off;
        off += scratch_al(bytes);
        return p;
    };
    sc.kp_all = (KP*)take(NOALLOC_MAX_MEMBERS * sizeof(KP));
    sc.ctl_all = (Ctl*)take(NOALLOC_MAX_MEMBERS * sizeof(Ctl));
    sc.kps = (const KP**)take(NOALLOC_MAX_MEMBERS * sizeof(const KP*));
    sc.order = (int*)take(NOALLOC_MAX_MEMBERS * sizeof(int));
    sc.queue = (int*)take(sizeof(int));
    sc.team_member = (int*)take(2 * NOALLOC_MAX_WORKGROUPS * sizeof(int));
    // nteams (1 + 2 ts) <= 3 gmax barrier records for every team size the launch can use
    sc.bars = (Bar*)take(3 * (size_t)NOALLOC_MAX_WORKGROUPS * sizeof(Bar));
    sc.args = (BatchArgs*)take(sizeof(BatchArgs));
    return off <= NOALLOC_SCRATCH_BYTES ? FI_OK : bfail(b, FI_ERR_INVALID, "static scratch layout overflow");
}

int batch_integrate(fi_batch* b, const double* t_targets, int32_t fsal, bool noalloc);

// a composite ensemble: each sub-ensemble in one launch sequence of its own kernel, then the
// single-context members one after another, all on this device and in this order; every member's
// state is checked first so that no part runs when another would be refused
int composite_integrate(fi_batch* b, const double* t_targets, int32_t fsal, bool noalloc) {
    const int nm = (int)b->mem.size();
    for (int m = 0; m < nm; ++m) {
        const fi_ctx* c = b->mem[m];
        if (!c->integ) return bfail(b, FI_ERR_STATE, mname(b, m) + ": fi_batch_integrate before fi_integrator_init");
        if (t_targets[m] < c->ctl.t) return bfail(b, FI_ERR_INVALID, mname(b, m) + ": target before current time");
    }
    for (fi_batch* sb : b->subs) {
        std::vector<double> tt;
        for (int g : sb->gid) tt.push_back(t_targets[g]);
        const int st = batch_integrate(sb, tt.data(), fsal, noalloc);
        if (st) return bfail(b, st, sb->err);
    }
    for (int g : b->solo) {
        fi_ctx* c = b->mem[g];
        if (!(c->ctl.t < t_targets[g])) continue;
        // as the ensemble kernels treat their members: the member's own output times are not
        // used, and with fsal the FSAL derivative is re-evaluated at the target, as after an
        // output callback (under the implicit step that call advances u, as it does there)
        std::vector<double> tout;
        tout.swap(c->tout);
        const long long kout = c->kout;
        int st = fi_integrate(c, t_targets[g], nullptr, nullptr);
        c->tout.swap(tout);
        c->kout = kout;
        if (!st && fsal) {
            st = rhs_on_device(c, c->ctl.t, c->ctl.cur, c->ctl.k1);
            c->ctl.last_u = c->ctl.cur;
        }
        if (st) return bfail(b, st, mname(b, g) + ": " + c->err);
    }
    return FI_OK;
}

int batch_integrate(fi_batch* b, const double* t_targets, int32_t fsal, bool noalloc) {
    if (!b || !t_targets) return FI_ERR_INVALID;
    if (!b->subs.empty() || !b->solo.empty()) return composite_integrate(b, t_targets, fsal, noalloc);
    const int nm = (int)b->mem.size();
    if (noalloc && nm > NOALLOC_MAX_MEMBERS)
        return bfail(b, FI_ERR_INVALID, "fi_batch_integrate_each_noalloc: " + std::to_string(nm) +
                                            " members exceed the static scratch maximum of " +
                                            std::to_string(NOALLOC_MAX_MEMBERS));
    const std::vector<double> t_target(t_targets, t_targets + nm);
    bool mt_double = true, dist = false;
    size_t lds_d = 0, lds_dd = 0, lds_f = 0;
    for (int m = 0; m < nm; ++m) {
        fi_ctx* c = b->mem[m];
        if (!c->integ) return bfail(b, FI_ERR_STATE, mname(b, m) + ": fi_batch_integrate before fi_integrator_init");
        if (t_target[m] < c->ctl.t) return bfail(b, FI_ERR_INVALID, mname(b, m) + ": target before current time");
        if (c->k1_stale) {
            // state written since the last step: re-evaluate the FSAL derivative as fi_integrate does
            const int st = rhs_on_device(c, c->ctl.t, c->ctl.cur, c->ctl.k1);
            if (st) return bfail(b, st, mname(b, m) + ": " + c->err);
            c->ctl.last_u = c->ctl.cur;
            c->k1_stale = false;
        }
        mt_double = mt_double && c->mt_double;
        lds_d = std::max(lds_d, c->lds_d);
        lds_dd = std::max(lds_dd, c->lds_dist);
        lds_f = std::max(lds_f, c->lds_f);
        dist = dist || c->has_dist;
    }
    const bool elra = b->mem[0]->elra;
    // with a distorted member every member runs the distortion kernel (K = 1 where unset); ELRA
    // members read K^2 at run time and need the padded-grid phases' LDS only
    size_t lds = !mt_double ? lds_f : dist ? lds_dd : lds_d;
    if (elra) {
        lds = 0;
        for (int m = 0; m < nm; ++m) lds = std::max(lds, b->mem[m]->lds);
    }
    // relaxed-mantle members batch row pairs and columns per item within the launch's LDS: offer
    // 64 KB (two 256-thread workgroups still share a CU), unless that lowers the occupancy
    const size_t lds_need = lds;
    if (elra) lds = std::max(lds, (size_t)64 * 1024);
    b->elra = elra;
    if (lds > 160 * 1024) return bfail(b, FI_ERR_UNSUPPORTED, "members too large for one workgroup's LDS");
    b->dist = dist;
    BHIPCHECK(b, hipSetDevice(b->dev));
    int ncu = 0;
    BHIPCHECK(b, hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, b->dev));
    // Members up to 256 cells wide on the FP64-moment path run the 256-thread build of the
    // ensemble kernel when two of its workgroups fit on a CU; the others the 512-thread build.
    // Both round every sum in the same order, so members do not depend on the choice.
    int maxnx = 0;
    for (int m = 0; m < nm; ++m) maxnx = std::max(maxnx, b->mem[m]->nx);
    const void* f = nullptr;
    int per_cu = 0;
    b->wg_threads = NT;
    const char* nt_env = getenv("FASTISO_BATCH_WG_THREADS");
    const bool allow256 = !(nt_env && atoi(nt_env) == NT);
    auto occupancy = [&](const void* fk, int threads, size_t l, int* out) -> hipError_t {
        hipError_t e = hipFuncSetAttribute(fk, hipFuncAttributeMaxDynamicSharedMemorySize, (int)l);
        if (e != hipSuccess) return e;
        return hipOccupancyMaxActiveBlocksPerMultiprocessor(out, fk, threads, l);
    };
    if (allow256 && (elra || mt_double) && maxnx <= 256) {
        const void* f2 = fi256::batch_kernel_nt256(dist, elra);
        BHIPCHECK(b, occupancy(f2, 256, lds, &per_cu));
        if (per_cu < 2 && lds > lds_need) {
            lds = lds_need;
            BHIPCHECK(b, occupancy(f2, 256, lds, &per_cu));
        }
        if (per_cu >= 2) {
            f = f2;
            b->wg_threads = 256;
        }
    }
    if (!f) {
        f = batch_kernel(mt_double, dist, elra);
        BHIPCHECK(b, occupancy(f, NT, lds, &per_cu));
        if (lds > lds_need) {
            int per_cu_need = 0;
            BHIPCHECK(b, occupancy(f, NT, lds_need, &per_cu_need));
            if (per_cu_need > per_cu) {
                lds = lds_need;
                BHIPCHECK(b, occupancy(f, NT, lds, &per_cu));
            }
        }
    }
    const int gmax = per_cu * ncu;
    if (gmax < 1) return bfail(b, FI_ERR_UNSUPPORTED, "ensemble kernel does not fit on the device");
    // queue order: largest members first (cost per attempt ~ cells)
    if ((int)b->order_h.size() != nm) {
        b->order_h.resize(nm);
        for (int m = 0; m < nm; ++m) b->order_h[m] = m;
        std::stable_sort(b->order_h.begin(), b->order_h.end(),
                         [&](int x, int y) { return b->mem[x]->N > b->mem[y]->N; });
    }
    unsigned char* sbase = nullptr;
    if (noalloc) {
        const int st = noalloc_scratch(b, &sbase);
        if (st) return st;
    }
    if (b->tsize_fixed == 0 && b->cands.empty()) {
        for (int ts : {64, 32, 16, 8, 4, 2, 1})
            if (ts <= gmax && (ts < 64 || b->wg_threads < NT)) b->cands.push_back(ts);
        b->cand_rate.assign(b->cands.size(), 0.0);
    }
    auto done = [&]() {
        for (int m = 0; m < nm; ++m)
            if (b->mem[m]->ctl.t < t_target[m]) return false;
        return true;
    };
    bool first_launch = true;
    while (!done()) {
        int ts;
        bool trial = false;
        if (b->tsize_fixed > 0) ts = std::min(b->tsize_fixed, gmax);
        else if (b->tsize_best > 0) ts = b->tsize_best;
        else {
            ts = b->cands[b->cand_next];
            trial = true;
        }
        float ms = 0;
        double work = 0, cap_share = 0;
        const int cap = trial && !b->tune_window ? TUNE_ATTEMPTS : LAUNCH_ATTEMPTS;
        BatchScratch sc;
        int st = noalloc ? batch_scratch_static(b, sbase, nm, gmax, sc)
                         : batch_scratch_alloc(b, nm, std::max(1, gmax / ts), ts, sc);
        if (st) return st;
        st = batch_launch(b, sc, f, lds, gmax, ts, cap, fsal, t_target, &ms, &work, &cap_share);
        if (st) return st;
        const bool window_start = first_launch;
        const bool whole_window = first_launch && done();
        first_launch = false;
        // A capped trial in which members reached their window end timed the window's tail: its
        // rate is not comparable with the other candidates'. One that straddled a window end is
        // run again at the next launch; when even a window's first launch falls short, the windows
        // are shorter than the cap and every candidate is timed over whole windows instead.
        bool timed = trial && (b->tune_window ? whole_window : cap_share >= TUNE_CAPPED_SHARE);
        if (trial && !b->tune_window && !timed && window_start) {
            b->tune_window = true;
            b->cand_next = 0;
        }
        if (timed) {
            const size_t k = b->cand_next;
            b->cand_rate[k] = ms > 0 ? work / ms : 0.0;
            b->cand_next += 1;
            size_t best = 0;
            for (size_t q = 1; q <= k; ++q)
                if (b->cand_rate[q] > b->cand_rate[best]) best = q;
            // candidates run from the largest team down; the rate rises to a peak and falls
            // again, so a smaller team clearly slower than the best so far ends the search
            if (b->cand_next == b->cands.size() || b->cand_rate[k] < TUNE_STOP * b->cand_rate[best])
                b->tsize_best = b->cands[best];
        }
    }
    return FI_OK;
}

}  // namespace

extern "C" {

int fi_batch_integrate_each(fi_batch* b, const double* t_targets, int32_t fsal) {
    return batch_integrate(b, t_targets, fsal, false);
}

int fi_batch_integrate_each_noalloc(fi_batch* b, const double* t_targets, int32_t fsal) {
    return batch_integrate(b, t_targets, fsal, true);
}

int fi_batch_autotune_fft(fi_batch* b, int32_t attempts) {
    if (!b || attempts < 1) return FI_ERR_INVALID;
    // one member of each kind and grid is tuned on its own context; the others take its plans
    std::vector<char> done(b->mem.size(), 0);
    for (size_t m = 0; m < b->mem.size(); ++m) {
        if (done[m]) continue;
        fi_ctx* c = b->mem[m];
        fi_tune_info info;
        int st = fi_autotune(c, FI_TUNE_FFT_PLANS, attempts, &info);
        if (st) return bfail(b, st, mname(b, (int)m) + ": " + c->err);
        for (size_t q = m; q < b->mem.size(); ++q) {
            fi_ctx* d = b->mem[q];
            if (d->nx != c->nx || d->ny != c->ny || d->elra != c->elra || d->implicit != c->implicit) continue;
            if ((st = fi_set_fft_plans(d, info.plans))) return bfail(b, st, mname(b, (int)q) + ": " + d->err);
            done[q] = 1;
        }
    }
    return FI_OK;
}

int fi_batch_get_info(const fi_batch* b, fi_batch_info* o) {
    if (!b || !o) return FI_ERR_INVALID;
    std::memset(o, 0, sizeof(*o));
    if (!b->subs.empty() || !b->solo.empty()) {
        // a composite: the team sizes of its largest sub-ensemble, launches and device time
        // summed over the sub-ensembles (the single-context members are not counted)
        const fi_batch* big = nullptr;
        for (const fi_batch* sb : b->subs)
            if (!big || sb->mem.size() > big->mem.size()) big = sb;
        if (big) fi_batch_get_info(big, o);
        o->device_ms = o->n_launch = 0;
        o->cell_attempts = 0;
        for (const fi_batch* sb : b->subs) {
            o->device_ms += sb->device_ms;
            o->n_launch += sb->n_launch;
            o->cell_attempts += sb->cell_attempts;
        }
        o->n_members = (int32_t)b->mem.size();
        return FI_OK;
    }
    o->n_members = (int32_t)b->mem.size();
    o->tuned = (b->tsize_fixed > 0 || b->tsize_best > 0) ? 1 : 0;
    // the size in use once tuned (a given size is capped by the co-resident workgroups at launch)
    o->team_size = b->tsize_fixed > 0 ? (b->last_tsize > 0 ? b->last_tsize : b->tsize_fixed)
                   : b->tsize_best > 0 ? b->tsize_best : b->last_tsize;
    o->last_team_size = b->last_tsize;
    o->n_teams = b->last_nteams;
    o->wg_threads = b->n_launch > 0 ? b->wg_threads : 0;
    o->device_ms = b->device_ms;
    o->n_launch = b->n_launch;
    o->cell_attempts = b->cell_attempts;
    const int nc = (int)std::min<size_t>(b->cands.size(), FI_BATCH_MAX_CANDIDATES);
    o->n_candidates = b->tsize_fixed > 0 ? 0 : nc;
    for (int k = 0; k < o->n_candidates; ++k) {
        o->cand_size[k] = b->cands[k];
        o->cand_rate[k] = b->cand_rate[k];
    }
    return FI_OK;
}

}  // extern "C"
