// Device-side data structures and phases of the LV-ELVA engine.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

// Threads per workgroup and the namespace of this build of the device code: 512 threads (eight
// wave64, two per SIMD) in namespace fi; the ensemble kernels are also built with 256 threads in
// namespace fi256 (fi_batch256.hip).
#ifndef FI_NT
#define FI_NT 512
#endif
#ifndef FI_NS
#define FI_NS fi
#endif

#include "fi_fft.hpp"

namespace FI_NS {

// Coherent load of device memory that another workgroup may have written earlier in the same
// launch: an agent-scope relaxed atomic load (global_load ... sc1), which reads the device's
// coherence point instead of a line the XCD's L2 may hold from before that write. With every such
// read coherent, the grid barrier needs no agent-scope acquire (buffer_inv sc1, an invalidate of
// the whole L2 of the XCD).
template <typename T>
__device__ __forceinline__ typename std::remove_const<T>::type ldc(T* a) {
    typedef typename std::remove_const<T>::type U;
    if constexpr (sizeof(U) == 8) {
        return __builtin_bit_cast(U, __hip_atomic_load((unsigned long long*)a, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
    } else if constexpr (sizeof(U) == 4) {
        return __builtin_bit_cast(U, __hip_atomic_load((unsigned int*)a, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
    } else {
        static_assert(sizeof(U) == 1, "coherent loads of 1, 4 or 8 bytes");
        return __builtin_bit_cast(U, __hip_atomic_load((unsigned char*)a, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
    }
}

// A buffer that the persistent kernels write while they run (state, derivative slots, spectra,
// partial sums, sparse-update fields): indexing loads through ldc, so no read of such a buffer can
// bypass it. Stores go through st_shared (write-through) on .p. Same layout as T*.
template <typename T>
struct G {
    T* p;
    G() = default;
    __host__ __device__ constexpr G(T* q) : p(q) {}
    template <typename U, typename = typename std::enable_if<std::is_same<const U, T>::value>::type>
    __host__ __device__ constexpr G(G<U> o) : p(o.p) {}
    __device__ __forceinline__ typename std::remove_const<T>::type operator[](long i) const { return ldc(p + i); }
    __host__ __device__ G operator+(long i) const { return G(p + i); }
    __host__ __device__ explicit operator bool() const { return p != nullptr; }
};

constexpr int NT = FI_NT;
// the canonical workgroup width of every block reduction: a build with fewer threads gives each
// thread VNT / NT virtual lanes, so sums round identically in both builds
constexpr int VNT = 512;
constexpr int NV = VNT / NT;
static_assert(VNT % NT == 0, "NT must divide 512");

enum { MODE_RHS = 0, MODE_INTEGRATE = 1, MODE_SPARSE_ONLY = 2 };
enum { ICE_EXTERNAL = 0, ICE_SNAPSHOTS = 1, ICE_SEPARABLE = 2 };
enum { ERR_TIME_RANGE = 3, ERR_NONFINITE = 4, ERR_ABORT = 5, ERR_DT_MIN = 7, ERR_BSL_RANGE = 8 };

// BS3 tableau (Bogacki & Shampine 1989) and PI-controller constants
#define BS3_A21 0.5
#define BS3_A32 0.75
#define BS3_A41 (2.0 / 9.0)
#define BS3_A42 (1.0 / 3.0)
#define BS3_A43 (4.0 / 9.0)
#define BS3_BT1 (2.0 / 9.0 - 7.0 / 24.0)
#define BS3_BT2 (1.0 / 3.0 - 0.25)
#define BS3_BT3 (4.0 / 9.0 - 1.0 / 3.0)
#define BS3_BT4 (-0.125)
#define PI_BETA1 (7.0 / 30.0)
#define PI_BETA2 (2.0 / 15.0)
#define PI_GAMMA 0.9
#define PI_QMIN 0.2
#define PI_QMAX 10.0
#define PI_QOLDINIT 1e-4

// Control block: integrator state shared by host and device. Every workgroup keeps an
// identical copy in LDS and updates it with identical arithmetic; workgroup 0 writes it back.
struct Ctl {
    double t, dt, qold, q11, EEst, t0;
    double stop;        // clip steps to land here (next output time or the target)
    double next_out;    // leave the loop after an accepted step with t >= next_out
    double t_rhs;       // RHS mode: evaluation time
    double t_last;      // time of the last RHS call
    double V_af, V_den, V_pov, delta_V, z_bsl;
    double A_bsl;       // current ocean surface of the BSL rule (m^2)
    long long count, nrhs, nacc, nrej, steps;
    int cur;       // U buffer holding u_n
    int k1;        // K slot of the FSAL derivative
    int last_k;    // K slot written by the last RHS call
    int last_u;    // U buffer holding the last RHS call's (corner-corrected) displacement
    int err;       // device error (ERR_*)
    int last_trig; // the last RHS call ran the sparse update
    int pad0, pad1;
};

// grid barrier words: the release generation and the abort flag, followed in memory by one
// 128-byte arrival slot per workgroup (BAR_SLOT unsigneds apart)
struct Bar {
    unsigned release;
    int abort;
    unsigned pad[30];
};
constexpr int BAR_SLOT = 32;

// The workgroups that cooperate on one simulation: the whole grid for a single context, one
// team of the grid for an ensemble member. rank/size take the place of blockIdx.x/gridDim.x in
// every phase; bar is the team's barrier region (a header Bar, then 128-byte arrival and
// release slots per workgroup).
struct Team {
    int rank, size;
    Bar* bar;
};

struct KP {
    int nx, ny, nhx, nfx, nfy, nfhx, nrowitems, ncolitems, colnb;
    int c2rnb, nc2ritems;   // row pairs per item of the standalone inverse-transform phase, its item count
    int erb, nerowitems;    // relaxed mantle: row pairs per item of its rows and inverse phases, item count
    int ecb, necolitems;    // relaxed mantle: spectral columns per item of its columns phase, item count
    int i1, i2, j1, j2;  // 1-based crop indices (src/convolutions.jl:196-204)
    double dx, dy, A, inv_dx, inv_dy, inv_dx2, inv_dy2;
    double g, spy, rho_ice, rho_sw_ice, rho_ice_sw, rho_seawater, rho_water, rho_um, rho_litho, nu;
    double dt_sparse, reltol, abstol, vbc_x, ebc_x, sbc_x, bsl_z;
    double N;  // nx*ny as double
    int rigid_litho, rigid_mantle, compressible_column, lv_surface, interactive, vol_goelzer, den_goelzer,
        pov_goelzer;
    int ice_mode, ice_nt, ice_flat, mt_double;
    // barystatic sea level: rule (FI_BSL_*), internal update flag, reference ocean surface and the
    // rule's table (z -> A for the piecewise-constant rule, t -> z for the imposed one)
    int bsl_rule, bsl_internal, bsl_n, bsl_flat;
    double bsl_A_ref;
    const double* bsl_x;
    const double* bsl_y;
    const double* ice_t;
    const double* ice_a;
    const float* ice_H;
    // ImposedSeaSurface: dz_ss snapshots [ss_nt][nx*ny] (FP64) at ascending times ss_t (yr)
    int imposed_ss, ss_nt, ss_flat;
    const double* ss_t;
    const double* ss_X;
    // state and fields (float unless noted); U buffer b at Ub + b*Nl, K slot k at Kb + k*Nl
    long Nl;
    G<double> Ub;
    G<float> Kb;
    G<float> kcorner;  // [4 slots][4 corners], raw
    G<float2> specb;   // two spectrum buffers [2][nhx][ny]; an RHS call writes buffer (call count & 1)
    long specN;      // nhx * ny
    const float* mult;
    const float* D;
    const float* inv2eta;   // K / (2 eta)
    const float* Kinv;      // 1/K per cell, null without map distortion
    const float* K2;        // K^2 per cell, null without map distortion
    const unsigned char* mask;
    const float* H_af_ref;
    const float* u_ref;
    const float* z_b_ref;
    const float* z_ss_ref;
    const float* ue_ref;
    float* H_ext;
    G<float> ue;
    G<float> dz_ss;
    G<float> z_ss;
    G<float> zoff;
    G<float> zoff_trig;
    G<float> LW;
    G<float> LT;
    G<float> H_water;
    G<unsigned char> mg;
    G<unsigned char> mo;
    G<float> sig_load_prev;
    G<float> sig_geo_prev;
    G<float2> pA;
    G<float2> pB;
    G<float2> oA;
    G<float2> oB;
    G<float> ue_raw;
    G<float> dz_raw;
    G<float> ext_corners;  // [2][4]
    const double2* ge_hat;
    const double2* gg_hat;
    G<double> part;    // [G][4] error moments
    G<double> vpart;   // [G][4] volume partial sums
    FftPlan px, py, pfx, pfy;
    const float2* tw32;
    Ctl* ctl;
    Bar* bar;
    int mode, rhs_ubuf, rhs_kslot, max_steps;
    int write_sigma_always;
    unsigned long long spin_limit;  // wall-clock ticks before a barrier wait gives up
    // optional phase timing (workgroup 0): [tag] work ticks, [8+tag] wait ticks, [16+tag] count
    unsigned long long* ptime;
    // Timer history (null when off): RHS call k since fi_init (k < hist_cap) stores its time and
    // the device wall clock at its end at index k, written by workgroup 0 alone
    double* hist_t;
    unsigned long long* hist_clk;
    long long hist_cap;
    // relaxed mantle (ELRA): viscous Green half spectrum [nfhx][nfy] (FP64, pre-scaled by
    // 1/(nfx nfy)), 1/tau per cell (1/yr) and the equilibrium displacement u_eq
    int elra;
    const double2* gv_hat;
    const float* inv_tau;
    G<float> ueq;
    // implicit step of a Maxwell mantle under a rigid or laterally constant lithosphere: the
    // even parts of (a - 1) and b of u <- a u + b F on the n x n spectrum, [nhx][ny], scaled by
    // 1/(nx ny); dt_min (yr) is the integrator step
    int implicit;
    const float* imp_am1;
    const float* imp_b;
    double dt_min;
};
enum { TAG_ROWS = 0, TAG_COLS = 1, TAG_C2R = 2, TAG_SPARSE = 3, TAG_KIN = 4 };

}  // namespace FI_NS
