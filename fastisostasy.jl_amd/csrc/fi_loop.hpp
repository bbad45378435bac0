// The persistent integration loop and the ensemble kernels, shared by two builds of the device
// code: workgroups of 512 threads (namespace fi, every kernel) and of 256 threads (namespace
// fi256, the ensemble kernels only, two workgroups per CU for members up to 256 cells wide).
// FI_NT / FI_NS select the build before fi_device.hpp is included; FI_BATCH_BOUNDS gives the
// ensemble kernels' launch bounds.
#pragma once
#include "fi_kernels.hpp"

#ifndef FI_BATCH_BOUNDS
#define FI_BATCH_BOUNDS __launch_bounds__(NT)
#endif

namespace FI_NS {
// The adaptive BS3 loop of one simulation on a team: from (s.t, U[s.cur], K[s.k1]) until t
// reaches s.next_out (after an accepted step) or s.stop, or p.max_steps attempts. Every team
// member holds the same Ctl copy in LDS and takes the same decisions.
template <typename MT, bool DIST, bool ELRA>
__device__ __forceinline__ void integrate_loop(const KP& p, const Team& tm, Ctl& s, unsigned& gen, const float2* twx,
                                               const float2* twy, char* smem, double* red, int* flag,
                                               unsigned long long& tl) {
    long long attempts = 0;
    for (;;) {
        if (s.err || s.t >= s.stop || attempts >= p.max_steps) break;
        const double tn = s.t;
        double dt = s.dt;
        bool clipped = false;
        if (tn + dt >= s.stop) {
            dt = s.stop - tn;
            clipped = true;
        }
        if (!(dt > 1e-14 * fmax(1.0, fabs(tn)))) {
            if (threadIdx.x == 0) s.err = ERR_DT_MIN;
            __syncthreads();
            break;
        }
        const int k1 = s.k1;
        const int s2 = (k1 + 1) & 3, s3 = (k1 + 2) & 3, s4 = (k1 + 3) & 3;
        const int cur = s.cur, unew = 1 - s.cur;
        bool ok = true;
        for (int stg = 0; stg < 3; ++stg) {
            StageDesc st{};
            st.ubase = cur;
            st.dt = dt;
            st.write_u = -1;
            st.kin = -1;
            if (stg == 0) {          // tmp = uprev + dt a21 k1
                st.ng = 1;
                st.g0 = k1;
                st.a0 = BS3_A21;
                st.kout = s2;
                st.t = tn + 0.5 * dt;
                st.t_next_max = tn + 0.75 * dt;
            } else if (stg == 1) {   // tmp = uprev + dt a32 k2 (k2 still spectral)
                st.ng = 0;
                st.kin = s2;
                st.ain = BS3_A32;
                if constexpr (ELRA) {   // the ELRA phases leave every derivative in its K slot
                    st.ng = 1;
                    st.g0 = s2;
                    st.a0 = BS3_A32;
                    st.kin = -1;
                }
                st.kout = s3;
                st.t = tn + 0.75 * dt;
                st.t_next_max = tn + dt;
            } else {                 // u = uprev + dt (a41 k1 + a42 k2 + a43 k3); error moments
                st.ng = 2;
                st.g0 = k1;
                st.a0 = BS3_A41;
                st.g1 = s2;
                st.a1 = BS3_A42;
                st.kin = s3;
                st.ain = BS3_A43;
                if constexpr (ELRA) {
                    st.ng = 3;
                    st.g2 = s3;
                    st.a2 = BS3_A43;
                    st.kin = -1;
                }
                st.write_u = unew;
                st.kout = s4;
                st.do_c2r = 1;
                st.t = tn + dt;
                st.t_next_max = tn + dt + 0.5 * PI_QMAX * dt;
                st.final_err = 1;
                st.e1 = k1;
                st.e2 = s2;
                st.e3 = s3;
                st.u_prev = cur;
            }
            bool good;
            if constexpr (ELRA) good = elra_call(p, tm, s, st, gen, twx, smem, red, flag, false, tl);
            else good = rhs_call<MT, DIST>(p, tm, s, st, gen, twx, twy, smem, red, flag, false, tl);
            if (!good) {
                ok = false;
                break;
            }
        }
        if (!ok) break;
        ++attempts;
#ifdef FI_SUBPHASE_TIMING
        unsigned long long te0 = wall_clock64();
#endif
        // error norm with the corner-BC offsets of k1..k4 folded in:
        // sum ((a - C) w)^2 = S2 - 2 C S1 + C^2 S0, from one partial per row pair
        // (one round of loads by the whole workgroup, fixed-order sum: identical in every workgroup);
        // ELRA derivatives carry no BC offset
        const double C = ELRA ? 0.0
                              : dt * (BS3_BT1 * k_offset(p, k1) + BS3_BT2 * k_offset(p, s2) +
                                      BS3_BT3 * k_offset(p, s3) + BS3_BT4 * k_offset(p, s4));
        double Smv[NV][3] = {};
        for (int w0 = threadIdx.x; w0 < p.nrowitems; w0 += VNT)
#pragma unroll
        for (int vv = 0; vv < NV; ++vv) {
            const int w = w0 + vv * NT;
            if (w >= p.nrowitems) break;
            Smv[vv][0] += p.part[4 * w];
            Smv[vv][1] += p.part[4 * w + 1];
            Smv[vv][2] += p.part[4 * w + 2];
        }
        block_sum<3>(red, Smv);
        const double* Sm = Smv[0];
        if (threadIdx.x == 0) {
            const double S2 = Sm[0], S1 = Sm[1], S0 = Sm[2];
            const double e2 = (S2 - 2.0 * C * S1 + C * C * S0) / p.N;
            // a NaN moment means a NaN in u_n, u_{n+1} or a derivative: OrdinaryDiffEq's PI
            // controller then proposes a NaN step and stops (retcode DtNaN); an infinite error
            // norm (overflow) is rejected with the smallest step factor
            const bool nan_state = isnan(S2) || isnan(S1) || isnan(S0);
            const double EEst = isnan(e2) ? INFINITY : sqrt(fmax(e2, 0.0));
            s.EEst = nan_state ? NAN : EEst;
            double qv;
            bool finite = isfinite(EEst);
            if (nan_state) {
                s.err = ERR_NONFINITE;
                s.nrej += 1;
                s.steps = 0;
            } else if (!finite) {
                s.q11 = PI_GAMMA / PI_QMIN;
                qv = 1.0 / PI_QMIN;
            } else if (EEst == 0.0) {
                qv = 1.0 / PI_QMAX;
            } else {
                const double q11 = pow(EEst, PI_BETA1);
                s.q11 = q11;
                qv = q11 / pow(s.qold, PI_BETA2);
                qv = fmax(1.0 / PI_QMAX, fmin(1.0 / PI_QMIN, qv / PI_GAMMA));
            }
            if (nan_state) {
            } else if (finite && EEst <= 1.0) {
                s.qold = fmax(EEst, PI_QOLDINIT);
                s.t = clipped ? s.stop : tn + dt;
                s.cur = unew;
                s.k1 = s4;
                s.last_u = unew;
                s.dt = dt / qv;
                s.nacc += 1;
                s.steps = 1;  // accepted flag for this attempt
            } else {
                s.dt = dt / fmin(1.0 / PI_QMIN, s.q11 / PI_GAMMA);
                s.nrej += 1;
                s.steps = 0;
            }
        }
        __syncthreads();
#ifdef FI_SUBPHASE_TIMING
        if (p.ptime && tm.rank == 0 && threadIdx.x == 0) p.ptime[22] += wall_clock64() - te0;
#endif
        if (s.steps == 1 && s.t >= s.next_out) break;
    }
}

// A single RHS evaluation on U[ubuf] at time t into K[kslot] (update_diagnostics!).
template <typename MT, bool DIST, bool ELRA>
__device__ __forceinline__ bool single_rhs(const KP& p, const Team& tm, Ctl& s, unsigned& gen, const float2* twx,
                                           const float2* twy, char* smem, double* red, int* flag,
                                           unsigned long long& tl, int ubuf, int kslot, double t, bool sparse_only,
                                           bool force_sigma) {
    StageDesc st{};
    st.ubase = ubuf;
    st.ng = 0;
    st.kin = -1;
    st.dt = 0;
    st.write_u = -1;
    st.kout = kslot;
    st.do_c2r = 1;
    st.t = t;
    st.t_next_max = t;
    st.force_sigma = force_sigma ? 1 : 0;
    if constexpr (ELRA) return elra_call(p, tm, s, st, gen, twx, smem, red, flag, sparse_only, tl);
    else return rhs_call<MT, DIST>(p, tm, s, st, gen, twx, twy, smem, red, flag, sparse_only, tl);
}

__device__ __forceinline__ void load_twiddles(const KP& p, float2* twx, float2* twy) {
    for (int k = threadIdx.x; k < p.nx; k += NT) twx[k] = p.tw32[p.px.tw + k];
    for (int k = threadIdx.x; k < p.ny; k += NT) twy[k] = p.tw32[p.py.tw + k];
}

// The parameter block is not written while a kernel runs. Reading it through the constant
// address space keeps its fields in scalar registers across stores, and lets the compiler treat
// the pointers it holds as global memory (global rather than flat loads).
using ConstKP = __attribute__((address_space(4))) KP;

// ------------------------------------------------------------------------------------------
// persistent ensemble kernel: the grid is split into nteams teams of tsize workgroups (team =
// blockIdx.x mod nteams, so with nteams a multiple of 8 a team shares one XCD); each team takes
// members from a work queue (ordered largest first) and integrates each to its own s.stop with
// team-local barriers. With fsal set, a member that reached s.stop re-evaluates its first-same-as-
// last derivative there, as after an output callback. Results do not depend on the team size:
// every cross-workgroup sum runs over fixed per-row-pair partials.
// ------------------------------------------------------------------------------------------
struct BatchArgs {
    const KP* const* kps;   // member parameter blocks (device memory)
    const int* order;       // queue position -> member
    int nmem, nteams, tsize, fsal;
    Bar* bars;              // nteams regions of (1 + 2 tsize) Bar
    int* queue;             // queue head (zeroed by the host)
    int* team_member;       // per team two slots (alternating pops): the member it is working on,
                            // written by its rank 0
    unsigned long long spin_limit;
};

// one member of the ensemble kernel. The parameter block is passed as a restrict pointer: nothing
// the integration writes aliases it, so its fields are read with scalar loads (as in the kernel of
// one context) instead of being reloaded into vector registers after every store.
template <typename MT, bool DIST, bool ELRA = false>
__device__ __forceinline__ bool batch_member(const ConstKP* __restrict__ pp, const Team& tm, Ctl& s, unsigned& gen,
                                             char* smem_all, double* red, int* flag, unsigned long long& tl,
                                             int fsal) {
    const KP& p = *(const KP*)pp;
    float2* twx = (float2*)smem_all;
    float2* twy = twx + p.nx;
    char* smem = (char*)(twy + p.ny);
    load_twiddles(p, twx, twy);
    if (threadIdx.x == 0) s = *p.ctl;
    __syncthreads();
    const bool todo = s.t < s.stop && !s.err;
    if (todo) {
        integrate_loop<MT, DIST, ELRA>(p, tm, s, gen, twx, twy, smem, red, flag, tl);
        if (*flag) return false;
        if (fsal && !s.err && s.t >= s.stop) {
            if (!single_rhs<MT, DIST, ELRA>(p, tm, s, gen, twx, twy, smem, red, flag, tl, s.cur, s.k1, s.t, false, true)) {
                if (*flag) return false;
            }
            if (threadIdx.x == 0) s.last_u = s.cur;
            __syncthreads();
        }
    }
    if (tm.rank == 0 && threadIdx.x == 0) *p.ctl = s;
    return true;
}

template <typename MT, bool DIST>
__global__ void FI_BATCH_BOUNDS fi_batch_persistent(const BatchArgs* __restrict__ ba) {
    extern __shared__ __align__(16) char smem_all[];
    __shared__ Ctl s;
    __shared__ double red[64];
    __shared__ int flag;
    __shared__ int s_m;
    const int team = (int)blockIdx.x % ba->nteams;
    const Team tm{(int)blockIdx.x / ba->nteams, ba->tsize, ba->bars + (long)team * (1 + 2 * ba->tsize)};
    unsigned gen = 0;
    unsigned long long tl = wall_clock64();
    for (unsigned pops = 0;; ++pops) {
        // rank 0 pops the next member; the team barrier publishes it. The slot alternates with
        // the pop count: a member with nothing left to do runs no team barrier, so rank 0 may pop
        // again before every workgroup has read the previous member, but it cannot pop twice
        // more before the next team barrier, which every workgroup reaches only after its read.
        int* slot = &ba->team_member[2 * team + (int)(pops & 1u)];
        if (tm.rank == 0 && threadIdx.x == 0) {
            const int q = atomicAdd(ba->queue, 1);
            __hip_atomic_store(slot, q < ba->nmem ? ba->order[q] : -1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        if (!team_sync(tm, gen, &flag, ba->spin_limit)) return;
        if (threadIdx.x == 0) s_m = __hip_atomic_load(slot, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __syncthreads();
        const int m = __builtin_amdgcn_readfirstlane(s_m);
        if (m < 0) return;
        // the member table is read as global memory (a load through a generic pointer counts as
        // divergent), so the member's parameter block is a uniform, constant pointer
        const unsigned long long pk =
            ((const __attribute__((address_space(1))) unsigned long long*)(unsigned long long)ba->kps)[m];
        if (!batch_member<MT, DIST>((const ConstKP*)pk, tm, s, gen, smem_all, red, &flag, tl, ba->fsal)) return;
    }
}

// the ensemble kernel for relaxed-mantle (ELRA) members: the same queue and teams as
// fi_batch_persistent (its body is repeated, as for the single-context kernels, so that the
// LV-ELVA instances keep their code), each member integrated by the ELRA loop
__global__ void FI_BATCH_BOUNDS fi_batch_persistent_elra(const BatchArgs* __restrict__ ba) {
    extern __shared__ __align__(16) char smem_all[];
    __shared__ Ctl s;
    __shared__ double red[64];
    __shared__ int flag;
    __shared__ int s_m;
    const int team = (int)blockIdx.x % ba->nteams;
    const Team tm{(int)blockIdx.x / ba->nteams, ba->tsize, ba->bars + (long)team * (1 + 2 * ba->tsize)};
    unsigned gen = 0;
    unsigned long long tl = wall_clock64();
    for (unsigned pops = 0;; ++pops) {
        // alternating slots, as in fi_batch_persistent
        int* slot = &ba->team_member[2 * team + (int)(pops & 1u)];
        if (tm.rank == 0 && threadIdx.x == 0) {
            const int q = atomicAdd(ba->queue, 1);
            __hip_atomic_store(slot, q < ba->nmem ? ba->order[q] : -1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        if (!team_sync(tm, gen, &flag, ba->spin_limit)) return;
        if (threadIdx.x == 0) s_m = __hip_atomic_load(slot, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __syncthreads();
        const int m = __builtin_amdgcn_readfirstlane(s_m);
        if (m < 0) return;
        const unsigned long long pk =
            ((const __attribute__((address_space(1))) unsigned long long*)(unsigned long long)ba->kps)[m];
        if (!batch_member<double, false, true>((const ConstKP*)pk, tm, s, gen, smem_all, red, &flag, tl, ba->fsal))
            return;
    }
}

}  // namespace FI_NS
