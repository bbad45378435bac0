// Phases of the LV-ELVA right-hand side and the persistent integration kernel (gfx950).
//
// One RHS evaluation = update_diagnostics!(dudt, u, sim, t) (src/simulation.jl:335-393) runs as
// three grid-wide phases separated by grid barriers:
//   rows  : stage displacement (BS3 combination + viscous corner BC, FP64), column loads,
//           the two-sweep plate stencil (src/deformation.jl:110-131, src/derivatives.jl CPU edge
//           rules) on LDS rows with +-2 halos, rhs * K/(2 eta), and a length-nx FFT of row pairs
//           packed as one complex sequence; the half spectrum is stored transposed [kx][j];
//   cols  : length-ny FFT of each spectral column, multiplication by spy/(nx ny) / (R kappa),
//           inverse FFT (src/deformation.jl:95-98);
//   c2r   : inverse real transform of row pairs -> dudt (m/yr) before the corner BC; the BC
//           offset (mean of the four raw corners) is published and applied by consumers.
// When the sparse trigger fires (src/simulation.jl:355-381) four more phases compute the elastic
// and sea-surface convolutions on the zero-padded nfx x nfy grid (both in one pass) and the
// pointwise sparse diagnostics.
#pragma once
#include "fi_device.hpp"

namespace fi {

// ------------------------------------------------------------------------------------------
// grid barrier (all workgroups co-resident: cooperative launch). Bounded wait: a barrier that
// does not complete within spin_limit wall-clock ticks raises the abort flag and every
// workgroup leaves the kernel.
// ------------------------------------------------------------------------------------------
__device__ __forceinline__ bool grid_sync(const KP& p, unsigned& gen, int* s_flag) {
    __syncthreads();
    if (threadIdx.x == 0) {
        Bar* bar = p.bar;
        __threadfence();
        const unsigned g = gen;
        const unsigned arrived = atomicAdd(&bar->count, 1u);
        if (arrived == gridDim.x - 1) {
            atomicExch(&bar->count, 0u);
            __threadfence();
            atomicExch(&bar->gen, g + 1u);
        } else {
            const unsigned long long t0 = wall_clock64();
            while (__hip_atomic_load(&bar->gen, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) == g) {
                if (__hip_atomic_load(&bar->abort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) break;
                if (wall_clock64() - t0 > p.spin_limit) {
                    atomicExch(&bar->abort, 1);
                    break;
                }
                __builtin_amdgcn_s_sleep(1);
            }
        }
        __threadfence();
        gen = g + 1u;
        *s_flag = __hip_atomic_load(&bar->abort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    __syncthreads();
    return *s_flag == 0;
}

// Block-wide sum of three doubles per thread (fixed order -> identical in every workgroup).
__device__ __forceinline__ void block_sum3(double* red, double a, double b, double c, double out[3]) {
    const int t = threadIdx.x;
    red[t] = a;
    red[NT + t] = b;
    red[2 * NT + t] = c;
    __syncthreads();
    for (int s = NT / 2; s > 0; s >>= 1) {
        if (t < s) {
            red[t] += red[t + s];
            red[NT + t] += red[NT + t + s];
            red[2 * NT + t] += red[2 * NT + t + s];
        }
        __syncthreads();
    }
    out[0] = red[0];
    out[1] = red[NT];
    out[2] = red[2 * NT];
    __syncthreads();
}

// ------------------------------------------------------------------------------------------
// stage description and per-call scalars
// ------------------------------------------------------------------------------------------
struct StageDesc {
    int ubase;         // U buffer of the base state
    int nk;            // number of derivative terms (0, 1 or 3)
    int slot0, slot1, slot2;
    double a0, a1, a2;
    double dt;
    int write_u;       // U buffer receiving the corner-corrected stage displacement, or -1
    int kout;          // K slot receiving this call's derivative
    double t;          // call time
    double t_next_max; // upper bound of the next call's time (for the stale-load store)
    int trig;          // set by rhs_call: this call ran the sparse update
    int final_err;     // accumulate the BS3 error moments in the c2r phase
    int e1, e2, e3;    // K slots of k1, k2, k3 for the error moments
    int u_prev;        // U buffer of u_n for the error moments
    int write_sigma;   // set by rhs_call: store this call's loads for a later sparse update
};

struct IceT {
    int i0, i1, exact, err;
    double tt, dtk, a;  // t - t_i0, t_i1 - t_i0, separable amplitude a(t)
};

// TimeInterpolation2D interpolate! bracket (src/interpolations.jl:51-74)
__device__ __forceinline__ IceT ice_time(const KP& p, double t) {
    IceT it;
    it.i0 = it.i1 = 0;
    it.exact = 1;
    it.err = 0;
    it.tt = 0;
    it.dtk = 1;
    it.a = 0;
    if (p.ice_mode == ICE_EXTERNAL) return it;
    const double* tv = p.ice_t;
    const int nt = p.ice_nt;
    if (t < tv[0]) {
        if (!p.ice_flat) it.err = 1;
        it.i0 = it.i1 = 0;
    } else if (t > tv[nt - 1]) {
        if (!p.ice_flat) it.err = 1;
        it.i0 = it.i1 = nt - 1;
    } else {
        int lo = 0, hi = nt - 1;  // first index with tv[i] >= t
        while (lo < hi) {
            int mid = (lo + hi) >> 1;
            if (tv[mid] >= t) hi = mid;
            else lo = mid + 1;
        }
        if (tv[lo] == t) {
            it.i0 = it.i1 = lo;
        } else {
            it.i0 = lo - 1;
            it.i1 = lo;
            it.exact = 0;
            it.tt = t - tv[lo - 1];
            it.dtk = tv[lo] - tv[lo - 1];
        }
    }
    if (p.ice_mode == ICE_SEPARABLE) {
        const double a0 = p.ice_a[it.i0], a1 = p.ice_a[it.i1];
        it.a = it.exact ? a0 : a0 + (a1 - a0) / it.dtk * it.tt;
    }
    return it;
}

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

__device__ __forceinline__ const float* kslot(const KP& p, int k) { return p.Kb + (long)k * p.Nl; }
__device__ __forceinline__ double* ubuf(const KP& p, int b) { return p.Ub + (long)b * p.Nl; }

__device__ __forceinline__ double u_raw(const KP& p, const StageDesc& st, long c) {
    const double u0 = ubuf(p, st.ubase)[c];
    if (st.nk == 0) return u0;
    double v = st.a0 * (double)kslot(p, st.slot0)[c];
    if (st.nk == 3) v += st.a1 * (double)kslot(p, st.slot1)[c] + st.a2 * (double)kslot(p, st.slot2)[c];
    return u0 + st.dt * v;
}

// apply_bc!(u, CornerBC) offset of the stage displacement (src/boundary_conditions.jl:144-148)
__device__ __forceinline__ double u_offset(const KP& p, const StageDesc& st) {
    const long n1 = p.nx - 1, n2 = (long)(p.ny - 1) * p.nx, n3 = n2 + p.nx - 1;
    const double a = 0.25 * u_raw(p, st, 0), b = 0.25 * u_raw(p, st, n1);
    const double c = 0.25 * u_raw(p, st, n2), d = 0.25 * u_raw(p, st, n3);
    return ((a + b) + (c + d)) - p.vbc_x;
}

// corner-BC offset of a raw derivative slot
__device__ __forceinline__ double k_offset(const KP& p, int slot) {
    const float* q = p.kcorner + 4 * slot;
    const double a = 0.25 * (double)q[0], b = 0.25 * (double)q[1], c = 0.25 * (double)q[2], d = 0.25 * (double)q[3];
    return ((a + b) + (c + d)) - p.vbc_x;
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

// ------------------------------------------------------------------------------------------
// rows phase
// ------------------------------------------------------------------------------------------
template <typename MT>
__device__ void phase_rows(const KP& p, const StageDesc& st, double uoff, const IceT& it, char* smem) {
    const int nx = p.nx, ny = p.ny;
    double* su = (double*)smem;
    MT* sMxx = (MT*)(su + 6 * nx);
    MT* sMyy = sMxx + 4 * nx;
    MT* sMxy = sMyy + 4 * nx;
    float2* sf = (float2*)(sMxy + 4 * nx);
    float2* sg = sf + nx;
    const double nu = p.nu;
    const bool wsig = st.write_sigma != 0;
    for (int item = blockIdx.x; item < p.nrowitems; item += gridDim.x) {
        const int j0 = 2 * item, je = min(j0 + 2, ny);
        const int jm0 = max(0, min(j0 - 1, ny - 3)), jm1 = min(ny, max(je + 1, 3));
        const int ju0 = max(0, min(jm0 - 1, ny - 3)), ju1 = min(ny, max(jm1 + 1, 3));
        // stage displacement with the corner BC applied
        for (int idx = threadIdx.x; idx < (ju1 - ju0) * nx; idx += NT) {
            const int r = idx / nx, i = idx - r * nx, j = ju0 + r;
            const long c = i + (long)j * nx;
            const double u = u_raw(p, st, c) - uoff;
            su[idx] = u;
            if (st.write_u >= 0 && j >= j0 && j < je) ubuf(p, st.write_u)[c] = u;
        }
        __syncthreads();
        // plate moments M = -D (...) on rows [jm0, jm1)
        for (int idx = threadIdx.x; idx < (jm1 - jm0) * nx; idx += NT) {
            const int r = idx / nx, i = idx - r * nx, j = jm0 + r;
            const long c = i + (long)j * nx;
            const double uxx = d2x(su + (j - ju0) * nx, i, nx, p.inv_dx2);
            const double uyy = d2y(su, ju0, j, i, nx, ny, p.inv_dy2);
            const double uxy = dxy(su, ju0, j, i, nx, ny, p.inv_dx, p.inv_dy);
            const double D = (double)p.D[c];
            sMxx[idx] = (MT)(-D * (nu * uyy + uxx));
            sMyy[idx] = (MT)(-D * (nu * uxx + uyy));
            sMxy[idx] = (MT)(-D * (1.0 - nu) * uxy);
        }
        __syncthreads();
        // rhs = -g sigma_full + dxx Mxx + 2 dxy Mxy + dyy Myy, times K/(2 eta)
        for (int idx = threadIdx.x; idx < 2 * nx; idx += NT) {
            const int r = idx / nx, i = idx - r * nx, j = j0 + r;
            float f = 0.0f;
            if (j < ny) {
                const long c = i + (long)j * nx;
                const double bxx = d2x(sMxx + (j - jm0) * nx, i, nx, p.inv_dx2);
                const double byy = d2y(sMyy, jm0, j, i, nx, ny, p.inv_dy2);
                const double bxy = dxy(sMxy, jm0, j, i, nx, ny, p.inv_dx, p.inv_dy);
                const double u = su[(j - ju0) * nx + i];
                const double H = ice_at(p, it, c);
                const double zd = (double)(st.trig ? p.zoff_trig[c] : p.zoff[c]) + u;
                const double Haf = hab(H, zd, p.rho_sw_ice);
                const double s_ice = p.rho_ice * (Haf - (double)p.H_af_ref[c]);
                const double m = p.mask[c] ? 1.0 : 0.0;
                const double s_load = m * (s_ice + (p.interactive ? (double)p.LW[c] : 0.0));
                const double s_mant = p.rho_um * u;
                const double s_full = s_load + m * ((double)p.LT[c] + s_mant);
                const double rhs = -p.g * s_full + (bxx + (2.0 * bxy + byy));
                f = (float)(rhs * (double)p.inv2eta[c]);
                if (wsig) {
                    p.sig_load_prev[c] = (float)s_load;
                    p.sig_geo_prev[c] = (float)(p.compressible_column
                                                    ? s_load + m * (s_mant - p.rho_um * (double)p.u_ref[c])
                                                    : s_full);
                }
            }
            if (r == 0) sf[i].x = f;
            else sf[i].y = f;
        }
        __syncthreads();
        float2* res = fft_lds<float2>(sf, sg, 1, nx, p.px, p.tw32, -1);
        for (int kx = threadIdx.x; kx < p.nhx; kx += NT) {
            const float2 Z = res[kx], Zm = res[kx == 0 ? 0 : nx - kx];
            float2 X0, X1;
            X0.x = 0.5f * (Z.x + Zm.x);
            X0.y = 0.5f * (Z.y - Zm.y);
            X1.x = 0.5f * (Z.y + Zm.y);
            X1.y = -0.5f * (Z.x - Zm.x);
            p.spec[(long)kx * ny + j0] = X0;
            if (j0 + 1 < ny) p.spec[(long)kx * ny + j0 + 1] = X1;
        }
        __syncthreads();
    }
}

// ------------------------------------------------------------------------------------------
// columns phase: FFT_y, * spy/(nx ny)/(R kappa), inverse FFT_y
// ------------------------------------------------------------------------------------------
__device__ void phase_cols(const KP& p, char* smem) {
    const int ny = p.ny, nb = p.colnb;
    float2* sa = (float2*)smem;
    float2* sb = sa + nb * ny;
    for (int item = blockIdx.x; item < p.ncolitems; item += gridDim.x) {
        const int kx0 = item * nb;
        const int nbb = min(nb, p.nhx - kx0);
        const long base = (long)kx0 * ny;
        for (int idx = threadIdx.x; idx < nbb * ny; idx += NT) sa[idx] = p.spec[base + idx];
        __syncthreads();
        float2* r = fft_lds<float2>(sa, sb, nbb, ny, p.py, p.tw32, -1);
        for (int idx = threadIdx.x; idx < nbb * ny; idx += NT) {
            const float m = p.mult[base + idx];
            r[idx].x *= m;
            r[idx].y *= m;
        }
        __syncthreads();
        float2* o = (r == sa) ? sb : sa;
        float2* r2 = fft_lds<float2>(r, o, nbb, ny, p.py, p.tw32, +1);
        for (int idx = threadIdx.x; idx < nbb * ny; idx += NT) p.spec[base + idx] = r2[idx];
        __syncthreads();
    }
}

// ------------------------------------------------------------------------------------------
// c2r phase: inverse real transform of row pairs; optional BS3 error moments
// ------------------------------------------------------------------------------------------
__device__ void phase_c2r(const KP& p, const StageDesc& st, char* smem, double acc[3]) {
    const int nx = p.nx, ny = p.ny, nhx = p.nhx;
    const bool even = (nx % 2) == 0;
    float2* sf = (float2*)smem;
    float2* sg = sf + nx;
    float* K = p.Kb + (long)st.kout * p.Nl;
    for (int item = blockIdx.x; item < p.nrowitems; item += gridDim.x) {
        const int j0 = 2 * item;
        const bool two = j0 + 1 < ny;
        for (int k = threadIdx.x; k < nx; k += NT) {
            float2 A, B;
            if (k < nhx) {
                A = p.spec[(long)k * ny + j0];
                if (two) B = p.spec[(long)k * ny + j0 + 1];
                else { B.x = 0; B.y = 0; }
                if (k == 0 || (even && k == nx / 2)) { A.y = 0; B.y = 0; }
            } else {
                const int kk = nx - k;
                A = p.spec[(long)kk * ny + j0];
                if (two) B = p.spec[(long)kk * ny + j0 + 1];
                else { B.x = 0; B.y = 0; }
                A.y = -A.y;
                B.y = -B.y;
            }
            float2 Z;
            Z.x = A.x - B.y;
            Z.y = A.y + B.x;
            sf[k] = Z;
        }
        __syncthreads();
        float2* res = fft_lds<float2>(sf, sg, 1, nx, p.px, p.tw32, +1);
        for (int idx = threadIdx.x; idx < 2 * nx; idx += NT) {
            const int r = idx / nx, i = idx - r * nx, j = j0 + r;
            if (j >= ny) continue;
            const long c = i + (long)j * nx;
            const float v = r ? res[i].y : res[i].x;
            K[c] = v;
            if ((j == 0 || j == ny - 1) && (i == 0 || i == nx - 1))
                p.kcorner[4 * st.kout + (j == 0 ? 0 : 2) + (i == 0 ? 0 : 1)] = v;
            if (st.final_err) {
                const double a = st.dt * (BS3_BT1 * (double)kslot(p, st.e1)[c] +
                                          BS3_BT2 * (double)kslot(p, st.e2)[c] +
                                          BS3_BT3 * (double)kslot(p, st.e3)[c] + BS3_BT4 * (double)v);
                const double un = ubuf(p, st.u_prev)[c], up = ubuf(p, st.write_u)[c];
                const double w = 1.0 / (p.abstol + fmax(fabs(un), fabs(up)) * p.reltol);
                acc[0] += (a * w) * (a * w);
                acc[1] += a * w * w;
                acc[2] += w * w;
            }
        }
        __syncthreads();
    }
}

// ------------------------------------------------------------------------------------------
// sparse update phases (src/simulation.jl:359-381)
// ------------------------------------------------------------------------------------------
// padded rows: the two convolution inputs of row j packed as one complex sequence
__device__ void phase_sp_rows(const KP& p, char* smem) {
    const int nx = p.nx, ny = p.ny, nfx = p.nfx;
    float2* sf = (float2*)smem;
    float2* sg = sf + nfx;
    for (int j = blockIdx.x; j < ny; j += gridDim.x) {
        for (int i = threadIdx.x; i < nfx; i += NT) {
            float2 z;
            z.x = 0;
            z.y = 0;
            if (i < nx) {
                const long c = i + (long)j * nx;
                if (!p.rigid_litho) z.x = p.sig_load_prev[c];  // sigma_load * K^2, K = 1
                if (p.lv_surface) z.y = p.sig_geo_prev[c];     // mass anomaly / A (A folded in the kernel)
            }
            sf[i] = z;
        }
        __syncthreads();
        float2* res = fft_lds<float2>(sf, sg, 1, nfx, p.pfx, p.tw32, -1);
        for (int kx = threadIdx.x; kx < p.nfhx; kx += NT) {
            const float2 Z = res[kx], Zm = res[kx == 0 ? 0 : nfx - kx];
            float2 A, B;
            A.x = 0.5f * (Z.x + Zm.x);
            A.y = 0.5f * (Z.y - Zm.y);
            B.x = 0.5f * (Z.y + Zm.y);
            B.y = -0.5f * (Z.x - Zm.x);
            p.pA[(long)kx * ny + j] = A;
            p.pB[(long)kx * ny + j] = B;
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
__device__ void phase_sp_cols(const KP& p, char* smem) {
    const int ny = p.ny, nfy = p.nfy, no = ny + 2;
    float2* sa = (float2*)smem;
    float2* sb = sa + 2 * nfy;
    for (int kx = blockIdx.x; kx < p.nfhx; kx += gridDim.x) {
        for (int idx = threadIdx.x; idx < 2 * nfy; idx += NT) {
            const int s = idx / nfy, j = idx - s * nfy;
            float2 v;
            v.x = 0;
            v.y = 0;
            if (j < ny) v = (s ? p.pB : p.pA)[(long)kx * ny + j];
            sa[idx] = v;
        }
        __syncthreads();
        float2* r = fft_lds<float2>(sa, sb, 2, nfy, p.pfy, p.tw32, -1);
        for (int idx = threadIdx.x; idx < 2 * nfy; idx += NT) {
            const int s = idx / nfy, ky = idx - s * nfy;
            const double2 G = (s ? p.gg_hat : p.ge_hat)[(long)kx * nfy + ky];
            const double xr = r[idx].x, xi = r[idx].y;
            r[idx].x = (float)(xr * G.x - xi * G.y);
            r[idx].y = (float)(xr * G.y + xi * G.x);
        }
        __syncthreads();
        float2* o = (r == sa) ? sb : sa;
        float2* r2 = fft_lds<float2>(r, o, 2, nfy, p.pfy, p.tw32, +1);
        for (int idx = threadIdx.x; idx < 2 * no; idx += NT) {
            const int s = idx / no, jj = idx - s * no;
            (s ? p.oB : p.oA)[(long)kx * no + jj] = r2[s * nfy + sp_row_of(p, jj)];
        }
        __syncthreads();
    }
}

// needed padded rows: inverse real transforms, crop [i1:i2] and the extended corners
__device__ void phase_sp_c2r(const KP& p, char* smem) {
    const int nx = p.nx, nfx = p.nfx, nfhx = p.nfhx, no = p.ny + 2;
    const bool even = (nfx % 2) == 0;
    float2* sf = (float2*)smem;
    float2* sg = sf + nfx;
    for (int jj = blockIdx.x; jj < no; jj += gridDim.x) {
        for (int k = threadIdx.x; k < nfx; k += NT) {
            float2 A, B;
            if (k < nfhx) {
                A = p.oA[(long)k * no + jj];
                B = p.oB[(long)k * no + jj];
                if (k == 0 || (even && k == nfx / 2)) { A.y = 0; B.y = 0; }
            } else {
                const int kk = nfx - k;
                A = p.oA[(long)kk * no + jj];
                B = p.oB[(long)kk * no + jj];
                A.y = -A.y;
                B.y = -B.y;
            }
            float2 Z;
            Z.x = A.x - B.y;
            Z.y = A.y + B.x;
            sf[k] = Z;
        }
        __syncthreads();
        float2* res = fft_lds<float2>(sf, sg, 1, nfx, p.pfx, p.tw32, +1);
        if (jj == 0 || jj == no - 1) {
            if (threadIdx.x == 0) {
                const int q = (jj == 0) ? 0 : 2;
                p.ext_corners[q] = res[0].x;
                p.ext_corners[q + 1] = res[2 * nx - 2].x;
                p.ext_corners[4 + q] = res[0].y;
                p.ext_corners[4 + q + 1] = res[2 * nx - 2].y;
            }
        } else {
            const int j = jj - 1;
            for (int ii = threadIdx.x; ii < nx; ii += NT) {
                const float2 v = res[p.i1 - 1 + ii];
                p.ue_raw[ii + (long)j * nx] = v.x;
                p.dz_raw[ii + (long)j * nx] = v.y;
            }
        }
        __syncthreads();
    }
}

// pointwise sparse diagnostics; volume partial sums for internal_update_bsl!
__device__ void phase_sp_point(const KP& p, const StageDesc& st, double uoff, const IceT& it, double acc[3]) {
    const long N = (long)p.nx * p.ny;
    const float* ec = p.ext_corners;
    const double oe = ((0.25 * ec[0] + 0.25 * ec[1]) + (0.25 * ec[2] + 0.25 * ec[3])) - p.ebc_x;
    const double og = ((0.25 * ec[4] + 0.25 * ec[5]) + (0.25 * ec[6] + 0.25 * ec[7])) - p.sbc_x;
    for (long c = (long)blockIdx.x * NT + threadIdx.x; c < N; c += (long)gridDim.x * NT) {
        const double u = u_raw(p, st, c) - uoff;
        const double H = ice_at(p, it, c);
        const double ue_old = p.ue[c], zss_old = p.z_ss[c];
        const double zb_old = (double)p.z_b_ref[c] + ue_old + u;
        const double haf_pre = hab(H, zb_old - zss_old, p.rho_sw_ice);
        acc[0] += haf_pre * p.A;
        acc[1] += H * p.A;
        acc[2] += fmax(zss_old - zb_old, 0.0) * p.A;
        const double ue_new = p.rigid_litho ? 0.0 : (double)p.ue_raw[c] - oe;
        const double dz_new = p.lv_surface ? (double)p.dz_raw[c] - og : (double)p.dz_ss[c];
        const double zss_new = (double)p.z_ss_ref[c] + dz_new + p.bsl_z;
        const double haf = hab(H, zb_old - zss_new, p.rho_sw_ice);
        const bool mgc = haf > 0;
        const bool moc = (zss_new - zb_old > 0) && !mgc;
        const double buf = fmax(zss_new - zb_old, 0.0);
        const double hw = (H <= 1 ? buf : 0.0) + ((!mgc && H > 1) ? (buf - H * p.rho_ice_sw) : 0.0);
        const double m = p.mask[c] ? 1.0 : 0.0;
        const double sw = p.interactive ? p.rho_seawater * (zss_new - (double)p.z_ss_ref[c]) * (moc ? 1.0 : 0.0) * m : 0.0;
        const double ca_litho = p.rho_litho * (ue_new - (double)p.ue_ref[c]);
        p.ue[c] = (float)ue_new;
        p.dz_ss[c] = (float)dz_new;
        p.z_ss[c] = (float)zss_new;
        p.mg[c] = mgc;
        p.mo[c] = moc;
        p.H_water[c] = (float)hw;
        p.LW[c] = (float)sw;
        p.LT[c] = (float)(ca_litho - p.rho_um * (double)p.u_ref[c]);
        p.zoff[c] = (float)((double)p.z_b_ref[c] + ue_new - zss_new);
        p.zoff_trig[c] = (float)((double)p.z_b_ref[c] + ue_old - zss_new);
    }
}

// ------------------------------------------------------------------------------------------
// one RHS call
// ------------------------------------------------------------------------------------------
template <typename MT>
__device__ bool rhs_call(const KP& p, Ctl& s, StageDesc& st, unsigned& gen, char* smem, double* red, int* flag,
                         bool sparse_only) {
    const IceT it = ice_time(p, st.t);
    if (it.err) {
        if (threadIdx.x == 0) s.err = ERR_TIME_RANGE;
        __syncthreads();
        return false;
    }
    const double uoff = u_offset(p, st);
    const bool trig = sparse_only || (((st.t - s.t0) / p.dt_sparse) >= (double)s.count);
    st.trig = trig ? 1 : 0;
    if (trig) {
        phase_sp_rows(p, smem);
        if (!grid_sync(p, gen, flag)) return false;
        phase_sp_cols(p, smem);
        if (!grid_sync(p, gen, flag)) return false;
        phase_sp_c2r(p, smem);
        if (!grid_sync(p, gen, flag)) return false;
        double acc[3] = {0, 0, 0}, tot[3];
        phase_sp_point(p, st, uoff, it, acc);
        block_sum3(red, acc[0], acc[1], acc[2], tot);
        if (threadIdx.x == 0) {
            p.vpart[4 * blockIdx.x + 0] = tot[0];
            p.vpart[4 * blockIdx.x + 1] = tot[1];
            p.vpart[4 * blockIdx.x + 2] = tot[2];
        }
        if (!grid_sync(p, gen, flag)) return false;
        if (threadIdx.x == 0) {
            double v0 = 0, v1 = 0, v2 = 0;
            for (int w = 0; w < (int)gridDim.x; ++w) {
                v0 += p.vpart[4 * w];
                v1 += p.vpart[4 * w + 1];
                v2 += p.vpart[4 * w + 2];
            }
            const double Vold = s.V_af + s.V_den + s.V_pov;
            s.V_af = p.vol_goelzer ? v0 * p.rho_ice_sw : 0.0;
            s.V_den = p.den_goelzer ? v1 * (p.rho_ice / p.rho_water - p.rho_ice / p.rho_seawater) : 0.0;
            s.V_pov = p.pov_goelzer ? v2 : 0.0;
            s.delta_V = (s.V_af + s.V_den + s.V_pov) - Vold;
            s.z_bsl = p.bsl_z;
            s.count += 1;
        }
        __syncthreads();
        if (sparse_only) return true;
    }
    // store this call's loads if the next call may trigger (the sparse update reads the
    // column anomalies of the previous call, src/simulation.jl:362-385)
    st.write_sigma = (p.write_sigma_always || (((st.t_next_max - s.t0) / p.dt_sparse) >= (double)s.count)) ? 1 : 0;
    phase_rows<MT>(p, st, uoff, it, smem);
    if (!grid_sync(p, gen, flag)) return false;
    phase_cols(p, smem);
    if (!grid_sync(p, gen, flag)) return false;
    double acc[3] = {0, 0, 0};
    phase_c2r(p, st, smem, acc);
    if (st.final_err) {
        double tot[3];
        block_sum3(red, acc[0], acc[1], acc[2], tot);
        if (threadIdx.x == 0) {
            p.part[4 * blockIdx.x + 0] = tot[0];
            p.part[4 * blockIdx.x + 1] = tot[1];
            p.part[4 * blockIdx.x + 2] = tot[2];
        }
    }
    if (!grid_sync(p, gen, flag)) return false;
    if (threadIdx.x == 0) {
        s.nrhs += 1;
        s.t_last = st.t;
        s.last_k = st.kout;
        s.last_trig = trig ? 1 : 0;
    }
    __syncthreads();
    return true;
}

}  // namespace fi
