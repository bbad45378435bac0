// =====================================================================================
// ORACLE -- TEST INFRASTRUCTURE ONLY.
//
// FP64 CPU restatement of FastIsostasy.jl's LV-ELVA deformation path, exporting the same
// C ABI as the HIP engine (include/fastiso.h) plus ref_* hooks for the unit checks.
// Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it, and only
// as the checker or as the timed CPU baseline; the product never links or calls it.
//
// Each routine follows the cited reference file:line (paths relative to the reference
// repository root) in operation order, including the reference's ordering quirks:
//   - the sparse elastic / sea-surface update reads the column anomalies of the PREVIOUS
//     RHS call (src/simulation.jl:355-385);
//   - apply_bc! mutates the solver's stage vector (src/simulation.jl:341);
//   - the sparse trigger is evaluated on every RHS call, rejected stages included.
//
// What pins each part (tests/test_oracle_*.py):
//   FFT                 numpy.fft on random data (library routine);
//   stencils            closed form of X^2 Y^2 (test/test_derivatives.jl:15-26) + exactness of
//                       the one-sided edge formulas on quadratics;
//   padded convolution  brute-force direct 2D convolution (test/test_convolution.jl:2-10 uses
//                       DSP.conv) and delta kernels;
//   Green tables        closed form of the geoid kernel; GL nodes (closed form n<=5); elastic
//                       table vs an independent numpy quadrature of the Farrell table;
//   spectral operator   eigenfunction test: cos modes are scaled by 1/kappa exactly;
//   BS3 integrator      order test on du/dt = lambda u and the published tableau values;
//   full trajectory     the analytic disc-load solution (src/analytic_solutions.jl:11-65)
//                       evaluated independently with scipy.special.j0/j1.
//   Parity unpinned: the step-size sequence (OrdinaryDiffEq controller internals are not in
//   the reference; constants chosen below are the published PI-controller defaults).
// =====================================================================================
#include <algorithm>
#include <chrono>
#include <cmath>
#include <complex>
#include <omp.h>
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "../include/fastiso.h"
#include "ref_fft.hpp"

using fiorc::cd;
using fiorc::FFT;

namespace {

typedef std::vector<double> vec;

// ---------------------------------------------------------------------------------------
// Small helpers
// ---------------------------------------------------------------------------------------

// nextprod((2,3,5,7), n)  (src/convolutions.jl:17-18)
int nextfastfft(int n) {
    for (int m = std::max(n, 1);; ++m) {
        int r = m;
        for (int p : {2, 3, 5, 7}) while (r % p == 0) r /= p;
        if (r == 1) return m;
    }
}

// fftint(N)  (src/derivatives.jl:216-222)
std::vector<double> fftint(int N) {
    std::vector<double> f;
    int N2 = N / 2;
    for (int k = 0; k <= N2; ++k) f.push_back(k);
    if (N % 2 == 0) for (int k = N2 - 1; k >= 1; --k) f.push_back(k);
    else for (int k = N2; k >= 1; --k) f.push_back(k);
    return f;
}

// samesize_conv_indices(N, M)  (src/convolutions.jl:196-204); 1-based as in the reference
void samesize_conv_indices(int N, int M, int& j1, int& j2) {
    j1 = (N % 2 == 0) ? M : M + 1;
    j2 = 2 * N - 1 - M;
}

// get_differential_fourier + DC fix (src/derivatives.jl:201-214, src/domain.jl:151-155)
vec pseudodiff(int nx, int ny, double Wx, double Wy) {
    double mux = M_PI / Wx, muy = M_PI / Wy;
    std::vector<double> fx = fftint(nx), fy = fftint(ny);
    vec p((size_t)nx * ny);
    for (int j = 0; j < ny; ++j)
        for (int i = 0; i < nx; ++i) {
            double X = mux * fx[i], Y = muy * fy[j];
            p[i + (size_t)j * nx] = std::sqrt(X * X + Y * Y);
        }
    // pseudodiff[1,1] = 1e-3 * mean([pseudodiff[1,2], pseudodiff[2,1]])
    p[0] = 1e-3 * ((p[(size_t)nx] + p[1]) / 2.0);
    return p;
}

// Gauss-Legendre nodes/weights on [-1,1], ascending (FastGaussQuadrature.gausslegendre;
// textbook Newton iteration on the Legendre recurrence).
void gauss_legendre(int n, vec& x, vec& w) {
    x.assign(n, 0.0);
    w.assign(n, 0.0);
    for (int i = 0; i < (n + 1) / 2; ++i) {
        double z = std::cos(M_PI * (i + 0.75) / (n + 0.5)), pp = 0.0;
        for (int it = 0; it < 100; ++it) {
            double p1 = 1.0, p2 = 0.0;
            for (int k = 1; k <= n; ++k) {
                double p3 = p2;
                p2 = p1;
                p1 = ((2.0 * k - 1.0) * z * p2 - (k - 1.0) * p3) / k;
            }
            pp = n * (z * p1 - p2) / (z * z - 1.0);
            double z1 = z;
            z = z1 - p1 / pp;
            if (std::fabs(z - z1) < 1e-16) break;
        }
        x[i] = -z;
        x[n - 1 - i] = z;
        w[i] = w[n - 1 - i] = 2.0 / ((1.0 - z * z) * pp * pp);
    }
    if (n % 2 == 1) x[n / 2] = 0.0;
}

// get_loadgreen with the linear interpolant (src/material.jl:299-332)
double loadgreen(double r, const vec& rm, const vec& ge) {
    if (r < 0.01) return ge[0] / (rm[1] * 1e12);
    if (r > rm.back()) return 0.0;
    size_t k = std::upper_bound(rm.begin(), rm.end(), r) - rm.begin();  // rm[k-1] <= r < rm[k]
    if (k >= rm.size()) return ge.back() / (r * 1e12);
    double t = (r - rm[k - 1]) / (rm[k] - rm[k - 1]);
    double g = ge[k - 1] + t * (ge[k] - ge[k - 1]);
    return g / (r * 1e12);
}

// get_elastic_green with quadrature2D (src/material.jl:340-364, src/utils.jl:146-186)
vec elastic_green(int nx, int ny, double Wx, double Wy, const vec& rm, const vec& ge, int nq) {
    double dx = 2 * Wx / nx, dy = 2 * Wy / ny;
    int mx = nx / 2, my = ny / 2;
    vec xq, wq;
    gauss_legendre(nq, xq, wq);
    vec G((size_t)nx * ny);
    for (int i = 1; i <= nx; ++i)
        for (int j = 1; j <= ny; ++j) {
            int p = i - mx - 1, q = j - my - 1;
            double x1 = p * dx, x2 = (p + 1) * dx, y1 = q * dy, y2 = (q + 1) * dy;
            double mxl = 2.0 / (x2 - x1), pxl = -1.0 - mxl * x1;
            double myl = 2.0 / (y2 - y1), pyl = -1.0 - myl * y1;
            double s = 0.0;
            for (int a = 0; a < nq; ++a)
                for (int b = 0; b < nq; ++b) {
                    double xx = (xq[a] - pxl) / mxl, yy = (xq[b] - pyl) / myl;
                    s = s + loadgreen(std::sqrt(xx * xx + yy * yy), rm, ge) * wq[a] * wq[b] / mxl / myl;
                }
            G[(i - 1) + (size_t)(j - 1) * nx] = s;
        }
    return G;
}

// get_dz_ss_green (src/sealevel.jl:208-219)
vec geoid_green(int nx, int ny, double Wx, double Wy, double r_pole, double mE) {
    double dx = 2 * Wx / nx, dy = 2 * Wy / ny;
    vec G((size_t)nx * ny);
    auto unb = [&](double R) { return r_pole / (2.0 * mE * std::sin(R / (2.0 * r_pole))); };
    double gmax = unb(dx / 2);
    for (int j = 0; j < ny; ++j)
        for (int i = 0; i < nx; ++i) {
            double X = -Wx + (i + 1) * dx, Y = -Wy + (j + 1) * dy;
            double R = std::sqrt(X * X + Y * Y);
            G[i + (size_t)j * nx] = std::min(unb(R), gmax);
        }
    return G;
}

// ---------------------------------------------------------------------------------------
// Finite differences, CPU rule (src/derivatives.jl:48-57,66-76,114-124,166-191)
// ---------------------------------------------------------------------------------------
// Dx, Dy: per-cell grid steps K dx, K dy (RegionalDomain.Dx/.Dy, src/domain.jl:167); each
// difference is divided by the step at its output cell
void dxx(double* du, const double* u, int nx, int ny, const double* Dx) {
#pragma omp parallel for schedule(static)
    for (int j = 0; j < ny; ++j) {
        const double* c = u + (size_t)j * nx;
        const double* h = Dx + (size_t)j * nx;
        double* o = du + (size_t)j * nx;
        for (int i = 1; i < nx - 1; ++i) o[i] = (c[i + 1] - 2 * c[i] + c[i - 1]) / (h[i] * h[i]);
        o[0] = (c[2] - 2 * c[1] + c[0]) / (h[0] * h[0]);
        o[nx - 1] = (c[nx - 1] - 2 * c[nx - 2] + c[nx - 3]) / (h[nx - 1] * h[nx - 1]);
    }
}
void dyy(double* du, const double* u, int nx, int ny, const double* Dy) {
    auto U = [&](int i, int j) { return u[i + (size_t)j * nx]; };
#pragma omp parallel for schedule(static)
    for (int j = 0; j < ny; ++j) {
        double* o = du + (size_t)j * nx;
        const double* h = Dy + (size_t)j * nx;
        if (j == 0) for (int i = 0; i < nx; ++i) o[i] = (U(i, 2) - 2 * U(i, 1) + U(i, 0)) / (h[i] * h[i]);
        else if (j == ny - 1)
            for (int i = 0; i < nx; ++i) o[i] = (U(i, ny - 1) - 2 * U(i, ny - 2) + U(i, ny - 3)) / (h[i] * h[i]);
        else for (int i = 0; i < nx; ++i) o[i] = (U(i, j + 1) - 2 * U(i, j) + U(i, j - 1)) / (h[i] * h[i]);
    }
}
void dx1(double* du, const double* u, int nx, int ny, const double* Dx) {
#pragma omp parallel for schedule(static)
    for (int j = 0; j < ny; ++j) {
        const double* c = u + (size_t)j * nx;
        const double* h = Dx + (size_t)j * nx;
        double* o = du + (size_t)j * nx;
        for (int i = 1; i < nx - 1; ++i) o[i] = (c[i + 1] - c[i - 1]) / (2 * h[i]);
        o[0] = (c[1] - c[0]) / h[0];
        o[nx - 1] = (c[nx - 1] - c[nx - 2]) / h[nx - 1];
    }
}
void dy1(double* du, const double* u, int nx, int ny, const double* Dy) {
    auto U = [&](int i, int j) { return u[i + (size_t)j * nx]; };
#pragma omp parallel for schedule(static)
    for (int j = 0; j < ny; ++j) {
        double* o = du + (size_t)j * nx;
        const double* h = Dy + (size_t)j * nx;
        if (j == 0) for (int i = 0; i < nx; ++i) o[i] = (U(i, 1) - U(i, 0)) / h[i];
        else if (j == ny - 1) for (int i = 0; i < nx; ++i) o[i] = (U(i, ny - 1) - U(i, ny - 2)) / h[i];
        else for (int i = 0; i < nx; ++i) o[i] = (U(i, j + 1) - U(i, j - 1)) / (2 * h[i]);
    }
}
// update_second_derivatives!(uxx, uyy, ux, uxy, u1, u2, u3)
void second_derivatives(double* uxx, double* uyy, double* ux, double* uxy, const double* u1,
                        const double* u2, const double* u3, int nx, int ny, const double* Dx, const double* Dy) {
    dxx(uxx, u1, nx, ny, Dx);
    dyy(uyy, u2, nx, ny, Dy);
    dx1(ux, u3, nx, ny, Dx);
    dy1(uxy, ux, nx, ny, Dy);
}

// ---------------------------------------------------------------------------------------
// 2D FFT on column-major (nx, ny) complex data
// ---------------------------------------------------------------------------------------
void fft2d(cd* a, int nx, int ny, const FFT& fx, const FFT& fy, int sign) {
#pragma omp parallel
    {
        std::vector<cd> work(std::max(nx, ny)), col(ny);
#pragma omp for schedule(static)
        for (int j = 0; j < ny; ++j) fx.exec(a + (size_t)j * nx, work.data(), sign);
#pragma omp for schedule(static)
        for (int i = 0; i < nx; ++i) {
            for (int j = 0; j < ny; ++j) col[j] = a[i + (size_t)j * nx];
            fy.exec(col.data(), work.data(), sign);
            for (int j = 0; j < ny; ++j) a[i + (size_t)j * nx] = col[j];
        }
    }
}

// Padded linear convolution (ConvolutionPlanHelpers / ConvolutionPlan / conv!,
// src/convolutions.jl:1-15,21-56,81-106). The rfft/irfft pair is restated as the full complex
// transform of the real padded array followed by the real part, which is the same mathematics.
struct ConvPlan {
    int nx = 0, ny = 0, nfx = 0, nfy = 0;
    FFT fx, fy;
    void init(int nx_, int ny_) {
        nx = nx_;
        ny = ny_;
        nfx = nextfastfft(2 * nx - 1);
        nfy = nextfastfft(2 * ny - 1);
        fx = FFT(nfx);
        fy = FFT(nfy);
    }
    std::vector<cd> spectrum(const double* kernel) const {
        std::vector<cd> k((size_t)nfx * nfy, cd(0.0, 0.0));
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i) k[i + (size_t)j * nfx] = kernel[i + (size_t)j * nx];
        fft2d(k.data(), nfx, nfy, fx, fy, -1);
        return k;
    }
    // output_cropped = (irfft(rfft(pad(input)) .* kernel_fft))[1:2nx-1, 1:2ny-1]
    void conv_cropped(const double* input, const std::vector<cd>& khat, vec& cropped) const {
        std::vector<cd> a((size_t)nfx * nfy, cd(0.0, 0.0));
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i) a[i + (size_t)j * nfx] = input[i + (size_t)j * nx];
        fft2d(a.data(), nfx, nfy, fx, fy, -1);
        for (size_t k = 0; k < a.size(); ++k) a[k] *= khat[k];
        fft2d(a.data(), nfx, nfy, fx, fy, +1);
        double s = 1.0 / ((double)nfx * nfy);
        int ex = 2 * nx - 1, ey = 2 * ny - 1;
        cropped.resize((size_t)ex * ey);
        for (int j = 0; j < ey; ++j)
            for (int i = 0; i < ex; ++i) cropped[i + (size_t)j * ex] = a[i + (size_t)j * nfx].real() * s;
    }
};

// apply_bc!(X, OffsetBC) for a CornerBC (src/boundary_conditions.jl:144-148,199-207,221-243):
// X .-= sum(W .* X) - x_border with W = 1/4 on the four corners.
void apply_corner_bc(double* X, int nx, int ny, double x_border) {
    double a = 0.25 * X[0], b = 0.25 * X[nx - 1];
    double c = 0.25 * X[(size_t)(ny - 1) * nx], d = 0.25 * X[(nx - 1) + (size_t)(ny - 1) * nx];
    double off = ((a + b) + (c + d)) - x_border;
    size_t N = (size_t)nx * ny;
    for (size_t k = 0; k < N; ++k) X[k] -= off;
}

// samesize_conv!(output, input, p, h, domain, bc, ExtendedBCSpace) (src/convolutions.jl:134-143)
void samesize_conv_ext(double* out, const double* input, const ConvPlan& cp, const std::vector<cd>& khat,
                       int i1, int i2, int j1, int j2, double x_border) {
    vec crop;
    cp.conv_cropped(input, khat, crop);
    int ex = 2 * cp.nx - 1, ey = 2 * cp.ny - 1;
    apply_corner_bc(crop.data(), ex, ey, x_border);
    int nx = cp.nx;
    for (int j = j1; j <= j2; ++j)
        for (int i = i1; i <= i2; ++i)
            out[(i - i1) + (size_t)(j - j1) * nx] = crop[(i - 1) + (size_t)(j - 1) * ex];
}

// MaxwellMantle under a rigid or laterally constant lithosphere: the implicit step
inline bool implicit_step(const fi_config& g) {
    return g.mantle == FI_MAXWELL_MANTLE && g.lithosphere != FI_LATERALLY_VARIABLE_LITHOSPHERE;
}

// height_above_floatation (src/topography.jl:20-23)
inline double hab(double H, double zb, double zss, double rho_sw_ice) {
    return std::max(H + std::min(zb - zss, 0.0) * rho_sw_ice, 0.0);
}

}  // namespace

// =======================================================================================
// Context
// =======================================================================================
struct fi_ctx {
    fi_config cfg;
    std::string err;
    int nx = 0, ny = 0, mx = 0, my = 0;
    size_t N = 0;
    double dx = 0, dy = 0, A = 0, rho_sw_ice = 0;
    // RegionalDomain.K and what derives from it (src/domain.jl:140-147,167): Dx = K dx, Dy = K dy,
    // cell area dx dy K^2
    vec K, Dx, Dy, Acell;
    int i1 = 0, i2 = 0, j1 = 0, j2 = 0;
    vec pd;
    FFT vfx, vfy;  // n x n complex transforms of the viscous operator (src/tools.jl:100-109)
    ConvPlan cp;
    std::vector<cd> ge_hat, gg_hat, gv_hat;
    // static inputs
    vec D, eta, spinv, maskactive, green_e, green_g, green_v, tau, litho_h, u_ref, ue_ref, dz_ss_ref, z_b_ref,
        H_ice_ref_in;
    bool has_D = false, has_eta = false, has_spinv = false, has_ge = false, has_gg = false, has_gv = false,
         has_tau = false, has_href = false;
    // ice forcing
    int ice_mode = 0;  // 0 external, 1 snapshots, 2 separable
    vec ice_t, ice_a, ice_H;
    int ice_flat = 0;
    // Timer history (src/simulation.jl:28-52): capacity, the recorded calls' times and wall seconds
    long long hist_cap = 0;
    vec hist_t, hist_wall;
    // ImposedSeaSurface snapshots of dz_ss (src/sealevel.jl:132-134)
    vec ss_t, ss_X;
    int ss_flat = 0;
    // barystatic sea-level rule (fi_set_bsl); A_OCEAN_PD (src/barystatic_sealevel.jl:1) by default
    int bsl_rule = FI_BSL_CONSTANT, bsl_update = FI_BSL_INTERNAL_UPDATE, bsl_flat = 0;
    double bsl_A_ref = 3.625e14;
    vec bsl_x, bsl_y;
    vec tout;
    // reference state
    vec H_ice_ref, H_af_ref, H_water_ref, z_ss_ref;
    std::vector<uint8_t> mg_ref, mo_ref;
    // current state (src/state.jl:49-77)
    vec u, ue, dudt, H_ice, H_af, H_water, z_b, dz_ss, z_ss, u_eq;
    std::vector<uint8_t> mg, mo;
    vec ca_ice, ca_sw, ca_sed, ca_litho, ca_mantle, ca_load, ca_full;
    double V_af = 0, V_pov = 0, V_den = 0, delta_V = 0, z_bsl = 0, A_bsl = 0;
    int64_t count = 0;
    bool k1_stale = false;
    double rkc_rho = 0;   // spectral radius estimate of fi_integrate_rkc (1/yr)
    uint64_t plans[4] = {0, 0, 0, 0};   // FFT plan codes pinned by fi_set_fft_plans (unused here)
    double t0 = 0, t_end = 0, t_timer = 0;
    bool initialized = false;
    // preallocated buffers (PreAllocated, src/tools.jl:4-16)
    vec rhs, bxx, byy, bx, bxy, Mxx, Myy, Mxy, buf;
    std::vector<cd> fftrhs;
    // integrator
    bool integ = false;
    double t = 0, dt = 0, qold = 1e-4, q11 = 1, EEst = 0;
    vec uprev, k1, k2, k3, k4, tmp;
    int64_t nrhs = 0, nacc = 0, nrej = 0, kout = 0;
    int fail = 0;
};

namespace {

int fail(fi_ctx* c, int code, const std::string& msg) {
    if (c) c->err = msg;
    return code;
}

// TimeInterpolation2D interpolate! (src/interpolations.jl:51-74)
int interpolate_ice(fi_ctx* c, double t, double* out) {
    const vec& tv = c->ice_t;
    int nt = (int)tv.size();
    size_t N = c->N;
    auto snap = [&](int k, size_t p) {
        return c->ice_mode == 1 ? c->ice_H[(size_t)k * N + p] : c->ice_a[k] * c->ice_H[p];
    };
    if (t < tv.front()) {
        if (!c->ice_flat) return fail(c, FI_ERR_TIME_RANGE, "TimeInterpolation2D out of range.");
        for (size_t p = 0; p < N; ++p) out[p] = snap(0, p);
    } else if (t > tv.back()) {
        if (!c->ice_flat) return fail(c, FI_ERR_TIME_RANGE, "TimeInterpolation2D out of range.");
        for (size_t p = 0; p < N; ++p) out[p] = snap(nt - 1, p);
    } else {
        int i = (int)(std::lower_bound(tv.begin(), tv.end(), t) - tv.begin());  // searchsortedfirst - 1 (0-based)
        if (i < nt && tv[i] == t) {
            for (size_t p = 0; p < N; ++p) out[p] = snap(i, p);
        } else {
            i -= 1;
            double f = (t - tv[i]);
            double dtk = tv[i + 1] - tv[i];
            for (size_t p = 0; p < N; ++p) {
                double a = snap(i, p), b = snap(i + 1, p);
                out[p] = a + (b - a) / dtk * f;
            }
        }
    }
    return FI_OK;
}

// TimeInterpolation2D interpolate! (src/interpolations.jl:51-74) over nt snapshots of N cells
int interpolate_snapshots(fi_ctx* c, const vec& tv, const vec& X, int flat, double t, double* out) {
    const int nt = (int)tv.size();
    const size_t N = c->N;
    if (t < tv.front() || t > tv.back()) {
        if (!flat) return fail(c, FI_ERR_TIME_RANGE, "TimeInterpolation2D out of range.");
        const size_t k = t < tv.front() ? 0 : (size_t)nt - 1;
        std::copy(X.begin() + k * N, X.begin() + (k + 1) * N, out);
        return FI_OK;
    }
    const int i = (int)(std::lower_bound(tv.begin(), tv.end(), t) - tv.begin());  // searchsortedfirst
    if (tv[i] == t) {
        std::copy(X.begin() + (size_t)i * N, X.begin() + (size_t)(i + 1) * N, out);
        return FI_OK;
    }
    const double f = t - tv[i - 1], dtk = tv[i] - tv[i - 1];
    const double* a = X.data() + (size_t)(i - 1) * N;
    const double* b = X.data() + (size_t)i * N;
    for (size_t p = 0; p < N; ++p) out[p] = a[p] + (b[p] - a[p]) / dtk * f;
    return FI_OK;
}

// watercolumn! (src/loads.jl:69-85)
void watercolumn(fi_ctx* c, const vec& H, const std::vector<uint8_t>& mg, const vec& zb, const vec& zss, vec& Hw) {
    double r = c->cfg.rho_ice / c->cfg.rho_seawater;
    for (size_t p = 0; p < c->N; ++p) {
        double b = std::max(zss[p] - zb[p], 0.0);
        Hw[p] = (H[p] <= 1 ? 1.0 : 0.0) * b + (mg[p] ? 0.0 : 1.0) * (H[p] > 1 ? 1.0 : 0.0) * (b - H[p] * r);
    }
}

double sum_times_A(fi_ctx* c, const vec& X) {
    double s = 0.0;
    for (size_t p = 0; p < c->N; ++p) s += X[p] * c->Acell[p];
    return s;
}

// update_V_af!/update_V_den!/update_V_pov! (src/sealevel.jl:273-331)
void update_volumes(fi_ctx* c) {
    const fi_config& g = c->cfg;
    if (g.volume_contribution == FI_GOELZER_CONTRIBUTION)
        c->V_af = sum_times_A(c, c->H_af) * g.rho_ice / g.rho_seawater;
    else c->V_af = 0.0;
    if (g.density_contribution == FI_GOELZER_CONTRIBUTION) {
        double f = g.rho_ice / g.rho_water - g.rho_ice / g.rho_seawater;
        c->V_den = sum_times_A(c, c->H_ice) * f;
    } else c->V_den = 0.0;
    if (g.adjustment_contribution == FI_GOELZER_CONTRIBUTION) {
        double s = 0.0;
        for (size_t p = 0; p < c->N; ++p) s += std::max(c->z_ss[p] - c->z_b[p], 0.0) * c->Acell[p];
        c->V_pov = s;
    } else c->V_pov = 0.0;
}

void update_Haf(fi_ctx* c) {
    for (size_t p = 0; p < c->N; ++p) c->H_af[p] = hab(c->H_ice[p], c->z_b[p], c->z_ss[p], c->rho_sw_ice);
}
void columnanom_ice(fi_ctx* c) {
    for (size_t p = 0; p < c->N; ++p) c->ca_ice[p] = c->cfg.rho_ice * (c->H_af[p] - c->H_af_ref[p]);
}

void set_distortion(fi_ctx* c, const vec& K) {
    c->K = K;
    c->Dx.resize(c->N);
    c->Dy.resize(c->N);
    c->Acell.resize(c->N);
    for (size_t p = 0; p < c->N; ++p) {
        c->Dx[p] = K[p] * c->dx;
        c->Dy[p] = K[p] * c->dy;
        c->Acell[p] = (c->dx * c->dy) * (K[p] * K[p]);
    }
}

// interpolate(t, ::TimeInterpolation0D) (src/interpolations.jl:28-49): the knot value on an
// exact hit (searchsortedfirst), else linear between the knots around t; outside them the end
// value when flat, else an error
double interpolate0d(const vec& x, const vec& y, int flat, double t, bool* err) {
    const size_t n = x.size();
    if (t < x.front()) {
        if (!flat) *err = true;
        return y.front();
    }
    if (t > x.back()) {
        if (!flat) *err = true;
        return y.back();
    }
    const size_t i = (size_t)(std::lower_bound(x.begin(), x.end(), t) - x.begin());
    if (i < n && x[i] == t) return y[i];
    return y[i - 1] + (y[i] - y[i - 1]) / (x[i] - x[i - 1]) * (t - x[i - 1]);
}

// The sparse block of update_diagnostics! (src/simulation.jl:359-381)
int sparse_block(fi_ctx* c) {
    const fi_config& g = c->cfg;
    size_t N = c->N;
    // update_elasticresponse! (src/deformation.jl:175-189)
    if (g.lithosphere == FI_RIGID_LITHOSPHERE) {
        std::fill(c->ue.begin(), c->ue.end(), 0.0);
    } else {
        for (size_t p = 0; p < N; ++p) c->buf[p] = c->ca_load[p] * (c->K[p] * c->K[p]);
        samesize_conv_ext(c->ue.data(), c->buf.data(), c->cp, c->ge_hat, c->i1, c->i2, c->j1, c->j2,
                          g.elastic_x_border);
    }
    // columnanom_litho! (src/loads.jl:41-44)
    for (size_t p = 0; p < N; ++p) c->ca_litho[p] = g.rho_litho * (c->ue[p] - c->ue_ref[p]);
    // internal_update_bsl! / external: update_delta_V! (src/sealevel.jl:242-262), then with an
    // internal update update_bsl!(bsl, -delta_V, t) (src/barystatic_sealevel.jl:215-234)
    double Vold = c->V_af + c->V_den + c->V_pov;
    update_volumes(c);
    double Vnew = c->V_af + c->V_den + c->V_pov;
    c->delta_V = Vnew - Vold;
    if (c->bsl_update == FI_BSL_INTERNAL_UPDATE) {
        const double dV = -c->delta_V;
        bool err = false;
        switch (c->bsl_rule) {
            case FI_BSL_CONSTANT_OCEAN_SURFACE:
                c->A_bsl = c->bsl_A_ref;
                c->z_bsl += dV / c->A_bsl;
                break;
            case FI_BSL_PIECEWISE_CONSTANT:
                c->z_bsl += dV / c->A_bsl;
                c->A_bsl = interpolate0d(c->bsl_x, c->bsl_y, c->bsl_flat, c->z_bsl, &err);
                break;
            case FI_BSL_IMPOSED:
                c->z_bsl = interpolate0d(c->bsl_x, c->bsl_y, c->bsl_flat, c->t_timer, &err);
                break;
            default:
                c->z_bsl = g.bsl_z;
        }
        if (err) return fail(c, FI_ERR_TIME_RANGE, "TimeInterpolation0D out of range.");
    }
    // update_dz_ss! (src/sealevel.jl:173-201)
    if (g.sea_surface == FI_LATERALLY_VARIABLE_SEA_SURFACE) {
        if (g.lithosphere_column == FI_COMPRESSIBLE_LITHOSPHERE_COLUMN) {
            for (size_t p = 0; p < N; ++p) c->buf[p] = c->ca_load[p] + c->maskactive[p] * c->ca_mantle[p];
            for (size_t p = 0; p < N; ++p) c->buf[p] = c->Acell[p] * c->buf[p];
        } else {
            for (size_t p = 0; p < N; ++p) c->buf[p] = c->Acell[p] * c->ca_full[p];
        }
        samesize_conv_ext(c->dz_ss.data(), c->buf.data(), c->cp, c->gg_hat, c->i1, c->i2, c->j1, c->j2,
                          g.ssh_x_border);
    } else if (g.sea_surface == FI_IMPOSED_SEA_SURFACE) {
        // update_dz_ss!(sim, ::ImposedSeaSurface) (src/sealevel.jl:188-191) at sim.timer.t
        const int st = interpolate_snapshots(c, c->ss_t, c->ss_X, c->ss_flat, c->t_timer, c->dz_ss.data());
        if (st) return st;
    }
    // update_z_ss! (src/sealevel.jl:228-231)
    for (size_t p = 0; p < N; ++p) c->z_ss[p] = c->z_ss_ref[p] + c->dz_ss[p] + c->z_bsl;
    update_Haf(c);
    // update_maskgrounded!, update_maskocean! (src/topography.jl:1-4,25-28)
    for (size_t p = 0; p < N; ++p) c->mg[p] = c->H_af[p] > 0;
    for (size_t p = 0; p < N; ++p) c->mo[p] = (c->z_ss[p] - c->z_b[p] > 0) && !c->mg[p];
    // columnanom_water! (src/loads.jl:56-67)
    watercolumn(c, c->H_ice, c->mg, c->z_b, c->z_ss, c->H_water);
    if (g.sealevel_load == FI_INTERACTIVE_SEALEVEL_LOAD) {
        for (size_t p = 0; p < N; ++p)
            c->ca_sw[p] = g.rho_seawater * (c->z_ss[p] - c->z_ss_ref[p]) * (c->mo[p] ? 1.0 : 0.0) * c->maskactive[p];
    } else {
        std::fill(c->ca_sw.begin(), c->ca_sw.end(), 0.0);
    }
    columnanom_ice(c);
    c->count += 1;
    return FI_OK;
}

// update_deformation_rhs! (src/deformation.jl:110-131) then
// update_dudt!(MaxwellMantle, LaterallyVariableLithosphere) (src/deformation.jl:91-103)
void update_dudt_lvelva(fi_ctx* c, double* u, double* dudt) {
    const fi_config& g = c->cfg;
    int nx = c->nx, ny = c->ny;
    size_t N = c->N;
    double nu = g.litho_poissonratio;
    for (size_t p = 0; p < N; ++p) c->rhs[p] = -g.g * c->ca_full[p];
    second_derivatives(c->bxx.data(), c->byy.data(), c->bx.data(), c->bxy.data(), u, u, u, nx, ny, c->Dx.data(),
                       c->Dy.data());
#pragma omp parallel for schedule(static)
    for (size_t p = 0; p < N; ++p) {
        c->Mxx[p] = -c->D[p] * (nu * c->byy[p] + c->bxx[p]);
        c->Myy[p] = -c->D[p] * (nu * c->bxx[p] + c->byy[p]);
        c->Mxy[p] = -c->D[p] * (1 - nu) * c->bxy[p];
    }
    second_derivatives(c->bxx.data(), c->byy.data(), c->bx.data(), c->bxy.data(), c->Mxx.data(), c->Myy.data(),
                       c->Mxy.data(), nx, ny, c->Dx.data(), c->Dy.data());
#pragma omp parallel for schedule(static)
    for (size_t p = 0; p < N; ++p) c->rhs[p] += c->bxx[p] + (2 * c->bxy[p] + c->byy[p]);
    // samesize_conv!(rhs, ..., EmptyConvolution) is a no-op (src/convolutions.jl:120-122)
#pragma omp parallel for schedule(static)
    for (size_t p = 0; p < N; ++p) c->fftrhs[p] = cd(c->rhs[p] * c->K[p] / (2 * c->eta[p]), 0.0);
    fft2d(c->fftrhs.data(), nx, ny, c->vfx, c->vfy, -1);
#pragma omp parallel for schedule(static)
    for (size_t p = 0; p < N; ++p) c->fftrhs[p] *= c->spinv[p];
    fft2d(c->fftrhs.data(), nx, ny, c->vfx, c->vfy, +1);
    double s = 1.0 / ((double)nx * ny);
#pragma omp parallel for schedule(static)
    for (size_t p = 0; p < N; ++p) dudt[p] = (c->fftrhs[p].real() * s) * g.seconds_per_year;
    apply_corner_bc(dudt, nx, ny, g.viscous_x_border);
}

// update_dudt!(RelaxedMantle, lithosphere) (src/deformation.jl:29-45): the equilibrium displacement
// of the load, u_eq = samesize_conv(-(sigma_load + sigma_litho) g K^2, green_viscous) in the
// regular BC space (crop to the domain, then the corner BC, src/convolutions.jl:145-154), and
// dudt = (u_eq - u) / tau. The update_deformation_rhs! call at the top of that method only fills
// scratch buffers that this method overwrites, so it has no effect on the result.
void update_dudt_elra(fi_ctx* c, double* dudt) {
    const fi_config& g = c->cfg;
    size_t N = c->N;
    for (size_t p = 0; p < N; ++p) c->buf[p] = -(c->ca_load[p] + c->ca_litho[p]) * g.g * (c->K[p] * c->K[p]);
    vec crop;
    c->cp.conv_cropped(c->buf.data(), c->gv_hat, crop);
    const int nx = c->nx, ex = 2 * nx - 1;
    for (int j = c->j1; j <= c->j2; ++j)
        for (int i = c->i1; i <= c->i2; ++i)
            c->u_eq[(i - c->i1) + (size_t)(j - c->j1) * nx] = crop[(i - 1) + (size_t)(j - 1) * ex];
    apply_corner_bc(c->u_eq.data(), nx, c->ny, g.viscous_x_border);
    for (size_t p = 0; p < N; ++p) dudt[p] = 1.0 / c->tau[p] * (c->u_eq[p] - c->u[p]);
}

// update_dudt!(MaxwellMantle, lithosphere) for a rigid or laterally constant lithosphere
// (src/deformation.jl:52-89): one implicit (trapezoidal) step of dt_min on the n x n spectrum,
// u <- Re IFFT[((nabla - dt/2 beta) FFT(u) + dt FFT(F)) / (nabla + dt/2 beta)] with
// nabla = 2 eta kappa R = 2 eta / scaled_pseudodiff_inv, beta = rho_um g + D kappa^4 and
// F = -(sigma_load + sigma_litho) g K^2, all taken elementwise on the Fourier layout; then the
// corner BC. u is overwritten (and becomes the state); dudt is not written.
void update_u_implicit(fi_ctx* c, double* u) {
    const fi_config& g = c->cfg;
    const int nx = c->nx, ny = c->ny;
    size_t N = c->N;
    const double dt = g.dt_min * g.seconds_per_year;
    std::vector<cd> F(N), U(N);
    for (size_t p = 0; p < N; ++p) F[p] = cd(-(c->ca_load[p] + c->ca_litho[p]) * g.g * (c->K[p] * c->K[p]), 0.0);
    for (size_t p = 0; p < N; ++p) U[p] = cd(u[p], 0.0);
    fft2d(F.data(), nx, ny, c->vfx, c->vfy, -1);
    fft2d(U.data(), nx, ny, c->vfx, c->vfy, -1);
    for (size_t p = 0; p < N; ++p) {
        const double kap = c->pd[p];
        const double nabla = 2.0 * c->eta[p] / c->spinv[p];
        const double beta = g.rho_uppermantle * g.g + c->D[p] * (kap * kap) * (kap * kap);
        U[p] = ((nabla - (dt / 2) * beta) * U[p] + dt * F[p]) / (nabla + (dt / 2) * beta);
    }
    fft2d(U.data(), nx, ny, c->vfx, c->vfy, +1);
    const double s = 1.0 / ((double)nx * ny);
    for (size_t p = 0; p < N; ++p) u[p] = U[p].real() * s;
    apply_corner_bc(u, nx, ny, g.viscous_x_border);
    std::memcpy(c->u.data(), u, N * sizeof(double));
}

// update_diagnostics!(dudt, u, sim, t) (src/simulation.jl:335-393)
int update_diagnostics(fi_ctx* c, double* dudt, double* u, double t) {
    const fi_config& g = c->cfg;
    size_t N = c->N;
    c->t_timer = t;
    c->nrhs += 1;
    apply_corner_bc(u, c->nx, c->ny, g.viscous_x_border);
    // update_bedrock! (src/topography.jl:30-34)
    if (u != c->u.data()) std::memcpy(c->u.data(), u, N * sizeof(double));
    for (size_t p = 0; p < N; ++p) c->z_b[p] = c->z_b_ref[p] + c->ue[p] + c->u[p];
    // columnanom_mantle! (src/loads.jl:36-39)
    for (size_t p = 0; p < N; ++p) c->ca_mantle[p] = g.rho_uppermantle * (c->u[p] - c->u_ref[p]);
    // apply_bc!(H_ice, t, ice_thickness) (src/boundary_conditions.jl:55-62)
    if (c->ice_mode != 0) {
        int st = interpolate_ice(c, t, c->H_ice.data());
        if (st) return st;
    }
    update_Haf(c);
    columnanom_ice(c);
    if (((t - c->t0) / g.dt_sparse) >= (double)c->count) {
        int st = sparse_block(c);
        if (st) return st;
    }
    // columnanom_load!, columnanom_full! (src/loads.jl:96-106)
    for (size_t p = 0; p < N; ++p) c->ca_load[p] = c->maskactive[p] * (c->ca_ice[p] + c->ca_sw[p] + c->ca_sed[p]);
    for (size_t p = 0; p < N; ++p)
        c->ca_full[p] = c->ca_load[p] + c->maskactive[p] * (c->ca_litho[p] + c->ca_mantle[p]);
    // update_dudt! (src/deformation.jl:20-27,91-103)
    if (g.mantle == FI_RIGID_MANTLE) {
        std::fill(dudt, dudt + N, 0.0);
    } else if (g.mantle == FI_RELAXED_MANTLE) {
        update_dudt_elra(c, dudt);
    } else if (implicit_step(g)) {
        update_u_implicit(c, u);
    } else {
        update_dudt_lvelva(c, u, dudt);
    }
    if (dudt != c->dudt.data()) std::memcpy(c->dudt.data(), dudt, N * sizeof(double));
    // t_computation! (src/simulation.jl:46-52): call nrhs - 1 since fi_init, kept below the capacity
    if (c->nrhs - 1 < c->hist_cap) {
        const size_t k = (size_t)(c->nrhs - 1);
        if (c->hist_t.size() <= k) {
            c->hist_t.resize(k + 1);
            c->hist_wall.resize(k + 1);
        }
        c->hist_t[k] = t;
        c->hist_wall[k] = std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
    }
    return FI_OK;
}

// ---------------------------------------------------------------------------------------
// BS3 (Bogacki-Shampine 3(2), FSAL) with a PI step-size controller, restated from the
// published method (Bogacki & Shampine 1989; Hairer, Norsett & Wanner II.4 initial step;
// PI control after Gustafsson / Soderlind) as OrdinaryDiffEq's BS3 uses it by default
// (src/simulation.jl:11-15,255-257). Controller constants: beta1 = 7/30, beta2 = 2/15,
// gamma = 0.9, qmin = 0.2, qmax = 10, qoldinit = 1e-4 -- parity of the step sequence is
// unpinned (no reference source holds them).
// ---------------------------------------------------------------------------------------
const double A21 = 0.5, A32 = 0.75, A41 = 2.0 / 9.0, A42 = 1.0 / 3.0, A43 = 4.0 / 9.0;
const double C2 = 0.5, C3 = 0.75;
const double BT1 = 2.0 / 9.0 - 7.0 / 24.0, BT2 = 1.0 / 3.0 - 0.25, BT3 = 4.0 / 9.0 - 1.0 / 3.0, BT4 = -0.125;
const double BETA1 = 7.0 / 30.0, BETA2 = 2.0 / 15.0, GAMMA = 0.9, QMIN = 0.2, QMAX = 10.0, QOLDINIT = 1e-4;

double rms(const vec& a) {
    double s = 0.0;
    for (double x : a) s += x * x;
    return std::sqrt(s / (double)a.size());
}

}  // namespace

// =======================================================================================
// C ABI
// =======================================================================================
extern "C" {

int32_t fi_abi_version(void) { return FI_ABI_VERSION; }
const char* fi_engine_name(void) { return "cpu-oracle-fp64"; }

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
    c->bsl_z = 0.0;
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
    fi_ctx* c = new fi_ctx();
    c->cfg = *cfg;
    c->nx = cfg->nx;
    c->ny = cfg->ny;
    c->N = (size_t)c->nx * c->ny;
    c->mx = c->nx / 2;
    c->my = c->ny / 2;
    c->dx = 2 * cfg->Wx / c->nx;
    c->dy = 2 * cfg->Wy / c->ny;
    c->A = c->dx * c->dy;
    c->rho_sw_ice = cfg->rho_seawater / cfg->rho_ice;
    samesize_conv_indices(c->nx, c->mx, c->i1, c->i2);
    samesize_conv_indices(c->ny, c->my, c->j1, c->j2);
    c->pd = pseudodiff(c->nx, c->ny, cfg->Wx, cfg->Wy);
    c->vfx = FFT(c->nx);
    c->vfy = FFT(c->ny);
    c->cp.init(c->nx, c->ny);
    size_t N = c->N;
    set_distortion(c, vec(N, 1.0));
    c->maskactive.assign(N, 1.0);
    c->u_ref.assign(N, 0.0);
    c->ue_ref.assign(N, 0.0);
    c->dz_ss_ref.assign(N, 0.0);
    c->z_b_ref.assign(N, 1e6);
    for (vec* v : {&c->rhs, &c->bxx, &c->byy, &c->bx, &c->bxy, &c->Mxx, &c->Myy, &c->Mxy, &c->buf}) v->assign(N, 0.0);
    c->fftrhs.assign(N, cd(0, 0));
    *out = c;
    return FI_OK;
}

void fi_destroy(fi_ctx* c) { delete c; }
const char* fi_last_error(const fi_ctx* c) { return c ? c->err.c_str() : "null context"; }

int fi_set_field(fi_ctx* c, int32_t f, const double* h, size_t n) {
    if (!c || !h || n != c->N) return fail(c, FI_ERR_INVALID, "fi_set_field: size mismatch");
    vec v(h, h + n);
    switch (f) {
        case FI_LITHO_RIGIDITY: c->D = v; c->has_D = true; break;
        case FI_EFFECTIVE_VISCOSITY: c->eta = v; c->has_eta = true; break;
        case FI_SCALED_PSEUDODIFF_INV: c->spinv = v; c->has_spinv = true; break;
        case FI_MASK_ACTIVE: c->maskactive = v; break;
        case FI_GREEN_ELASTIC: c->green_e = v; c->has_ge = true; break;
        case FI_GREEN_GEOID: c->green_g = v; c->has_gg = true; break;
        case FI_GREEN_VISCOUS: c->green_v = v; c->has_gv = true; break;
        case FI_LITHO_THICKNESS: c->litho_h = v; break;
        case FI_TAU:
            for (double x : v)
                if (!(x > 0) || !std::isfinite(x)) return fail(c, FI_ERR_INVALID, "tau must be positive");
            c->tau = v;
            c->has_tau = true;
            break;
        case FI_REF_U: c->u_ref = v; break;
        case FI_REF_UE: c->ue_ref = v; break;
        case FI_REF_DZ_SS: c->dz_ss_ref = v; break;
        case FI_REF_Z_B: c->z_b_ref = v; break;
        case FI_REF_H_ICE: c->H_ice_ref_in = v; c->has_href = true; break;
        case FI_DISTORTION:
            for (double k : v)
                if (!(k > 0) || !std::isfinite(k)) return fail(c, FI_ERR_INVALID, "distortion K must be positive");
            set_distortion(c, v);
            break;
        case FI_U:
            if (!c->initialized) return fail(c, FI_ERR_STATE, "fi_set_field(FI_U) before fi_init");
            c->u = v;
            if (c->integ) {
                c->uprev = v;
                c->k1_stale = true;   // the next fi_integrate re-evaluates the FSAL derivative
            }
            break;
        case FI_H_ICE:
            if (!c->initialized) return fail(c, FI_ERR_STATE, "fi_set_field(FI_H_ICE) before fi_init");
            c->H_ice = v;
            break;
        default: return fail(c, FI_ERR_INVALID, "fi_set_field: field not writable");
    }
    return FI_OK;
}

int fi_get_field(fi_ctx* c, int32_t f, double* h, size_t n) {
    if (!c || !h || n != c->N) return fail(c, FI_ERR_INVALID, "fi_get_field: size mismatch");
    if (!c->initialized && f >= FI_U) return fail(c, FI_ERR_STATE, "fi_get_field before fi_init");
    if (f == FI_U_X || f == FI_U_Y) {
        // thinplate_horizontal_displacement!(u_x, u_y, u + ue, litho_thickness, domain)
        // (src/deformation.jl:141-164) as the output callbacks call it (src/simulation.jl:218-221)
        if (c->litho_h.size() != c->N) return fail(c, FI_ERR_STATE, "fi_get_field: litho thickness not set");
        const vec& u = c->integ ? c->uprev : c->u;
        vec w(c->N);
        for (size_t p = 0; p < c->N; ++p) w[p] = u[p] + c->ue[p];
        if (f == FI_U_X) dx1(h, w.data(), c->nx, c->ny, c->Dx.data());
        else dy1(h, w.data(), c->nx, c->ny, c->Dy.data());
        for (size_t p = 0; p < c->N; ++p) h[p] *= -c->litho_h[p] / 2;
        return FI_OK;
    }
    const vec* v = nullptr;
    const std::vector<uint8_t>* m = nullptr;
    switch (f) {
        case FI_LITHO_RIGIDITY: v = &c->D; break;
        case FI_EFFECTIVE_VISCOSITY: v = &c->eta; break;
        case FI_SCALED_PSEUDODIFF_INV: v = &c->spinv; break;
        case FI_MASK_ACTIVE: v = &c->maskactive; break;
        case FI_GREEN_ELASTIC: v = &c->green_e; break;
        case FI_GREEN_GEOID: v = &c->green_g; break;
        case FI_GREEN_VISCOUS: v = &c->green_v; break;
        case FI_TAU: v = &c->tau; break;
        case FI_LITHO_THICKNESS: v = &c->litho_h; break;
        case FI_REF_U: v = &c->u_ref; break;
        case FI_REF_UE: v = &c->ue_ref; break;
        case FI_REF_DZ_SS: v = &c->dz_ss_ref; break;
        case FI_REF_Z_B: v = &c->z_b_ref; break;
        case FI_REF_H_ICE: v = &c->H_ice_ref; break;
        case FI_DISTORTION: v = &c->K; break;
        case FI_U: v = c->integ ? &c->uprev : &c->u; break;
        case FI_DUDT: v = &c->dudt; break;
        case FI_UE: v = &c->ue; break;
        case FI_DZ_SS: v = &c->dz_ss; break;
        case FI_Z_SS: v = &c->z_ss; break;
        case FI_Z_B: v = &c->z_b; break;
        case FI_H_ICE: v = &c->H_ice; break;
        case FI_H_AF: v = &c->H_af; break;
        case FI_H_WATER: v = &c->H_water; break;
        case FI_MASK_GROUNDED: m = &c->mg; break;
        case FI_MASK_OCEAN: m = &c->mo; break;
        case FI_CANOM_ICE: v = &c->ca_ice; break;
        case FI_CANOM_SEAWATER: v = &c->ca_sw; break;
        case FI_CANOM_LITHO: v = &c->ca_litho; break;
        case FI_CANOM_MANTLE: v = &c->ca_mantle; break;
        case FI_CANOM_LOAD: v = &c->ca_load; break;
        case FI_CANOM_FULL: v = &c->ca_full; break;
        case FI_U_EQ: v = &c->u_eq; break;
        default: return fail(c, FI_ERR_INVALID, "fi_get_field: unknown field");
    }
    if (v) {
        if (v->size() != n) return fail(c, FI_ERR_STATE, "fi_get_field: field not set");
        std::memcpy(h, v->data(), n * sizeof(double));
    } else {
        for (size_t p = 0; p < n; ++p) h[p] = (*m)[p] ? 1.0 : 0.0;
    }
    return FI_OK;
}

int fi_set_ice_snapshots(fi_ctx* c, const double* t, int32_t nt, const double* H, int32_t flat) {
    if (!c || !t || !H || nt < 1) return fail(c, FI_ERR_INVALID, "fi_set_ice_snapshots: bad args");
    for (int k = 1; k < nt; ++k)
        if (!(t[k] > t[k - 1])) return fail(c, FI_ERR_INVALID, "ice times must increase");
    c->ice_mode = 1;
    c->ice_t.assign(t, t + nt);
    c->ice_H.assign(H, H + (size_t)nt * c->N);
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
    // t_computation! keeps the calls past t_span[0]; run! subtracts the first one's time
    const size_t m = std::min(c->hist_t.size(), (size_t)std::max<long long>(c->nrhs, 0));
    int64_t j = 0;
    double w0 = 0.0;
    for (size_t k = 0; k < m; ++k) {
        if (!(c->hist_t[k] > c->t0)) continue;
        if (j == 0) w0 = c->hist_wall[k];
        if (j < capacity) {
            t[j] = c->hist_t[k];
            seconds[j] = c->hist_wall[k] - w0;
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
    c->ice_mode = 2;
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
    c->bsl_x = table ? vec(x, x + n) : vec();
    c->bsl_y = table ? vec(y, y + n) : vec();
    return FI_OK;
}

int fi_set_output_times(fi_ctx* c, const double* t, int32_t n) {
    if (!c || (n > 0 && !t) || n < 0) return fail(c, FI_ERR_INVALID, "fi_set_output_times: bad args");
    c->tout.assign(t, t + n);
    std::sort(c->tout.begin(), c->tout.end());
    c->kout = 0;
    return FI_OK;
}

// Simulation(...) (src/simulation.jl:105-156)
int fi_init(fi_ctx* c, double t0, double t_end) {
    if (!c) return FI_ERR_INVALID;
    const fi_config& g = c->cfg;
    size_t N = c->N;
    if (g.mantle == FI_MAXWELL_MANTLE && (!c->has_D || !c->has_eta || !c->has_spinv))
        return fail(c, FI_ERR_INVALID, "litho rigidity, effective viscosity and scaled_pseudodiff_inv required");
    if (g.mantle == FI_RELAXED_MANTLE && (!c->has_gv || !c->has_tau))
        return fail(c, FI_ERR_INVALID, "viscous Green kernel and tau required");
    if (implicit_step(g) && !(g.dt_min > 0))
        return fail(c, FI_ERR_INVALID, "the implicit step needs dt_min (DiffEqOptions.dt_min with Euler)");
    if (g.lithosphere != FI_RIGID_LITHOSPHERE && !c->has_ge)
        return fail(c, FI_ERR_INVALID, "elastic Green kernel required");
    if (g.sea_surface == FI_LATERALLY_VARIABLE_SEA_SURFACE && !c->has_gg)
        return fail(c, FI_ERR_INVALID, "sea-surface Green kernel required");
    if (g.sea_surface == FI_IMPOSED_SEA_SURFACE && c->ss_t.empty())
        return fail(c, FI_ERR_INVALID, "ImposedSeaSurface needs dz_ss snapshots (fi_set_sea_surface_snapshots)");
    c->hist_t.clear();
    c->hist_wall.clear();
    if (g.sealevel_load == FI_INTERACTIVE_SEALEVEL_LOAD) {
        double s = 0;
        for (double m : c->maskactive) s += m;
        if (s > 0.6 * c->nx * c->ny)
            return fail(c, FI_ERR_INVALID, "Mask defining regions of active load must not cover more than 60% "
                                           "of the cells when using an interactive sea level.");
    }
    if (c->has_ge) c->ge_hat = c->cp.spectrum(c->green_e.data());
    if (c->has_gg) c->gg_hat = c->cp.spectrum(c->green_g.data());
    if (g.mantle == FI_RELAXED_MANTLE) c->gv_hat = c->cp.spectrum(c->green_v.data());
    c->t0 = t0;
    c->t_end = t_end;
    c->t_timer = t0;
    // H_ice_ref = apply_bc!(zeros, t0, ice_thickness)
    c->H_ice_ref.assign(N, 0.0);
    if (c->has_href) c->H_ice_ref = c->H_ice_ref_in;
    else if (c->ice_mode != 0) {
        int st = interpolate_ice(c, t0, c->H_ice_ref.data());
        if (st) return st;
    }
    c->z_ss_ref.resize(N);
    for (size_t p = 0; p < N; ++p) c->z_ss_ref[p] = g.bsl_z + c->dz_ss_ref[p];
    c->mg_ref.resize(N);
    c->mo_ref.resize(N);
    c->H_af_ref.resize(N);
    for (size_t p = 0; p < N; ++p) {
        double h = hab(c->H_ice_ref[p], c->z_b_ref[p], c->z_ss_ref[p], g.rho_seawater / g.rho_ice);
        c->mg_ref[p] = h > 0;
        c->mo_ref[p] = (c->z_ss_ref[p] - c->z_b_ref[p] > 0) && !c->mg_ref[p];
        c->H_af_ref[p] = h;
    }
    c->H_water_ref.resize(N);
    watercolumn(c, c->H_ice_ref, c->mg_ref, c->z_b_ref, c->z_ss_ref, c->H_water_ref);
    // CurrentState(domain, ref, z_bsl) (src/state.jl:80-103)
    c->u = c->u_ref;
    c->ue = c->ue_ref;
    c->u_eq = c->u_ref;
    c->dudt.assign(N, 0.0);
    c->H_ice = c->H_ice_ref;
    c->H_af = c->H_af_ref;
    c->H_water = c->H_water_ref;
    for (vec* v : {&c->ca_ice, &c->ca_sw, &c->ca_sed, &c->ca_litho, &c->ca_mantle, &c->ca_load, &c->ca_full})
        v->assign(N, 0.0);
    c->z_b = c->z_b_ref;
    c->dz_ss.assign(N, 0.0);
    c->z_ss = c->z_ss_ref;
    c->V_af = c->V_pov = c->V_den = c->delta_V = 0.0;
    c->z_bsl = g.bsl_z;
    c->A_bsl = c->bsl_A_ref;
    c->mg = c->mg_ref;
    c->mo = c->mo_ref;
    c->count = 0;
    c->initialized = true;
    c->integ = false;
    c->nrhs = c->nacc = c->nrej = 0;
    c->kout = 0;
    return FI_OK;
}

int fi_rhs(fi_ctx* c, double t, double* u, double* dudt) {
    if (!c || !u || !dudt) return FI_ERR_INVALID;
    if (!c->initialized) return fail(c, FI_ERR_STATE, "fi_rhs before fi_init");
    return update_diagnostics(c, dudt, u, t);
}

int fi_rhs_device(fi_ctx* c, double t, void* u, void* dudt) {
    return fi_rhs(c, t, (double*)u, (double*)dudt);
}

int fi_sparse_update(fi_ctx* c, double t) {
    if (!c) return FI_ERR_INVALID;
    if (!c->initialized) return fail(c, FI_ERR_STATE, "fi_sparse_update before fi_init");
    c->t_timer = t;
    return sparse_block(c);
}

// init_problem! (src/simulation.jl:305-311), FSAL initialisation and the initial step
int fi_integrator_init(fi_ctx* c) {
    if (!c) return FI_ERR_INVALID;
    if (!c->initialized) return fail(c, FI_ERR_STATE, "fi_integrator_init before fi_init");
    const fi_config& g = c->cfg;
    size_t N = c->N;
    update_volumes(c);
    int st = update_diagnostics(c, c->dudt.data(), c->u.data(), c->t0);
    if (st) return st;
    c->t = c->t0;
    c->uprev = c->u;
    for (vec* v : {&c->k1, &c->k2, &c->k3, &c->k4, &c->tmp}) v->assign(N, 0.0);
    st = update_diagnostics(c, c->k1.data(), c->uprev.data(), c->t);  // FSAL: fsalfirst = f(u0, t0)
    if (st) return st;
    if (implicit_step(g)) {
        // Euler with dt = dt_min (src/simulation.jl:244-250): both calls above advanced u (the
        // implicit step writes u and leaves the derivative at zero)
        c->dt = g.dt_min;
        c->integ = true;
        c->k1_stale = false;
        c->kout = 0;
        return FI_OK;
    }
    // initial step (Hairer, Norsett & Wanner, Solving ODEs I, II.4)
    vec sk(N), tq(N);
    for (size_t p = 0; p < N; ++p) sk[p] = g.abstol + std::fabs(c->uprev[p]) * g.reltol;
    for (size_t p = 0; p < N; ++p) tq[p] = c->uprev[p] / sk[p];
    double d0 = rms(tq);
    for (size_t p = 0; p < N; ++p) tq[p] = c->k1[p] / sk[p];
    double d1 = rms(tq);
    double dtmax = c->t_end - c->t0;
    double dt0 = (d0 < 1e-5 || d1 < 1e-5) ? 1e-6 : 0.01 * (d0 / d1);
    dt0 = std::min(dt0, dtmax);
    for (size_t p = 0; p < N; ++p) c->tmp[p] = c->uprev[p] + dt0 * c->k1[p];
    st = update_diagnostics(c, c->k2.data(), c->tmp.data(), c->t + dt0);
    if (st) return st;
    for (size_t p = 0; p < N; ++p) tq[p] = (c->k2[p] - c->k1[p]) / sk[p];
    double d2 = rms(tq) / dt0;
    double mx = std::max(d1, d2);
    double dt1 = (mx <= 1e-15) ? std::max(1e-6, dt0 * 1e-3) : std::pow(10.0, -(2.0 + std::log10(mx)) / 3.0);
    c->dt = std::min(std::min(100 * dt0, dt1), dtmax);
    c->qold = QOLDINIT;
    c->integ = true;
    c->k1_stale = false;
    c->kout = 0;
    return FI_OK;
}

int fi_integrate(fi_ctx* c, double t_target, fi_output_cb cb, void* user) {
    if (!c) return FI_ERR_INVALID;
    if (!c->integ) return fail(c, FI_ERR_STATE, "fi_integrate before fi_integrator_init");
    const fi_config& g = c->cfg;
    size_t N = c->N;
    if (t_target < c->t) return fail(c, FI_ERR_INVALID, "fi_integrate: target before current time");
    if (c->k1_stale) {
        int st = update_diagnostics(c, c->k1.data(), c->uprev.data(), c->t);
        if (st) return st;
        c->k1_stale = false;
    }
    while (c->t < t_target) {
        // next stop: the first output time after t, or t_target (modify_dt_for_tstops!)
        double stop = t_target;
        for (double to : c->tout)
            if (to > c->t) { stop = std::min(stop, to); break; }
        if (implicit_step(g)) {
            // Euler step: u = uprev + dt * 0, then f(u, t + dt) replaces u by one implicit step of
            // dt_min; an output callback is followed by the FSAL re-evaluation, one more step
            const bool clip = c->t + g.dt_min >= stop;
            c->t = clip ? stop : c->t + g.dt_min;
            int st = update_diagnostics(c, c->k1.data(), c->uprev.data(), c->t);
            if (st) return st;
            c->nacc += 1;
            if (c->kout < (int64_t)c->tout.size() && c->t >= c->tout[c->kout]) {
                c->u = c->uprev;
                if (cb) cb(user, c->t, (int32_t)c->kout);
                c->kout += 1;
                if ((st = update_diagnostics(c, c->k1.data(), c->uprev.data(), c->t))) return st;
            }
            continue;
        }
        double dt = c->dt;
        bool clipped = false;
        if (c->t + dt >= stop) { dt = stop - c->t; clipped = true; }
        if (!(dt > 1e-14 * std::max(1.0, std::fabs(c->t)))) return fail(c, FI_ERR_DT_MIN, "dt below minimum");
        int st;
        double t = c->t;
        for (size_t p = 0; p < N; ++p) c->tmp[p] = c->uprev[p] + dt * A21 * c->k1[p];
        if ((st = update_diagnostics(c, c->k2.data(), c->tmp.data(), t + C2 * dt))) return st;
        for (size_t p = 0; p < N; ++p) c->tmp[p] = c->uprev[p] + dt * A32 * c->k2[p];
        if ((st = update_diagnostics(c, c->k3.data(), c->tmp.data(), t + C3 * dt))) return st;
        vec unew(N);
        for (size_t p = 0; p < N; ++p) unew[p] = c->uprev[p] + dt * (A41 * c->k1[p] + A42 * c->k2[p] + A43 * c->k3[p]);
        if ((st = update_diagnostics(c, c->k4.data(), unew.data(), t + dt))) return st;
        double s = 0.0;
        for (size_t p = 0; p < N; ++p) {
            double ut = dt * (BT1 * c->k1[p] + BT2 * c->k2[p] + BT3 * c->k3[p] + BT4 * c->k4[p]);
            double sc = g.abstol + std::max(std::fabs(c->uprev[p]), std::fabs(unew[p])) * g.reltol;
            double e = ut / sc;
            s += e * e;
        }
        double EEst = std::sqrt(s / (double)N);
        c->EEst = EEst;
        // OrdinaryDiffEq: a NaN error norm makes the PI controller propose a NaN step, and the
        // integration stops with retcode DtNaN; an infinite one is rejected with factor 1/qmin
        if (std::isnan(EEst)) {
            c->nrej += 1;
            return fail(c, FI_ERR_NONFINITE, "NaN dt detected: NaN in the state or its derivative");
        }
        double q;
        if (!std::isfinite(EEst)) {
            c->q11 = 1.0 / QMIN * GAMMA * 10;
            q = 1.0 / QMIN;
        } else if (EEst == 0.0) {
            q = 1.0 / QMAX;
        } else {
            double q11 = std::pow(EEst, BETA1);
            c->q11 = q11;
            q = q11 / std::pow(c->qold, BETA2);
            q = std::max(1.0 / QMAX, std::min(1.0 / QMIN, q / GAMMA));
        }
        if (std::isfinite(EEst) && EEst <= 1.0) {
            // accept (step_accept_controller!)
            c->qold = std::max(EEst, QOLDINIT);
            c->t = clipped ? stop : t + dt;
            c->uprev.swap(unew);
            c->k1.swap(c->k4);
            c->dt = dt / q;
            c->nacc += 1;
            // output callback (nout_condition / nout_affect!) then the FSAL re-evaluation that
            // a DiscreteCallback's u_modified! triggers
            if (c->kout < (int64_t)c->tout.size() && c->t >= c->tout[c->kout]) {
                c->u = c->uprev;
                if (cb) cb(user, c->t, (int32_t)c->kout);
                c->kout += 1;
                if ((st = update_diagnostics(c, c->k1.data(), c->uprev.data(), c->t))) return st;
            }
        } else {
            // reject (step_reject_controller!)
            c->dt = dt / std::min(1.0 / QMIN, c->q11 / GAMMA);
            c->nrej += 1;
        }
    }
    c->u = c->uprev;
    return FI_OK;
}

int fi_set_time(fi_ctx* c, double t) {
    if (!c) return FI_ERR_INVALID;
    if (!c->integ) return fail(c, FI_ERR_STATE, "fi_set_time before fi_integrator_init");
    // its load-storing call would advance u under the implicit step
    if (implicit_step(c->cfg)) return fail(c, FI_ERR_UNSUPPORTED, "fi_set_time: not with the implicit step");
    if (!(t >= c->t0)) return fail(c, FI_ERR_INVALID, "fi_set_time: t before the start time");
    const int64_t due = (int64_t)std::floor((t - c->t0) / c->cfg.dt_sparse);
    c->count = due + 1;
    vec u = c->uprev, d(c->N);
    int st = update_diagnostics(c, d.data(), u.data(), t);
    if (st) return st;
    if ((st = sparse_block(c))) return st;
    c->count = due + 1;
    c->t = t;
    c->u = c->uprev;
    c->k1_stale = true;
    return FI_OK;
}

}  // extern "C"

namespace {

// RKC2 (Sommeijer, Shampine & Verwer 1997, Appl. Numer. Math. 22, 315-326): damped Chebyshev
// stability polynomial R_s(z) = a_s + b_s T_s(w0 + w1 z), w0 = 1 + eps/s^2, w1 = T_s'(w0)/T_s''(w0),
// b_j = T_j''(w0)/T_j'(w0)^2 (b_0 = b_1 = b_2), a_j = 1 - b_j T_j(w0), and the three-term stage
// recursion
//   Y_1 = Y_0 + mut_1 h F_0,   mut_1 = b_1 w1,
//   Y_j = (1 - mu_j - nu_j) Y_0 + mu_j Y_{j-1} + nu_j Y_{j-2} + mut_j h F_{j-1} + gt_j h F_0,
//   mu_j = 2 b_j w0 / b_{j-1}, nu_j = -b_j / b_{j-2}, mut_j = 2 b_j w1 / b_{j-1},
//   gt_j = -(1 - b_{j-1} T_{j-1}(w0)) mut_j,
// at stage times c_1 = mut_1, c_j = mu_j c_{j-1} + nu_j c_{j-2} + mut_j + gt_j.
struct RkcCoef {
    std::vector<double> mu, nu, mut, gt, c;
    double beta = 0;
};

static RkcCoef rkc_coef(int s) {
    const double eps = 2.0 / 13.0, w0 = 1.0 + eps / ((double)s * s);
    std::vector<double> T(s + 1), dT(s + 1), ddT(s + 1);
    T[0] = 1.0; T[1] = w0;
    dT[0] = 0.0; dT[1] = 1.0;
    ddT[0] = 0.0; ddT[1] = 0.0;
    for (int j = 2; j <= s; ++j) {
        T[j] = 2.0 * w0 * T[j - 1] - T[j - 2];
        dT[j] = 2.0 * T[j - 1] + 2.0 * w0 * dT[j - 1] - dT[j - 2];
        ddT[j] = 4.0 * dT[j - 1] + 2.0 * w0 * ddT[j - 1] - ddT[j - 2];
    }
    const double w1 = dT[s] / ddT[s];
    std::vector<double> b(s + 1);
    for (int j = 2; j <= s; ++j) b[j] = ddT[j] / (dT[j] * dT[j]);
    b[0] = b[1] = b[2];
    RkcCoef k;
    k.mu.assign(s + 1, 0.0); k.nu.assign(s + 1, 0.0); k.mut.assign(s + 1, 0.0); k.gt.assign(s + 1, 0.0);
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
    k.beta = (w0 + 1.0) * ddT[s] / dT[s];
    return k;
}

// fewest stages (>= 2) whose stability interval covers x = 1.2 rho h
static int rkc_stages(double x) {
    int s = 2;
    while (rkc_coef(s).beta < x && s < 100000) s = std::max(s + 1, (int)(s * 1.05));
    while (s > 2 && rkc_coef(s - 1).beta >= x) --s;
    return s;
}

// the pseudo-random start vector of the power iteration: uniform in [-1, 1) per cell (splitmix64)
static double rkc_start_value(uint64_t p) {
    uint64_t z = 0x9E3779B97F4A7C15ull * (p + 1) + 0x5EEDull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    return (double)(z >> 11) * (2.0 / 9007199254740992.0) - 1.0;
}

}  // namespace

extern "C" {

int fi_integrate_rkc(fi_ctx* c, double t_target, double dt_max, double rho, fi_rkc_info* info) {
    if (!c || !(dt_max > 0)) return FI_ERR_INVALID;
    if (!c->integ) return fail(c, FI_ERR_STATE, "fi_integrate_rkc before fi_integrator_init");
    if (implicit_step(c->cfg)) return fail(c, FI_ERR_UNSUPPORTED, "fi_integrate_rkc: not with the implicit step");
    if (!(t_target >= c->t)) return fail(c, FI_ERR_INVALID, "fi_integrate_rkc: target before current time");
    const size_t N = c->N;
    const int64_t r0 = c->nrhs;
    int st;
    vec F0(N), F1(N), w(N);
    auto rms = [&](const vec& a) {
        double q = 0;
        for (double x : a) q += x * x;
        return std::sqrt(q / (double)N);
    };
    if (!(rho > 0)) {
        if (!(c->rkc_rho > 0)) {
            // 40 power iterations on J v = f(u + v) - f(u) (the right-hand side is affine in u)
            vec v(N);
            for (size_t p = 0; p < N; ++p) v[p] = rkc_start_value(p);
            const double n0 = rms(v);
            for (double& x : v) x /= n0;
            w = c->uprev;
            if ((st = update_diagnostics(c, F0.data(), w.data(), c->t))) return st;
            double lam = 0;
            for (int it = 0; it < 40; ++it) {
                for (size_t p = 0; p < N; ++p) w[p] = c->uprev[p] + v[p];
                if ((st = update_diagnostics(c, F1.data(), w.data(), c->t))) return st;
                for (size_t p = 0; p < N; ++p) v[p] = F1[p] - F0[p];
                lam = rms(v);
                if (!(lam > 0) || !std::isfinite(lam)) return fail(c, FI_ERR_NONFINITE, "fi_integrate_rkc: power iteration");
                for (double& x : v) x /= lam;
            }
            c->rkc_rho = lam;
        }
        rho = c->rkc_rho;
    }
    const double t_start = c->t, span = t_target - t_start;
    int nsub = span > 0 ? (int)std::ceil(span / dt_max - 1e-12) : 0;
    const double h = nsub > 0 ? span / nsub : 0.0;
    const int s = nsub > 0 ? rkc_stages(1.2 * rho * h) : 0;
    if (nsub > 0) {
        const RkcCoef k = rkc_coef(s);
        vec Y0 = c->uprev, Ym1(N), Ym2(N), Y(N);
        for (int sub = 0; sub < nsub; ++sub) {
            const double t0 = t_start + sub * h;
            w = Y0;
            if ((st = update_diagnostics(c, F0.data(), w.data(), t0))) return st;
            for (size_t p = 0; p < N; ++p) Ym1[p] = Y0[p] + k.mut[1] * h * F0[p];
            Ym2 = Y0;
            for (int j = 2; j <= s; ++j) {
                w = Ym1;
                if ((st = update_diagnostics(c, F1.data(), w.data(), t0 + k.c[j - 1] * h))) return st;
                const double a0 = 1.0 - k.mu[j] - k.nu[j];
                for (size_t p = 0; p < N; ++p)
                    Y[p] = a0 * Y0[p] + k.mu[j] * Ym1[p] + k.nu[j] * Ym2[p] + k.mut[j] * h * F1[p] +
                           k.gt[j] * h * F0[p];
                Ym2.swap(Ym1);
                Ym1.swap(Y);
            }
            // corner BC on the new state (src/boundary_conditions.jl:144-148)
            const int nx = c->nx, ny = c->ny;
            const double off = 0.25 * (Ym1[0] + Ym1[nx - 1] + Ym1[(size_t)(ny - 1) * nx] + Ym1[(size_t)(ny - 1) * nx + nx - 1])
                               - c->cfg.viscous_x_border;
            for (size_t p = 0; p < N; ++p) Y0[p] = Ym1[p] - off;
        }
        c->uprev = Y0;
    }
    c->t = t_target;
    c->u = c->uprev;
    c->k1_stale = true;
    if (info) {
        info->rho = rho;
        info->h = h;
        info->substeps = nsub;
        info->stages = s;
        info->n_rhs = c->nrhs - r0;
        info->graph = 0;
        info->reserved = 0;
    }
    return FI_OK;
}

// Checkpoints: the dynamic state of the restatement (the fields of sim.now, src/state.jl:49-104,
// the loads of the last call and the integrator), walked in a fixed order to measure, write or
// read a host blob (mode 0/1/2).
static int ck_walk(fi_ctx* c, int mode, unsigned char* blob, size_t& off) {
    auto raw = [&](void* p, size_t n) {
        if (mode == 1) std::memcpy(blob + off, p, n);
        if (mode == 2) std::memcpy(p, blob + off, n);
        off += n;
    };
    auto sc = [&](auto& v) { raw(&v, sizeof(v)); };
    auto vv = [&](vec& v) {
        if (mode == 2) v.resize(c->N);
        raw(v.data(), c->N * sizeof(double));
    };
    auto vb = [&](std::vector<uint8_t>& v) {
        if (mode == 2) v.resize(c->N);
        raw(v.data(), c->N);
    };
    int32_t hdr[4] = {0x46494b52, c->nx, c->ny, c->ice_mode}, hr[4] = {hdr[0], hdr[1], hdr[2], hdr[3]};
    raw(hr, sizeof(hr));
    if (mode == 2 && std::memcmp(hr, hdr, sizeof(hdr)) != 0)
        return fail(c, FI_ERR_INVALID, "checkpoint of another grid, forcing or library");
    vv(c->u); vv(c->ue); vv(c->dudt); vv(c->H_ice); vv(c->H_af); vv(c->H_water); vv(c->z_b); vv(c->dz_ss);
    vv(c->z_ss); vb(c->mg); vb(c->mo); vv(c->u_eq);
    vv(c->ca_ice); vv(c->ca_sw); vv(c->ca_sed); vv(c->ca_litho); vv(c->ca_mantle); vv(c->ca_load); vv(c->ca_full);
    sc(c->V_af); sc(c->V_pov); sc(c->V_den); sc(c->delta_V); sc(c->z_bsl); sc(c->A_bsl); sc(c->count); sc(c->k1_stale);
    sc(c->t_timer); sc(c->integ); sc(c->t); sc(c->dt); sc(c->qold); sc(c->q11); sc(c->EEst);
    vv(c->uprev); vv(c->k1);
    sc(c->nrhs); sc(c->nacc); sc(c->nrej); sc(c->kout);
    if (mode == 2) {   // stage work vectors of fi_integrate (sized by fi_integrator_init otherwise)
        c->k2.resize(c->N);
        c->k3.resize(c->N);
        c->k4.resize(c->N);
        c->tmp.resize(c->N);
    }
    return FI_OK;
}

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
    size_t off = 0;
    return ck_walk(c, 1, (unsigned char*)buf, off);
}

int fi_checkpoint_load(fi_ctx* c, const void* buf, size_t bytes) {
    size_t need = 0;
    int st = fi_checkpoint_size(c, &need);
    if (st) return st;
    if (!buf || bytes != need) return fail(c, FI_ERR_INVALID, "fi_checkpoint_load: buffer size");
    size_t off = 0;
    return ck_walk(c, 2, (unsigned char*)const_cast<void*>(buf), off);
}

// The device engine's autotuner (persistent grid size, FFT radix orders) has nothing to tune in
// this FP64 CPU restatement: fi_autotune leaves the run unchanged and reports the default plans;
// pinned plan codes are validated and stored, and change nothing.
namespace {
uint64_t default_plan_code(int L) {
    uint64_t code = 0;
    int nr = 0, m = L;
    for (int r : {8, 4, 2, 3, 5, 7})
        while (m % r == 0 && nr < 16) {
            code |= (uint64_t)r << (4 * nr++);
            m /= r;
        }
    return m == 1 ? code : 0;
}
bool plan_code_ok(int L, uint64_t code) {
    long long prod = 1;
    int nr = 0;
    for (; nr < 16; ++nr) {
        const int r = (int)((code >> (4 * nr)) & 15u);
        if (r == 0) break;
        if (r != 2 && r != 3 && r != 4 && r != 5 && r != 7 && r != 8) return false;
        prod *= r;
    }
    return nr > 0 && (nr == 16 || (code >> (4 * nr)) == 0) && prod == L;
}
void plan_lengths(const fi_ctx* c, int L[4]) {
    L[0] = c->nx;
    L[1] = c->ny;
    L[2] = nextfastfft(2 * c->nx - 1);
    L[3] = nextfastfft(2 * c->ny - 1);
}
}  // namespace

int fi_set_fft_plans(fi_ctx* c, const uint64_t plans[4]) {
    if (!c || !plans) return FI_ERR_INVALID;
    if (!c->initialized) return fail(c, FI_ERR_STATE, "fi_set_fft_plans before fi_init");
    int L[4];
    plan_lengths(c, L);
    for (int q = 0; q < 4; ++q)
        if (plans[q] && !plan_code_ok(L[q], plans[q]))
            return fail(c, FI_ERR_INVALID, "fi_set_fft_plans: radices (2, 3, 4, 5, 7, 8) must multiply to the length");
    for (int q = 0; q < 4; ++q)
        if (plans[q]) c->plans[q] = plans[q];
    return FI_OK;
}

int fi_get_fft_plans(fi_ctx* c, uint64_t plans[4]) {
    if (!c || !plans) return FI_ERR_INVALID;
    if (!c->initialized) return fail(c, FI_ERR_STATE, "fi_get_fft_plans before fi_init");
    int L[4];
    plan_lengths(c, L);
    for (int q = 0; q < 4; ++q) plans[q] = c->plans[q] ? c->plans[q] : default_plan_code(L[q]);
    return FI_OK;
}

int fi_set_workgroups(fi_ctx* c, int32_t g) {
    if (!c) return FI_ERR_INVALID;
    if (!c->initialized) return fail(c, FI_ERR_STATE, "fi_set_workgroups before fi_init");
    return g >= 1 ? FI_OK : fail(c, FI_ERR_INVALID, "fi_set_workgroups: g >= 1");
}

int fi_autotune(fi_ctx* c, int32_t what, int32_t attempts, fi_tune_info* info) {
    if (!c || !info || attempts < 1 || (what & ~(FI_TUNE_WORKGROUPS | FI_TUNE_FFT_PLANS))) return FI_ERR_INVALID;
    std::memset(info, 0, sizeof(*info));
    if (!c->integ) return fail(c, FI_ERR_STATE, "fi_autotune before fi_integrator_init");
    info->workgroups = 1;
    return fi_get_fft_plans(c, info->plans);
}

int fi_batch_autotune_fft(fi_batch* b, int32_t attempts) { return (b && attempts >= 1) ? FI_OK : FI_ERR_INVALID; }

int fi_get_stats(fi_ctx* c, fi_stats* o) {
    if (!c || !o) return FI_ERR_INVALID;
    std::memset(o, 0, sizeof(*o));
    o->t = c->integ ? c->t : c->t_timer;
    o->dt = c->dt;
    o->n_rhs = c->nrhs;
    o->n_accept = c->nacc;
    o->n_reject = c->nrej;
    o->count_sparse_updates = c->count;
    o->n_output = c->kout;
    o->V_af = c->V_af;
    o->V_den = c->V_den;
    o->V_pov = c->V_pov;
    o->delta_V = c->delta_V;
    o->z_bsl = c->z_bsl;
    o->A_ocean = c->A_bsl;
    o->last_EEst = c->EEst;
    return FI_OK;
}

// ---------------------------------------------------------------------------------------
// ref_* hooks for the unit checks
// ---------------------------------------------------------------------------------------
void ref_fft(int n, const double* re, const double* im, double* ore, double* oim, int sign) {
    FFT f(n);
    std::vector<cd> a(n), w(n);
    for (int k = 0; k < n; ++k) a[k] = cd(re[k], im[k]);
    f.exec(a.data(), w.data(), sign);
    for (int k = 0; k < n; ++k) { ore[k] = a[k].real(); oim[k] = a[k].imag(); }
}

void ref_derivatives(int nx, int ny, double dx, double dy, const double* u, double* uxx, double* uyy,
                     double* ux, double* uxy) {
    const vec Dx((size_t)nx * ny, dx), Dy((size_t)nx * ny, dy);
    second_derivatives(uxx, uyy, ux, uxy, u, u, u, nx, ny, Dx.data(), Dy.data());
}

// the same with per-cell steps K dx, K dy
void ref_derivatives_k(int nx, int ny, double dx, double dy, const double* K, const double* u, double* uxx,
                       double* uyy, double* ux, double* uxy) {
    vec Dx((size_t)nx * ny), Dy((size_t)nx * ny);
    for (size_t p = 0; p < Dx.size(); ++p) {
        Dx[p] = K[p] * dx;
        Dy[p] = K[p] * dy;
    }
    second_derivatives(uxx, uyy, ux, uxy, u, u, u, nx, ny, Dx.data(), Dy.data());
}

// conv! -> output_cropped ((2nx-1) x (2ny-1)), src/convolutions.jl:94-106
void ref_conv_cropped(int nx, int ny, const double* kernel, const double* input, double* out) {
    ConvPlan cp;
    cp.init(nx, ny);
    std::vector<cd> kh = cp.spectrum(kernel);
    vec crop;
    cp.conv_cropped(input, kh, crop);
    std::memcpy(out, crop.data(), crop.size() * sizeof(double));
}

// samesize_conv! with an extended-space CornerBC (src/convolutions.jl:134-143)
void ref_samesize_conv(int nx, int ny, const double* kernel, const double* input, double x_border, double* out) {
    ConvPlan cp;
    cp.init(nx, ny);
    std::vector<cd> kh = cp.spectrum(kernel);
    int i1, i2, j1, j2;
    samesize_conv_indices(nx, nx / 2, i1, i2);
    samesize_conv_indices(ny, ny / 2, j1, j2);
    samesize_conv_ext(out, input, cp, kh, i1, i2, j1, j2, x_border);
}

int ref_nextfastfft(int n) { return nextfastfft(n); }

void ref_gauss_legendre(int n, double* x, double* w) {
    vec xx, ww;
    gauss_legendre(n, xx, ww);
    for (int k = 0; k < n; ++k) { x[k] = xx[k]; w[k] = ww[k]; }
}

void ref_pseudodiff(int nx, int ny, double Wx, double Wy, double* out) {
    vec p = pseudodiff(nx, ny, Wx, Wy);
    std::memcpy(out, p.data(), p.size() * sizeof(double));
}

void ref_elastic_green(int nx, int ny, double Wx, double Wy, const double* rm, const double* ge, int nrm,
                       int nq, double* out) {
    vec r(rm, rm + nrm), g(ge, ge + nrm);
    vec G = elastic_green(nx, ny, Wx, Wy, r, g, nq);
    std::memcpy(out, G.data(), G.size() * sizeof(double));
}

// OpenMP threads of the oracle's parallel loops (the CPU baseline times all cores and one core)
void ref_set_threads(int n) { omp_set_num_threads(n > 0 ? n : 1); }
int ref_max_threads(void) { return omp_get_max_threads(); }

void ref_geoid_green(int nx, int ny, double Wx, double Wy, double r_pole, double mE, double* out) {
    vec G = geoid_green(nx, ny, Wx, Wy, r_pole, mE);
    std::memcpy(out, G.data(), G.size() * sizeof(double));
}

// SolidEarth(domain; layer_boundaries, layer_viscosities, ...) derived fields
// (src/solidearth.jl:175-233, src/material.jl:10,46-51,154-178,221-262); nl layers of nx*ny.
// compressible != 0 applies CompressibleMantle. Outputs D, eta_eff, R, scaled_pseudodiff_inv.
void ref_solidearth(int nx, int ny, double Wx, double Wy, int nl, const double* lb, const double* lv,
                    const double* maskactive, double E, double nu, double nu_m, int compressible,
                    double* D, double* eta, double* R, double* spinv) {
    size_t N = (size_t)nx * ny;
    vec kappa = pseudodiff(nx, ny, Wx, Wy);
    for (size_t p = 0; p < N; ++p) {
        double h = lb[p];
        D[p] = (E * h * h * h) / (12 * (1 - nu * nu));
        eta[p] = lv[(size_t)(nl - 1) * N + p];
        R[p] = 1.0;
    }
    for (int l = 1; l <= nl - 1; ++l) {
        // channel_thickness = lb[:,:,end-l+1] - lb[:,:,end-l]; visc_ratio = lv[:,:,end-l] ./ eta
        int a = nl - l, b = nl - l - 1;  // 0-based layer indices of end-l+1 and end-l
        double sth = 0, svr = 0, cnt = 0;
        for (size_t p = 0; p < N; ++p) {
            sth += lb[(size_t)a * N + p] - lb[(size_t)b * N + p];
            if (maskactive[p] != 0) { svr += lv[(size_t)b * N + p] / eta[p]; cnt += 1; }
        }
        double th = sth / (double)N, vr = svr / cnt;
        for (size_t p = 0; p < N; ++p) {
            // channel_scaling (src/material.jl:232-262)
            double k = kappa[p];
            double C = std::cosh(th * k), S = std::sinh(th * k);
            double num = 0, den = 0;
            num += 2 * vr * C * S;
            num += (1 - vr * vr) * th * th * k * k;
            num += vr * vr * S * S + C * C;
            den += (vr + 1 / vr) * C * S;
            den += (vr - 1 / vr) * th * k;
            den += S * S + C * C;
            R[p] *= num / den;
        }
    }
    if (compressible) {
        double f = (1 + 0.5) / (1 + nu_m);
        for (size_t p = 0; p < N; ++p) eta[p] *= f;
    }
    for (size_t p = 0; p < N; ++p) spinv[p] = 1.0 / (R[p] * kappa[p]);
}

// analytic_solution(r, t, c, solidearth, H0, R0; n_quad_support) (src/analytic_solutions.jl:11-65);
// t in seconds, D and eta are the (laterally constant) rigidity and effective viscosity.
double ref_analytic_solution(double r, double t, double rho_ice, double g, double H0, double R0,
                             double rho_um, double D, double eta, int nq) {
    vec xq, wq;
    gauss_legendre(nq, xq, wq);
    std::vector<double> support;
    support.push_back(1e-14);
    for (int k = 0; k <= 140; ++k) support.push_back(std::pow(10.0, -10.0 + 0.05 * k));
    support.push_back(1.0);
    double scaling = rho_ice * g * H0 * R0;
    double integral = 0.0;
    for (size_t s = 0; s + 1 < support.size(); ++s) {
        double x1 = support[s], x2 = support[s + 1];
        double m = 2.0 / (x2 - x1), p = -1.0 - m * x1;
        double sum = 0.0;
        for (int i = 0; i < nq; ++i) {
            double kappa = (xq[i] - p) / m;
            double beta = rho_um * g + D * std::pow(kappa, 4);
            double j0v = ::j0(kappa * r), j1v = ::j1(kappa * R0);
            double f = std::isinf(t) ? -j0v * j1v / beta
                                     : (std::exp(-beta * t / (2 * eta * kappa)) - 1) * j0v * j1v / beta;
            sum = sum + f * wq[i] / m;
        }
        integral += sum;
    }
    return scaling * integral;
}

// BS3 + PI controller on the scalar test problem du/dt = lam * u, u(0) = u0, to t_end (no
// tstops): the same stage, error and controller code path as fi_integrate, used for the
// order and tableau checks. Returns u(t_end); nsteps receives the accepted step count.
double ref_bs3_linear(double lam, double u0, double t_end, double reltol, double abstol, double dt_fixed,
                      int64_t* nsteps) {
    double t = 0, u = u0, k1 = lam * u0, qold = QOLDINIT, dt = dt_fixed > 0 ? dt_fixed : 1e-3, q11 = 1;
    int64_t n = 0;
    while (t < t_end) {
        double h = std::min(dt, t_end - t);
        double k2 = lam * (u + h * A21 * k1);
        double k3 = lam * (u + h * A32 * k2);
        double un = u + h * (A41 * k1 + A42 * k2 + A43 * k3);
        double k4 = lam * un;
        if (dt_fixed > 0) { t += h; u = un; k1 = k4; ++n; continue; }
        double ut = h * (BT1 * k1 + BT2 * k2 + BT3 * k3 + BT4 * k4);
        double e = std::fabs(ut / (abstol + std::max(std::fabs(u), std::fabs(un)) * reltol));
        double q;
        if (e == 0) q = 1 / QMAX;
        else {
            q11 = std::pow(e, BETA1);
            q = std::max(1 / QMAX, std::min(1 / QMIN, q11 / std::pow(qold, BETA2) / GAMMA));
        }
        if (e <= 1) { qold = std::max(e, QOLDINIT); t += h; u = un; k1 = k4; dt = h / q; ++n; }
        else dt = h / std::min(1 / QMIN, q11 / GAMMA);
    }
    if (nsteps) *nsteps = n;
    return u;
}

}  // extern "C"

// ------------------------------------------------------------------------------------------
// Ensembles: the reference integrates one deep copy of the Simulation per member in a threaded
// loop (ext/FastIsostasyEnsembleKalmanProcessesExt.jl:103-122); restated here as a loop over
// independent member contexts (OpenMP inside each member's RHS).
// ------------------------------------------------------------------------------------------
struct fi_batch {
    std::vector<fi_ctx*> mem;
    std::string err;
    int64_t n_launch = 0;
    double cell_attempts = 0;
};

extern "C" {

int fi_create_batch(const fi_config* cfgs, int32_t nmem, fi_batch** out) {
    if (!cfgs || !out || nmem < 1) return FI_ERR_INVALID;
    *out = nullptr;
    fi_batch* b = new fi_batch();
    for (int m = 0; m < nmem; ++m) {
        fi_ctx* c = nullptr;
        int st = fi_create(&cfgs[m], &c);
        if (st) {
            fi_batch_destroy(b);
            return st;
        }
        b->mem.push_back(c);
    }
    *out = b;
    return FI_OK;
}

fi_ctx* fi_batch_member(fi_batch* b, int32_t m) {
    if (!b || m < 0 || m >= (int32_t)b->mem.size()) return nullptr;
    return b->mem[m];
}

void fi_batch_destroy(fi_batch* b) {
    if (!b) return;
    for (fi_ctx* c : b->mem) fi_destroy(c);
    delete b;
}

const char* fi_batch_last_error(const fi_batch* b) { return b ? b->err.c_str() : "null batch"; }

int fi_batch_set_team_size(fi_batch* b, int32_t ts) { return (b && ts >= 0) ? FI_OK : FI_ERR_INVALID; }

int fi_batch_integrate(fi_batch* b, double t_target, int32_t fsal) {
    if (!b) return FI_ERR_INVALID;
    std::vector<double> tt(b->mem.size(), t_target);
    return fi_batch_integrate_each(b, tt.data(), fsal);
}

int fi_batch_integrate_each(fi_batch* b, const double* t_targets, int32_t fsal) {
    if (!b || !t_targets) return FI_ERR_INVALID;
    for (size_t m = 0; m < b->mem.size(); ++m)
        if (b->mem[m]->integ && t_targets[m] < b->mem[m]->t) {
            b->err = "member " + std::to_string(m) + ": target before current time";
            return FI_ERR_INVALID;
        }
    for (size_t m = 0; m < b->mem.size(); ++m) {
        fi_ctx* c = b->mem[m];
        const double t_target = t_targets[m];
        if (!c->integ) {
            b->err = "member " + std::to_string(m) + ": fi_batch_integrate before fi_integrator_init";
            return FI_ERR_STATE;
        }
        if (c->t >= t_target) continue;
        const int64_t a0 = c->nacc + c->nrej;
        std::vector<double> tout;
        tout.swap(c->tout);   // member output times are not used
        const int64_t kout = c->kout;
        int st = fi_integrate(c, t_target, nullptr, nullptr);
        c->tout.swap(tout);
        c->kout = kout;
        if (!st && fsal) {
            st = update_diagnostics(c, c->k1.data(), c->uprev.data(), c->t);
            c->u = c->uprev;
        }
        if (st) {
            b->err = "member " + std::to_string(m) + ": " + c->err;
            return st;
        }
        b->cell_attempts += (double)(c->nacc + c->nrej - a0) * (double)c->N;
    }
    b->n_launch += 1;
    return FI_OK;
}

// the oracle's "device" arrays are host arrays (as for fi_rhs_device)
int fi_set_field_dev(fi_ctx* c, int32_t f, const void* d, size_t n) {
    return fi_set_field(c, f, (const double*)d, n);
}
int fi_get_field_dev(fi_ctx* c, int32_t f, void* d, size_t n) { return fi_get_field(c, f, (double*)d, n); }

// the CPU oracle allocates no device memory in either entry point
int fi_batch_integrate_each_noalloc(fi_batch* b, const double* t_targets, int32_t fsal) {
    return fi_batch_integrate_each(b, t_targets, fsal);
}

int fi_batch_get_info(const fi_batch* b, fi_batch_info* o) {
    if (!b || !o) return FI_ERR_INVALID;
    std::memset(o, 0, sizeof(*o));
    o->n_members = (int32_t)b->mem.size();
    o->team_size = 0;
    o->n_teams = 0;
    o->tuned = 1;
    o->n_launch = b->n_launch;
    o->cell_attempts = b->cell_attempts;
    return FI_OK;
}

}  // extern "C"
