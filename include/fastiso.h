/*
 * fastiso.h -- C ABI of the LV-ELVA deformation engine.
 *
 * One ABI, two implementations:
 *   libfastiso_hip.so  the MI355X (gfx950) engine: hand-written HIP kernels, device-resident state;
 *   libfastiso_ref.so  the FP64 CPU restatement used as the test oracle (built from oracle/).
 *
 * Conventions (all entry points):
 *   - plain C types only; fields are FP64 host arrays in Julia column-major order (nx, ny),
 *     x fastest: element (i, j) at i + j*nx (0-based);
 *   - every function returns an int status (FI_OK = 0, negative = error); no exception crosses
 *     the ABI; fi_last_error() gives the text of the last error on that context;
 *   - the library owns its device memory, the caller owns host arrays;
 *   - one fi_ctx per host thread; calls on one context are not thread-safe.
 *
 * Each entry point names the reference function it stands in for (FastIsostasy.jl, file:line).
 */
#ifndef FASTISO_H
#define FASTISO_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FI_ABI_VERSION 6

/* ---- status codes ------------------------------------------------------------------------ */
enum {
    FI_OK = 0,
    FI_ERR_INVALID = -1,      /* bad argument / size mismatch / missing required field        */
    FI_ERR_UNSUPPORTED = -2,  /* model choice not implemented by this engine                   */
    FI_ERR_TIME_RANGE = -3,   /* ice-thickness interpolation outside its time vector           */
                              /* (src/interpolations.jl:51-74 throws there)                      */
    FI_ERR_NONFINITE = -4,    /* non-finite state or error norm that the step control cannot fix */
    FI_ERR_DEVICE = -5,       /* HIP runtime error or device-side abort                         */
    FI_ERR_STATE = -6,        /* call out of order (e.g. fi_rhs before fi_init)                  */
    FI_ERR_DT_MIN = -7        /* adaptive step fell below the minimum step                        */
};

/* ---- model choices (singleton types of the reference) ------------------------------------ */
enum { /* src/solidearth.jl:13-36 */
    FI_RIGID_LITHOSPHERE = 0,
    FI_LATERALLY_CONSTANT_LITHOSPHERE = 1,
    FI_LATERALLY_VARIABLE_LITHOSPHERE = 2
};
enum { /* src/solidearth.jl:50-82 */
    FI_RIGID_MANTLE = 0,
    FI_RELAXED_MANTLE = 1,
    FI_MAXWELL_MANTLE = 2
};
enum { /* src/solidearth.jl:95-109 */
    FI_INCOMPRESSIBLE_LITHOSPHERE_COLUMN = 0,
    FI_COMPRESSIBLE_LITHOSPHERE_COLUMN = 1
};
enum { /* src/sealevel.jl:114-134 */
    FI_LATERALLY_CONSTANT_SEA_SURFACE = 0,
    FI_LATERALLY_VARIABLE_SEA_SURFACE = 1,
    FI_IMPOSED_SEA_SURFACE = 2          /* dz_ss from fi_set_sea_surface_snapshots */
};
enum { /* src/loads.jl:12-26 */
    FI_NO_SEALEVEL_LOAD = 0,
    FI_INTERACTIVE_SEALEVEL_LOAD = 1
};
enum { /* src/barystatic_sealevel.jl:21-101: the No- and Goelzer- contributions */
    FI_NO_CONTRIBUTION = 0,
    FI_GOELZER_CONTRIBUTION = 1
};

/* ---- configuration (the scalar part of RegionalDomain, PhysicalConstants, SolidEarth,
 *      RegionalSeaLevel, BoundaryConditions and SolverOptions) ----------------------------- */
typedef struct fi_config {
    int32_t nx, ny;              /* RegionalDomain(Wx, Wy, nx, ny)      src/domain.jl:76-83   */
    double Wx, Wy;               /* half-widths (m)                                            */
    /* PhysicalConstants                                                src/constants.jl:16-28 */
    double g;                    /* 9.8                                                        */
    double seconds_per_year;     /* 31557600 (src/utils.jl:7)                                  */
    double rho_ice;              /* 910                                                        */
    double rho_water;            /* 1000                                                       */
    double rho_seawater;         /* 1023                                                       */
    double r_pole;               /* 6357e3                                                     */
    double mE;                   /* 5.972e24                                                   */
    /* SolidEarth scalars                                               src/solidearth.jl:175-193 */
    double litho_poissonratio;   /* 0.28                                                       */
    double rho_uppermantle;      /* 3400                                                       */
    double rho_litho;            /* 3200                                                       */
    int32_t lithosphere;         /* FI_*_LITHOSPHERE                                           */
    int32_t mantle;              /* FI_*_MANTLE: FI_MAXWELL_MANTLE with the laterally variable
                                    lithosphere is LV-ELVA (src/deformation.jl:91-103), with a
                                    rigid or laterally constant one the implicit step (dt_min,
                                    src/deformation.jl:52-89); FI_RELAXED_MANTLE needs a rigid or
                                    laterally constant one (ELRA, src/deformation.jl:29-45)    */
    int32_t lithosphere_column;  /* FI_*_LITHOSPHERE_COLUMN                                    */
    /* RegionalSeaLevel                                                 src/sealevel.jl:150-166 */
    int32_t sea_surface;         /* FI_*_SEA_SURFACE                                           */
    int32_t sealevel_load;       /* FI_*_SEALEVEL_LOAD                                         */
    int32_t volume_contribution;     /* FI_GOELZER_CONTRIBUTION by default                     */
    int32_t density_contribution;    /* FI_GOELZER_CONTRIBUTION by default                     */
    int32_t adjustment_contribution; /* FI_NO_CONTRIBUTION by default                          */
    double bsl_z;                /* ConstantBSL: z = ref.z (src/barystatic_sealevel.jl:215-218) */
    /* BoundaryConditions: x_border of the three CornerBCs   src/boundary_conditions.jl:308-325 */
    double viscous_x_border, elastic_x_border, ssh_x_border;
    /* SolverOptions / DiffEqOptions                                    src/simulation.jl:11-26 */
    double dt_sparse;            /* dt_sparse_diagnostics, 10 yr                               */
    double reltol;               /* 1e-5                                                       */
    double abstol;               /* 1e-6                                                       */
    int32_t device;              /* HIP device ordinal (ignored by the CPU oracle)             */
    int32_t reserved;
    double dt_min;               /* DiffEqOptions.dt_min (yr) with alg = Euler: required by, and
                                    only used by, the implicit step of a Maxwell mantle under a
                                    rigid or laterally constant lithosphere (src/deformation.jl:
                                    52-89, src/simulation.jl:244-253); 0 otherwise             */
} fi_config;

/* ---- fields ------------------------------------------------------------------------------- */
enum {
    /* static inputs, set before fi_init (each nx*ny doubles) */
    FI_LITHO_RIGIDITY = 1,        /* SolidEarth.litho_rigidity          src/solidearth.jl:211  */
    FI_EFFECTIVE_VISCOSITY = 2,   /* SolidEarth.effective_viscosity     src/solidearth.jl:212-216 */
    FI_SCALED_PSEUDODIFF_INV = 3, /* SolidEarth.scaled_pseudodiff_inv   src/solidearth.jl:222  */
    FI_MASK_ACTIVE = 4,           /* SolidEarth.maskactive (0/1)        src/solidearth.jl:183  */
    FI_GREEN_ELASTIC = 5,         /* elastic Green kernel               src/tools.jl:64-71     */
    FI_GREEN_GEOID = 6,           /* sea-surface Green kernel           src/sealevel.jl:208-219 */
    FI_REF_U = 7,                 /* Simulation kwargs u_ref, ue_ref, dz_ss_ref, z_b_ref        */
    FI_REF_UE = 8,                /*                                    src/simulation.jl:112-116 */
    FI_REF_DZ_SS = 9,
    FI_REF_Z_B = 10,
    FI_REF_H_ICE = 11,            /* reference ice; default: forcing at t0 (src/simulation.jl:133-134) */
    FI_DISTORTION = 12,           /* RegionalDomain.K, the map scale factor (src/domain.jl:140-147, 167):
                                     grid steps K dx, K dy and cell area dx dy K^2 per cell, the RHS
                                     factor K / (2 eta) (src/deformation.jl:95) and K^2 on the
                                     convolution inputs (src/deformation.jl:177); default 1. The HIP
                                     engine supports it on its FP64-moment path (nx <= 512, ny >= 6) */
    FI_GREEN_VISCOUS = 13,        /* ELRA Green kernel (green_viscous, src/material.jl:375-380, built
                                     from mean(litho_rigidity) and rho_uppermantle in src/tools.jl:
                                     55-60); required with FI_RELAXED_MANTLE                   */
    FI_TAU = 14,                  /* SolidEarth.tau, relaxation time (yr) of the relaxed mantle
                                     (src/solidearth.jl:167,189); required with FI_RELAXED_MANTLE */
    FI_LITHO_THICKNESS = 15,      /* SolidEarth.litho_thickness (m), src/solidearth.jl:210; needed
                                     only to read FI_U_X / FI_U_Y                              */
    /* state, readable after fi_init; FI_U and FI_H_ICE are also writable */
    FI_U = 32,                    /* CurrentState fields                src/state.jl:49-77     */
    FI_DUDT = 33,
    FI_UE = 34,
    FI_DZ_SS = 35,
    FI_Z_SS = 36,
    FI_Z_B = 37,
    FI_H_ICE = 38,                /* writable: ExternallyUpdatedIceThickness (src/boundary_conditions.jl:48,60-62) */
    FI_H_AF = 39,
    FI_H_WATER = 40,
    FI_MASK_GROUNDED = 41,
    FI_MASK_OCEAN = 42,
    FI_CANOM_ICE = 43,            /* ColumnAnomalies                    src/state.jl:1-9       */
    FI_CANOM_SEAWATER = 44,
    FI_CANOM_LITHO = 45,
    FI_CANOM_MANTLE = 46,
    FI_CANOM_LOAD = 47,
    FI_CANOM_FULL = 48,
    FI_U_EQ = 49,                 /* CurrentState.u_eq, equilibrium displacement of the relaxed
                                     mantle (src/state.jl:60, src/deformation.jl:37-40); the
                                     reference state u_ref until the first RHS evaluation      */
    FI_U_X = 50,                  /* CurrentState.u_x, u_y: thin-plate horizontal displacement of
                                     u + ue, -h/2 times its x / y derivative (CPU edge rules),
                                     as the output callbacks compute it (src/simulation.jl:
                                     204-207,218-221; src/deformation.jl:141-164); needs
                                     FI_LITHO_THICKNESS                                        */
    FI_U_Y = 51
};

/* ---- run statistics ------------------------------------------------------------------------ */
typedef struct fi_stats {
    double t;                     /* current time (yr)                                          */
    double dt;                    /* step the controller proposes next (yr)                     */
    int64_t n_rhs;                /* RHS evaluations (update_diagnostics! calls)                */
    int64_t n_accept, n_reject;   /* accepted / rejected steps                                  */
    int64_t count_sparse_updates; /* CurrentState.count_sparse_updates  src/state.jl:76        */
    int64_t n_output;             /* output callbacks fired (NativeOutput.k - 1)               */
    double V_af, V_den, V_pov, delta_V, z_bsl;   /* src/state.jl:69-73                         */
    double last_EEst;             /* error norm of the last attempted step                      */
    double device_ms;             /* integration-kernel time, HIP events on the context's stream */
    int64_t n_launch;             /* integration-kernel launches timed in device_ms             */
    int64_t n_accept_timed;       /* accepted steps inside those launches                       */
    double A_ocean;               /* ocean surface of the BSL rule (bsl.A, m^2)                 */
} fi_stats;

/* Output callback, fired after an accepted step at which t >= the next output time
 * (nout_condition/nout_affect!, src/simulation.jl:194-225); k is the 0-based output index.
 * The callback may call fi_get_field / fi_get_stats / fi_set_field(FI_H_ICE). */
typedef void (*fi_output_cb)(void* user, double t, int32_t k);

typedef struct fi_ctx fi_ctx;

/* ---- entry points --------------------------------------------------------------------------- */

/* ABI version (FI_ABI_VERSION) and engine name ("hip-gfx950" or "cpu-oracle-fp64"). */
int32_t fi_abi_version(void);
const char* fi_engine_name(void);

/* Fill cfg with the reference defaults (PhysicalConstants(), SolidEarth(), RegionalSeaLevel(),
 * SolverOptions(); src/constants.jl:16-28, src/solidearth.jl:115-121,175-193,
 * src/sealevel.jl:150-166, src/simulation.jl:11-26) for an nx x ny grid of half-widths Wx, Wy. */
void fi_default_config(fi_config* cfg, int32_t nx, int32_t ny, double Wx, double Wy);

/* Create a context: RegionalDomain(Wx, Wy, nx, ny) (src/domain.jl:76-169) plus the buffers of
 * GIATools (src/tools.jl:49-98). Allocates device memory. */
int fi_create(const fi_config* cfg, fi_ctx** out);
void fi_destroy(fi_ctx* ctx);
const char* fi_last_error(const fi_ctx* ctx);

/* Set / get one field (count must be nx*ny). Static inputs go before fi_init. */
int fi_set_field(fi_ctx* ctx, int32_t field, const double* host, size_t count);
int fi_get_field(fi_ctx* ctx, int32_t field, double* host, size_t count);

/* fi_set_field / fi_get_field on device-resident FP64 arrays (HIP engine: device pointers on the
 * context's device; CPU oracle: host pointers): the array-type seam of the reference
 * (KernelMatrix and the CuMatrix methods, src/convenience_types.jl:2-29, src/domain.jl:66-67) for
 * a caller that holds its fields on the GPU. FI_U moves device to device in both directions (no
 * host round trip; the read applies the pending corner BC on the device); FI_H_ICE of an
 * externally updated ice thickness is converted to FP32 on the device. Other fields are staged
 * through host memory: static inputs are preprocessed on the host once, before fi_init. */
int fi_set_field_dev(fi_ctx* ctx, int32_t field, const void* dev, size_t count);
int fi_get_field_dev(fi_ctx* ctx, int32_t field, void* dev, size_t count);

/* TimeInterpolatedIceThickness(t_vec, H_vec) (src/boundary_conditions.jl:25-40) with the
 * piecewise-linear TimeInterpolation2D (src/interpolations.jl:51-74): H holds nt snapshots of
 * nx*ny each, t ascending (yr). flat_bc != 0 clamps outside [t_1, t_nt], else FI_ERR_TIME_RANGE. */
int fi_set_ice_snapshots(fi_ctx* ctx, const double* t, int32_t nt, const double* H, int32_t flat_bc);

/* ImposedSeaSurface(dz_ss_itp) (src/sealevel.jl:132-134): with cfg.sea_surface =
 * FI_IMPOSED_SEA_SURFACE every sparse update sets dz_ss to the TimeInterpolation2D of these nt
 * snapshots (nx*ny each, t ascending, yr) at the update's time (update_dz_ss!,
 * src/sealevel.jl:188-191), instead of the sea-surface convolution. flat_bc as for the ice. */
int fi_set_sea_surface_snapshots(fi_ctx* ctx, const double* t, int32_t nt, const double* dz_ss, int32_t flat_bc);

/* The same forcing when every snapshot is a(t_k) * H_shape: H(t) = a(t) H_shape with a(t)
 * piecewise linear over the knots (t_k, a_k). Mathematically identical to fi_set_ice_snapshots
 * on the snapshots a_k * H_shape; reads one field per RHS instead of two. */
int fi_set_ice_separable(fi_ctx* ctx, const double* t, const double* a, int32_t nt,
                         const double* H_shape, int32_t flat_bc);

/* Barystatic sea level: the rule of RegionalSeaLevel.bsl (update_bsl!, src/barystatic_sealevel.jl:
 * 215-234) and the mechanism of RegionalSeaLevel.update_bsl (internal_update_bsl!,
 * src/sealevel.jl:242-262), set before fi_init. At every sparse update the engine sums the volume
 * contributions of the state before the update, delta_V = V_new - V_old, and then, with an
 * internal update, applies update_bsl!(bsl, -delta_V, t):
 *   FI_BSL_CONSTANT                 z = cfg.bsl_z (the default)
 *   FI_BSL_CONSTANT_OCEAN_SURFACE   A = A_ref; z += -delta_V / A
 *   FI_BSL_PIECEWISE_CONSTANT       z += -delta_V / A; A = interp(z) over the table (x = z, y = A)
 *   FI_BSL_IMPOSED                  z = interp(t) over the table (x = t, y = z)
 * The ocean surface starts at A_ref and the BSL at cfg.bsl_z. interp is TimeInterpolation0D
 * (src/interpolations.jl:28-49): linear between ascending knots x[0..n-1]; outside them the end
 * value if flat_bc != 0, else FI_ERR_TIME_RANGE from the integration. FI_BSL_EXTERNAL_UPDATE
 * computes delta_V and leaves z unchanged. */
enum {
    FI_BSL_CONSTANT = 0,
    FI_BSL_CONSTANT_OCEAN_SURFACE = 1,
    FI_BSL_PIECEWISE_CONSTANT = 2,
    FI_BSL_IMPOSED = 3
};
enum { FI_BSL_INTERNAL_UPDATE = 0, FI_BSL_EXTERNAL_UPDATE = 1 };
int fi_set_bsl(fi_ctx* ctx, int32_t rule, int32_t update, double A_ref, const double* x, const double* y,
               int32_t n, int32_t flat_bc);

/* Output times = tstops of the solver and the times of the output callbacks
 * (tstops = sort(vcat(nout.t, ncout.t)), src/simulation.jl:255-257). */
int fi_set_output_times(fi_ctx* ctx, const double* t, int32_t n);

/* Finish the Simulation constructor (src/simulation.jl:105-156): the padded Green spectra
 * (ConvolutionPlan, src/convolutions.jl:41-56,89-92), the reference state (H_ice_ref at t0,
 * z_ss_ref, H_af_ref, masks, H_water_ref) and CurrentState(domain, ref, z_bsl). */
int fi_init(fi_ctx* ctx, double t0, double t_end);

/* The ODE right-hand side update_diagnostics!(dudt, u, sim, t) (src/simulation.jl:335-393):
 * corner BC on u (u is mutated, as the reference mutates the solver's stage vector), load
 * update, the sparse elastic / sea-surface update when (t - t0)/dt_sparse >= count, and
 * update_dudt!(MaxwellMantle, LaterallyVariableLithosphere) (src/deformation.jl:91-131).
 * u and dudt are nx*ny host FP64 arrays; dudt is in m/yr. */
int fi_rhs(fi_ctx* ctx, double t, double* u, double* dudt);

/* Same on device-resident FP64 arrays (HIP engine: device pointers, stream-ordered on the
 * context's stream; CPU oracle: host pointers). The form a Julia caller holding a device
 * array uses per solver stage. */
int fi_rhs_device(fi_ctx* ctx, double t, void* u, void* dudt);

/* Only the sparse block of update_diagnostics! (src/simulation.jl:359-381): elastic response,
 * barystatic volumes, sea-surface perturbation, masks, counted as one sparse update. */
int fi_sparse_update(fi_ctx* ctx, double t);

/* init_integrator / the first half of run! (src/simulation.jl:236-238,293-311): init_problem!
 * (volumes + one RHS at t0), FSAL initialisation and the initial step size. */
int fi_integrator_init(fi_ctx* ctx);

/* Advance the adaptive BS3 integration (src/simulation.jl:255-257; OrdinaryDiffEq BS3 with
 * reltol, abstol) until t == t_target, stopping exactly at the output times and firing cb after
 * accepted steps that reach one. run!(sim) = fi_integrator_init + fi_integrate(t_end);
 * step!(integrator, dt, true) (src/simulation.jl:319-322) = fi_integrate(t + dt). */
int fi_integrate(fi_ctx* ctx, double t_target, fi_output_cb cb, void* user);

int fi_get_stats(fi_ctx* ctx, fi_stats* out);

/* Autotuner. No counterpart in the reference, whose transforms are planned by FFTW/CUFFT
 * (plan_fft / plan_rfft in GIATools, src/tools.jl:100-109). After fi_integrator_init, with the
 * run not at t_end: for each candidate, `attempts` integration attempts (BS3 step attempts, or
 * implicit steps) from the current state are timed on the device (fastest of 2) and the state
 * restored, so the run then continues as if the call had not been made (the tuning launches
 * are not counted in fi_stats).
 *   FI_TUNE_WORKGROUPS: the persistent grid size among the row-pair and column-pair counts,
 *     3/4 and 1/2 of the default and the co-resident maximum. Results do not depend on it:
 *     every cross-workgroup sum runs over fixed per-row-pair partials.
 *   FI_TUNE_FFT_PLANS: the radix order of the hot path's transforms (nx and ny; the padded
 *     lengths under ELRA): radices 8, 4, 2 for the powers of two grouped in every way, and 3,
 *     5, 7, in every distinct order (at most 24 per length, the default first). Another order
 *     rounds differently: fi_set_fft_plans with the chosen codes reproduces a run bit for bit.
 * Plan codes hold the radix of pass s in bits [4s, 4s+4), first pass lowest, zero after the last. */
enum { FI_TUNE_WORKGROUPS = 1, FI_TUNE_FFT_PLANS = 2 };
typedef struct fi_tune_info {
    int32_t workgroups;          /* grid size in use after the call */
    int32_t n_wg_candidates;
    int32_t wg_candidates[8];
    double wg_us[8];             /* us per attempt of each grid size */
    uint64_t plans[4];           /* plan codes in use after the call: nx, ny, nfx, nfy */
    int32_t n_plan_candidates[4];
    double plan_default_us[4];   /* us per attempt with the default plan of that length */
    double plan_best_us[4];      /* ... with the fastest */
} fi_tune_info;
int fi_autotune(fi_ctx* ctx, int32_t what, int32_t attempts, fi_tune_info* info);
/* Pin plans (after fi_init; 0 keeps a length's plan) and read them back. */
int fi_set_fft_plans(fi_ctx* ctx, const uint64_t plans[4]);
int fi_get_fft_plans(fi_ctx* ctx, uint64_t plans[4]);
/* Pin the persistent grid size (after fi_init; 1 <= g <= the co-resident maximum). */
int fi_set_workgroups(fi_ctx* ctx, int32_t g);

/* Timer history (Timer.t_vec and Timer.t_computation, src/simulation.jl:28-52, pushed by
 * t_computation! at the end of every update_diagnostics!, src/simulation.jl:391). With
 * fi_set_timer_history(capacity > 0) before fi_init, right-hand-side call k of the context
 * (0-based since fi_init, k < capacity) stores its time and the device wall clock at its end, on
 * the device. fi_get_timer_history returns, as t_computation! keeps them, the recorded calls whose
 * time is past t0: *n_total of them, the first `capacity` copied to t (yr) and seconds (since the
 * first of them, as run! leaves t_computation, src/simulation.jl:260). A checkpoint load rewinds
 * the call count, so a rolled-back period is overwritten by its rerun. Single contexts only
 * (ensemble members record nothing). */
int fi_set_timer_history(fi_ctx* ctx, int64_t capacity);
int fi_get_timer_history(fi_ctx* ctx, double* t, double* seconds, int64_t capacity, int64_t* n_total);

/* Sets the integrator's time to t >= t0 without integrating (after fi_integrator_init), for
 * drivers that set the displacement themselves (the surrogate macro step and its rollback:
 * fi_set_field(FI_U) then fi_set_time). The column loads are stored at (t, u) and the sparse
 * update (elastic response, sea surface, barystatic sea level; src/simulation.jl:355-381) runs
 * with them at once; the next one is due one dt_sparse later. The next fi_integrate re-evaluates
 * the first-same-as-last derivative. No counterpart in the reference. FI_ERR_UNSUPPORTED under the
 * implicit step, whose every call advances u. */
int fi_set_time(fi_ctx* ctx, double t);

/* Second-order Runge-Kutta-Chebyshev integration (RKC2: Sommeijer, Shampine & Verwer 1997,
 * Appl. Numer. Math. 22; damping eps = 2/13) of the same right-hand side (fi_rhs semantics,
 * sparse trigger included) from the current time to t_target, in equal sub-steps of at most
 * dt_max years. Each sub-step takes the fewest stages s whose real stability interval
 * beta(s) = (w0 + 1) T_s''(w0) / T_s'(w0) covers 1.2 rho h, rho being the spectral radius of the
 * right-hand side's Jacobian (1/yr): the caller's value if rho > 0, else an estimate by 40
 * power iterations from a fixed pseudo-random field, made once per context. Stage vectors are
 * not corner-corrected (the right-hand side corrects its input); the state after each sub-step
 * is. No counterpart in the reference, which integrates with OrdinaryDiffEq's BS3: it is the
 * macro-step predictor of the surrogate-with-fallback mode, whose result the plain solver checks.
 * The next fi_integrate re-evaluates the first-same-as-last derivative. LV-ELVA and ELRA
 * contexts; FI_ERR_UNSUPPORTED under the implicit step. The call's kernels (one persistent
 * launch per right-hand side, one stage kernel between two of them, which also prepares the next
 * launch on the device) are captured once into a HIP graph per sub-step count, stage count,
 * sub-step and starting buffer, and the graph is replayed by later calls with the same ones;
 * FASTISO_RKC_GRAPH=0 in the environment launches them one by one instead. */
typedef struct fi_rkc_info {
    double rho;          /* spectral radius used (1/yr) */
    double h;            /* sub-step (yr) */
    int32_t substeps;
    int32_t stages;      /* stages per sub-step */
    int64_t n_rhs;       /* right-hand-side evaluations of the call (power iterations included) */
    int32_t graph;       /* 1: replayed from a HIP graph, 2: captured by this call, 0: launched directly */
    int32_t reserved;
} fi_rkc_info;
int fi_integrate_rkc(fi_ctx* ctx, double t_target, double dt_max, double rho, fi_rkc_info* info);

/* Checkpoints of a context's dynamic state: the displacement, the first-same-as-last derivative,
 * the step-size controller, the sparse-update state (elastic displacement, sea surface, masks,
 * stored loads, BSL volumes, counters) and the output counter. fi_checkpoint_size gives the byte
 * size, fi_checkpoint_save writes the state into a host buffer, fi_checkpoint_load restores it
 * into a context of the same library, created and initialised (fi_init) with the same
 * configuration and static fields. Restoring and integrating continues bit for bit like the
 * run that was saved. The reference keeps this state in its Simulation object
 * (src/state.jl:49-104) and has no checkpoint file of its own. */
int fi_checkpoint_size(fi_ctx* ctx, size_t* bytes);
int fi_checkpoint_save(fi_ctx* ctx, void* buf, size_t bytes);
int fi_checkpoint_load(fi_ctx* ctx, const void* buf, size_t bytes);

/* ---- ensembles of independent simulations (uneven grids) ----------------------------------
 * The reference runs an ensemble as one deep copy of the Simulation per member, integrated in a
 * threaded loop (ext/FastIsostasyEnsembleKalmanProcessesExt.jl:103-122). Here one persistent
 * launch advances every member: the device is split into teams of workgroups and each team
 * integrates members taken from a queue, largest first. */
typedef struct fi_batch fi_batch;

#define FI_BATCH_MAX_CANDIDATES 8
typedef struct fi_batch_info {
    int32_t n_members;
    int32_t team_size;     /* workgroups per team: the size in use once tuned (given, or chosen
                              by timing), else the size of the last launch */
    int32_t n_teams;       /* teams of the last launch */
    int32_t tuned;         /* 1 once the team size is fixed (given, or chosen by timing) */
    double device_ms;      /* kernel time of all ensemble launches (HIP events) */
    int64_t n_launch;
    double cell_attempts;  /* sum over members of BS3 attempts x grid cells, all launches */
    int32_t last_team_size;              /* workgroups per team of the last launch */
    int32_t n_candidates;                /* team sizes the tuner considers (0 with a given size) */
    int32_t cand_size[FI_BATCH_MAX_CANDIDATES];
    double cand_rate[FI_BATCH_MAX_CANDIDATES];  /* timed cell-attempts per ms, 0 if not timed */
    int32_t wg_threads;    /* threads per workgroup of the last launch: 256 (two workgroups per CU,
                              members up to 256 cells wide) or 512; 0 before any launch */
    int32_t reserved;
} fi_batch_info;

/* Creates nmem member contexts, member m as fi_create(&cfgs[m]). Members are set up, initialised
 * and read through fi_batch_member(m) with the single-context calls (fi_set_field, fi_set_ice_*,
 * fi_init, fi_integrator_init, fi_get_field, fi_get_stats); they are owned by the batch. Any
 * mix of member kinds is accepted, as the reference's ensemble loop runs any Simulation
 * (ext/FastIsostasyEnsembleKalmanProcessesExt.jl:114-122): LV-ELVA members share one ensemble
 * kernel, ELRA members another, and implicit-step members advance one after another through
 * their own integration; fi_batch_get_info then reports the largest group's team sizes and the
 * launches and device time of both ensemble kernels. */
int fi_create_batch(const fi_config* cfgs, int32_t nmem, fi_batch** out);
fi_ctx* fi_batch_member(fi_batch* b, int32_t m);
void fi_batch_destroy(fi_batch* b);
const char* fi_batch_last_error(const fi_batch* b);

/* Workgroups per member team. 0 (the default) times the candidate sizes (64 with 256-thread
 * workgroups, then 32, 16, ..., 1) on the first launches of fi_batch_integrate and keeps the
 * fastest, stopping once a smaller team is timed below 0.8 of the best rate: a trial of 64
 * attempts per member counts when at least 90% of its capped work ran; when windows are shorter
 * than that, each candidate is timed over a whole window. Ensembles whose members are at most 256
 * cells wide run workgroups of 256 threads, two per CU (FASTISO_BATCH_WG_THREADS=512 forces the
 * 512-thread build); fi_batch_info.wg_threads reports which ran. Results do not depend on the
 * team size or the workgroup width. */
int fi_batch_set_team_size(fi_batch* b, int32_t team_size);

/* Advances every member (after fi_integrator_init) to t_target with its own adaptive BS3 steps.
 * Member output times are not used. fsal = 1 then re-evaluates each member's first-same-as-last
 * derivative at t_target, as fi_integrate does after an output callback (src/simulation.jl:
 * 194-225): a member then matches a single context integrated with t_target among its output
 * times. */
int fi_batch_integrate(fi_batch* b, double t_target, int32_t fsal);

/* As fi_batch_integrate with one target time per member (t_targets[m] for member m). */
int fi_batch_integrate_each(fi_batch* b, const double* t_targets, int32_t fsal);

/* FI_TUNE_FFT_PLANS of fi_autotune for an ensemble: one member of each kind and grid is tuned
 * on its own context (after fi_integrator_init), the members of the same kind and grid take its
 * plans. */
int fi_batch_autotune_fft(fi_batch* b, int32_t attempts);

/* fi_batch_integrate_each without device allocation. No counterpart in the reference: it is for
 * callers that cannot allocate device memory while they run (tight serving loops, harnesses that
 * forbid hipMalloc). fi_batch_integrate_each may allocate and grow per-launch buffers (member
 * parameter blocks, barrier regions, queue); this variant takes them from one static device
 * array per GPU of FI_NOALLOC_SCRATCH_MAX_MEMBERS members and FI_NOALLOC_SCRATCH_MAX_WORKGROUPS
 * co-resident workgroups (about 2 MB, held until the process ends) and returns FI_ERR_INVALID
 * beyond those sizes. Set-up (fi_create_batch, fi_init, fi_integrator_init) still allocates.
 * Results equal fi_batch_integrate_each's bit for bit. Every call on one GPU shares that array:
 * calls that overlap on one GPU (from different host threads) race. */
#define FI_NOALLOC_SCRATCH_MAX_MEMBERS 1024
#define FI_NOALLOC_SCRATCH_MAX_WORKGROUPS 1024
int fi_batch_integrate_each_noalloc(fi_batch* b, const double* t_targets, int32_t fsal);
int fi_batch_get_info(const fi_batch* b, fi_batch_info* out);

#ifdef __cplusplus
}
#endif
#endif /* FASTISO_H */
