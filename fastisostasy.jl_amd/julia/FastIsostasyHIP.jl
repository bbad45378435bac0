# FastIsostasyHIP -- ccall shim that runs FastIsostasy.jl's LV-ELVA hot path on the MI355X engine
# (libfastiso_hip.so, include/fastiso.h). Julia is not available where this shim was written: it
# is the binding a maintainer adds next to the reference package, not code exercised by the tests.
#
# Usage (mirrors the reference API):
#     using FastIsostasy, FastIsostasyHIP
#     sim = Simulation(domain, bcs, sealevel, solidearth, t_span; nout = nout)   # reference ctor
#     hsim = HipSimulation(sim)            # uploads the fields the hot path reads
#     run!(hsim)                           # = run!(sim) on the GPU; fills sim.nout
#     # or, for coupling: integ = init_integrator(hsim); step!(integ, dt, true)
#     # or, RHS only inside the reference's own OrdinaryDiffEq loop: update_diagnostics!(dudt, u, hsim, t)
module FastIsostasyHIP

using FastIsostasy
import FastIsostasy: run!, init_integrator, step!, update_diagnostics!

const libfi = get(ENV, "FASTISO_HIP_LIB", joinpath(@__DIR__, "..", "lib", "libfastiso_hip.so"))

# fi_config (include/fastiso.h), same field order and C layout
mutable struct FiConfig
    nx::Int32; ny::Int32
    Wx::Float64; Wy::Float64
    g::Float64; seconds_per_year::Float64; rho_ice::Float64; rho_water::Float64
    rho_seawater::Float64; r_pole::Float64; mE::Float64
    litho_poissonratio::Float64; rho_uppermantle::Float64; rho_litho::Float64
    lithosphere::Int32; mantle::Int32; lithosphere_column::Int32
    sea_surface::Int32; sealevel_load::Int32; volume_contribution::Int32
    density_contribution::Int32; adjustment_contribution::Int32
    bsl_z::Float64
    viscous_x_border::Float64; elastic_x_border::Float64; ssh_x_border::Float64
    dt_sparse::Float64; reltol::Float64; abstol::Float64
    device::Int32; reserved::Int32; dt_min::Float64
end
FiConfig() = FiConfig(zeros(Int32, 2)..., zeros(12)..., zeros(Int32, 8)..., zeros(7)..., Int32(0), Int32(0), 0.0)

struct FiStats
    t::Float64; dt::Float64
    n_rhs::Int64; n_accept::Int64; n_reject::Int64; count_sparse_updates::Int64; n_output::Int64
    V_af::Float64; V_den::Float64; V_pov::Float64; delta_V::Float64; z_bsl::Float64; last_EEst::Float64
    device_ms::Float64; n_launch::Int64; n_accept_timed::Int64
    A_ocean::Float64
end

# field ids of include/fastiso.h (FI_*), keyed by the reference's field names
const FI = (litho_rigidity = 1, effective_viscosity = 2, scaled_pseudodiff_inv = 3, maskactive = 4,
            green_elastic = 5, green_geoid = 6, ref_u = 7, ref_ue = 8, ref_dz_ss = 9, ref_z_b = 10,
            ref_H_ice = 11, distortion = 12, green_viscous = 13, tau = 14, litho_thickness = 15,
            u = 32, dudt = 33, ue = 34, dz_ss = 35, z_ss = 36, z_b = 37, H_ice = 38, H_af = 39,
            H_water = 40, maskgrounded = 41, maskocean = 42, canom_ice = 43, canom_seawater = 44,
            canom_litho = 45, canom_mantle = 46, canom_load = 47, canom_full = 48, u_eq = 49,
            u_x = 50, u_y = 51)

# FI_ABI_VERSION of include/fastiso.h: the layouts of FiConfig and FiStats above follow it
const FI_ABI_VERSION = 6

function __init__()
    v = ccall((:fi_abi_version, libfi), Int32, ())
    v == FI_ABI_VERSION || error("$(libfi): ABI version $v, this binding expects $(FI_ABI_VERSION)")
    name = unsafe_string(ccall((:fi_engine_name, libfi), Cstring, ()))
    name == "hip-gfx950" || error("$(libfi) is the $(name) engine, not hip-gfx950")
    return nothing
end

check(ctx, st) = st == 0 || error("fastiso HIP engine error $st: " *
    unsafe_string(ccall((:fi_last_error, libfi), Cstring, (Ptr{Cvoid},), ctx)))

# Timer history kept per context: the first TIMER_CAPACITY right-hand-side calls
const TIMER_CAPACITY = Int64(1) << 20

mutable struct HipSimulation
    sim::Any                 # the reference Simulation (host copy of setup and outputs)
    ctx::Ptr{Cvoid}
end

setfield!(ctx, id, A) = check(ctx, ccall((:fi_set_field, libfi), Cint,
    (Ptr{Cvoid}, Int32, Ptr{Float64}, Csize_t), ctx, id, Float64.(Array(A)), length(A)))

function getfield!(A::Matrix{Float64}, ctx, id)
    check(ctx, ccall((:fi_get_field, libfi), Cint, (Ptr{Cvoid}, Int32, Ptr{Float64}, Csize_t),
        ctx, id, A, length(A)))
    return A
end

function HipSimulation(sim)
    d, c, se, sl, bcs = sim.domain, sim.c, sim.solidearth, sim.sealevel, sim.bcs
    cfg = FiConfig()
    ccall((:fi_default_config, libfi), Cvoid, (Ref{FiConfig}, Int32, Int32, Float64, Float64),
        cfg, d.nx, d.ny, d.Wx, d.Wy)
    cfg.g, cfg.seconds_per_year, cfg.rho_ice = c.g, c.seconds_per_year, c.rho_ice
    cfg.rho_water, cfg.rho_seawater, cfg.r_pole, cfg.mE = c.rho_water, c.rho_seawater, c.r_pole, c.mE
    cfg.litho_poissonratio, cfg.rho_uppermantle, cfg.rho_litho =
        se.litho_poissonratio, se.rho_uppermantle, se.rho_litho
    cfg.lithosphere = se.lithosphere isa RigidLithosphere ? 0 :
        se.lithosphere isa LaterallyConstantLithosphere ? 1 : 2
    cfg.mantle = se.mantle isa RigidMantle ? 0 : se.mantle isa RelaxedMantle ? 1 : 2
    cfg.sea_surface = sl.surface isa LaterallyVariableSeaSurface ? 1 : sl.surface isa ImposedSeaSurface ? 2 : 0
    cfg.sealevel_load = sl.load isa InteractiveSealevelLoad ? 1 : 0
    cfg.lithosphere_column = se.lithosphere_column isa CompressibleLithosphereColumn ? 1 : 0
    cfg.bsl_z = sl.bsl.z
    cfg.viscous_x_border = bcs.viscous_displacement.x_border
    cfg.elastic_x_border = bcs.elastic_displacement.x_border
    cfg.ssh_x_border = bcs.sea_surface_perturbation.x_border
    cfg.dt_sparse = sim.opts.dt_sparse_diagnostics
    cfg.reltol = sim.opts.diffeq.reltol
    sim.opts.diffeq.dt_min isa Real && (cfg.dt_min = sim.opts.diffeq.dt_min)
    ctxr = Ref{Ptr{Cvoid}}(C_NULL)
    st = ccall((:fi_create, libfi), Cint, (Ref{FiConfig}, Ref{Ptr{Cvoid}}), cfg, ctxr)
    st == 0 || error("fi_create failed: $st")
    ctx = ctxr[]
    setfield!(ctx, FI.litho_rigidity, se.litho_rigidity)
    setfield!(ctx, FI.effective_viscosity, se.effective_viscosity)
    setfield!(ctx, FI.scaled_pseudodiff_inv, se.scaled_pseudodiff_inv)
    setfield!(ctx, FI.maskactive, se.maskactive)
    setfield!(ctx, FI.green_elastic, sim.tools.elastic_convo.kernel)
    setfield!(ctx, FI.green_geoid, sim.tools.dz_ss_convo.kernel)
    if se.mantle isa RelaxedMantle   # ELRA: the viscous Green kernel of GIATools and tau
        setfield!(ctx, FI.green_viscous, sim.tools.viscous_convo.kernel)
        setfield!(ctx, FI.tau, se.tau .* ones(d.nx, d.ny))
    end
    d.correct_distortion && setfield!(ctx, FI.distortion, d.K)
    setfield!(ctx, FI.litho_thickness, se.litho_thickness)   # read by the u_x / u_y outputs
    setfield!(ctx, FI.ref_u, sim.ref.u)
    setfield!(ctx, FI.ref_ue, sim.ref.ue)
    setfield!(ctx, FI.ref_z_b, sim.ref.z_b)
    setfield!(ctx, FI.ref_dz_ss, sim.ref.z_ss .- sl.bsl.ref.z)
    setfield!(ctx, FI.ref_H_ice, sim.ref.H_ice)
    it = bcs.ice_thickness
    if it isa TimeInterpolatedIceThickness
        H = reduce(vcat, vec.(Array.(it.H_vec)))
        check(ctx, ccall((:fi_set_ice_snapshots, libfi), Cint,
            (Ptr{Cvoid}, Ptr{Float64}, Int32, Ptr{Float64}, Int32), ctx, Float64.(it.t_vec),
            length(it.t_vec), H, Int32(it.H_itp.flat_bc)))
    end
    if sl.surface isa ImposedSeaSurface        # TimeInterpolation2D (src/interpolations.jl:12-26)
        itp = sl.surface.dz_ss_itp
        X = reduce(vcat, vec.(Array.(itp.X)))
        check(ctx, ccall((:fi_set_sea_surface_snapshots, libfi), Cint,
            (Ptr{Cvoid}, Ptr{Float64}, Int32, Ptr{Float64}, Int32), ctx, Float64.(itp.t),
            length(itp.t), X, Int32(itp.flat_bc)))
    end
    check(ctx, ccall((:fi_set_timer_history, libfi), Cint, (Ptr{Cvoid}, Int64), ctx, TIMER_CAPACITY))
    tout = Float64.(sort(vcat(sim.nout.t, sim.ncout.t)))
    check(ctx, ccall((:fi_set_output_times, libfi), Cint, (Ptr{Cvoid}, Ptr{Float64}, Int32),
        ctx, tout, length(tout)))
    check(ctx, ccall((:fi_init, libfi), Cint, (Ptr{Cvoid}, Float64, Float64), ctx,
        sim.timer.t_span[1], sim.timer.t_span[2]))
    h = HipSimulation(sim, ctx)
    finalizer(x -> ccall((:fi_destroy, libfi), Cvoid, (Ptr{Cvoid},), x.ctx), h)
    return h
end

# output callback: NativeOutput snapshots (src/io.jl:349-353)
function _output(user::Ptr{Cvoid}, t::Float64, k::Int32)::Cvoid
    h = unsafe_pointer_to_objref(user)::HipSimulation
    nout = h.sim.nout
    for var in nout.vars
        A = zeros(Float64, h.sim.domain.nx, h.sim.domain.ny)
        push!(nout.vals[var], getfield!(A, h.ctx, getproperty(FI, var)))
    end
    nout.k += 1
    return nothing
end

function run!(h::HipSimulation)
    check(h.ctx, ccall((:fi_integrator_init, libfi), Cint, (Ptr{Cvoid},), h.ctx))
    cb = @cfunction(_output, Cvoid, (Ptr{Cvoid}, Float64, Int32))
    GC.@preserve h check(h.ctx, ccall((:fi_integrate, libfi), Cint,
        (Ptr{Cvoid}, Float64, Ptr{Cvoid}, Ptr{Cvoid}), h.ctx, h.sim.timer.t_span[2], cb,
        pointer_from_objref(h)))
    read_timer!(h)
    return nothing
end

# sim.timer.t_vec / t_computation as t_computation! and run! leave them (src/simulation.jl:46-52,260)
function read_timer!(h::HipSimulation)
    t = Vector{Float64}(undef, TIMER_CAPACITY)
    s = similar(t)
    n = Ref{Int64}(0)
    check(h.ctx, ccall((:fi_get_timer_history, libfi), Cint,
        (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Int64, Ref{Int64}), h.ctx, t, s, TIMER_CAPACITY, n))
    m = min(n[], TIMER_CAPACITY)
    tm = h.sim.timer
    empty!(tm.t_vec); append!(tm.t_vec, t[1:m])
    empty!(tm.t_computation); append!(tm.t_computation, s[1:m])
    return nothing
end

struct HipIntegrator
    h::HipSimulation
end

function init_integrator(h::HipSimulation)
    check(h.ctx, ccall((:fi_integrator_init, libfi), Cint, (Ptr{Cvoid},), h.ctx))
    return HipIntegrator(h)
end

function step!(integ::HipIntegrator, dt, force_dt)
    force_dt || error("step! is only defined with force_dt = true")
    s = Ref{FiStats}()
    check(integ.h.ctx, ccall((:fi_get_stats, libfi), Cint, (Ptr{Cvoid}, Ref{FiStats}), integ.h.ctx, s))
    cb = @cfunction(_output, Cvoid, (Ptr{Cvoid}, Float64, Int32))
    GC.@preserve integ check(integ.h.ctx, ccall((:fi_integrate, libfi), Cint,
        (Ptr{Cvoid}, Float64, Ptr{Cvoid}, Ptr{Cvoid}), integ.h.ctx, s[].t + dt, cb,
        pointer_from_objref(integ.h)))
    return nothing
end

# the ODE right-hand side on host arrays (src/simulation.jl:335-393); u is mutated by the BC
function update_diagnostics!(dudt::Matrix{Float64}, u::Matrix{Float64}, h::HipSimulation, t)
    check(h.ctx, ccall((:fi_rhs, libfi), Cint, (Ptr{Cvoid}, Float64, Ptr{Float64}, Ptr{Float64}),
        h.ctx, Float64(t), u, dudt))
    return nothing
end

# the same on device arrays (a ROCArray's pointer), one call per solver stage
function update_diagnostics_device!(dudt_ptr::Ptr{Float64}, u_ptr::Ptr{Float64}, h::HipSimulation, t)
    check(h.ctx, ccall((:fi_rhs_device, libfi), Cint, (Ptr{Cvoid}, Float64, Ptr{Cvoid}, Ptr{Cvoid}),
        h.ctx, Float64(t), u_ptr, dudt_ptr))
    return nothing
end

export HipSimulation, update_diagnostics_device!

end # module
