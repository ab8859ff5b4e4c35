"""
    ShaktiHIP

Shakti.jl plug-in that solves the hydraulic-head system on an MI355X through
`libshakti_hip.so` (C ABI in `include/shakti_hip.h`).  It adds new subtypes of
Shakti's `AbstractLinearSystem` / `AbstractIterativeSolver` and the two solve
methods the reference dispatches on (src/linear_solver.jl:915-945), so
`PicardSolver`, `EllipticHeadScheme` / `ParabolicHeadScheme` and
`Simulation` / `run!` drive it unchanged:

    ls = ShaktiHIP.HIPPCGSolver(grid; chebyshev_degree = 4, rtol = 1e-11, atol = 0.0)
    ps = PicardSolver(500, 1e-6, ls, grid)
    sim = Simulation(grid, state, tsteps, dt, p, "implicit", String[], mi, sl; ps = ps)
    run!(sim)

Each solve uploads the assembly inputs of the host State (mask, zb, h, K,
A_visc, N, b, mdot, beta, abs_ub, ieb), assembles and runs the PCG on the GPU,
and downloads `h`, exactly the contract of `solve_elliptic_linear_system!`
(overwrite `s.h`, everything else untouched).  `ls.ws.stats.niter` mirrors
Krylov's `ws.stats.niter` (test/linear_solver_test.jl:173-189); `ls.precond`
holds the Chebyshev bounds of the last solve.

A breakdown of the Krylov recurrence (r'z < 0: the preconditioner or the
operator is not positive definite) raises an `ErrorException`, as the
reference's `sqrt(r'z)` does inside `cg!`, and leaves `s.h` unchanged.

`HIPSimulation(sim)` keeps the whole State on the GPU instead: its `step!`
and `run!` advance it with whole-step library calls and refresh the host State
only for observers, so nothing crosses PCIe inside a step.

The learned form: `HIPUNet(params)` uploads a trained U-Net (the Float32
parameter vector, e.g. `NPZ.npzread(path)["params"]` of
`shakti_hip/weights/unet_16_32_64_128.npz`), and
`HIPPCGSolver(g; multigrid = true, form = :learned, network = net)` runs FP64
CG from the network's warm start.
"""
module ShaktiHIP

using Shakti
import Shakti: solve_elliptic_linear_system!, solve_parabolic_linear_system!

const libshk = joinpath(@__DIR__, "..", "shakti_hip", "libshakti_hip.so")

# ---- C structs (layout of include/shakti_hip.h) -------------------------------

struct ShkGrid
    nx::Int32; ny::Int32; lx::Float64; ly::Float64; dx::Float64; dy::Float64; dx2::Float64; dy2::Float64
end
struct ShkParams
    rho_w::Float64; rho_sw::Float64; rho_i::Float64; g::Float64; nu::Float64; n::Float64; omega::Float64
    L::Float64; br::Float64; lr::Float64; ct::Float64; cw::Float64; p_atm::Float64; b_min::Float64
    b_max::Float64; e_v::Float64
end
struct ShkModel
    kfs::Int32; gap_scheme::Int32; sliding_law::Int32; sliding_C::Float64; sliding_v0::Float64
end
const SHK_ABI_VERSION = 2

struct ShkSolverCfg
    precond::Int32; cheb_degree::Int32; cheb_nsteps::Int32; form::Int32; itmax::Int32; check_every::Int32
    atol::Float64; rtol::Float64; inner_rtol::Float64
    cheb_bounds::Int32; flags::Int32
end
mutable struct ShkSolveStats
    niter::Int32; solved::Int32; zero_curvature::Int32; bounds_warning::Int32; refinements::Int32; fallback::Int32
    rnorm0::Float64; rnorm::Float64; lambda_min::Float64; lambda_max::Float64
    tired::Int32; breakdown::Int32; breakdown_value::Float64; lz_lambda_min::Float64; lz_lambda_max::Float64
    gershgorin::Float64; true_residual::Float64
    ShkSolveStats() = new(0, 0, 0, 0, 0, 0, 0.0, 0.0, 1.0, 1.0, 0, 0, 0.0, NaN, NaN, NaN, NaN)
end
struct ShkUnetCfg
    max_batch::Int32; max_h::Int32; max_w::Int32; leaky_slope::Float32; loss_kind::Int32
end

# field ids (enum shk_field)
const SHK_H, SHK_B, SHK_MDOT, SHK_K, SHK_ZB, SHK_IEB, SHK_A_VISC, SHK_N, SHK_MASK = 0, 3, 6, 11, 13, 16, 18, 19, 20
const SHK_BETA, SHK_ABS_UB = 4, 5

function check(rc, ctx, what)
    if rc != 0
        msg = ctx == C_NULL ? "" : unsafe_string(ccall((:shk_last_error, libshk), Cstring, (Ptr{Cvoid},), ctx))
        error("$what failed ($rc): $msg")   # the reference's ErrorException contract
    end
end

# ---- the new linear-system / solver subtypes ----------------------------------

"The device-resident matrix-free system (aP, aE, aN, rhs in the symmetric face layout)."
struct HIPMatrixFreeLinearSystem <: Shakti.AbstractLinearSystem end

"Krylov's `ws.stats` of the last solve, plus the FP64 residual of the returned head."
mutable struct HIPStats
    niter::Int
    solved::Bool
    zero_curvature::Bool
    true_residual::Float64
end
HIPStats() = HIPStats(0, false, false, NaN)

"The solver's workspace view: `ls.ws.stats` as Krylov's `ws.stats`."
struct HIPWorkspace
    stats::HIPStats
end

"The Chebyshev bounds of the last solve (`ls.precond`, as the reference's ChebyshevPreconditioner)."
mutable struct HIPChebyshevBounds
    degree::Int
    lambda_min::Float64
    lambda_max::Float64
end

"""
    HIPUNet(params::Vector{Float32}; max_h, max_w, slope = 0.1f0)

A trained U-Net on the device (shk_unet_create + shk_unet_set_params,
include/shakti_learned.h), with buffers for one image of up to
`max_h` x `max_w` cells (rounded up to multiples of 32 and 64: the grid's
ny and nx).
"""
mutable struct HIPUNet
    handle::Ptr{Cvoid}
end

function HIPUNet(params::Vector{Float32}; max_h::Integer, max_w::Integer, slope::Real = 0.1)
    H8 = Int32(cld(max_h, 32) * 32)
    W8 = Int32(cld(max_w, 64) * 64)
    n = ccall((:shk_unet_num_params, libshk), Int64, ())
    length(params) == n || error("HIPUNet: expected $n parameters, got $(length(params))")
    out = Ref{Ptr{Cvoid}}(C_NULL)
    rc = ccall((:shk_unet_create, libshk), Cint, (Ref{ShkUnetCfg}, Ref{Ptr{Cvoid}}),
               Ref(ShkUnetCfg(1, H8, W8, Float32(slope), 0)), out)
    rc == 0 || error("shk_unet_create failed ($rc)")
    net = HIPUNet(out[])
    finalizer(x -> x.handle != C_NULL && ccall((:shk_unet_destroy, libshk), Cint, (Ptr{Cvoid},), x.handle), net)
    rc = GC.@preserve params ccall((:shk_unet_set_params, libshk), Cint, (Ptr{Cvoid}, Ptr{Float32}, Int64),
                                   net.handle, params, n)
    rc == 0 || error("shk_unet_set_params failed ($rc): " *
                     unsafe_string(ccall((:shk_unet_last_error, libshk), Cstring, (Ptr{Cvoid},), net.handle)))
    return net
end

"""
    HIPPCGSolver(g; chebyshev_degree = nothing, chebyshev_nsteps_estimate = 15,
                 atol = sqrt(eps()), rtol = sqrt(eps()), itmax = 0, check_every = 8,
                 form = :fp64, inner_rtol = 1e-3, multigrid = false,
                 chebyshev_bounds = :gershgorin, network = nothing)

CG on the GPU with Jacobi (`chebyshev_degree === nothing`) or Chebyshev
preconditioning; same options and errors as `CGIterativeSolver(g,
MatrixFreeLinearSystem; ...)` (src/linear_solver.jl:832-843).

`form = :fp32_refine` solves in FP32 inside an FP64 iterative-refinement loop
that stops on the same FP64 residual test; `multigrid = true` (no Chebyshev
degree) preconditions with the FP32 geometric-multigrid V-cycle: the FP64 CG
itself with `:fp64` (the fastest form), the FP32 inner solves with
`:fp32_refine`.  `form = :learned` with `network = HIPUNet(...)` warm-starts
FP64 CG from the network's correction (multigrid or Chebyshev).

`chebyshev_bounds = :reference` keeps the reference's bounds (0.9 / 1.1 x the
extreme Ritz values of the estimate); the default `:gershgorin` takes
lambda_max from the Gershgorin bound of D^-1 A, which keeps the Chebyshev
polynomial positive on the whole spectrum (a stated deviation).
"""
mutable struct HIPPCGSolver <: Shakti.AbstractIterativeSolver
    ctx::Ptr{Cvoid}
    grid::Shakti.Grid
    cfg::ShkSolverCfg
    ws::HIPWorkspace
    precond::Union{Nothing, HIPChebyshevBounds}
    network::Union{Nothing, HIPUNet}
end

# `ls.stats` as before `ls.ws.stats` existed
Base.getproperty(ls::HIPPCGSolver, f::Symbol) = f === :stats ? getfield(ls, :ws).stats : getfield(ls, f)

function HIPPCGSolver(g::Shakti.Grid; chebyshev_degree::Union{Nothing, Int} = nothing,
                      chebyshev_nsteps_estimate::Int = 15, amg::Bool = false, atol = sqrt(eps(Float64)),
                      rtol = sqrt(eps(Float64)), itmax::Int = 0, check_every::Int = 8,
                      form::Symbol = :fp64, inner_rtol = 1e-3, multigrid::Bool = false,
                      chebyshev_bounds::Symbol = :gershgorin, network::Union{Nothing, HIPUNet} = nothing)
    amg && error("AMGPreconditioner is CPU/SparseMatrixCSC-only; use chebyshev_degree with HIPPCGSolver.")
    form in (:fp64, :fp32_refine, :learned) || error("HIPPCGSolver: unknown form $form")
    (form === :learned) == (network !== nothing) ||
        error("HIPPCGSolver: form = :learned takes the trained network as network = HIPUNet(...), and only it does")
    chebyshev_bounds in (:gershgorin, :reference) ||
        error("HIPPCGSolver: chebyshev_bounds must be :gershgorin or :reference (got $chebyshev_bounds)")
    multigrid && chebyshev_degree !== nothing &&
        error("HIPPCGSolver: multigrid = true replaces chebyshev_degree: give one of them")
    chebyshev_degree !== nothing && chebyshev_degree < 1 &&
        error("ChebyshevPreconditioner: degree must be >= 1 (got $chebyshev_degree)")
    chebyshev_nsteps_estimate >= 2 ||
        error("ChebyshevPreconditioner: nsteps_estimate must be >= 2 (got $chebyshev_nsteps_estimate)")
    precond = multigrid ? 2 : (chebyshev_degree === nothing ? 0 : 1)
    ccall((:shk_abi_version, libshk), Int32, ()) == SHK_ABI_VERSION ||
        error("libshakti_hip.so has another ABI version than this binding ($SHK_ABI_VERSION); rebuild it")
    formcode = form === :fp32_refine ? 1 : (form === :learned ? 2 : 0)
    cfg = ShkSolverCfg(precond, something(chebyshev_degree, 1), chebyshev_nsteps_estimate, formcode, itmax,
                       check_every, atol, rtol, inner_rtol, chebyshev_bounds === :reference ? 1 : 0, 0)
    sg = Ref(ShkGrid(g.nx, g.ny, g.lx, g.ly, g.dx, g.dy, g.dx2, g.dy2))
    sp = Ref(params_c(Shakti.ModelParameters()))
    sm = Ref(ShkModel(0, 0, 0, 0.5, 0.0))
    sc = Ref(cfg)
    out = Ref{Ptr{Cvoid}}(C_NULL)
    check(ccall((:shk_create, libshk), Cint,
                (Ref{ShkGrid}, Ref{ShkParams}, Ref{ShkModel}, Ref{ShkSolverCfg}, Ref{Ptr{Cvoid}}),
                sg, sp, sm, sc, out), C_NULL, "shk_create")
    pc = chebyshev_degree === nothing ? nothing : HIPChebyshevBounds(chebyshev_degree, 1.0, 1.0)
    ls = HIPPCGSolver(out[], g, cfg, HIPWorkspace(HIPStats()), pc, network)
    finalizer(x -> x.ctx != C_NULL && ccall((:shk_destroy, libshk), Cint, (Ptr{Cvoid},), x.ctx), ls)
    if network !== nothing
        # the context holds the network's handle; ls.network keeps it alive
        check(ccall((:shk_attach_unet, libshk), Cint, (Ptr{Cvoid}, Ptr{Cvoid}), ls.ctx, network.handle), ls.ctx,
              "shk_attach_unet")
    end
    return ls
end

params_c(p) = ShkParams(p.rho_w, p.rho_sw, p.rho_i, p.g, p.nu, p.n, p.omega, p.L, p.br, p.lr, p.ct, p.cw,
                        p.p_atm, p.b_min, p.b_max, p.e_v)
kfs_code(::Shakti.Arithmetic) = Int32(0)
kfs_code(::Shakti.Harmonic) = Int32(1)

function upload(ls, field, a::AbstractArray{Float64})
    h = Array(a)  # device arrays of other backends come back to the host first
    GC.@preserve h check(ccall((:shk_upload, libshk), Cint, (Ptr{Cvoid}, Int32, Ptr{Float64}, Csize_t),
                               ls.ctx, field, h, length(h)), ls.ctx, "shk_upload")
end

function download!(a::AbstractArray{Float64}, ls, field)
    h = Array{Float64}(undef, size(a))
    GC.@preserve h check(ccall((:shk_download, libshk), Cint, (Ptr{Cvoid}, Int32, Ptr{Float64}, Csize_t),
                               ls.ctx, field, h, length(h)), ls.ctx, "shk_download")
    copyto!(a, h)
end

function sync_inputs!(ls::HIPPCGSolver, s, p, kfs)
    check(ccall((:shk_set_params, libshk), Cint, (Ptr{Cvoid}, Ref{ShkParams}), ls.ctx, Ref(params_c(p))),
          ls.ctx, "shk_set_params")
    check(ccall((:shk_set_model, libshk), Cint, (Ptr{Cvoid}, Ref{ShkModel}), ls.ctx,
                Ref(ShkModel(kfs_code(kfs), 0, 0, 0.5, 0.0))), ls.ctx, "shk_set_model")
    check(ccall((:shk_set_solver_cfg, libshk), Cint, (Ptr{Cvoid}, Ref{ShkSolverCfg}), ls.ctx, Ref(ls.cfg)),
          ls.ctx, "shk_set_solver_cfg")
    for (f, a) in ((SHK_MASK, s.mask), (SHK_ZB, s.zb), (SHK_H, s.h), (SHK_K, s.K), (SHK_A_VISC, s.A_visc),
                   (SHK_N, s.N), (SHK_B, s.b), (SHK_MDOT, s.mdot), (SHK_BETA, s.beta), (SHK_ABS_UB, s.abs_ub),
                   (SHK_IEB, s.ieb))
        upload(ls, f, a)
    end
end

function record!(ls::HIPPCGSolver, st::ShkSolveStats)
    s = ls.ws.stats
    s.niter, s.solved, s.zero_curvature, s.true_residual = st.niter, st.solved != 0, st.zero_curvature != 0,
                                                           st.true_residual
    if ls.precond !== nothing
        ls.precond.lambda_min, ls.precond.lambda_max = st.lambda_min, st.lambda_max
        st.bounds_warning != 0 && @warn "ChebyshevPreconditioner: eigenvalue bound estimate was invalid " *
                                        "(lambda_min=$(st.lz_lambda_min), lambda_max=$(st.lz_lambda_max)); " *
                                        "reusing previous bounds for this solve" maxlog = 10
    end
    return nothing
end

"solve_elliptic_linear_system! for the GPU solver (same contract as src/linear_solver.jl:915-925)."
function solve_elliptic_linear_system!(ls::HIPPCGSolver, s::Shakti.State, g::Shakti.Grid,
                                       p::Shakti.ModelParameters, kfs::Shakti.AbstractKFaceScheme)
    sync_inputs!(ls, s, p, kfs)
    st = ShkSolveStats()
    rc = ccall((:shk_solve_elliptic, libshk), Cint, (Ptr{Cvoid}, Ref{ShkSolveStats}), ls.ctx, st)
    record!(ls, st)
    check(rc, ls.ctx, "shk_solve_elliptic")
    download!(s.h, ls, SHK_H)
    return nothing
end

# ---- the whole timestep on a device-resident State ------------------------------

# every State field in the order of enum shk_field (src/state.jl:13-57)
const STATE_FIELDS = (:h, :pw, :po, :b, :beta, :abs_ub, :mdot, :shear, :potential, :sensible, :Re, :K, :G,
                      :zb, :zs, :H, :ieb, :lambda, :A_visc, :N, :mask, :dhdx, :q_x, :Re_x, :b_x, :ub_x,
                      :taub_x, :dpwdx, :valid_x, :dhdy, :q_y, :Re_y, :b_y, :ub_y, :taub_y, :dpwdy, :valid_y)

gap_code(::Shakti.ImplicitGapScheme) = Int32(0)
gap_code(::Shakti.ExplicitGapScheme) = Int32(1)
sliding_model(sl::Shakti.RegularizedCoulombSlidingLaw) = (Int32(0), Float64(sl.C), 0.0, false)
sliding_model(::Shakti.PrescribedSlidingLaw) = (Int32(1), 0.5, 0.0, false)
sliding_model(::Shakti.LinearSlidingLaw) = (Int32(2), 0.5, 0.0, true)
sliding_model(::Shakti.RegularizedCoulombFieldSlidingLaw) = (Int32(3), 0.5, 0.0, true)
sliding_model(sl::Shakti.RegularizedCoulombV0SlidingLaw) = (Int32(4), 0.5, Float64(sl.v0), true)
relax_alpha(::Shakti.NoHeadRelaxation) = 0.0
relax_alpha(hr::Shakti.UnderHeadRelaxation) = Float64(hr.alpha)

"""
    HIPSimulation(sim::Shakti.Simulation; solver = HIPPCGSolver(sim.grid; multigrid = true, rtol = 1e-10,
                  atol = 0.0, check_every = 32), sliding_C = nothing, moulin = nothing)

`sim` with its State kept on the GPU: `step!` / `run!` below advance it with
whole-step library calls (`shk_step`: update_ieb!, step_h!, step_b!,
src/run.jl:82-164), so no field crosses PCIe inside a step.  The host State
`sim.state` is uploaded once here and refreshed from the device only where
something reads it: an observer that records (src/observer.jl), the verbose
diagnostics line, and `sync_host!`.  The head scheme (`EllipticHeadScheme`
with its `PicardSolver`'s iters, tol, check_every and relaxation, or
`ParabolicHeadScheme`, which takes the parabolic solve of `shk_step` through
`p.e_v`), the gap scheme, the K_face scheme and the sliding law are read from
`sim`; `ps.converged` / `ps.last_iter` are set as Picard_loop! sets them.

The field sliding laws store only their face values (`Cx`, `Cx2`); give
their cell-centred coefficient as `sliding_C` (the library staggers it as
their constructors do, src/sliding_law.jl:73-90, 284-301, 436-452).  The
melt input is refreshed on the host by the reference's own `update_ieb!`,
which writes one uniform value (src/melt_input.jl:86-94), and the library
writes that value on the device; `moulin` (an nx x ny field of point-moulin
input) is added on top of it every step (no reference counterpart).
"""
mutable struct HIPSimulation
    sim::Shakti.Simulation
    ls::HIPPCGSolver
    moulin::Bool
end

function HIPSimulation(sim::Shakti.Simulation;
                       solver::HIPPCGSolver = HIPPCGSolver(sim.grid; multigrid = true, rtol = 1e-10, atol = 0.0,
                                                           check_every = 32),
                       sliding_C::Union{Nothing, AbstractMatrix} = nothing,
                       moulin::Union{Nothing, AbstractMatrix} = nothing)
    ls, s = solver, sim.state
    code, C, v0, field = sliding_model(sim.sl)
    field && sliding_C === nothing &&
        error("HIPSimulation: $(typeof(sim.sl)) needs its cell-centred coefficient as sliding_C")
    check(ccall((:shk_set_params, libshk), Cint, (Ptr{Cvoid}, Ref{ShkParams}), ls.ctx, Ref(params_c(sim.p))),
          ls.ctx, "shk_set_params")
    check(ccall((:shk_set_model, libshk), Cint, (Ptr{Cvoid}, Ref{ShkModel}), ls.ctx,
                Ref(ShkModel(kfs_code(sim.kfs), gap_code(sim.gs), code, C, v0))), ls.ctx, "shk_set_model")
    check(ccall((:shk_set_solver_cfg, libshk), Cint, (Ptr{Cvoid}, Ref{ShkSolverCfg}), ls.ctx, Ref(ls.cfg)),
          ls.ctx, "shk_set_solver_cfg")
    for (i, f) in enumerate(STATE_FIELDS)
        upload(ls, Int32(i - 1), getfield(s, f))
    end
    if field
        Ch = Array{Float64}(sliding_C)
        GC.@preserve Ch check(ccall((:shk_set_sliding_coefficient, libshk), Cint,
                                    (Ptr{Cvoid}, Ptr{Float64}, Csize_t), ls.ctx, Ch, length(Ch)),
                              ls.ctx, "shk_set_sliding_coefficient")
    end
    if moulin !== nothing
        M = Array{Float64}(moulin)
        GC.@preserve M check(ccall((:shk_set_moulin_field, libshk), Cint, (Ptr{Cvoid}, Ptr{Float64}, Csize_t),
                                   ls.ctx, M, length(M)), ls.ctx, "shk_set_moulin_field")
    end
    return HIPSimulation(sim, ls, moulin !== nothing)
end

"Download every State field of the device into `h.sim.state`."
function sync_host!(h::HIPSimulation)
    for (i, f) in enumerate(STATE_FIELDS)
        download!(getfield(h.sim.state, f), h.ls, Int32(i - 1))
    end
    return h.sim.state
end

picard_args(hs::Shakti.EllipticHeadScheme) = (Int32(hs.ps.iters), Float64(hs.ps.tol), Int32(hs.ps.check_every),
                                              relax_alpha(hs.ps.hr))
picard_args(::Shakti.ParabolicHeadScheme) = (Int32(1), 0.0, Int32(1), 0.0)
set_picard!(hs::Shakti.EllipticHeadScheme, conv, last) = (hs.ps.converged = conv; hs.ps.last_iter = last)
set_picard!(::Shakti.ParabolicHeadScheme, conv, last) = nothing

"step! (src/run.jl:82-88) on the device: update_ieb!, step_h!, step_b! in one library call."
function Shakti.step!(h::HIPSimulation)
    sim = h.sim
    if sim.mi isa Shakti.ConstantMeltInput
        mode, uniform = h.moulin ? Int32(2) : Int32(0), 0.0
    else
        # the reference's update_ieb! writes one value everywhere; the device takes that value
        Shakti.update_ieb!(sim.mi, sim.state, sim.total_time[])
        mode, uniform = h.moulin ? Int32(2) : Int32(1), Float64(Array(sim.state.ieb)[1])
    end
    iters, tol, check_every, alpha = picard_args(sim.hs)
    conv, last, cg = Ref{Int32}(0), Ref{Int32}(0), Ref{Int64}(0)
    check(ccall((:shk_step, libshk), Cint,
                (Ptr{Cvoid}, Float64, Int32, Float64, Int32, Float64, Int32, Float64, Ref{Int32}, Ref{Int32},
                 Ref{Int64}), h.ls.ctx, Float64(sim.dt), mode, uniform, iters, tol, check_every, alpha, conv, last,
                cg), h.ls.ctx, "shk_step")
    set_picard!(sim.hs, conv[] != 0, Int(last[]))
    return h
end

observes(::Shakti.NoObserver) = false
observes(::Shakti.AbstractObserver) = true

"""
    run!(h::HIPSimulation; checkpoint_every, checkpoint_path, restart_path)

run! (src/run.jl:21-60) on the device State: the same loop, observer calls
and verbose line, with checkpoints written and read by the library
(shk_checkpoint_save / shk_checkpoint_load: the whole device State, written
asynchronously while the next steps run, atomically renamed, checksummed).
"""
function Shakti.run!(h::HIPSimulation; checkpoint_every::Union{Nothing, Int} = nothing,
                     checkpoint_path::Union{Nothing, String} = nothing,
                     restart_path::Union{Nothing, String} = nothing)
    sim, ctx = h.sim, h.ls.ctx
    (checkpoint_every === nothing) != (checkpoint_path === nothing) &&
        error("checkpoint_every and checkpoint_path must be given together")
    obs = observes(sim.observer)
    if restart_path === nothing
        sim.total_time[] = zero(sim.dt)
        Shakti.prepare!(sim.observer, sim.state)
        Shakti.observe!(sim.observer, sim.state, 0, sim.total_time[])
        start_t = 0
    else
        t_ref, tt_ref = Ref{Int64}(0), Ref{Float64}(0.0)
        check(ccall((:shk_checkpoint_load, libshk), Cint, (Ptr{Cvoid}, Cstring, Ref{Int64}, Ref{Float64}),
                    ctx, restart_path, t_ref, tt_ref), ctx, "shk_checkpoint_load")
        start_t = Int(t_ref[])
        sim.total_time[] = tt_ref[]
        sync_host!(h)
        Shakti.resume!(sim.observer, sim.state, start_t)
    end
    for t in (start_t + 1):sim.tsteps
        step_time = @elapsed Shakti.step!(h)
        sim.total_time[] += sim.dt
        obs && sync_host!(h)
        Shakti.observe!(sim.observer, sim.state, t, sim.total_time[])
        if checkpoint_every !== nothing && t % checkpoint_every == 0
            check(ccall((:shk_checkpoint_save, libshk), Cint, (Ptr{Cvoid}, Cstring, Int64, Float64, Int32),
                        ctx, checkpoint_path, t, sim.total_time[], 1), ctx, "shk_checkpoint_save")
        end
        if sim.verbose
            converged, last_iter = Shakti.picard_status(sim.hs)
            println("$t / $(sim.tsteps) completed in $(round(step_time; digits = 4))s. Picard converged: $converged in $last_iter iterations")
            if converged === false
                s = obs ? sim.state : sync_host!(h)
                println("  diagnostics: N=$(extrema(Array(s.N))) Re=$(extrema(Array(s.Re))) b=$(extrema(Array(s.b))) h=$(extrema(Array(s.h)))")
            end
            flush(stdout)
        end
    end
    checkpoint_every !== nothing &&
        check(ccall((:shk_checkpoint_wait, libshk), Cint, (Ptr{Cvoid},), ctx), ctx, "shk_checkpoint_wait")
    Shakti.finalize!(sim.observer, sim.state)
    return h
end

"solve_parabolic_linear_system! for the GPU solver (src/linear_solver.jl:935-945)."
function solve_parabolic_linear_system!(ls::HIPPCGSolver, s::Shakti.State, g::Shakti.Grid,
                                        p::Shakti.ModelParameters, kfs::Shakti.AbstractKFaceScheme, dt)
    sync_inputs!(ls, s, p, kfs)
    st = ShkSolveStats()
    rc = ccall((:shk_solve_parabolic, libshk), Cint, (Ptr{Cvoid}, Float64, Ref{ShkSolveStats}), ls.ctx,
               Float64(dt), st)
    record!(ls, st)
    check(rc, ls.ctx, "shk_solve_parabolic")
    download!(s.h, ls, SHK_H)
    return nothing
end

end # module
