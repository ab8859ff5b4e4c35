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
