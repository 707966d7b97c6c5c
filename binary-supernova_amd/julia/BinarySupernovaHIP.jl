# BinarySupernovaHIP.jl -- Julia bindings of libbsn_hip (include/bsn.h).
#
# The drop-in for the reference's uniform-grid and FMR entry points on AMD
# MI355X.  Methods keep the reference's names, positional arguments and
# keywords (src/euler3d.jl:33-260, src/rk3.jl:14, src/fmr3d.jl:451,616,
# src/fmr_gravity.jl:146, src/gravity_bh.jl:58, src/sinks.jl:32,
# src/self_gravity.jl:35):
#
#   * on a `HIPGrid` (state resident in GPU memory, the analogue of a
#     CuArray U) every call runs on the device without host transfers;
#   * on a plain `Array{Float64,4}` the `*_host` C entry points upload,
#     run and download, mutating U in place like the reference.
#
# Not exercised in this repository's test suite: the build machine has no
# Julia toolchain.  The C ABI underneath is tested through Python ctypes.
module BinarySupernovaHIP

export HIPGrid, euler3d_step!, euler3d_rhs!, cfl_dt_3d, apply_floors_3d!,
       fill_ghost_3d_outflow!, fill_ghost_3d_periodic!, rk3_step!, HIPFMR,
       fmr3d_step!, cfl_dt_fmr3d, fmr_run!, fmr_autotune!, fmr_set_mode!, set_phi!, download,
       solve_fmr_poisson, restrict_density_to_coarse!, _prolong_scalar_trilinear!,
       fmr_set_self_gravity!,
       add_bh_gravity_source!,
       add_sink_sources!, set_point_masses!, set_sinks!, solve_poisson_isolated,
       set_self_gravity!, gas_force_on_bh, accrete!, gas_diagnostics, gas_energy_total,
       gas_kinetic_total, gas_momentum_total, gas_angular_momentum_total, bound_gas_mass,
       write_snapshot, read_snapshot!, save_checkpoint_uniform, load_checkpoint_uniform!,
       save_checkpoint, load_checkpoint!, init_trajectory_file, append_trajectory,
       read_trajectory

const libbsn = joinpath(@__DIR__, "..", "libbsn_hip.so")
const libbsn_io = joinpath(@__DIR__, "..", "libbsn_io.so")
const NG = 3
const BSN_HOST = Cint(0)

bcmask(bc::Symbol) = bc === :periodic ? Cint(7) : bc === :outflow ? Cint(0) :
                     throw(ArgumentError("bc must be :outflow or :periodic"))
precision_id(p::Symbol) = p === :fp64 ? Cint(0) : p === :split ? Cint(1) :
                          throw(ArgumentError("precision must be :fp64 or :split"))

function check(rc::Cint)
    rc == 0 && return nothing
    msg = unsafe_string(ccall((:bsn_last_error, libbsn), Cstring, ()))
    error("libbsn_hip status $rc: $msg")
end

mutable struct Context
    handle::Ptr{Cvoid}
end
const _ctx = Ref{Union{Nothing,Context}}(nothing)
function context()
    if _ctx[] === nothing
        h = Ref{Ptr{Cvoid}}(C_NULL)
        check(ccall((:bsn_ctx_create, libbsn), Cint, (Cint, Ref{Ptr{Cvoid}}), 0, h))
        _ctx[] = Context(h[])
    end
    return _ctx[]
end

"""
    HIPGrid(U; precision=:fp64)

Device-resident copy of a reference-layout state `U` of size
`(5, nx+6, ny+6, nz+6)`.
"""
mutable struct HIPGrid
    handle::Ptr{Cvoid}
    nx::Int
    ny::Int
    nz::Int
    function HIPGrid(U::Array{Float64,4}; precision::Symbol = :fp64)
        size(U, 1) == 5 || throw(ArgumentError("U must be (5, nx+6, ny+6, nz+6)"))
        nx, ny, nz = size(U, 2) - 2NG, size(U, 3) - 2NG, size(U, 4) - 2NG
        h = Ref{Ptr{Cvoid}}(C_NULL)
        check(ccall((:bsn_grid_create, libbsn), Cint,
                    (Ptr{Cvoid}, Int64, Int64, Int64, Cint, Ref{Ptr{Cvoid}}),
                    context().handle, nx, ny, nz, precision_id(precision), h))
        g = new(h[], nx, ny, nz)
        finalizer(x -> ccall((:bsn_grid_destroy, libbsn), Cint, (Ptr{Cvoid},), x.handle), g)
        check(ccall((:bsn_grid_set_state, libbsn), Cint, (Ptr{Cvoid}, Ptr{Float64}, Cint),
                    g.handle, U, BSN_HOST))
        return g
    end
end

"""Copy the device state into a host array of the reference layout."""
function download(g::HIPGrid, U::Array{Float64,4} = zeros(5, g.nx + 2NG, g.ny + 2NG, g.nz + 2NG))
    check(ccall((:bsn_grid_get_state, libbsn), Cint, (Ptr{Cvoid}, Ptr{Float64}, Cint),
                g.handle, U, BSN_HOST))
    return U
end

"""Static potential (nx, ny, nz) added as the ρ∇Φ source after the flux
divergence of every stage (src/gpu_kernels.jl:291-319)."""
function set_phi!(g::HIPGrid, Φ::Union{Nothing,Array{Float64,3}})
    check(ccall((:bsn_grid_set_phi, libbsn), Cint, (Ptr{Cvoid}, Ptr{Float64}, Cint),
                g.handle, Φ === nothing ? C_NULL : Φ, BSN_HOST))
end

# --- src/euler3d.jl drop-ins --------------------------------------------------

function fill_ghost_3d_outflow!(U::HIPGrid, nx::Int, ny::Int, nz::Int)
    check(ccall((:bsn_fill_ghost_3d, libbsn), Cint, (Ptr{Cvoid}, Cint), U.handle, Cint(0)))
end
function fill_ghost_3d_periodic!(U::HIPGrid, nx::Int, ny::Int, nz::Int)
    check(ccall((:bsn_fill_ghost_3d, libbsn), Cint, (Ptr{Cvoid}, Cint), U.handle, Cint(7)))
end

function apply_floors_3d!(U::HIPGrid, nx::Int, ny::Int, nz::Int, ρ_floor::Real, P_floor::Real,
                          γ::Real)
    check(ccall((:bsn_apply_floors_3d, libbsn), Cint, (Ptr{Cvoid}, Float64, Float64, Float64),
                U.handle, ρ_floor, P_floor, γ))
end

function euler3d_rhs!(dU::Array{Float64,4}, Fx::Array{Float64,4}, Fy::Array{Float64,4},
                      Fz::Array{Float64,4}, U::HIPGrid, nx::Int, ny::Int, nz::Int,
                      dx::Real, dy::Real, dz::Real, γ::Real; bc::Symbol = :outflow,
                      ρ_floor::Real = 0.0, P_floor::Real = 0.0, fill_ghosts::Bool = true)
    check(ccall((:bsn_euler3d_rhs, libbsn), Cint,
                (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Cint,
                 Float64, Float64, Float64, Float64, Cint, Float64, Float64, Cint),
                U.handle, dU, Fx, Fy, Fz, BSN_HOST, dx, dy, dz, γ, bcmask(bc), ρ_floor, P_floor,
                Cint(fill_ghosts)))
end

function euler3d_step!(U::HIPGrid, nx::Int, ny::Int, nz::Int, dx::Real, dy::Real, dz::Real,
                       dt::Real, γ::Real; bc::Symbol = :outflow, ρ_floor::Real = 0.0,
                       P_floor::Real = 0.0)
    check(ccall((:bsn_euler3d_step, libbsn), Cint,
                (Ptr{Cvoid}, Float64, Float64, Float64, Float64, Float64, Cint, Float64, Float64),
                U.handle, dx, dy, dz, dt, γ, bcmask(bc), ρ_floor, P_floor))
end

function euler3d_step!(U::Array{Float64,4}, nx::Int, ny::Int, nz::Int, dx::Real, dy::Real,
                       dz::Real, dt::Real, γ::Real; bc::Symbol = :outflow, ρ_floor::Real = 0.0,
                       P_floor::Real = 0.0, precision::Symbol = :fp64)
    check(ccall((:bsn_euler3d_step_host, libbsn), Cint,
                (Ptr{Float64}, Int64, Int64, Int64, Float64, Float64, Float64, Float64, Float64,
                 Cint, Float64, Float64, Cint),
                U, nx, ny, nz, dx, dy, dz, dt, γ, bcmask(bc), ρ_floor, P_floor,
                precision_id(precision)))
end

function cfl_dt_3d(U::HIPGrid, nx::Int, ny::Int, nz::Int, dx::Real, dy::Real, dz::Real,
                   γ::Real, cfl::Real)
    dt = Ref{Float64}(0.0)
    check(ccall((:bsn_cfl_dt_3d, libbsn), Cint,
                (Ptr{Cvoid}, Float64, Float64, Float64, Float64, Float64, Ref{Float64}),
                U.handle, dx, dy, dz, γ, cfl, dt))
    return dt[]
end

function cfl_dt_3d(U::Array{Float64,4}, nx::Int, ny::Int, nz::Int, dx::Real, dy::Real,
                   dz::Real, γ::Real, cfl::Real)
    dt = Ref{Float64}(0.0)
    check(ccall((:bsn_cfl_dt_3d_host, libbsn), Cint,
                (Ptr{Float64}, Int64, Int64, Int64, Float64, Float64, Float64, Float64, Float64,
                 Ref{Float64}),
                U, nx, ny, nz, dx, dy, dz, γ, cfl, dt))
    return dt[]
end

# --- src/rk3.jl ---------------------------------------------------------------

"""
    rk3_step!(U::HIPGrid, L!, dt, tmp1, tmp2)

Shu-Osher SSP-RK3 on a device grid; `L!(tmp1, U)` fills the host array
`tmp1` with L(U).  `tmp2` is unused: the base state lives on the device.
"""
function rk3_step!(U::HIPGrid, L!, dt::Real, tmp1::Array{Float64,4}, tmp2 = nothing)
    check(ccall((:bsn_rk3_save, libbsn), Cint, (Ptr{Cvoid},), U.handle))
    for stage in 1:3
        L!(tmp1, U)
        check(ccall((:bsn_rk3_combine, libbsn), Cint,
                    (Ptr{Cvoid}, Cint, Ptr{Float64}, Cint, Float64),
                    U.handle, Cint(stage), tmp1, BSN_HOST, dt))
    end
    return nothing
end

# --- src/fmr3d.jl -------------------------------------------------------------

"""
    HIPFMR(levels, boxes, ratios)

Nested levels (root first) with `boxes[l] = (ci_lo, ci_hi, cj_lo, cj_hi, ck_lo,
ck_hi)` of level l+1 inside level l; two levels reproduce FMRGrid3D.  Each
level's parameters must have been set with `set_params!`.
"""
mutable struct HIPFMR
    handle::Ptr{Cvoid}
    levels::Vector{HIPGrid}
    function HIPFMR(levels::Vector{HIPGrid}, boxes::Vector{NTuple{6,Int}}, ratios::Vector{Int})
        h = Ref{Ptr{Cvoid}}(C_NULL)
        lh = Int64[v for b in boxes for v in b]
        check(ccall((:bsn_fmr_create, libbsn), Cint,
                    (Ptr{Cvoid}, Cint, Ptr{Ptr{Cvoid}}, Ptr{Int64}, Ptr{Cint}, Ref{Ptr{Cvoid}}),
                    context().handle, length(levels), [g.handle for g in levels], lh,
                    Cint.(ratios), h))
        m = new(h[], levels)
        finalizer(x -> ccall((:bsn_fmr_destroy, libbsn), Cint, (Ptr{Cvoid},), x.handle), m)
        return m
    end
end

function set_params!(g::HIPGrid, dx, dy, dz, γ; bc::Symbol = :outflow, ρ_floor = 0.0, P_floor = 0.0)
    check(ccall((:bsn_grid_set_params, libbsn), Cint,
                (Ptr{Cvoid}, Float64, Float64, Float64, Float64, Cint, Float64, Float64),
                g.handle, dx, dy, dz, γ, bcmask(bc), ρ_floor, P_floor))
end

"""fmr3d_step!(G, dt_c; self_gravity = false, G_grav = 1.0) -- with self_gravity
the composite potential (solve_fmr_poisson) times G_grav is frozen for the root
step; without it a level's phi (set_phi!), when set, is a static potential."""
function fmr3d_step!(G::HIPFMR, dt_c::Float64; self_gravity::Bool = false, G_grav::Real = 1.0)
    check(ccall((:bsn_fmr3d_step_gravity, libbsn), Cint, (Ptr{Cvoid}, Float64, Cint, Float64),
                G.handle, dt_c, self_gravity ? 1 : 0, G_grav))
end

"""solve_fmr_poisson(G) -- (Φ per level, root first), each (nx, ny, nz); the
parents' density inside each child box is restricted first."""
function solve_fmr_poisson(G::HIPFMR)
    Φ = [zeros(g.nx, g.ny, g.nz) for g in G.levels]
    check(ccall((:bsn_solve_fmr_poisson, libbsn), Cint, (Ptr{Cvoid}, Ptr{Ptr{Float64}}, Cint),
                G.handle, [pointer(p) for p in Φ], BSN_HOST))
    return Tuple(Φ)
end

"""restrict_density_to_coarse!(G; level = 1)"""
function restrict_density_to_coarse!(G::HIPFMR; level::Int = 1)
    check(ccall((:bsn_fmr_restrict_density, libbsn), Cint, (Ptr{Cvoid}, Cint), G.handle, level))
end

"""_prolong_scalar_trilinear!(Φ_f, Φ_c, G; level = 1)"""
function _prolong_scalar_trilinear!(Φ_f::Array{Float64,3}, Φ_c::Array{Float64,3}, G::HIPFMR;
                                    level::Int = 1)
    check(ccall((:bsn_fmr_prolong_scalar, libbsn), Cint,
                (Ptr{Cvoid}, Cint, Ptr{Float64}, Cint, Ptr{Float64}, Cint),
                G.handle, level, Φ_c, BSN_HOST, Φ_f, BSN_HOST))
    return Φ_f
end

"""fmr_set_self_gravity!(G, on; G_grav = 1.0) -- the composite solve at every
root step of fmr_run!."""
function fmr_set_self_gravity!(G::HIPFMR, on::Bool; G_grav::Real = 1.0)
    check(ccall((:bsn_fmr_set_self_gravity, libbsn), Cint, (Ptr{Cvoid}, Cint, Float64),
                G.handle, on ? 1 : 0, G_grav))
end

function cfl_dt_fmr3d(G::HIPFMR, cfl::Real)
    dt = Ref{Float64}(0.0)
    check(ccall((:bsn_cfl_dt_fmr3d, libbsn), Cint, (Ptr{Cvoid}, Float64, Ref{Float64}),
                G.handle, cfl, dt))
    return dt[]
end

"""fmr_run!(G, cfl, nsteps; t_end = 0, graph = true) -- nsteps CFL-limited root
steps with dt computed on the device (one hipGraph per step when graph)."""
function fmr_run!(G::HIPFMR, cfl::Real, nsteps::Int; t_end::Real = 0.0, graph::Bool = true)
    check(ccall((:bsn_fmr_run, libbsn), Cint, (Ptr{Cvoid}, Float64, Float64, Cint, Cint),
                G.handle, cfl, t_end, nsteps, graph ? 1 : 0))
end

"""fmr_autotune!(G) -- kernel shapes of every level timed on its current state;
returns the JSON report."""
function fmr_autotune!(G::HIPFMR)
    buf = zeros(UInt8, 16384)
    check(ccall((:bsn_fmr_autotune, libbsn), Cint, (Ptr{Cvoid}, Ptr{UInt8}, Cint),
                G.handle, buf, length(buf)))
    return unsafe_string(pointer(buf))
end

"""fmr_set_mode!(G, inplace) -- level step form: false fuses the z sweep with the
stage update (the default), true steps in place; both give the same bits."""
function fmr_set_mode!(G::HIPFMR, inplace::Bool)
    check(ccall((:bsn_fmr_set_mode, libbsn), Cint, (Ptr{Cvoid}, Cint), G.handle, inplace ? 1 : 0))
end

# ---------------------------------------------------------------------------
# Point masses and sinks (src/gravity_bh.jl:58-78, src/sinks.jl:32-58).
# `bhs` is the reference's Vector{BlackHole} (fields pos, vel, mass, eps);
# r_sink and t_sink are the reference's functions (src/BinarySupernova.jl:159-171),
# passed in so that this module does not depend on BinarySupernova.

bh_pos(bhs) = Float64[c for bh in bhs for c in bh.pos[1:3]]
bh_vel(bhs) = Float64[c for bh in bhs for c in bh.vel[1:3]]

function add_bh_gravity_source!(dU::Array{Float64,4}, U::HIPGrid, nx::Int, ny::Int, nz::Int,
                                dx::Real, dy::Real, dz::Real, bhs, x0::Real, y0::Real, z0::Real)
    check(ccall((:bsn_add_bh_gravity_source, libbsn), Cint,
                (Ptr{Cvoid}, Ptr{Float64}, Cint, Float64, Float64, Float64, Cint, Ptr{Float64},
                 Ptr{Float64}, Ptr{Float64}, Float64, Float64, Float64),
                U.handle, dU, BSN_HOST, dx, dy, dz, length(bhs), bh_pos(bhs),
                Float64[bh.mass for bh in bhs], Float64[bh.eps for bh in bhs], x0, y0, z0))
    return dU
end

function add_sink_sources!(dU::Array{Float64,4}, U::HIPGrid, nx::Int, ny::Int, nz::Int,
                           dx::Real, dy::Real, dz::Real, bhs, x0::Real, y0::Real, z0::Real,
                           r_sink, t_sink; f_sink::Float64 = 1.0, torque_free::Bool = true,
                           ρ_sink_min::Float64 = 0.0)
    check(ccall((:bsn_add_sink_sources, libbsn), Cint,
                (Ptr{Cvoid}, Ptr{Float64}, Cint, Float64, Float64, Float64, Cint, Ptr{Float64},
                 Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Cint, Float64, Float64, Float64,
                 Float64),
                U.handle, dU, BSN_HOST, dx, dy, dz, length(bhs), bh_pos(bhs), bh_vel(bhs),
                Float64[r_sink(bh) for bh in bhs], Float64[t_sink(bh, f_sink) for bh in bhs],
                Cint(torque_free), ρ_sink_min, x0, y0, z0))
    return dU
end

"""Point masses every later step adds to each stage's RHS (coupled step order:
divergence, point masses, phi, sinks).  An empty `bhs` removes them."""
function set_point_masses!(g::HIPGrid, bhs, x0::Real, y0::Real, z0::Real)
    check(ccall((:bsn_grid_set_point_masses, libbsn), Cint,
                (Ptr{Cvoid}, Cint, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Float64, Float64,
                 Float64),
                g.handle, length(bhs), bh_pos(bhs), Float64[bh.mass for bh in bhs],
                Float64[bh.eps for bh in bhs], x0, y0, z0))
    return g
end

function set_sinks!(g::HIPGrid, bhs, x0::Real, y0::Real, z0::Real, r_sink, t_sink;
                    f_sink::Float64 = 1.0, torque_free::Bool = true, ρ_sink_min::Float64 = 0.0)
    check(ccall((:bsn_grid_set_sinks, libbsn), Cint,
                (Ptr{Cvoid}, Cint, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Cint,
                 Float64, Float64, Float64, Float64),
                g.handle, length(bhs), bh_pos(bhs), bh_vel(bhs),
                Float64[r_sink(bh) for bh in bhs], Float64[t_sink(bh, f_sink) for bh in bhs],
                Cint(torque_free), ρ_sink_min, x0, y0, z0))
    return g
end

# ---------------------------------------------------------------------------
# Self-gravity (src/self_gravity.jl:35-76)

"""solve_poisson_isolated(ρ, nx, ny, nz, dx, dy, dz) on the GPU (hipFFT).  The
grid `g` (of active shape nx, ny, nz) holds the cached plans and spectrum."""
function solve_poisson_isolated(g::HIPGrid, ρ::Array{Float64,3}, nx::Int, ny::Int, nz::Int,
                                dx::Real, dy::Real, dz::Real)
    size(ρ) == (nx, ny, nz) || throw(DimensionMismatch("ρ must be (nx, ny, nz)"))
    Φ = similar(ρ)
    check(ccall((:bsn_solve_poisson_isolated, libbsn), Cint,
                (Ptr{Cvoid}, Ptr{Float64}, Cint, Ptr{Float64}, Cint, Float64, Float64, Float64),
                g.handle, ρ, BSN_HOST, Φ, BSN_HOST, dx, dy, dz))
    return Φ
end

"""Every later step solves phi from each RK3 stage's density and adds the
self-gravity source (the coupled step with self_gravity on)."""
function set_self_gravity!(g::HIPGrid, on::Bool = true)
    check(ccall((:bsn_grid_set_self_gravity, libbsn), Cint, (Ptr{Cvoid}, Cint), g.handle,
                Cint(on)))
    return g
end

# ---------------------------------------------------------------------------
# What the bodies and the diagnostics read from the gas (src/gravity_bh.jl:91,
# src/sinks.jl:73, src/diagnostics.jl): device reductions, sums only to the host.

function gas_force_on_bh(U::HIPGrid, nx::Int, ny::Int, nz::Int, dx::Real, dy::Real, dz::Real,
                         bh, x0::Real, y0::Real, z0::Real)
    F = zeros(3)
    check(ccall((:bsn_gas_force_on_bh, libbsn), Cint,
                (Ptr{Cvoid}, Float64, Float64, Float64, Cint, Ptr{Float64}, Ptr{Float64},
                 Ptr{Float64}, Float64, Float64, Float64, Ptr{Float64}),
                U.handle, dx, dy, dz, Cint(1), bh_pos([bh]), Float64[bh.mass], Float64[bh.eps],
                x0, y0, z0, F))
    return F
end

function accrete!(bh, U::HIPGrid, nx::Int, ny::Int, nz::Int, dx::Real, dy::Real, dz::Real,
                  x0::Real, y0::Real, z0::Real, dt::Float64, r_sink, t_sink;
                  f_sink::Float64 = 1.0, ρ_sink_min::Float64 = 0.0)
    vel = Float64[bh.vel[1], bh.vel[2], bh.vel[3]]
    mass = Float64[bh.mass]
    check(ccall((:bsn_accrete, libbsn), Cint,
                (Ptr{Cvoid}, Float64, Float64, Float64, Cint, Ptr{Float64}, Ptr{Float64},
                 Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Float64, Float64, Float64, Float64,
                 Float64, Ptr{Float64}),
                U.handle, dx, dy, dz, Cint(1), bh_pos([bh]), vel, mass, Float64[r_sink(bh)],
                Float64[t_sink(bh, f_sink)], ρ_sink_min, x0, y0, z0, dt, C_NULL))
    bh.vel[1:3] .= vel
    bh.mass = mass[1]
    return nothing
end

"""All five gas diagnostics of src/diagnostics.jl in one pass."""
function gas_diagnostics(U::HIPGrid, nx::Int, ny::Int, nz::Int, dx::Real, dy::Real, dz::Real,
                         x0::Real = 0.0, y0::Real = 0.0, z0::Real = 0.0, bhs = (),
                         γ::Real = 5 / 3)
    out = zeros(9)
    check(ccall((:bsn_gas_diagnostics, libbsn), Cint,
                (Ptr{Cvoid}, Float64, Float64, Float64, Float64, Float64, Float64, Cint,
                 Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Float64, Ptr{Float64}),
                U.handle, dx, dy, dz, x0, y0, z0, Cint(length(bhs)), bh_pos(bhs),
                Float64[bh.mass for bh in bhs], Float64[bh.eps for bh in bhs], γ, out))
    return out
end

gas_energy_total(U::HIPGrid, nx, ny, nz, dx, dy, dz) = gas_diagnostics(U, nx, ny, nz, dx, dy, dz)[1]
gas_kinetic_total(U::HIPGrid, nx, ny, nz, dx, dy, dz) = gas_diagnostics(U, nx, ny, nz, dx, dy, dz)[2]
gas_momentum_total(U::HIPGrid, nx, ny, nz, dx, dy, dz) = gas_diagnostics(U, nx, ny, nz, dx, dy, dz)[3:5]
gas_angular_momentum_total(U::HIPGrid, nx, ny, nz, dx, dy, dz, x0, y0, z0) =
    gas_diagnostics(U, nx, ny, nz, dx, dy, dz, x0, y0, z0)[6:8]
bound_gas_mass(U::HIPGrid, nx, ny, nz, dx, dy, dz, x0, y0, z0, bhs, γ) =
    gas_diagnostics(U, nx, ny, nz, dx, dy, dz, x0, y0, z0, bhs, γ)[9]

# ---------------------------------------------------------------------------
# Snapshot and checkpoint files (src/io.jl:29-84, src/checkpoint.jl:244-318),
# streamed between the device state and HDF5 through pinned slabs.

function checkio(rc::Cint)
    rc == 0 && return nothing
    msg = unsafe_string(ccall((:bsn_io_last_error, libbsn_io), Cstring, ()))
    error("libbsn_io: status $rc: $msg")
end

function write_snapshot(filename::String, U::HIPGrid, nx::Int, ny::Int, nz::Int,
                        dx::Real, dy::Real, dz::Real, t::Real, γ::Real)
    checkio(ccall((:bsn_write_snapshot, libbsn_io), Cint,
                  (Ptr{Cvoid}, Cstring, Float64, Float64, Float64, Float64, Float64),
                  U.handle, filename, dx, dy, dz, t, γ))
    return nothing
end

"""The active cells of U from a snapshot (its ghost cells are left as they are)."""
function read_snapshot!(U::HIPGrid, filename::String)
    checkio(ccall((:bsn_read_snapshot, libbsn_io), Cint, (Ptr{Cvoid}, Cstring),
                  U.handle, filename))
    return U
end

function save_checkpoint_uniform(path::AbstractString, U::HIPGrid, bhs;
                                 t::Real, step::Integer, dt_last::Real, t_snap::Real,
                                 t_bh2_sink_on::Real, nx::Integer, ny::Integer, nz::Integer,
                                 dx::Real, dy::Real, dz::Real, γ::Real, x0::Real, y0::Real,
                                 z0::Real, ρ_floor::Real, P_floor::Real)
    meta = Float64[t, dt_last, t_snap, t_bh2_sink_on]
    params = Float64[dx, dy, dz, γ, x0, y0, z0, ρ_floor, P_floor]
    props = Float64[p for bh in bhs for p in (bh.mass, bh.eps, bh.c_code, bh.r_floor)]
    checkio(ccall((:bsn_save_checkpoint_uniform, libbsn_io), Cint,
                  (Ptr{Cvoid}, Ptr{Float64}, Int64, Int64, Int64, Cstring, Int64, Ptr{Float64},
                   Ptr{Float64}, Cint, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}),
                  U.handle, C_NULL, nx, ny, nz, path, step, meta, params, Cint(length(bhs)),
                  bh_pos(bhs), bh_vel(bhs), props))
    return path
end

"""The state of a uniform checkpoint into U; returns (bodies as (pos, vel, mass,
eps, c_code, r_floor) tuples, grid params, meta)."""
function load_checkpoint_uniform!(U::HIPGrid, path::AbstractString; max_bodies::Int = 8)
    step = Ref{Int64}(0)
    meta = zeros(5); params = zeros(9); nb = Ref{Cint}(0)
    pos = zeros(3 * max_bodies); vel = zeros(3 * max_bodies); props = zeros(4 * max_bodies)
    checkio(ccall((:bsn_load_checkpoint_uniform, libbsn_io), Cint,
                  (Ptr{Cvoid}, Ptr{Float64}, Cstring, Ref{Int64}, Ptr{Float64}, Ptr{Float64}, Cint,
                   Ref{Cint}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}),
                  U.handle, C_NULL, path, step, meta, params, Cint(max_bodies), nb, pos, vel,
                  props))
    bodies = [(pos[3i-2:3i], vel[3i-2:3i], props[4i-3], props[4i-2], props[4i-1], props[4i])
              for i in 1:nb[]]
    grid = (; nx = U.nx, ny = U.ny, nz = U.nz, dx = params[1], dy = params[2], dz = params[3],
            γ = params[4], x0 = params[5], y0 = params[6], z0 = params[7],
            ρ_floor = params[8], P_floor = params[9])
    m = (; t = meta[1], step = step[], dt_last = meta[2], t_snap = meta[3],
         t_bh2_sink_on = meta[4], format_version = Int(meta[5]))
    return bodies, grid, m
end

"""save_checkpoint(path, G, bhs; t, step, dt_last, box, ratio, dxs, γ, bc, ρ_floor, P_floor)
(src/checkpoint.jl:33-117) for a two-level HIPFMR: both levels stream from
the device.  box = (ci_lo, ci_hi, cj_lo, cj_hi, ck_lo, ck_hi); dxs = the
(dx, dy, dz) of the coarse and the fine level."""
function save_checkpoint(path::AbstractString, G::HIPFMR, bhs; t::Real = 0.0,
                         step::Integer = 0, dt_last::Real = 0.0, box::NTuple{6,Int},
                         ratio::Integer, dxs::NTuple{2,NTuple{3,Float64}}, γ::Real,
                         bc::Symbol = :outflow, ρ_floor::Real = 0.0, P_floor::Real = 0.0)
    length(G.levels) == 2 || error("the checkpoint format holds a two-level FMR grid")
    c, f = G.levels
    n = Int64[c.nx, c.ny, c.nz, f.nx, f.ny, f.nz]
    d = Float64[dxs[1]..., dxs[2]...]
    props = Float64[p for bh in bhs for p in (bh.mass, bh.eps, bh.c_code, bh.r_floor)]
    checkio(ccall((:bsn_save_checkpoint_fmr, libbsn_io), Cint,
                  (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Ptr{Int64}, Ptr{Float64},
                   Cstring, Int64, Ptr{Float64}, Ptr{Int64}, Cint, Ptr{Float64}, Cstring, Cint,
                   Ptr{Float64}, Ptr{Float64}, Ptr{Float64}),
                  c.handle, f.handle, C_NULL, C_NULL, n, d, path, step, Float64[t, dt_last],
                  Int64[box...], Cint(ratio), Float64[γ, ρ_floor, P_floor], String(bc),
                  Cint(length(bhs)), bh_pos(bhs), bh_vel(bhs), props))
    return path
end

"""load_checkpoint!(G, path) (src/checkpoint.jl:119-172): both saved level
states into the two-level HIPFMR G (which must have the stored shapes);
returns (bodies as (pos, vel, mass, eps, c_code, r_floor) tuples, meta)."""
function load_checkpoint!(G::HIPFMR, path::AbstractString; max_bodies::Int = 8)
    length(G.levels) == 2 || error("the checkpoint format holds a two-level FMR grid")
    step = Ref{Int64}(0)
    meta = zeros(3); nb = Ref{Cint}(0)
    pos = zeros(3 * max_bodies); vel = zeros(3 * max_bodies); props = zeros(4 * max_bodies)
    checkio(ccall((:bsn_load_checkpoint_fmr, libbsn_io), Cint,
                  (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Cstring, Ref{Int64},
                   Ptr{Float64}, Cint, Ref{Cint}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}),
                  G.levels[1].handle, G.levels[2].handle, C_NULL, C_NULL, path, step, meta,
                  Cint(max_bodies), nb, pos, vel, props))
    bodies = [(pos[3i-2:3i], vel[3i-2:3i], props[4i-3], props[4i-2], props[4i-1], props[4i])
              for i in 1:nb[]]
    return bodies, (; t = meta[1], step = step[], dt_last = meta[2],
                    format_version = Int(meta[3]))
end

"""init_trajectory_file(filename, n_bh) (src/io.jl:94-108)."""
function init_trajectory_file(filename::String, n_bh::Int)
    checkio(ccall((:bsn_init_trajectory_file, libbsn_io), Cint, (Cstring, Cint), filename,
                  Cint(n_bh)))
    return nothing
end

"""append_trajectory(filename, t, bhs, r_sink) (src/io.jl:115-133); r_sink is the
reference's function (src/BinarySupernova.jl:159)."""
function append_trajectory(filename::String, t::Real, bhs, r_sink)
    checkio(ccall((:bsn_append_trajectory, libbsn_io), Cint,
                  (Cstring, Float64, Cint, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}),
                  filename, t, Cint(length(bhs)), Float64[bh.mass for bh in bhs],
                  Float64[r_sink(bh) for bh in bhs], bh_pos(bhs), bh_vel(bhs)))
    return nothing
end

"""read_trajectory(filename) (src/io.jl:144-161) -> (t, bh_data), bh_data one
NamedTuple (mass, rsink, pos (N×3), vel (N×3)) per body."""
function read_trajectory(filename::String)
    nb = Ref{Cint}(0); n = Ref{Int64}(0)
    checkio(ccall((:bsn_trajectory_info, libbsn_io), Cint, (Cstring, Ref{Cint}, Ref{Int64}),
                  filename, nb, n))
    N, B = Int(n[]), Int(nb[])
    t = zeros(N); mass = zeros(max(B, 1) * N); rs = zeros(max(B, 1) * N)
    pos = zeros(3 * max(B, 1) * N); vel = zeros(3 * max(B, 1) * N)
    checkio(ccall((:bsn_read_trajectory, libbsn_io), Cint,
                  (Cstring, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}),
                  filename, t, mass, rs, pos, vel))
    # the library writes body b, record r, axis a at (b*N + r)*3 + a
    bh_data = [(mass = mass[b*N+1:(b+1)*N], rsink = rs[b*N+1:(b+1)*N],
                pos = permutedims(reshape(pos[3b*N+1:3(b+1)*N], 3, N)),
                vel = permutedims(reshape(vel[3b*N+1:3(b+1)*N], 3, N))) for b in 0:B-1]
    return t, bh_data
end

end # module
