# MHDKernelsHIPExt: the MI355X backend of lib/MHDKernels as a package extension.
#
# The reference picks GPU kernels through its backend registry (MHDKernels.jl:46-66) and
# the launch methods that its CUDA and Metal extensions specialise
# (ext/MHDKernelsCUDAExt.jl:6-26, ext/MHDKernelsMetalExt.jl:136-168). This extension does
# the same for AMDGPU.jl's ROCBackend: it registers `:hip`, allocates the state's fields as
# ROCArrays, and routes every step the production driver calls to libpmfmhd.so through
# its C ABI (include/pmfmhd.h). The state, its arrays and its swap discipline stay the
# reference's own: each launch reads s.U and writes s.scratch (or s.U and the CT packed
# buffer in place), and the callers in integrator_ref.jl / integrator_cube.jl swap.
#
#   _cube_launch!           integrator_cube.jl:467-476  -> pmf_step_into (fused kernel)
#   _cube_drag_launch!      integrator_cube.jl:492-503  -> pmf_step_sources_into
#   _cube_radiation_launch! integrator_cube.jl:522-540  -> pmf_step_sources_into
#   max_wavespeed!          timestep.jl:17-23           -> pmf_max_wavespeed_of
#   step_ref!               integrator_ref.jl:210-220   -> pmf_step_into (+ ct_update!)
#   step_ref_drag!          integrator_ref.jl:222-235   -> pmf_step_sources_into
#   step_ref_radiation!     integrator_ref.jl:238-258   -> pmf_step_sources_into
#
# For CT states the library runs ct_update! (ct.jl:158-173) on the output set itself, so
# the step_ref methods here swap and return without calling the Julia ct_update!.
#
# Every ccall below is written as `ccall((:name, LIB), Ret, (ArgTypes...), args...)` with
# the argument types of its prototype in include/pmfmhd.h; a CPU test of this repository
# parses them against the header.
module MHDKernelsHIPExt

using MHDKernels
using AMDGPU

const LIB = get(ENV, "PMFMHD_LIB", "libpmfmhd.so")

# include/pmfmhd.h: pmf_params, field for field
struct PmfParams
    nx::Int32
    ny::Int32
    nz::Int32
    riemann::Int32
    recon::Int32
    bc::NTuple{3,Int32}
    solver::Int32
    precision::Int32
    dx::Float64
    gamma::Float64
    smallr::Float64
    pfl::Float64
    llf_dmin::Float64
    llf_pmin::Float64
end

# include/pmfmhd.h: PMF_INTEGRATOR_*, PMF_SRC_*
const INTEGRATOR_REF = Int32(0)
const INTEGRATOR_CUBE = Int32(1)
const SRC_TRACK_DENSITY = Int32(1)
const SRC_TRACK_MAGNETIC = Int32(2)
const SRC_TRACK_ENERGY = Int32(4)
const SRC_CHECK_ADMISSIBILITY = Int32(8)
const SRC_FALLBACK = Int32(16)
const SRC_GLOBAL_FALLBACK = Int32(32)
const SRC_ROBUST_FALLBACK = Int32(64)

function __init__()
    if AMDGPU.functional()
        MHDKernels.register_backend!(:hip, ROCBackend())
    end
end

MHDKernels.device_zeros(::ROCBackend, ::Type{T}, dims::Dims) where {T} = AMDGPU.zeros(T, dims)

pmf_error() = unsafe_string(ccall((:pmf_last_error, LIB), Cstring, ()))
# The library reuses the reference's error texts; rethrow them as the reference would.
check(rc::Integer) = rc == 0 ? nothing : error(pmf_error())

const HipState = MHDKernels.MHDState{T,<:ROCArray} where {T}

function params_of(s::MHDKernels.MHDState{T}) where {T}
    T === Float32 || T === Float64 || error("libpmfmhd steps Float32 and Float64 states (got $T)")
    cx, cy, cz = MHDKernels.bc_codes(s)
    PmfParams(Int32(s.dims[1]), Int32(s.dims[2]), Int32(s.dims[3]),
              Int32(MHDKernels.riemann_code_of(s)), Int32(MHDKernels.recon_code_of(s)),
              (Int32(cx), Int32(cy), Int32(cz)), MHDKernels.uses_ct(s) ? Int32(1) : Int32(0),
              T === Float64 ? Int32(1) : Int32(0), Float64(s.dx), Float64(s.γ), Float64(s.smallr),
              Float64(s.pfl), Float64(s.llf_dmin), Float64(s.llf_pmin))
end

# One field-less library workspace (parameters, stream binding, reduction buffers) per
# state, created on first use, rebuilt if the state's parameters change, destroyed
# with the state.
mutable struct Workspace
    handle::Ptr{Cvoid}
    params::PmfParams
end
const WORKSPACES = IdDict{Any,Workspace}()
const WS_LOCK = ReentrantLock()

function destroy!(w::Workspace)
    if w.handle != C_NULL
        ccall((:pmf_destroy, LIB), Cint, (Ptr{Cvoid},), w.handle)
        w.handle = C_NULL
    end
    nothing
end

function workspace(s::MHDKernels.MHDState)
    p = params_of(s)
    w = lock(WS_LOCK) do
        w = get(WORKSPACES, s, nothing)
        if w === nothing || w.params != p
            w === nothing || destroy!(w)
            h = Ref{Ptr{Cvoid}}(C_NULL)
            check(ccall((:pmf_create_workspace, LIB), Cint, (Ref{PmfParams}, Ref{Ptr{Cvoid}}), p, h))
            w = Workspace(h[], p)
            WORKSPACES[s] = w
            finalizer(s) do st
                lock(WS_LOCK) do
                    ws = pop!(WORKSPACES, st, nothing)
                    ws === nothing || destroy!(ws)
                end
            end
        end
        w
    end
    # launch on the task's current HIP stream, so the library's work is ordered with
    # AMDGPU.jl's (default priority, as every stream here)
    check(ccall((:pmf_set_stream, LIB), Cint, (Ptr{Cvoid}, Ptr{Cvoid}), w.handle, AMDGPU.stream().stream))
    w.handle
end

fieldptrs(t) = Ptr{Cvoid}[Ptr{Cvoid}(pointer(a)) for a in t]
ctptr(s) = MHDKernels.uses_ct(s) ? Ptr{Cvoid}(pointer(s.ct.packed)) : C_NULL

# Every launcher of the reference ends with KA.synchronize (integrator_ref.jl:216,
# integrator_cube.jl:475): wait for the workspace's stream.
finish(ws::Ptr{Cvoid}) = check(ccall((:pmf_synchronize, LIB), Cint, (Ptr{Cvoid},), ws))

function step_into!(s, ins, outs, dt, ch, decay, integrator::Int32)
    ws = workspace(s)
    check(ccall((:pmf_step_into, LIB), Cint,
                (Ptr{Cvoid}, Ptr{Ptr{Cvoid}}, Ptr{Ptr{Cvoid}}, Ptr{Cvoid}, Cdouble, Cdouble, Cdouble, Int32),
                ws, fieldptrs(ins), fieldptrs(outs), ctptr(s), Float64(dt), Float64(ch), Float64(decay), integrator))
    finish(ws)
end

function sources_into!(s, dt, ch, decay, q, correction, flags::Int32, integrator::Int32;
                       target_velocity = (0.0, 0.0, 0.0))
    ws = workspace(s)
    v0 = Float64[target_velocity[1], target_velocity[2], target_velocity[3]]
    corr = correction === nothing ? C_NULL : Ptr{Cvoid}(pointer(correction))
    check(ccall((:pmf_step_sources_into, LIB), Cint,
                (Ptr{Cvoid}, Ptr{Ptr{Cvoid}}, Ptr{Ptr{Cvoid}}, Ptr{Cvoid}, Ptr{Cvoid}, Cdouble, Cdouble, Cdouble,
                 Cdouble, Ptr{Cdouble}, Int32, Int32),
                ws, fieldptrs(s.U), fieldptrs(s.scratch), ctptr(s), corr, Float64(dt), Float64(ch),
                Float64(decay), Float64(q), v0, flags, integrator))
    finish(ws)
end

flag(b::Bool, bit::Int32) = b ? bit : Int32(0)

# ---- :cube seams (step_cube!, step_cube_drag!, step_cube_radiation! check the grid and
# swap; the library's fused kernel takes any grid those checks let through)

function MHDKernels._cube_launch!(::Val, be::ROCBackend, s::MHDKernels.MHDState, dt::Real, ch::Real, decay::Real)
    step_into!(s, s.U, s.scratch, dt, ch, decay, INTEGRATOR_CUBE)
end

function MHDKernels._cube_drag_launch!(be::ROCBackend, s::MHDKernels.MHDState{T}, dt::Real, ch::Real,
                                       decay::Real, q::T) where {T}
    sources_into!(s, dt, ch, decay, q, nothing, Int32(0), INTEGRATOR_CUBE)
end

function MHDKernels._cube_radiation_launch!(be::ROCBackend, s::MHDKernels.MHDState{T}, dt::Real, ch::Real,
                                            decay::Real, q::T, correction, track_density::Val{TD},
                                            track_magnetic::Val{TB}, track_energy::Val{TE}, admissibility::Val{CK},
                                            fallback::Val{FB}, global_fallback::Val{GF},
                                            robust_fallback::Val{RF}) where {T,TD,TB,TE,CK,FB,GF,RF}
    flags = flag(TD, SRC_TRACK_DENSITY) | flag(TB, SRC_TRACK_MAGNETIC) | flag(TE, SRC_TRACK_ENERGY) |
            flag(CK, SRC_CHECK_ADMISSIBILITY) | flag(FB, SRC_FALLBACK) | flag(GF, SRC_GLOBAL_FALLBACK) |
            flag(RF, SRC_ROBUST_FALLBACK)
    sources_into!(s, dt, ch, decay, q, correction, flags, INTEGRATOR_CUBE)
end

# ---- CFL reduction

function MHDKernels.max_wavespeed!(w, s::HipState{T}) where {T}
    length(w) == MHDKernels.ncells(s) || error("wavespeed work length $(length(w)) != ncells $(MHDKernels.ncells(s))")
    smax = Ref{Cdouble}(0.0)
    check(ccall((:pmf_max_wavespeed_of, LIB), Cint, (Ptr{Cvoid}, Ptr{Ptr{Cvoid}}, Ref{Cdouble}),
                workspace(s), fieldptrs(s.U), smax))
    T(smax[])
end

# ---- :ref steps (the production CT runs and the :ref radiation runs call these)

function MHDKernels.step_ref!(s::HipState{T}, dt::Real; ch::Real, decay::Real) where {T}
    step_into!(s, s.U, s.scratch, dt, ch, decay, INTEGRATOR_REF)
    s.U, s.scratch = s.scratch, s.U
    s
end

function MHDKernels.step_ref_drag!(s::HipState{T}, dt::Real; ch::Real, decay::Real, drag_impulse::Real,
                                   target_velocity = (0, 0, 0)) where {T}
    sources_into!(s, dt, ch, decay, drag_impulse, nothing, Int32(0), INTEGRATOR_REF;
                  target_velocity = Float64.(target_velocity))
    s.U, s.scratch = s.scratch, s.U
    s
end

function MHDKernels.step_ref_radiation!(s::HipState{T}, dt::Real; ch::Real, decay::Real, drag_impulse::Real,
                                        correction, track_density::Bool = true, track_magnetic::Bool = true,
                                        track_energy::Bool = true, robust_fallback::Bool = false) where {T}
    flags = flag(track_density, SRC_TRACK_DENSITY) | flag(track_magnetic, SRC_TRACK_MAGNETIC) |
            flag(track_energy, SRC_TRACK_ENERGY) | flag(robust_fallback, SRC_ROBUST_FALLBACK)
    sources_into!(s, dt, ch, decay, drag_impulse, correction, flags, INTEGRATOR_REF)
    s.U, s.scratch = s.scratch, s.U
    s
end

end # module
