"""
    ChmyMI355X

MI355X (gfx950) backend for Chmy.jl's staggered-grid hot path.  Every kernel on
the path is a `ccall` into `libchmy_mi355x.so` (C ABI: `include/chmy_mi355x.h`).

What stays Chmy's own: `Arch`, `UniformGrid`, `Field`, `VectorField`,
`TensorField`, `Launcher`, `batch`, `bc!`, `interior`, `set!`.  What this module
adds, all dispatched on its own types so that no method of Chmy or of
ChmyAMDGPUExt is replaced:

  - `MI355XBackend <: KernelAbstractions.GPU` with `MI355XDevice`, and the hooks
    Chmy asks of a backend (`get_device`, `set_device!`, `heuristic_groupsize`,
    `pointertype`, `unsafe_wrap`, `gpu_aware_compat`; src/Architectures.jl:46-89,
    ext/ChmyAMDGPUExt/ChmyAMDGPUExt.jl:7-19);
  - `MI355XArray`, device memory from `chmy_malloc_zeros`, which
    `KernelAbstractions.zeros(MI355XBackend(), T, dims)` returns, so
    `Field(arch, grid, loc)` (src/Fields/field.jl:56-62) allocates it unchanged;
  - `Kernels`: one callable per hot-path kernel of the examples, passed to a
    `Launcher` in place of the `@kernel` function.  The Launcher calls it as it
    calls a KernelAbstractions kernel, `k(backend, groupsize, worksize)(args..., offset)`
    (src/KernelLaunch.jl:121-126,152-158), with `worksize` a tuple or a
    `StaticSize` (the `bc=` path);
  - `bc!` for a whole `BatchSet` and for one `(Side, Dim, FieldBatch)` on an
    `SingleDeviceArchitecture{MI355XBackend}` (src/BoundaryConditions/batch.jl:20-29,180-184);
  - `set!` on MI355X fields: Chmy's own CPU `set!` into a host field, then one
    upload (bitwise Chmy's CPU values);
  - `iterate!` (fused PT iterations with the V batch, in two-iteration passes
    for free-slip batches), `iter2!` (one such pass) and `iter2_first!` (one
    such pass that keeps the first iteration's V, for the thermal updates
    inside the PT loop); `stokes_residual!` and
    `stokes_accumulate!` for FP32 sweeps with FP64 defect correction;
    `maxabs` for the residual norms; `apply_op!` for any GridOperator (plain,
    masked by a FieldMask component, on uniform or FunctionAxis grids);
  - `rrule`s for the SIA time loop and the steady Stokes solve.

Usage, against examples/stokes_3d_inc_ve_T.jl:

    using Chmy, ChmyMI355X
    using ChmyMI355X.Kernels: update_stress!, update_velocity!, update_old!,
                              update_thermal_flux!, update_thermal!
    arch = Arch(MI355XBackend())
    ...   # the example's body, unchanged

Field storage is Chmy's layout (size n .+ 4H per dim, x fastest), so
descriptors pass device pointers with no copies.
"""
module ChmyMI355X

using Chmy
using KernelAbstractions
import KernelAbstractions.NDIteration: StaticSize
import Chmy.Architectures: get_device, set_device!, heuristic_groupsize, pointertype, gpu_aware_compat
import Chmy.BoundaryConditions: bc!
import Chmy.Fields: set!
using ChainRulesCore

export MI355XBackend, MI355XDevice, MI355XArray, Kernels

const LIB = get(ENV, "CHMY_MI355X_LIB", joinpath(@__DIR__, "..", "lib", "libchmy_mi355x.so"))

# ------------------------------------------------------------------ descriptors
# Layouts of chmy_field_t, chmy_grid_t, chmy_box_t and chmy_bc_t.
struct ChmyField
    data::Ptr{Cvoid}
    nd::Int32
    dtype::Int32
    n::NTuple{3,Int64}
    halo::NTuple{3,Int32}
    loc::NTuple{3,Int32}
end

struct ChmyGrid
    nd::Int32
    reserved::Int32
    n::NTuple{3,Int64}
    origin::NTuple{3,Float64}
    extent::NTuple{3,Float64}
    spacing::NTuple{3,Float64}
    inv_spacing::NTuple{3,Float64}
    axis::NTuple{3,Ptr{Float64}}
end

struct ChmyBox
    lo::NTuple{3,Int64}
    len::NTuple{3,Int64}
end

struct ChmyBC
    kind::Int32
    vkind::Int32
    value::Float64
    vfield::Ptr{ChmyField}
end

const CHMY_F64 = Int32(0)
const CHMY_F32 = Int32(1)
const CHMY_DIRICHLET = Int32(0)
const CHMY_NEUMANN = Int32(1)
const CHMY_VALUE_ZERO = Int32(0)
const CHMY_VALUE_SCALAR = Int32(1)
const CHMY_VALUE_FIELD = Int32(2)
const CHMY_ITER_RESIDUAL = Cint(1)
const CHMY_ITER_EXACT = Cint(2)

"""Error raised for a non-zero status of the library, with `chmy_last_error()`."""
struct ChmyError <: Exception
    code::Cint
    msg::String
end
Base.showerror(io::IO, e::ChmyError) = print(io, "libchmy_mi355x (status ", e.code, "): ", e.msg)

function check(rc::Cint)
    rc == 0 && return nothing
    msg = unsafe_string(ccall((:chmy_last_error, LIB), Cstring, ()))
    throw(ChmyError(rc, msg))
end

# ------------------------------------------------------------------ backend, device, streams
struct MI355XBackend <: KernelAbstractions.GPU end

"""A GPU by Chmy's 1-based `device_id` (`Arch(backend; device_id)`)."""
struct MI355XDevice
    id::Int
end

get_device(::MI355XBackend, device_id::Integer) = MI355XDevice(Int(device_id))
set_device!(dev::MI355XDevice) = check(ccall((:chmy_init, LIB), Cint, (Cint,), dev.id - 1))
# the library's kernels choose their own launch shapes; the Launcher only
# forwards this value (wrapped in StaticSize on the bc= path)
heuristic_groupsize(::MI355XBackend, ::Val{N}) where {N} = ntuple(d -> d == 1 ? 64 : 1, Val(N))
pointertype(::MI355XBackend, T::DataType) = Ptr{T}
gpu_aware_compat(::MI355XBackend) = false

# The stream of the running task: NULL (the device's default stream) unless
# the task activated the architecture with a priority, which gives it a
# stream of its own (Launcher workers, src/KernelLaunch.jl:40-54).  The
# library supports the default priority only, so every priority maps to it.
current_stream() = get(task_local_storage(), :chmy_mi355x_stream, C_NULL)::Ptr{Cvoid}

function KernelAbstractions.priority!(::MI355XBackend, prio::Symbol)
    prio in (:normal, :low, :high) || throw(ArgumentError("priority must be :normal, :low or :high"))
    if !haskey(task_local_storage(), :chmy_mi355x_stream)
        s = Ref{Ptr{Cvoid}}(C_NULL)
        check(ccall((:chmy_stream_create, LIB), Cint, (Cint, Ref{Ptr{Cvoid}}), 0, s))
        task_local_storage(:chmy_mi355x_stream, s[])
    end
    return nothing
end

KernelAbstractions.synchronize(::MI355XBackend) =
    check(ccall((:chmy_stream_sync, LIB), Cint, (Ptr{Cvoid},), current_stream()))

# ------------------------------------------------------------------ device array
"""
    MI355XArray{T,N}

Device memory of one MI355X, zero-filled at allocation.  Scalar indexing is
an error; `Array(a)` and `copyto!` move data to and from the host.
"""
mutable struct MI355XArray{T,N} <: AbstractArray{T,N}
    ptr::Ptr{T}
    dims::NTuple{N,Int}
    owned::Bool
end

function MI355XArray{T}(::UndefInitializer, dims::NTuple{N,Integer}) where {T,N}
    nbytes = Csize_t(prod(dims; init=1) * sizeof(T))
    p = Ref{Ptr{Cvoid}}(C_NULL)
    check(ccall((:chmy_malloc_zeros, LIB), Cint, (Ref{Ptr{Cvoid}}, Csize_t, Ptr{Cvoid}), p, nbytes, current_stream()))
    a = MI355XArray{T,N}(Ptr{T}(p[]), Int.(dims), true)
    finalizer(free!, a)
    return a
end

function free!(a::MI355XArray)
    if a.owned && a.ptr != C_NULL
        ccall((:chmy_free, LIB), Cint, (Ptr{Cvoid},), a.ptr)
        a.ptr = C_NULL
    end
    return nothing
end

Base.size(a::MI355XArray) = a.dims
Base.pointer(a::MI355XArray) = a.ptr
Base.getindex(::MI355XArray, ::Int...) = error("MI355XArray: no scalar indexing of device memory; copy with Array(a)")
Base.setindex!(::MI355XArray, v, ::Int...) = error("MI355XArray: no scalar indexing of device memory")

KernelAbstractions.get_backend(::MI355XArray) = MI355XBackend()
KernelAbstractions.allocate(::MI355XBackend, ::Type{T}, dims::Tuple; unified::Bool=false) where {T} =
    MI355XArray{T}(undef, dims)
# chmy_malloc_zeros returns zeroed memory (field.jl:59's zeros)
KernelAbstractions.zeros(::MI355XBackend, ::Type{T}, dims::Tuple; unified::Bool=false) where {T} =
    MI355XArray{T}(undef, dims)
Base.unsafe_wrap(::MI355XBackend, ptr::Ptr{T}, dims) where {T} =
    MI355XArray{T,length(dims)}(ptr, Int.(Tuple(dims)), false)

function Base.fill!(a::MI355XArray{T}, x) where {T}
    iszero(x) || error("MI355XArray: fill! with a non-zero value goes through set!")
    check(ccall((:chmy_memset_zero, LIB), Cint, (Ptr{Cvoid}, Csize_t, Ptr{Cvoid}), a.ptr,
                Csize_t(length(a) * sizeof(T)), current_stream()))
    return a
end

function Base.copyto!(dst::Array{T}, src::MI355XArray{T}) where {T}
    length(dst) == length(src) || throw(DimensionMismatch("copyto!: lengths differ"))
    s = current_stream()
    GC.@preserve dst begin
        check(ccall((:chmy_memcpy_d2h, LIB), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Csize_t, Ptr{Cvoid}), pointer(dst),
                    src.ptr, Csize_t(length(src) * sizeof(T)), s))
        check(ccall((:chmy_stream_sync, LIB), Cint, (Ptr{Cvoid},), s))
    end
    return dst
end

function Base.copyto!(dst::MI355XArray{T}, src::Array{T}) where {T}
    length(dst) == length(src) || throw(DimensionMismatch("copyto!: lengths differ"))
    s = current_stream()
    GC.@preserve src begin
        check(ccall((:chmy_memcpy_h2d, LIB), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Csize_t, Ptr{Cvoid}), dst.ptr,
                    pointer(src), Csize_t(length(src) * sizeof(T)), s))
        check(ccall((:chmy_stream_sync, LIB), Cint, (Ptr{Cvoid},), s))
    end
    return dst
end

Base.Array(a::MI355XArray{T,N}) where {T,N} = copyto!(Array{T,N}(undef, a.dims), a)
# interior(f) |> Array (the examples' copy-back): the storage, then the view's indices
Base.Array(v::SubArray{T,N,<:MI355XArray}) where {T,N} = Array(parent(v))[parentindices(v)...]

const MI355XField{T,N} = Field{T,N,L,H,<:MI355XArray} where {L,H}
const MI355XArch = SingleDeviceArchitecture{MI355XBackend}

# ------------------------------------------------------------------ descriptors of Chmy objects
dtcode(::Type{Float64}) = CHMY_F64
dtcode(::Type{Float32}) = CHMY_F32
loccode(::Center) = Int32(0)
loccode(::Vertex) = Int32(1)
pad3(t::Tuple, z) = ntuple(i -> i <= length(t) ? t[i] : z, Val(3))

function desc(f::MI355XField{T,N}) where {T,N}
    H = halo(f)
    h = H isa Integer ? ntuple(_ -> H, Val(N)) : H
    return ChmyField(Ptr{Cvoid}(pointer(parent(f))), Int32(N), dtcode(T), pad3(Int64.(size(f)), Int64(0)),
                     pad3(Int32.(h), Int32(0)), pad3(map(loccode, location(f)), Int32(0)))
end

const NOAXIS = Ptr{Float64}(C_NULL)

function desc(g::UniformGrid{N}) where {N}
    ax = g.axes
    return ChmyGrid(Int32(N), Int32(0), pad3(ntuple(d -> Int64(ax[d].length), Val(N)), Int64(0)),
                    pad3(ntuple(d -> Float64(ax[d].origin), Val(N)), 0.0),
                    pad3(ntuple(d -> Float64(ax[d].extent), Val(N)), 0.0),
                    pad3(ntuple(d -> Float64(ax[d].spacing), Val(N)), 0.0),
                    pad3(ntuple(d -> Float64(ax[d].inv_spacing), Val(N)), 0.0),
                    (NOAXIS, NOAXIS, NOAXIS))
end

# FunctionAxis (src/Grids/function_axis.jl): the metric table chmy_grid_t.axis
# points to, built once per axis from vertex(ax, i) with Chmy's own formulas
# (abstract_axis.jl:18,26-29, interpolation.jl:29) and kept on the device for
# as long as the axis lives.
const AXIS_PAD = 4
const AXIS_ROWS = 7
const AXIS_TABLES = IdDict{Any,MI355XArray{Float64,2}}()

function axis_table(ax::Chmy.Grids.FunctionAxis)
    get!(AXIS_TABLES, ax) do
        n = ax.length
        is = (-AXIS_PAD):(n + 1 + AXIS_PAD)
        v(i) = Float64(Chmy.Grids.vertex(ax, i))
        c(i) = 0.5 * (v(i) + v(i + 1))
        dc(i) = v(i + 1) - v(i)
        dv(i) = c(i) - c(i - 1)
        t = Matrix{Float64}(undef, length(is), AXIS_ROWS)   # column r+1 = row r of the C table
        for (q, i) in enumerate(is)
            t[q, 1] = v(i)
            t[q, 2] = c(i)
            t[q, 3] = dc(i)
            t[q, 4] = dv(i)
            t[q, 5] = inv(dc(i))
            t[q, 6] = inv(dv(i))
            t[q, 7] = 0.5 * dc(i - 1) * inv(dv(i))
        end
        dev = MI355XArray{Float64}(undef, size(t))
        copyto!(dev, t)
        dev
    end
end
axis_ptr(::Chmy.Grids.UniformAxis) = NOAXIS
axis_ptr(ax::Chmy.Grids.FunctionAxis) = pointer(axis_table(ax))
_uni(ax::Chmy.Grids.UniformAxis, f) = Float64(getfield(ax, f))
_uni(::Chmy.Grids.FunctionAxis, f) = 0.0

# a StructuredGrid with FunctionAxis dims: accepted by the operators, bc! and set!
function desc(g::Chmy.Grids.StructuredGrid{N}) where {N}
    ax = g.axes
    return ChmyGrid(Int32(N), Int32(0), pad3(ntuple(d -> Int64(ax[d].length), Val(N)), Int64(0)),
                    pad3(ntuple(d -> Float64(Chmy.Grids.origin(ax[d], Vertex())), Val(N)), 0.0),
                    pad3(ntuple(d -> Float64(Chmy.Grids.extent(ax[d], Vertex())), Val(N)), 0.0),
                    pad3(ntuple(d -> _uni(ax[d], :spacing), Val(N)), 0.0),
                    pad3(ntuple(d -> _uni(ax[d], :inv_spacing), Val(N)), 0.0),
                    pad3(ntuple(d -> axis_ptr(ax[d]), Val(N)), NOAXIS))
end

# a grid descriptor for kernels launched without the grid (update_old!): the
# cell counts of a Center field; the copy reads no spacing
function center_grid(f::MI355XField{T,N}) where {T,N}
    n = pad3(Int64.(size(f)), Int64(0))
    o = pad3(ntuple(_ -> 0.0, Val(N)), 0.0)
    one3 = pad3(ntuple(_ -> 1.0, Val(N)), 0.0)
    return ChmyGrid(Int32(N), Int32(0), n, o, Float64.(n), one3, one3, (NOAXIS, NOAXIS, NOAXIS))
end

descs(fs::NamedTuple) = ChmyField[desc(f) for f in values(fs)]

# the Launcher's worksize (a tuple, or StaticSize on the bc= path,
# src/KernelLaunch.jl:154-157) and Offset (KA index 1 -> logical 1 + O)
_ws(ws::Tuple) = ws
_ws(ws::Integer) = (ws,)
_ws(::StaticSize{S}) where {S} = S

function box(ws, ::Offset{O}, ::Val{N}) where {O,N}
    w = _ws(ws)
    lo = ntuple(d -> Int64(1 + (length(O) == 1 ? O[1] : O[d])), Val(N))
    return ChmyBox(pad3(lo, Int64(0)), pad3(Int64.(w), Int64(1)))
end

# ------------------------------------------------------------------ kernels
"""
    Kernels

Callables replacing the examples' `@kernel` functions on the hot path.  Each
is called by a `Launcher` (or directly, for sia.jl's kernels) exactly as the
KernelAbstractions kernel it replaces.
"""
module Kernels
    import ..ChmyMI355X: HipKernel
    const compute_q! = HipKernel{:compute_q!}()
    const update_C! = HipKernel{:update_C!}()
    const update_old! = HipKernel{:update_old!}()
    const update_stress! = HipKernel{:update_stress!}()
    const update_velocity! = HipKernel{:update_velocity!}()
    const update_thermal_flux! = HipKernel{:update_thermal_flux!}()
    const update_thermal! = HipKernel{:update_thermal!}()
    const update_q! = HipKernel{:update_q!}()
    const update_H! = HipKernel{:update_H!}()
end

struct HipKernel{name} end

struct BoundKernel{name,W}
    ws::W
end

(::HipKernel{name})(::MI355XBackend, groupsize, ws) where {name} = BoundKernel{name,typeof(ws)}(ws)

# k(args..., offset) from the Launcher; sia.jl calls its kernels with no offset
function (k::BoundKernel{name})(args...) where {name}
    if !isempty(args) && last(args) isa Offset
        return run_kernel(HipKernel{name}(), k.ws, Base.front(args), last(args))
    end
    return run_kernel(HipKernel{name}(), k.ws, args, Offset(0))
end

_nd(::UniformGrid{N}) where {N} = N

# compute_q!(q, C, χ, g) (examples/diffusion_2d.jl:8-13)
function run_kernel(::HipKernel{:compute_q!}, ws, (q, C, χ, g), off)
    N = _nd(g)
    check(ccall((:chmy_diffusion_flux, LIB), Cint,
                (Ref{ChmyGrid}, Ref{ChmyBox}, Ptr{ChmyField}, Ref{ChmyField}, Cdouble, Ptr{Cvoid}),
                desc(g), box(ws, off, Val(N)), descs(q), desc(C), Float64(χ), current_stream()))
end

# update_C!(C, q, Δt, g) (examples/diffusion_2d.jl:15-19)
function run_kernel(::HipKernel{:update_C!}, ws, (C, q, Δt, g), off)
    N = _nd(g)
    check(ccall((:chmy_diffusion_update, LIB), Cint,
                (Ref{ChmyGrid}, Ref{ChmyBox}, Ref{ChmyField}, Ptr{ChmyField}, Cdouble, Ptr{Cvoid}),
                desc(g), box(ws, off, Val(N)), desc(C), descs(q), Float64(Δt), current_stream()))
end

# update_old!(T, τ, T_old, τ_old) (examples/stokes_3d_inc_ve_T.jl:11-21); no grid argument
function run_kernel(::HipKernel{:update_old!}, ws, (T, τ, T_old, τ_old), off)
    N = ndims(T)
    check(ccall((:chmy_stokes_update_old, LIB), Cint,
                (Ref{ChmyGrid}, Ref{ChmyBox}, Ref{ChmyField}, Ptr{ChmyField}, Ref{ChmyField}, Ptr{ChmyField},
                 Ptr{Cvoid}),
                center_grid(T), box(ws, off, Val(N)), desc(T), descs(τ), desc(T_old), descs(τ_old),
                current_stream()))
end

# η: the reference's number (eta == NULL), or a Center Field (build extension)
_eta_s(η::Number) = Float64(η)
_eta_s(η::Field) = 1.0

# update_stress!(τ, Pr, ∇V, V, τ_old, η, η_ve, G, dt, dτ_Pr, dτ_r, g) (stokes_3d_inc_ve_T.jl:23-46)
function run_kernel(::HipKernel{:update_stress!}, ws, (τ, Pr, ∇V, V, τ_old, η, η_ve, G, dt, dτ_Pr, dτ_r, g), off)
    N = _nd(g)
    ηd = η isa Number ? nothing : Ref(desc(η))
    GC.@preserve ηd begin
        ηp = ηd === nothing ? Ptr{ChmyField}(C_NULL) : Base.unsafe_convert(Ptr{ChmyField}, ηd)
        check(ccall((:chmy_stokes_stress, LIB), Cint,
                    (Ref{ChmyGrid}, Ref{ChmyBox}, Ptr{ChmyField}, Ref{ChmyField}, Ref{ChmyField}, Ptr{ChmyField},
                     Ptr{ChmyField}, Ptr{ChmyField}, Cdouble, Cdouble, Cdouble, Cdouble, Cdouble, Cdouble, Ptr{Cvoid}),
                    desc(g), box(ws, off, Val(N)), descs(τ), desc(Pr), desc(∇V), descs(V), descs(τ_old), ηp,
                    _eta_s(η), Float64(η_ve), Float64(G), Float64(dt), Float64(dτ_Pr), Float64(dτ_r),
                    current_stream()))
    end
end

# update_velocity!(V, r_V, Pr, τ, ρg, η_ve, νdτ, g) (stokes_3d_inc_ve_T.jl:48-57; 2D :36-43);
# ρg a Field at V's last component's location, or 0
function run_kernel(::HipKernel{:update_velocity!}, ws, (V, r_V, Pr, τ, ρg, η_ve, νdτ, g), off)
    N = _nd(g)
    ρd = ρg isa Number ? nothing : Ref(desc(ρg))
    (ρg isa Number && !iszero(ρg)) && throw(ArgumentError("update_velocity!: ρg is a Field or 0"))
    GC.@preserve ρd begin
        ρp = ρd === nothing ? Ptr{ChmyField}(C_NULL) : Base.unsafe_convert(Ptr{ChmyField}, ρd)
        check(ccall((:chmy_stokes_velocity, LIB), Cint,
                    (Ref{ChmyGrid}, Ref{ChmyBox}, Ptr{ChmyField}, Ptr{ChmyField}, Ref{ChmyField}, Ptr{ChmyField},
                     Ptr{ChmyField}, Ptr{ChmyField}, Cdouble, Cdouble, Cdouble, Cdouble, Ptr{Cvoid}),
                    desc(g), box(ws, off, Val(N)), descs(V), descs(r_V), desc(Pr), descs(τ), ρp,
                    Ptr{ChmyField}(C_NULL), Float64(η_ve), 1.0, 1.0, Float64(νdτ), current_stream()))
    end
end

# update_thermal_flux!(qT, T, V, λ_ρCp, g) (stokes_3d_inc_ve_T.jl:59-71)
function run_kernel(::HipKernel{:update_thermal_flux!}, ws, (qT, T, V, λ_ρCp, g), off)
    N = _nd(g)
    check(ccall((:chmy_thermal_flux, LIB), Cint,
                (Ref{ChmyGrid}, Ref{ChmyBox}, Ptr{ChmyField}, Ref{ChmyField}, Ptr{ChmyField}, Cdouble, Ptr{Cvoid}),
                desc(g), box(ws, off, Val(N)), descs(qT), desc(T), descs(V), Float64(λ_ρCp), current_stream()))
end

# update_thermal!(T, T_old, qT, dt, g) (stokes_3d_inc_ve_T.jl:73-77)
function run_kernel(::HipKernel{:update_thermal!}, ws, (T, T_old, qT, dt, g), off)
    N = _nd(g)
    check(ccall((:chmy_thermal_update, LIB), Cint,
                (Ref{ChmyGrid}, Ref{ChmyBox}, Ref{ChmyField}, Ref{ChmyField}, Ptr{ChmyField}, Cdouble, Ptr{Cvoid}),
                desc(g), box(ws, off, Val(N)), desc(T), desc(T_old), descs(qT), Float64(dt), current_stream()))
end

# ------------------------------------------------------------------ SIA (examples/sia.jl)
# sia.jl builds B, S, ∇S, D and M as FunctionFields of Julia closures.  The
# device evaluates them inside its kernels from H, the bed B (a Field on the
# full storage, halos included) and the constants, so the binding needs those
# numbers: build D and M with `sia_diffusivity` and `mass_balance`, which
# return the same FunctionFields (Chmy evaluates them on the CPU as before)
# and record their constants.
struct SIAConsts
    As::Float64
    npow::Int
end
struct MBConsts
    a::Float64
    b::Float64
    c::Float64
end
const SIA_D = IdDict{Any,SIAConsts}()
const SIA_M = IdDict{Any,MBConsts}()
const BED_CACHE = IdDict{Any,Any}()

"""D = (H^(n+2) + As H^n) |∇S|^(n-1) at centres (sia.jl:46-49)."""
function sia_diffusivity(grid, H, ∇S; As, n)
    D = FunctionField(grid, Center(); discrete=true, parameters=(H, ∇S, As, n)) do g, l, ix, iy, H, ∇S, As, n
        @inline
        return (H[ix, iy]^(n + 2) + As * H[ix, iy]^n) * vmag(∇S, g, ix, iy)^(n - 1)
    end
    SIA_D[D] = SIAConsts(Float64(As), Int(n))
    return D
end

"""M = min(a (S - b), c) at centres (sia.jl:51-54)."""
function mass_balance(grid, S; a=0.5, b=0.3, c=0.25)
    M = FunctionField(grid, Center(); discrete=true, parameters=(S,)) do g, l, ix, iy, S
        @inline
        return min(a * (S[ix, iy] - b), c)
    end
    SIA_M[M] = MBConsts(Float64(a), Float64(b), Float64(c))
    return M
end

# the bed of S = B + H on the device, over the full padded storage (halos
# included: the flux stencil reads them), evaluated by Chmy on the CPU once
function device_bed(B, H::MI355XField{T,N}) where {T,N}
    B isa MI355XField && return B
    return get!(BED_CACHE, B) do
        grid = B.grid
        hb = Field(CPU(), grid, Center(), T)
        h = halo(hb)
        for iy in (1 - 2h):(size(hb, 2) + 2h), ix in (1 - 2h):(size(hb, 1) + 2h)
            hb[ix, iy] = B[ix, iy]
        end
        db = Field(MI355XBackend(), grid, Center(), T)
        copyto!(parent(db), parent(hb))
        db
    end
end

# update_q!(q, D, S, g) (sia.jl:12-16): q at vertices 1..n
function run_kernel(::HipKernel{:update_q!}, ws, (q, D, S, g), off)
    k = get(SIA_D, D) do
        throw(ArgumentError("update_q!: build D with ChmyMI355X.sia_diffusivity"))
    end
    B, H = S.parameters
    Bd = device_bed(B, H)
    check(ccall((:chmy_sia_flux, LIB), Cint,
                (Ref{ChmyGrid}, Ptr{ChmyField}, Ref{ChmyField}, Ref{ChmyField}, Ptr{ChmyField}, Cdouble, Cint,
                 Ptr{Cvoid}),
                desc(g), descs(q), desc(H), desc(Bd), Ptr{ChmyField}(C_NULL), k.As, Cint(k.npow), current_stream()))
end

# update_H!(H, q, M, Δt, g) (sia.jl:7-10): cells 1..n, in place
function run_kernel(::HipKernel{:update_H!}, ws, (H, q, M, Δt, g), off)
    m = get(SIA_M, M) do
        throw(ArgumentError("update_H!: build M with ChmyMI355X.mass_balance"))
    end
    S = M.parameters[1]
    B, _ = S.parameters
    Bd = device_bed(B, H)
    check(ccall((:chmy_sia_update, LIB), Cint,
                (Ref{ChmyGrid}, Ref{ChmyField}, Ref{ChmyField}, Ptr{ChmyField}, Ref{ChmyField}, Cdouble, Cdouble,
                 Cdouble, Cdouble, Ptr{Cvoid}),
                desc(g), desc(H), desc(H), descs(q), desc(Bd), m.a, m.b, m.c, Float64(Δt), current_stream()))
end

# ------------------------------------------------------------------ bc!
kindcode(::FirstOrderBC{T,Chmy.BoundaryConditions.DirichletKind}) where {T} = CHMY_DIRICHLET
kindcode(::FirstOrderBC{T,Chmy.BoundaryConditions.NeumannKind}) where {T} = CHMY_NEUMANN

# a BoundaryFunction value (boundary_function.jl:34-44), evaluated by Chmy on the
# CPU at the point and location first_order_boundary_condition.jl:42-84 uses,
# over the tangential indices 0..n+2, into a device field of the tangential grid
function materialise(grid::UniformGrid{N}, f::Field, D::Int, S::Int, bc) where {N}
    locs = location(f)
    m = size(f, D)
    isdir = kindcode(bc) == CHMY_DIRICHLET
    if isdir && locs[D] isa Vertex
        loc, normal = locs, (S == 1 ? 1 : m)
    elseif isdir
        loc, normal = ntuple(_ -> Center(), Val(N)), (S == 1 ? 0 : m + 1)
    else
        loc, normal = map(flip, locs), (S == 1 ? 0 : m + 1)
    end
    tang = Tuple(t for t in 1:N if t != D)
    tgrid = UniformGrid(Arch(CPU()); origin=ntuple(q -> grid.axes[tang[q]].origin, N - 1),
                        extent=ntuple(q -> grid.axes[tang[q]].extent, N - 1),
                        dims=ntuple(q -> grid.axes[tang[q]].length, N - 1))
    hv = Field(CPU(), tgrid, Center(), eltype(f))
    for J in CartesianIndices(ntuple(q -> 0:(grid.axes[tang[q]].length + 2), N - 1))
        I = ntuple(d -> d == D ? normal : J[findfirst(==(d), tang)], Val(N))
        hv[Tuple(J)...] = Chmy.BoundaryConditions.value(bc, grid, loc, Dim(D), I...)
    end
    dv = Field(MI355XBackend(), tgrid, Center(), eltype(f))
    copyto!(parent(dv), parent(hv))
    return dv
end

function bcdesc!(keep, grid, f, D, S, bc::FirstOrderBC)
    v = bc.value
    if v === nothing
        return ChmyBC(kindcode(bc), CHMY_VALUE_ZERO, 0.0, Ptr{ChmyField}(C_NULL))
    elseif v isa Number
        return ChmyBC(kindcode(bc), CHMY_VALUE_SCALAR, Float64(v), Ptr{ChmyField}(C_NULL))
    end
    vf = v isa BoundaryFunction ? materialise(grid, f, D, S, bc) : v
    r = Ref(desc(vf))
    push!(keep, r, vf)
    return ChmyBC(kindcode(bc), CHMY_VALUE_FIELD, 0.0, Base.unsafe_convert(Ptr{ChmyField}, r))
end

# bc!(arch, grid, ::BatchSet) (batch.jl:20-29): the whole set in one call
function bc!(arch::MI355XArch, grid::UniformGrid{N}, batch::BatchSet{N}) where {N}
    counts = zeros(Int32, 6)
    fs = ChmyField[]
    cs = ChmyBC[]
    keep = Any[]
    for D in 1:N, S in 1:2
        b = batch[D][S]
        b isa FieldBatch || continue   # EmptyBatch; ExchangeBatch needs a distributed grid
        for (f, c) in zip(b.fields, b.conditions)
            push!(fs, desc(f))
            push!(cs, bcdesc!(keep, grid, f, D, S, c))
        end
        counts[2D+S-2] = length(b.fields)
    end
    GC.@preserve keep begin
        check(ccall((:chmy_bc_batchset, LIB), Cint,
                    (Ref{ChmyGrid}, Ptr{Int32}, Ptr{ChmyField}, Ptr{ChmyBC}, Ptr{Cvoid}),
                    desc(grid), counts, fs, cs, current_stream()))
    end
    return nothing
end

# bc!(side, dim, arch, grid, ::FieldBatch) (batch.jl:180-184), the Launcher's
# outer_width path (KernelLaunch.jl:173)
function bc!(::Side{S}, ::Dim{D}, arch::MI355XArch, grid::UniformGrid, batch::FieldBatch) where {S,D}
    keep = Any[]
    fs = ChmyField[desc(f) for f in batch.fields]
    cs = ChmyBC[bcdesc!(keep, grid, f, D, S, c) for (f, c) in zip(batch.fields, batch.conditions)]
    GC.@preserve keep begin
        check(ccall((:chmy_bc_side, LIB), Cint,
                    (Ref{ChmyGrid}, Cint, Cint, Cint, Ptr{ChmyField}, Ptr{ChmyBC}, Ptr{Cvoid}),
                    desc(grid), Cint(D - 1), Cint(S), Cint(length(fs)), fs, cs, current_stream()))
    end
    return nothing
end

# ------------------------------------------------------------------ set! (field.jl:86-145)
# Chmy's own CPU set! into a host copy of the storage, then one upload: the
# values are Chmy's CPU values bit for bit.
function _via_host!(f::MI355XField{T,N}, setter) where {T,N}
    hf = Field{typeof(location(f)),halo(f)}(Array(parent(f)), size(f))
    setter(hf)
    copyto!(parent(f), parent(hf))
    return nothing
end

set!(f::MI355XField, val::Number) = _via_host!(f, hf -> set!(hf, val))
set!(f::MI355XField, A::AbstractArray) = _via_host!(f, hf -> set!(hf, A))
set!(f::MI355XField, other::AbstractField) = _via_host!(f, hf -> set!(hf, other))
set!(f::MI355XField{T,N}, grid::StructuredGrid{N}, fun::F; discrete=false, parameters=()) where {T,N,F} =
    _via_host!(f, hf -> set!(hf, grid, fun; discrete=discrete, parameters=parameters))

# ------------------------------------------------------------------ fused paths and adjoints
"""
    iterate!(state0, state1, τ_old, η, η_ml, grid, bc_V, niter, p; flags=0) -> parity

`niter` fused PT iterations (update_stress! then update_velocity! then the V
batch, stokes_3d_inc_ve_T.jl:164-165) in one call, double-buffered between
`state0 = (V, τ, Pr)` and `state1`; returns the set (0 or 1) holding the new
state.  `bc_V` is the example's `(V.x => (x=Dirichlet(..), ...), ...)` list;
`p` holds G, dt, dτ_Pr, dτ_r, νdτ and, for a scalar viscosity, η and η_ve.
"""
# the V batch of bc_V as (counts per (dim, side), V component per condition,
# condition descriptors, value fields to keep alive)
function vbatch(grid::UniformGrid{N}, bc_V, V0) where {N}
    comps = Int32[]
    cs = ChmyBC[]
    counts = zeros(Int32, 6)
    keep = Any[]
    bs = batch(grid, bc_V...)
    for D in 1:N, S in 1:2
        b = bs[D][S]
        b isa FieldBatch || continue
        for (f, c) in zip(b.fields, b.conditions)
            push!(comps, Int32(findfirst(x -> x === f, values(V0)) - 1))
            push!(cs, bcdesc!(keep, grid, f, D, S, c))
        end
        counts[2D+S-2] = length(b.fields)
    end
    return counts, comps, cs, keep
end

function iterate!(state0, state1, τ_old, η, η_ml, grid::UniformGrid{N}, bc_V, niter::Integer, p;
                  flags::Integer=0, flags_last::Integer=flags) where {N}
    (V0, τ0, P0), (V1, τ1, P1) = state0, state1
    counts, comps, cs, keep = vbatch(grid, bc_V, V0)
    par = Ref{Cint}(0)
    ηd = η isa Number ? nothing : Ref(desc(η))
    md = η_ml === nothing ? nothing : Ref(desc(η_ml))
    GC.@preserve keep ηd md begin
        ηp = ηd === nothing ? Ptr{ChmyField}(C_NULL) : Base.unsafe_convert(Ptr{ChmyField}, ηd)
        mp = md === nothing ? Ptr{ChmyField}(C_NULL) : Base.unsafe_convert(Ptr{ChmyField}, md)
        check(ccall((:chmy_stokes_iterate, LIB), Cint,
                    (Ref{ChmyGrid}, Ptr{ChmyField}, Ptr{ChmyField}, Ref{ChmyField}, Ptr{ChmyField}, Ptr{ChmyField},
                     Ref{ChmyField}, Ptr{ChmyField}, Ptr{ChmyField}, Ptr{ChmyField}, Ptr{ChmyField}, Ptr{ChmyField},
                     Ptr{ChmyField}, Ptr{ChmyField}, Ptr{ChmyField}, Ptr{Int32}, Ptr{Int32}, Ptr{ChmyBC}, Cint, Cint,
                     Cint, Cdouble, Cdouble, Cdouble, Cdouble, Cdouble, Cdouble, Cdouble, Ref{Cint}, Ptr{Cvoid}),
                    desc(grid), descs(V0), descs(τ0), desc(P0), descs(V1), descs(τ1), desc(P1), descs(τ_old), ηp, mp,
                    Ptr{ChmyField}(C_NULL), Ptr{ChmyField}(C_NULL), Ptr{ChmyField}(C_NULL),
                    Ptr{ChmyField}(C_NULL), Ptr{ChmyField}(C_NULL), counts, comps, cs, Cint(niter), Cint(flags),
                    Cint(flags_last), _eta_s(η), Float64(p.η_ve), Float64(p.G), Float64(p.dt), Float64(p.dτ_Pr),
                    Float64(p.dτ_r), Float64(p.νdτ), par, current_stream()))
    end
    return Int(par[])
end

"""
    iter2!(state0, state1, τ_old, η, η_ml, grid, bc_V, p; flags=0)

Two PT iterations in one pass from `state0` into `state1`, the free-slip V
batch applied in-kernel between them (bitwise two `iterate!` steps); the
caller applies the V batch to `state1` and swaps, as with one iteration.
`bc_V` must be the free-slip batch (each V component Dirichlet on its own
dim, Neumann on the others, scalar values).
"""
function iter2!(state0, state1, τ_old, η, η_ml, grid::UniformGrid{N}, bc_V, p; flags::Integer=0) where {N}
    (V0, τ0, P0), (V1, τ1, P1) = state0, state1
    counts, comps, cs, keep = vbatch(grid, bc_V, V0)
    ηd = η isa Number ? nothing : Ref(desc(η))
    md = η_ml === nothing ? nothing : Ref(desc(η_ml))
    GC.@preserve keep ηd md begin
        ηp = ηd === nothing ? Ptr{ChmyField}(C_NULL) : Base.unsafe_convert(Ptr{ChmyField}, ηd)
        mp = md === nothing ? Ptr{ChmyField}(C_NULL) : Base.unsafe_convert(Ptr{ChmyField}, md)
        check(ccall((:chmy_stokes_iter2, LIB), Cint,
                    (Ref{ChmyGrid}, Ptr{ChmyField}, Ptr{ChmyField}, Ref{ChmyField}, Ptr{ChmyField}, Ptr{ChmyField},
                     Ref{ChmyField}, Ptr{ChmyField}, Ptr{ChmyField}, Ptr{ChmyField}, Ptr{ChmyField}, Ptr{ChmyField},
                     Ptr{ChmyField}, Ptr{Int32}, Ptr{Int32}, Ptr{ChmyBC}, Cint, Cdouble, Cdouble, Cdouble, Cdouble,
                     Cdouble, Cdouble, Cdouble, Ptr{Cvoid}),
                    desc(grid), descs(V0), descs(τ0), desc(P0), descs(V1), descs(τ1), desc(P1), descs(τ_old), ηp, mp,
                    Ptr{ChmyField}(C_NULL), Ptr{ChmyField}(C_NULL), Ptr{ChmyField}(C_NULL), counts, comps, cs,
                    Cint(flags), _eta_s(η), Float64(p.η_ve), Float64(p.G), Float64(p.dt), Float64(p.dτ_Pr),
                    Float64(p.dτ_r), Float64(p.νdτ), current_stream()))
    end
    return state1
end

"""
    iter2_first!(state0, state1, V_first, τ_old, η, η_ml, grid, bc_V, p; rhog=nothing, flags=0)

`iter2!` that also stores the first iteration's V (before its V batch) into
`V_first` (2D or 3D, Float64; build extension).  After the V batch on `V_first`
and on `state1`'s V, `V_first` is bitwise V after the first of the two
iterations: the V that the first iteration's `update_thermal_flux!` reads in
the physical time loop (stokes_3d_inc_ve_T.jl:164-169, stokes_2d_inc_ve_T.jl:145-151).
`rhog` is the `ρgz` (2D: `ρgy`) field update_velocity! subtracts.
"""
function iter2_first!(state0, state1, V_first, τ_old, η, η_ml, grid::UniformGrid{N}, bc_V, p;
                      rhog=nothing, flags::Integer=0) where {N}
    (V0, τ0, P0), (V1, τ1, P1) = state0, state1
    counts, comps, cs, keep = vbatch(grid, bc_V, V0)
    ηd = η isa Number ? nothing : Ref(desc(η))
    md = η_ml === nothing ? nothing : Ref(desc(η_ml))
    rd = rhog === nothing ? nothing : Ref(desc(rhog))
    GC.@preserve keep ηd md rd begin
        ηp = ηd === nothing ? Ptr{ChmyField}(C_NULL) : Base.unsafe_convert(Ptr{ChmyField}, ηd)
        mp = md === nothing ? Ptr{ChmyField}(C_NULL) : Base.unsafe_convert(Ptr{ChmyField}, md)
        rp = rd === nothing ? Ptr{ChmyField}(C_NULL) : Base.unsafe_convert(Ptr{ChmyField}, rd)
        check(ccall((:chmy_stokes_iter2_first, LIB), Cint,
                    (Ref{ChmyGrid}, Ptr{ChmyField}, Ptr{ChmyField}, Ref{ChmyField}, Ptr{ChmyField}, Ptr{ChmyField},
                     Ref{ChmyField}, Ptr{ChmyField}, Ptr{ChmyField}, Ptr{ChmyField}, Ptr{ChmyField}, Ptr{ChmyField},
                     Ptr{Int32}, Ptr{Int32}, Ptr{ChmyBC}, Cint, Cdouble, Cdouble, Cdouble, Cdouble, Cdouble, Cdouble,
                     Cdouble, Ptr{Cvoid}),
                    desc(grid), descs(V0), descs(τ0), desc(P0), descs(V1), descs(τ1), desc(P1), descs(V_first),
                    descs(τ_old), ηp, mp, rp, counts, comps, cs, Cint(flags), _eta_s(η), Float64(p.η_ve),
                    Float64(p.G), Float64(p.dt), Float64(p.dτ_Pr), Float64(p.dτ_r), Float64(p.νdτ),
                    current_stream()))
    end
    return state1
end

"""
    stokes_residual!(x, τ_old, η, grid, p, forcing; rhog=nothing, divV=nothing, r_V=nothing)

Defect correction, first half (build extension): the f64 residual of the
state `x = (V, τ, Pr)` in the reference's expression order
(stokes_3d_inc_ve_T.jl:23-57), written as the correction sweeps' forcing
`forcing = (τ_old_c, f_V, f_P)` (τ_old_c = G dt R_τ, f_V = R_V, f_P = R_P; f32
or bf16 fields); with `divV` and `r_V` also ∇V and r_V in f64 for the
convergence check.  The sweeps are `iterate!` on a correction state from 0
with that forcing (flags as usual, the forcing in `τ_old`).
"""
function stokes_residual!(x, τ_old, η, grid::UniformGrid, p, forcing; rhog=nothing, divV=nothing, r_V=nothing)
    V, τ, Pr = x
    τc, fV, fP = forcing
    ηd = η isa Number ? nothing : Ref(desc(η))
    gd = rhog === nothing ? nothing : Ref(desc(rhog))
    dd = divV === nothing ? nothing : Ref(desc(divV))
    rd = r_V === nothing ? nothing : descs(r_V)
    GC.@preserve ηd gd dd rd begin
        ηp = ηd === nothing ? Ptr{ChmyField}(C_NULL) : Base.unsafe_convert(Ptr{ChmyField}, ηd)
        gp = gd === nothing ? Ptr{ChmyField}(C_NULL) : Base.unsafe_convert(Ptr{ChmyField}, gd)
        dp = dd === nothing ? Ptr{ChmyField}(C_NULL) : Base.unsafe_convert(Ptr{ChmyField}, dd)
        rp = rd === nothing ? Ptr{ChmyField}(C_NULL) : pointer(rd)
        check(ccall((:chmy_stokes_residual, LIB), Cint,
                    (Ref{ChmyGrid}, Ptr{ChmyField}, Ptr{ChmyField}, Ref{ChmyField}, Ptr{ChmyField}, Ptr{ChmyField},
                     Ptr{ChmyField}, Ptr{ChmyField}, Cdouble, Cdouble, Cdouble, Ptr{ChmyField}, Ptr{ChmyField},
                     Ref{ChmyField}, Ptr{ChmyField}, Ptr{ChmyField}, Ptr{Cvoid}),
                    desc(grid), descs(V), descs(τ), desc(Pr), descs(τ_old), ηp, gp, Ptr{ChmyField}(C_NULL),
                    _eta_s(η), Float64(p.G), Float64(p.dt), descs(τc), descs(fV), desc(fP), dp, rp,
                    current_stream()))
    end
    return forcing
end

"""
    stokes_accumulate!(x, d, grid)

Defect correction, second half: `x .+= d` over the Launcher box for
`x = (V, τ, Pr)` in f64 and the correction `d` (f32 or f64).  The real V
batch on `x` follows, as after any iteration.
"""
function stokes_accumulate!(x, d, grid::UniformGrid)
    (V, τ, Pr), (Vc, τc, Pc) = x, d
    check(ccall((:chmy_stokes_accumulate, LIB), Cint,
                (Ref{ChmyGrid}, Ptr{ChmyField}, Ptr{ChmyField}, Ref{ChmyField}, Ptr{ChmyField}, Ptr{ChmyField},
                 Ref{ChmyField}, Ptr{Cvoid}),
                desc(grid), descs(V), descs(τ), desc(Pr), descs(Vc), descs(τc), desc(Pc), current_stream()))
    return x
end

"""
    maxabs(f; box=nothing) -> Float64

`maximum(abs.(interior(f)))` on the device (the examples' residual norms,
stokes_3d_inc_ve_T.jl:172-175), or over a box of logical indices.
"""
function maxabs(f::MI355XField; box=nothing)
    out = Ref{Cdouble}(0.0)
    bd = box === nothing ? nothing : Ref(box)
    GC.@preserve bd begin
        bp = bd === nothing ? Ptr{ChmyBox}(C_NULL) : Base.unsafe_convert(Ptr{ChmyBox}, bd)
        check(ccall((:chmy_maxabs, LIB), Cint, (Ref{ChmyField}, Ptr{ChmyBox}, Ref{Cdouble}, Ptr{Cvoid}),
                    desc(f), bp, out, current_stream()))
    end
    return out[]
end

# ------------------------------------------------------------------ operators
const CHMY_OP_PARTIAL = Cint(0)
const CHMY_OP_LERP = Cint(1)
const CHMY_OP_DIVG = Cint(2)
const CHMY_OP_VMAG = Cint(3)
const CHMY_OP_D2 = Cint(4)
const CHMY_OP_DKD = Cint(5)
const CHMY_OP_LAPL = Cint(6)
const CHMY_OP_DIVG_GRAD = Cint(7)
const CHMY_OP_LEFT = Cint(8)
const CHMY_OP_RIGHT = Cint(9)
const CHMY_OP_HLERP = Cint(10)
const CHMY_OP_DELTA = Cint(11)

"""
    apply_op!(op, dst, src, grid; dim=1, masks=nothing, k=nothing, box=nothing)

One GridOperator evaluated into `dst` over a box (the GPU form of the
operator evaluations of test/test_grid_operators.jl): `op` a `CHMY_OP_*`
code, `src` a Field or a tuple of Fields (divg, vmag: the vector components),
`dim` 1-based, `masks` the FieldMask component at each source's location
(`at(ω, location(src))`, masked_operators.jl:11-19) for the masked forms,
`k` the coefficient of ∂k∂ and divg_grad.  `grid` may be a UniformGrid or a
StructuredGrid with FunctionAxis dims.  The default box is the Launcher's
[0, n+1] per dim.
"""
function apply_op!(op::Cint, dst::MI355XField, src, grid; dim::Integer=1, masks=nothing, k=nothing, box=nothing)
    sd = ChmyField[desc(f) for f in (src isa Tuple ? src : (src,))]
    md = masks === nothing ? nothing : ChmyField[desc(m) for m in (masks isa Tuple ? masks : (masks,))]
    kd = k === nothing ? nothing : Ref(desc(k))
    bd = box === nothing ? nothing : Ref(box)
    GC.@preserve md kd bd begin
        bp = bd === nothing ? Ptr{ChmyBox}(C_NULL) : Base.unsafe_convert(Ptr{ChmyBox}, bd)
        if md === nothing && kd === nothing
            check(ccall((:chmy_op, LIB), Cint,
                        (Cint, Ref{ChmyGrid}, Ptr{ChmyBox}, Ref{ChmyField}, Cint, Ptr{ChmyField}, Cint, Ptr{Cvoid}),
                        op, desc(grid), bp, desc(dst), Cint(length(sd)), sd, Cint(dim - 1), current_stream()))
        else
            mp = md === nothing ? Ptr{ChmyField}(C_NULL) : pointer(md)
            kp = kd === nothing ? Ptr{ChmyField}(C_NULL) : Base.unsafe_convert(Ptr{ChmyField}, kd)
            check(ccall((:chmy_op_masked, LIB), Cint,
                        (Cint, Ref{ChmyGrid}, Ptr{ChmyBox}, Ref{ChmyField}, Cint, Ptr{ChmyField}, Ptr{ChmyField},
                         Ptr{ChmyField}, Cint, Ptr{Cvoid}),
                        op, desc(grid), bp, desc(dst), Cint(length(sd)), sd, mp, kp, Cint(dim - 1),
                        current_stream()))
        end
    end
    return dst
end

"""
    sia_run!(states, B, logA, As, npow, mb, dts, grid)

The SIA time loop (sia.jl:71-96) with the given Δt sequence: `states[q+1]`
from `states[q]` by one step and the H Neumann batch; with the checkpoints of
a reverse sweep as some of the states, they are written in place.
"""
function sia_run!(states::Vector, B, A, As, npow, mb, dts::Vector{Float64}, grid)
    Ad = A === nothing ? nothing : Ref(desc(A))
    GC.@preserve Ad begin
        Ap = Ad === nothing ? Ptr{ChmyField}(C_NULL) : Base.unsafe_convert(Ptr{ChmyField}, Ad)
        check(ccall((:chmy_sia_run, LIB), Cint,
                    (Ref{ChmyGrid}, Ptr{ChmyField}, Cint, Ref{ChmyField}, Ptr{ChmyField}, Cdouble, Cint, Cdouble,
                     Cdouble, Cdouble, Ptr{Cdouble}, Ptr{Cvoid}),
                    desc(grid), ChmyField[desc(s) for s in states], Cint(length(states)), desc(B), Ap, Float64(As),
                    Cint(npow), mb.a, mb.b, mb.c, dts, current_stream()))
    end
    return states[end]
end

"""
    sia_reverse!(ckpt, stride, seg, λa, λb, grad_logA, B, A, As, npow, mb, dts, grid) -> λ

Reverse sweep of a checkpointed SIA run: on entry `λa` holds dJ/dH_N; returns
the field holding dJ/dH_0 and adds dJ/dlog A into `grad_logA`.
"""
function sia_reverse!(ckpt::Vector, stride::Integer, seg::Vector, λa, λb, grad_logA, B, A, As, npow, mb,
                      dts::Vector{Float64}, grid)
    slot = Ref{Cint}(0)
    check(ccall((:chmy_sia_reverse, LIB), Cint,
                (Ref{ChmyGrid}, Ptr{ChmyField}, Cint, Cint, Cint, Ptr{ChmyField}, Ref{ChmyField}, Ref{ChmyField},
                 Ref{ChmyField}, Ref{ChmyField}, Ref{ChmyField}, Cdouble, Cint, Cdouble, Cdouble, Cdouble,
                 Ptr{Cdouble}, Ref{Cint}, Ptr{Cvoid}),
                desc(grid), ChmyField[desc(c) for c in ckpt], Cint(length(ckpt)), Cint(stride), Cint(length(dts)),
                ChmyField[desc(s) for s in seg], desc(λa), desc(λb), desc(grad_logA), desc(B), desc(A), Float64(As),
                Cint(npow), mb.a, mb.b, mb.c, dts, slot, current_stream()))
    return slot[] == 0 ? λa : λb
end

"""
    SIAProblem

Everything the SIA time loop reads besides the state: the grid, the bed
(a Field with halos), the constants, the Δt sequence, and caller-owned
device buffers for checkpoints, the segment recompute and the adjoint.
"""
struct SIAProblem{G,F}
    grid::G
    B::F
    As::Float64
    npow::Int
    mb::MBConsts
    dts::Vector{Float64}
    stride::Int
    ckpt::Vector{F}      # ceil(nsteps / stride) states
    seg::Vector{F}       # 2 (stride - 1) states
    work::Vector{F}      # 2 states for the steps between checkpoints
    λ::Vector{F}         # 2 adjoint states
end

# the forward states: checkpoint s at step s*stride, work fields between
function sia_states(p::SIAProblem, H0)
    n = length(p.dts)
    states = Vector{Any}(undef, n + 1)
    for k in 0:n
        states[k+1] = k == 0 ? H0 : (k % p.stride == 0 && k ÷ p.stride < length(p.ckpt)) ?
                      p.ckpt[k÷p.stride+1] : p.work[k%2+1]
    end
    return states
end

"""
    sia_solve(H0, logA, p::SIAProblem) -> H_N

The SIA run from H0 with rate factor A = exp(logA) (a Center Field), for
`ChainRulesCore.rrule`.  H0 is copied into the first checkpoint.
"""
function sia_solve(H0, A, p::SIAProblem)
    copyto!(parent(p.ckpt[1]), Array(parent(H0)))
    states = sia_states(p, p.ckpt[1])
    return sia_run!(states, p.B, A, p.As, p.npow, p.mb, p.dts, p.grid)
end

function ChainRulesCore.rrule(::typeof(sia_solve), H0, A, p::SIAProblem)
    HN = sia_solve(H0, A, p)
    function sia_pullback(H̄N)
        λa, λb = p.λ
        copyto!(parent(λa), Array(parent(unthunk(H̄N))))
        fill!(parent(λb), 0)
        ḡ = Field(MI355XBackend(), p.grid, Center(), eltype(H0))
        λ0 = sia_reverse!(p.ckpt, p.stride, p.seg, λa, λb, ḡ, p.B, A, p.As, p.npow, p.mb, p.dts, p.grid)
        # ḡ is dJ/dlog A; the cotangent of A itself is ḡ ./ A, formed by the caller
        # when A is not parameterised by its logarithm
        return NoTangent(), λ0, ḡ, NoTangent()
    end
    return HN, sia_pullback
end

"""
    stokes_grad_eta!(grad, τ, τ_a, η, grid; logscale=true)

dJ/dlog η (or dJ/dη) at centres from the forward stresses τ and the adjoint
solve's stresses τ_a (the same PT problem with forcing f_V = -dJ/dV,
homogeneous boundary values and τ_old = 0).
"""
function stokes_grad_eta!(grad, τ, τ_a, η, grid; logscale::Bool=true)
    check(ccall((:chmy_stokes_grad_eta, LIB), Cint,
                (Ref{ChmyGrid}, Ptr{ChmyField}, Ptr{ChmyField}, Ref{ChmyField}, Ref{ChmyField}, Cint, Ptr{Cvoid}),
                desc(grid), descs(τ), descs(τ_a), desc(η), desc(grad), Cint(logscale), current_stream()))
    return grad
end

"""
    StokesSteady

A steady PT Stokes problem for `ChainRulesCore.rrule`: `solve` runs the
forward PT solve for a viscosity field and returns the converged state;
`adjoint_solve` runs the same solver with forcing f_V = -V̄, homogeneous
boundary values and τ_old = 0 and returns its stresses.  Both are caller
functions built on `iterate!` (the solvers' residual checks are the
caller's, as in stokes_3d_inc_ve_T.jl:170-180).
"""
struct StokesSteady{G,FS,FA}
    grid::G
    solve::FS
    adjoint_solve::FA
end

stokes_solve(η, s::StokesSteady) = s.solve(η)

function ChainRulesCore.rrule(::typeof(stokes_solve), η, s::StokesSteady)
    st = stokes_solve(η, s)
    function stokes_pullback(V̄)
        τ_a = s.adjoint_solve(η, unthunk(V̄))
        ḡ = Field(MI355XBackend(), s.grid, Center(), eltype(η))
        stokes_grad_eta!(ḡ, st.τ, τ_a, η, s.grid; logscale=false)
        return NoTangent(), ḡ, NoTangent()
    end
    return st, stokes_pullback
end

end # module
