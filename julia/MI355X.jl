"""
    MI355X

Julia binding of the MI355X (gfx950) stepper library `libvi_mi355x.so` through `ccall`:
the drop-in backend for ViscoelasticInversion's time-stepping path.

    using ViscoelasticInversion, MI355X
    p = MI355X.Parameters(Nx=1024, Nz=512, dx=5.0, dz=5.0, dt=4e-4, model=model,
                          qmodel=q, cpml=CPMLConfig(thickness=20, vmax=4500.0, dx=5.0),
                          free_surface=true, source=RickerSource(12.0, 1/12, 513, 7),
                          receivers=(ix, iz, :vz))
    sol = solve_problem(make_problem(p, tspan=(0.0, 1.6)))

What each function replaces (reference file:line):
  Parameters(; Nx, Nz, dx, dz, dt, fd_order_x, fd_order_z, bc, source, device, model, cpml,
             free_surface)                          src/sciml_solver/Parameters.jl:47-86
  make_problem(p; s0, v0, tspan)                    src/sciml_solver/ViscoelasticProblem.jl:273-292
  solve_problem(prob; solver)                       src/sciml_solver/ViscoelasticProblem.jl:307-309
  stress_eq!(ds, s, v, p, t), velocity_eq!(dv, s, v, p, t)
                                                    src/sciml_solver/ViscoelasticProblem.jl:62-174
  Fields(s, M, v), Parameters(λs, μs, τs, τns), leapfrog_step!(fields, parameters, dt)
                                                    src/visco-kernels.jl:87-113
and the build-defined additions: receivers and traces, the discrete-adjoint gradient with its
ChainRulesCore.rrule, batches of shots.

Layouts: a field is a Julia `Array{T}(undef, Nx, Nz)` (x fastest), which is the library's
[nz][nx] layout; s is (Nx, Nz, 3) = [σxx σzz σxz], v is (Nx, Nz, 2) = [vx vz], M is
(Nx, Nz, L, 3) = [xx zz xz]; traces and observed data are (nt, nrec). Grid indices are 1-based
here and converted to the library's 0-based indices. `T` is Float32 (the reference's GPU
precision) or Float64.

The library path defaults to the built package next to this file and can be set with the
environment variable VI_MI355X_LIB.
"""
module MI355X

using ViscoelasticInversion: ViscoelasticInversion, RickerSource
import ViscoelasticInversion: make_problem, solve_problem, stress_eq!, velocity_eq!
import ChainRulesCore

export MI355XBackend, Sim, Fields, StepParameters, leapfrog_step!, trace_misfit,
       misfit_gradient, gradient_bytes, checkpoint_for_memory, get_traces, set_receivers!,
       reset!, cached_bytes, release_cached!

const LIB = get(ENV, "VI_MI355X_LIB",
                joinpath(@__DIR__, "..", "viscoelasticinversion_amd", "libvi_mi355x.so"))

"The device value: `Parameters(...; device=MI355XBackend())`."
struct MI355XBackend end

# ---- C structs (include/vi_mi355x.h) --------------------------------------------------
struct GridDesc
    nx::Cint; nz::Cint
    dx::Cdouble; dz::Cdouble; dt::Cdouble
    precision::Cint; nrelax::Cint; fd_order_x::Cint; fd_order_z::Cint
end

struct CpmlDesc
    thickness::Cint
    d0::Cdouble; alpha_max::Cdouble; kappa_max::Cdouble; power::Cdouble
    left::Cint; right::Cint; top::Cint; bottom::Cint
end

const VX, VZ, SXX, SZZ, SXZ, MXX, MZZ, MXZ = 0, 1, 2, 3, 4, 5, 6, 7
const SOURCE_KINDS = Dict(:force_z => 0, :explosive => 1, :force_x => 2)
const COMPONENTS = Dict(:vx => 0, :vz => 1, :sxx => 2, :szz => 3, :sxz => 4)

"A non-zero status of the library becomes an error with its message (vi_last_error)."
function check(rc::Integer)
    rc == 0 && return nothing
    msg = unsafe_string(ccall((:vi_last_error, LIB), Cstring, ()))
    error("libvi_mi355x status $rc: $msg")
end

precision_code(::Type{Float32}) = Cint(0)
precision_code(::Type{Float64}) = Cint(1)

# ---- one shot on the device ----------------------------------------------------------
"""
    Sim{T}(Nx, Nz, dx, dz, dt; nrelax=0)

One simulation on the GPU (a library handle): the state (v, s, M, CPML memory), its model,
sources and receivers live in device memory until the object is finalised.
"""
mutable struct Sim{T<:Union{Float32,Float64}}
    h::Ptr{Cvoid}
    Nx::Int
    Nz::Int
    nrelax::Int
    dt::Float64
    nsrc::Int
    nrec::Int
    rec_nt::Int
    model::Any      # host copy of the arrays last uploaded (to detect changes)
    τns::Vector{Float64}
    function Sim{T}(Nx, Nz, dx, dz, dt; nrelax=0, fd_order_x=8,
                    fd_order_z=8) where {T<:Union{Float32,Float64}}
        g = Ref(GridDesc(Nx, Nz, dx, dz, dt, precision_code(T), nrelax, fd_order_x, fd_order_z))
        h = Ref{Ptr{Cvoid}}(C_NULL)
        check(ccall((:vi_create, LIB), Cint, (Ref{GridDesc}, Ref{Ptr{Cvoid}}), g, h))
        s = new{T}(h[], Nx, Nz, nrelax, Float64(dt), 0, 0, 0, nothing, Float64[])
        finalizer(destroy!, s)
    end
end

function destroy!(s::Sim)
    if s.h != C_NULL
        ccall((:vi_destroy, LIB), Cint, (Ptr{Cvoid},), s.h)
        s.h = C_NULL
    end
    return nothing
end

field_array(s::Sim{T}, a) where {T} = convert(Array{T}, a)

"Isotropic model: density ρ, Lamé λ, μ (Nx, Nz) and, with memory variables, τ."
function set_model!(s::Sim{T}, ρ, λ, μ, τ=nothing) where {T}
    s.nrelax > 0 && τ === nothing && error("τ is required when nrelax > 0")
    ρ, λ, μ = field_array(s, ρ), field_array(s, λ), field_array(s, μ)
    τa = τ === nothing ? nothing : field_array(s, τ)
    check(ccall((:vi_set_model, LIB), Cint, (Ptr{Cvoid}, Ptr{T}, Ptr{T}, Ptr{T}, Ptr{T}),
                s.h, ρ, λ, μ, τa === nothing ? Ptr{T}(C_NULL) : τa))
    s.model = (ρ=copy(ρ), λ=copy(λ), μ=copy(μ), τ=τa === nothing ? nothing : copy(τa))
    return nothing
end

"VTI model (VTIModel, src/sciml_solver/Models.jl:27-33): ρ and C11, C13, C33, C55."
function set_model_vti!(s::Sim{T}, ρ, C11, C13, C33, C55, τ=nothing) where {T}
    s.nrelax > 0 && τ === nothing && error("τ is required when nrelax > 0")
    a = map(x -> field_array(s, x), (ρ, C11, C13, C33, C55))
    τa = τ === nothing ? nothing : field_array(s, τ)
    check(ccall((:vi_set_model_vti, LIB), Cint,
                (Ptr{Cvoid}, Ptr{T}, Ptr{T}, Ptr{T}, Ptr{T}, Ptr{T}, Ptr{T}),
                s.h, a..., τa === nothing ? Ptr{T}(C_NULL) : τa))
    s.model = (ρ=a[1], C11=a[2], C13=a[3], C33=a[4], C55=a[5], τ=τa)
    return nothing
end

"Model from a reference model object: VTIModel (b, C11, C13, C33, C55) or an isotropic one
with fields (b, λ, μ); b = 1/ρ is the buoyancy."
function set_model!(s::Sim, m, τ=nothing)
    if hasproperty(m, :C11)
        set_model_vti!(s, 1 ./ m.b, m.C11, m.C13, m.C33, m.C55, τ)
    else
        set_model!(s, 1 ./ m.b, m.λ, m.μ, τ)
    end
end

"Relaxation times τns of the L standard linear solids (QModel, Models.jl:56-62)."
function set_relaxation!(s::Sim, τns)
    t = Vector{Float64}(τns)
    length(t) == s.nrelax || error("expected $(s.nrelax) relaxation times")
    check(ccall((:vi_set_relaxation, LIB), Cint, (Ptr{Cvoid}, Ptr{Float64}), s.h, t))
    s.τns = t
    return nothing
end

"CPML from a CPMLConfig (src/sciml_solver/CPML.jl:28-63); nothing removes it."
function set_cpml!(s::Sim, c)
    if c === nothing
        check(ccall((:vi_set_cpml, LIB), Cint, (Ptr{Cvoid}, Ptr{CpmlDesc}), s.h, C_NULL))
        return nothing
    end
    d = Ref(CpmlDesc(c.thickness, c.d0, c.alpha_max, c.kappa_max, c.power,
                     c.left, c.right, c.top, c.bottom))
    check(ccall((:vi_set_cpml, LIB), Cint, (Ptr{Cvoid}, Ref{CpmlDesc}), s.h, d))
    return nothing
end

"""Free surface: false (none), true / :image (stress imaging, CPML.jl:366-393), :afda
(H-AFDA one-sided stencils, free-surface-kernels.jl) or :hermitian (H-AFDA with the
Hermitian formula #3 for vz,z on the first row below the surface,
free-surface-kernels.jl:154-214)."""
function set_free_surface!(s::Sim, mode)
    m = mode === false || mode === nothing || mode === :none ? 0 :
        mode === :afda ? 2 : mode === :hermitian ? 3 : 1
    check(ccall((:vi_set_free_surface, LIB), Cint, (Ptr{Cvoid}, Cint), s.h, m))
    return nothing
end

"Half points of formula #3 (2..8; the library's default is 5)."
set_afda_hermitian!(s::Sim, points::Integer) =
    check(ccall((:vi_set_afda_hermitian, LIB), Cint, (Ptr{Cvoid}, Cint), s.h, points))

"""The library's formula #3 weights (_stencil_hermitian, Stencil.jl:119-189) as
(alpha, beta) with vz,z(h) = alpha vz,z(0) + sum beta[m] vz((m - 1/2) h)."""
function afda_hermitian_weights(h::Real, points::Integer)
    out = zeros(Float64, 9)
    check(ccall((:vi_afda_hermitian_weights, LIB), Cint, (Cdouble, Cint, Ptr{Cdouble}),
                h, points, out))
    return out[1], out[2:1+points]
end

"""
    set_sources!(s, ix, iz, kinds, w)

Point sources at 1-based cells (ix[q], iz[q]) of kinds :force_z (the reference's
RickerSource: added to dvz before the buoyancy), :force_x or :explosive; w is (nt, nsrc) with
sample k used by step k (already multiplied by the source coefficient).
"""
function set_sources!(s::Sim{T}, ix, iz, kinds, w::AbstractMatrix) where {T}
    n = length(ix)
    wa = convert(Matrix{T}, w)
    size(wa, 2) == n || error("w must be (nt, nsrc)")
    k = Cint[k isa Symbol ? SOURCE_KINDS[k] : k for k in kinds]
    check(ccall((:vi_set_sources, LIB), Cint,
                (Ptr{Cvoid}, Cint, Ptr{Cint}, Ptr{Cint}, Ptr{Cint}, Cint, Ptr{T}),
                s.h, n, Cint.(ix .- 1), Cint.(iz .- 1), k, size(wa, 1), wa))
    s.nsrc = n
    return nothing
end

"Receivers at 1-based cells sampling :vx, :vz, :sxx, :szz or :sxz after every step, for up
to nt_max steps."
function set_receivers!(s::Sim, ix, iz, comps, nt_max)
    n = length(ix)
    c = comps isa Symbol ? fill(Cint(COMPONENTS[comps]), n) :
        Cint[x isa Symbol ? COMPONENTS[x] : x for x in comps]
    check(ccall((:vi_set_receivers, LIB), Cint,
                (Ptr{Cvoid}, Cint, Ptr{Cint}, Ptr{Cint}, Ptr{Cint}, Cint),
                s.h, n, Cint.(ix .- 1), Cint.(iz .- 1), c, nt_max))
    s.nrec, s.rec_nt = n, nt_max
    return nothing
end

"Traces of the first nt steps since set_receivers!, (nt, nrec)."
function get_traces(s::Sim{T}, nt=s.rec_nt) where {T}
    out = Matrix{T}(undef, nt, s.nrec)
    check(ccall((:vi_get_traces, LIB), Cint, (Ptr{Cvoid}, Ptr{T}, Cint), s.h, out, nt))
    return out
end

function set_field!(s::Sim{T}, field, a) where {T}
    b = field_array(s, a)
    check(ccall((:vi_set_state, LIB), Cint, (Ptr{Cvoid}, Cint, Ptr{T}), s.h, field, b))
    return nothing
end

function get_field(s::Sim{T}, field) where {T}
    dims = field < MXX ? (s.Nx, s.Nz) : (s.Nx, s.Nz, s.nrelax)
    a = Array{T}(undef, dims)
    check(ccall((:vi_get_state, LIB), Cint, (Ptr{Cvoid}, Cint, Ptr{T}), s.h, field, a))
    return a
end

"Upload s (Nx, Nz, 3), v (Nx, Nz, 2) and, when given, M (Nx, Nz, L, 3)."
function set_state!(s::Sim, sv::AbstractArray, v::AbstractArray, M=nothing)
    for k in 1:3
        set_field!(s, SXX + k - 1, sv[:, :, k])
    end
    for k in 1:2
        set_field!(s, VX + k - 1, v[:, :, k])
    end
    if M !== nothing && s.nrelax > 0
        for k in 1:3
            set_field!(s, MXX + k - 1, M[:, :, :, k])
        end
    end
    return nothing
end

function get_state(s::Sim{T}) where {T}
    sv = Array{T}(undef, s.Nx, s.Nz, 3)
    v = Array{T}(undef, s.Nx, s.Nz, 2)
    for k in 1:3
        sv[:, :, k] = get_field(s, SXX + k - 1)
    end
    for k in 1:2
        v[:, :, k] = get_field(s, VX + k - 1)
    end
    M = s.nrelax > 0 ? cat((get_field(s, MXX + k - 1) for k in 1:3)...; dims=4) : nothing
    return sv, v, M
end

"Zero the state, the traces and the step counter."
reset!(s::Sim) = check(ccall((:vi_reset, LIB), Cint, (Ptr{Cvoid},), s.h))

"Advance n full steps (velocity kernel, then stress kernel) and wait."
function step!(s::Sim, n::Integer=1)
    check(ccall((:vi_step, LIB), Cint, (Ptr{Cvoid}, Cint), s.h, n))
    check(ccall((:vi_sync, LIB), Cint, (Ptr{Cvoid},), s.h))
    return nothing
end

"""
    rates(s, src_values) -> (dv, ds, dM)

Right-hand side at the state held by `s` (vi_rates): dv (Nx, Nz, 2), ds (Nx, Nz, 3) and
dM (Nx, Nz, L, 3) (or nothing); src_values holds one sample per source.
"""
function rates(s::Sim{T}, src_values=nothing) where {T}
    dv = Array{T}(undef, s.Nx, s.Nz, 2)
    ds = Array{T}(undef, s.Nx, s.Nz, 3)
    dM = s.nrelax > 0 ? Array{T}(undef, s.Nx, s.Nz, s.nrelax, 3) : nothing
    sv = src_values === nothing ? nothing : Vector{T}(src_values)
    check(ccall((:vi_rates, LIB), Cint, (Ptr{Cvoid}, Ptr{T}, Ptr{T}, Ptr{T}, Ptr{T}),
                s.h, sv === nothing ? Ptr{T}(C_NULL) : sv, dv, ds,
                dM === nothing ? Ptr{T}(C_NULL) : dM))
    return dv, ds, dM
end

"""
    misfit_gradient(s, dobs; checkpoint_every=0) -> (J, g)

J = ½ Σ (d - dobs)² of the traces of size(dobs, 1) steps from a zero state, and its gradient
by the discrete adjoint with √nt checkpointing (vi_gradient): g = (ρ=…, λ=…, μ=…, τ=…) for an
isotropic model, (ρ=…, C11=…, C13=…, C33=…, C55=…, τ=…) for a VTI model, each (Nx, Nz).
"""
function misfit_gradient(s::Sim{T}, dobs::AbstractMatrix; checkpoint_every=0) where {T}
    d = convert(Matrix{T}, dobs)
    nt = size(d, 1)
    size(d, 2) == s.nrec || error("dobs must be (nt, nrec)")
    J = Ref{Cdouble}(0)
    P = Ptr{T}
    gτ = s.nrelax > 0 ? Array{T}(undef, s.Nx, s.Nz) : nothing
    pτ = gτ === nothing ? P(C_NULL) : gτ
    if s.model !== nothing && hasproperty(s.model, :C11)
        g = [Array{T}(undef, s.Nx, s.Nz) for _ in 1:5]
        check(ccall((:vi_gradient_vti, LIB), Cint,
                    (Ptr{Cvoid}, Cint, P, Cint, Ref{Cdouble}, P, P, P, P, P, P),
                    s.h, nt, d, checkpoint_every, J, g..., pτ))
        return J[], (ρ=g[1], C11=g[2], C13=g[3], C33=g[4], C55=g[5], τ=gτ)
    end
    g = [Array{T}(undef, s.Nx, s.Nz) for _ in 1:3]
    check(ccall((:vi_gradient, LIB), Cint,
                (Ptr{Cvoid}, Cint, P, Cint, Ref{Cdouble}, P, P, P, P),
                s.h, nt, d, checkpoint_every, J, g..., pτ))
    return J[], (ρ=g[1], λ=g[2], μ=g[3], τ=gτ)
end

"""
    gradient_bytes(s, nt; checkpoint_every=0) -> Int

Temporary device bytes of one `misfit_gradient` call of `nt` steps (vi_gradient_bytes).
"""
function gradient_bytes(s::Sim, nt::Integer; checkpoint_every::Integer=0)
    b = Ref{Csize_t}(0)
    check(ccall((:vi_gradient_bytes, LIB), Cint, (Ptr{Cvoid}, Cint, Cint, Ref{Csize_t}),
                s.h, nt, checkpoint_every, b))
    return Int(b[])
end

"""
    checkpoint_for_memory(s, nt, max_bytes) -> Int

The largest checkpoint interval K = cld(nt, k), k = 1 … ceil(√nt), whose gradient needs at most
`max_bytes` of device memory, or 0 (checkpoints every ceil(√nt) steps) when none fits. A
gradient recomputes about nt - K forward steps, so the largest interval that fits is the
fastest; the gradient does not depend on K.
"""
function checkpoint_for_memory(s::Sim, nt::Integer, max_bytes::Integer)
    for k in 1:ceil(Int, sqrt(nt))
        K = cld(nt, k)
        gradient_bytes(s, nt; checkpoint_every=K) <= max_bytes && return K
    end
    return 0
end

"""
    cached_bytes() -> Int

Device bytes the library keeps for reuse on the current device: buffers of 64 MB or more
that simulations released, handed to the next allocation they fit instead of back to the
runtime (an allocation that does not fit frees them first).
"""
function cached_bytes()
    b = Ref{Csize_t}(0)
    check(ccall((:vi_cached_bytes, LIB), Cint, (Ref{Csize_t},), b))
    return Int(b[])
end

"Free the buffers the library keeps for reuse (e.g. before another library allocates)."
release_cached!() = (check(ccall((:vi_release_cached, LIB), Cint, ())); nothing)

# ---- the reference's solver API on the device -----------------------------------------
"""
    Parameters(; Nx=64, Nz=64, dx=10.0, dz=10.0, dt=0.001, fd_order_x=8, fd_order_z=8,
               bc=:dirichlet, source, device=MI355XBackend(), model, cpml=nothing,
               free_surface=false, qmodel=nothing, receivers=nothing, precision=Float32)

The keyword constructor of src/sciml_solver/Parameters.jl:47-86 for the MI355X backend, with
the same names and defaults (a 40 Hz Ricker source at the centre, ρ = 2000, λ = μ = 1e9).
Operators of order 2, 4, 6 or 8 per axis and the Dirichlet ring with an optional CPML are
available (the reference's periodic default is defective). `qmodel` adds L SLS mechanisms (τ and τns);
`receivers` is (ix, iz, components).
"""
struct GPUParameters
    Nx::Int; Nz::Int
    dx::Float64; dz::Float64; dt::Float64
    fd_order_x::Int; fd_order_z::Int
    model::Any
    qmodel::Any
    cpml::Any
    free_surface::Any
    source::Any
    receivers::Any
    precision::DataType
end

function Parameters(; Nx=64, Nz=64, dx=10.0, dz=10.0, dt=0.001, fd_order_x=8, fd_order_z=8,
                    bc=:dirichlet, source=nothing, device=MI355XBackend(), model=nothing,
                    cpml=nothing, free_surface=false, qmodel=nothing, receivers=nothing,
                    precision=Float32)
    (fd_order_x in (2, 4, 6, 8) && fd_order_z in (2, 4, 6, 8)) ||
        error("staggered operators of order 2, 4, 6 or 8 are implemented on MI355X")
    bc === :dirichlet || error("boundary condition $bc is not implemented on MI355X; " *
                               "use :dirichlet with an optional CPMLConfig")
    device isa MI355XBackend || error("MI355X.Parameters needs device=MI355XBackend()")
    src = source === nothing ? RickerSource(40.0, 0.2, Nx ÷ 2, Nz ÷ 2) : source
    m = model === nothing ?
        (b=fill(1 / 2000.0, Nx, Nz), λ=fill(1e9, Nx, Nz), μ=fill(1e9, Nx, Nz)) : model
    return GPUParameters(Nx, Nz, dx, dz, dt, fd_order_x, fd_order_z, m, qmodel, cpml,
                         free_surface, src, receivers, precision)
end

nrelax(p::GPUParameters) = p.qmodel === nothing ? 0 : length(p.qmodel.τns)
sources(p::GPUParameters) = p.source isa AbstractVector ? p.source : [p.source]

"1-based cell of a RickerSource (its `grid` holds one (nx, nz) tuple) or of an (ix, iz) pair."
source_cell(src) = hasproperty(src, :grid) ? Tuple(Array(src.grid)[1]) : (src[1], src[2])
source_kind(src) = hasproperty(src, :kind) ? src.kind : :force_z
source_coef(src) = hasproperty(src, :coefs) ? Float64(Array(src.coefs)[1]) : 1.0

"The reference's RickerSource value at time t: c r(t - tc) inside its open window, else 0
(Source.jl:42-58)."
function source_value(src, t)
    lo, hi = src.tspan
    return lo < t < hi ? source_coef(src) * ViscoelasticInversion._ricker(src.freq, t - src.tc) : 0.0
end

"""
    Sim{T}(p::GPUParameters)
    Sim(p::GPUParameters)

A device simulation configured from Parameters: model, relaxation, CPML, free surface, source
positions (samples are given per run) and receivers (capacity set by the run).
"""
function Sim{T}(p::GPUParameters) where {T}
    s = Sim{T}(p.Nx, p.Nz, p.dx, p.dz, p.dt; nrelax=nrelax(p), fd_order_x=p.fd_order_x,
               fd_order_z=p.fd_order_z)
    τ = p.qmodel === nothing ? nothing : p.qmodel.τ
    set_model!(s, p.model, τ)
    p.qmodel === nothing || set_relaxation!(s, Array(p.qmodel.τns))
    set_cpml!(s, p.cpml)
    set_free_surface!(s, p.free_surface)
    return s
end
Sim(p::GPUParameters) = Sim{p.precision}(p)

"Source samples (nt, nsrc) of nt steps from t0: force sources at t0 + k dt, explosive ones
at t0 + (k + ½) dt (the stress half step's time)."
function source_samples(p::GPUParameters, nt, t0)
    srcs = sources(p)
    w = zeros(nt, length(srcs))
    for (q, src) in enumerate(srcs)
        off = source_kind(src) === :explosive ? 0.5 : 0.0
        for k in 0:nt-1
            w[k+1, q] = source_value(src, t0 + (k + off) * p.dt)
        end
    end
    return w
end

function configure_run!(s::Sim, p::GPUParameters, nt, t0)
    srcs = sources(p)
    cells = source_cell.(srcs)
    set_sources!(s, first.(cells), last.(cells), source_kind.(srcs), source_samples(p, nt, t0))
    if p.receivers !== nothing
        ix, iz, comps = p.receivers
        set_receivers!(s, ix, iz, comps, nt)
    end
    return nothing
end

struct GPUProblem
    u0::Tuple
    tspan::Tuple
    p::GPUParameters
end

"make_problem(p; s0, v0, tspan) (ViscoelasticProblem.jl:273-292): zero state by default."
function make_problem(p::GPUParameters; s0=nothing, v0=nothing, tspan=(0.0, 2.0))
    T = p.precision
    s = s0 === nothing ? zeros(T, p.Nx, p.Nz, 3) : s0
    v = v0 === nothing ? zeros(T, p.Nx, p.Nz, 2) : v0
    size(s) == (p.Nx, p.Nz, 3) && size(v) == (p.Nx, p.Nz, 2) ||
        error("s0 must be (Nx, Nz, 3) and v0 (Nx, Nz, 2)")
    return GPUProblem((s, v), Tuple(tspan), p)
end

"""
    solve_problem(prob::GPUProblem; solver=:leapfrog, saveat=nothing)

nt = round((t1 - t0)/dt) fused leapfrog steps on the GPU (ViscoelasticProblem.jl:307-309);
returns (t, s, v, M, traces, saved) with the final state; `saveat` (steps) stores snapshots.
"""
function solve_problem(prob::GPUProblem; solver=:leapfrog, saveat=nothing, kwargs...)
    solver in (:leapfrog, nothing) || error("only the fused leapfrog integrator runs on MI355X")
    p = prob.p
    t0, t1 = prob.tspan
    nt = round(Int, (t1 - t0) / p.dt)
    s = Sim(p)
    try
        configure_run!(s, p, nt, t0)
        set_state!(s, prob.u0[1], prob.u0[2])
        saved = Any[]
        done = 0
        while done < nt
            n = saveat === nothing ? nt - done : min(saveat, nt - done)
            step!(s, n)
            done += n
            saveat === nothing || push!(saved, (t0 + done * p.dt, get_state(s)...))
        end
        sv, v, M = get_state(s)
        tr = p.receivers === nothing ? nothing : get_traces(s, nt)
        return (t=t0 .+ (1:nt) .* p.dt, s=sv, v=v, M=M, traces=tr, saved=saved,
                retcode=:Success)
    finally
        destroy!(s)
    end
end

# right-hand side: one device simulation per Parameters, kept for repeated calls
const RHS_SIMS = IdDict{GPUParameters,Any}()

function rhs_sim(p::GPUParameters)
    get!(RHS_SIMS, p) do
        s = Sim(p)
        srcs = sources(p)
        cells = source_cell.(srcs)
        set_sources!(s, first.(cells), last.(cells), source_kind.(srcs), zeros(1, length(srcs)))
        s
    end
end

"Stress imaging on the state (apply_free_surface!, CPML.jl:382-393, with the 4 image rows of
the 8th-order stencil above the surface row 5 in 1-based indices)."
function apply_images!(s::AbstractArray, Nx)
    for m in 1:4
        s[5:Nx-4, 5-m, 2] .= .-s[5:Nx-4, 5+m, 2]
        s[5:Nx-4, 5-m, 3] .= .-s[5:Nx-4, 4+m, 3]
    end
    return s
end

function device_rates(s, v, p::GPUParameters, t, M)
    sim = rhs_sim(p)
    set_state!(sim, s, v, M === nothing && nrelax(p) > 0 ?
                          zeros(p.precision, p.Nx, p.Nz, nrelax(p), 3) : M)
    return rates(sim, [source_value(src, t) for src in sources(p)])
end

"stress_eq!(ds, s, v, p, t) (ViscoelasticProblem.jl:62-102) on the GPU; M (Nx, Nz, L, 3) are
the memory variables (zero when not given)."
function stress_eq!(ds, s, v, p::GPUParameters, t; M=nothing)
    _, r, _ = device_rates(s, v, p, t, M)
    ds .= r
    return nothing
end

"velocity_eq!(dv, s, v, p, t) (ViscoelasticProblem.jl:110-174) on the GPU; with the imaging
surface the images are written into s first, as the reference's velocity_eq! does."
function velocity_eq!(dv, s, v, p::GPUParameters, t; M=nothing)
    (p.free_surface === true || p.free_surface === :image) && apply_images!(s, p.Nx)
    r, _, _ = device_rates(s, v, p, t, M)
    dv .= r
    return nothing
end

# ---- the fused step's own API (src/visco-kernels.jl:87-113) ----------------------------
"Fields(s, M, v) held on the device by a Sim."
struct Fields{T}
    sim::Sim{T}
end

"Parameters(λs, μs, τs, τns) of the fused step (visco-kernels.jl:108-113); ρ, when given,
replaces the density (visco-kernels.jl:10-11 reads it), else the loaded one is kept."
struct StepParameters
    λs::Any
    μs::Any
    τs::Any
    τns::Any
    ρ::Any
end
StepParameters(λs, μs, τs, τns; ρ=nothing) = StepParameters(λs, μs, τs, τns, ρ)

changed(old, new, T) = old === nothing || new === nothing ? new !== nothing :
                       convert(Array{T}, new) != old

"""
    leapfrog_step!(fields, parameters, dt)

One full step (velocity kernel, then stress kernel) of the device state with the model of
`parameters` (StepParameters or GPUParameters), uploaded only where it differs from what the
simulation holds; dt must be the simulation's.
"""
function leapfrog_step!(f::Fields{T}, p::StepParameters, dt) where {T}
    s = f.sim
    dt == s.dt || error("dt differs from the simulation's dt")
    m = s.model
    m === nothing && error("load a model first")
    hasproperty(m, :C11) && error("VTI simulation: pass GPUParameters with a VTIModel")
    ρ = p.ρ === nothing ? m.ρ : p.ρ
    τ = p.τs === nothing ? m.τ : p.τs
    if changed(m.ρ, ρ, T) || changed(m.λ, p.λs, T) || changed(m.μ, p.μs, T) ||
       (s.nrelax > 0 && changed(m.τ, τ, T))
        set_model!(s, ρ, p.λs, p.μs, s.nrelax > 0 ? τ : nothing)
    end
    if s.nrelax > 0 && p.τns !== nothing && Vector{Float64}(Array(p.τns)) != s.τns
        set_relaxation!(s, Array(p.τns))
    end
    step!(s, 1)
    return nothing
end

function leapfrog_step!(f::Fields, p::GPUParameters, dt)
    s = f.sim
    τ = p.qmodel === nothing ? nothing : p.qmodel.τ
    set_model!(s, p.model, τ)  # uploads (a GPUParameters holds no change record)
    p.qmodel === nothing || set_relaxation!(s, Array(p.qmodel.τns))
    dt == s.dt || error("dt differs from the simulation's dt")
    step!(s, 1)
    return nothing
end

# ---- trace misfit and its rrule --------------------------------------------------------
"""
    trace_misfit(p, dobs; tspan=(0.0, size(dobs, 1) * p.dt))

½ Σ (d - dobs)² of the receiver traces (nt, nrec) of p from a zero state.
"""
function trace_misfit(p::GPUParameters, dobs; tspan=(0.0, size(dobs, 1) * p.dt))
    sol = solve_problem(make_problem(p; tspan=tspan))
    return 0.5 * sum(abs2, Float64.(sol.traces) .- Float64.(dobs))
end

"""
ChainRulesCore.rrule of trace_misfit with respect to the model: the pullback gives
dJ/d(ρ, λ, μ[, τ]) (isotropic) or dJ/d(ρ, C11, C13, C33, C55[, τ]) (VTI) from the discrete
adjoint with √nt checkpointing (vi_gradient); the chain rule to (vp, vs, ρ, Q) uses
τ = L/(K̄ Q) (Models.jl:67), dτ/dQ = -τ/Q.
"""
function ChainRulesCore.rrule(::typeof(trace_misfit), p::GPUParameters, dobs;
                              tspan=(0.0, size(dobs, 1) * p.dt))
    nt = round(Int, (tspan[2] - tspan[1]) / p.dt)
    s = Sim(p)
    J, g = try
        configure_run!(s, p, nt, tspan[1])
        misfit_gradient(s, dobs[1:nt, :])
    finally
        destroy!(s)
    end
    function trace_misfit_pullback(J̄)
        ḡ = map(x -> x === nothing ? nothing : J̄ .* x, g)
        return (ChainRulesCore.NoTangent(), ChainRulesCore.Tangent{GPUParameters}(model=ḡ),
                ChainRulesCore.NoTangent())
    end
    return J, trace_misfit_pullback
end

end # module
