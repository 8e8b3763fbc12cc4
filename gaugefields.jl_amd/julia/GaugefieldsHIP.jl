"""
    GaugefieldsHIP

MI355X backend for the SU(3) gauge-link hot path of Gaugefields.jl, over libgfhip.so
(`include/gfhip.h`).  It plugs into Gaugefields.jl by dispatch, as the reference's own
backends do:

  * `HIPBackend <: Gaugefields.AbstractGaugeBackend` (API.jl:6-27) selects it in
    `gauge_configuration(lattice; backend=HIPBackend())` (API.jl:178-251);
  * `Gaugefields_4D_HIP{3} <: Gaugefields_4D{3}` (src/4D/gaugefields_4D.jl:10) is one link
    direction of a device-resident configuration, `TA_Gaugefields_4D_HIP{3} <:
    TA_Gaugefields_4D{3}` (src/4D/TA_gaugefields_4D.jl:1) one direction of its momenta;
  * methods of the reference's functions for `Vector`s of these fields, with the reference's
    signatures: calculate_Plaquette / measure_plaquette, evaluate_GaugeAction, calc_dSdUμ!,
    the MD provider trio md_action_workspace / md_potential / md_force!
    (molecular_dynamics.jl:13-43, 243-267) also for an MDActionSet and an MDForceGroup
    (:45-236; the reference's SextonWeingarten step then runs unchanged), md_hamiltonian,
    update_gaugefields!, update_momenta!, md_trajectory! (:494-730), initialize_TA_Gaugefields /
    gauss_distribution! / gaussian_momenta!, flow! (gradientflow.jl:171-238), calc_smearedU /
    back_prop (Abstractsmearing.jl:237-411), forward! / layer_pullback! of STOUT_Layer
    (stout_fast.jl:223-274), heatbath! / overrelaxation!, topological_charge,
    load_gaugefield! / save_binarydata (ildg_format.jl:797-969).

The AD rules of the stout map (a ChainRulesCore rrule and an Enzyme reverse rule whose
pullbacks call gf_stout_backward) are the package extensions in `ext/`.

A configuration's four directions share one `gf_ctx` (links, momenta, force, stout levels).
Fields made by `similar` get their own context, created on first use, so the temporaries the
reference's constructors allocate (md_driver, Gradientflow, Heatbath) cost no device memory
unless a method of the reference's generic path touches them.

This file is not executed in this repository (the image has no Julia); the ccall sequence of
every method is replayed through ctypes by tests/test_julia_binding.py.
"""
module GaugefieldsHIP

using LinearAlgebra
import Gaugefields
import Gaugefields: AbstractGaugeBackend, GaugeAction, MDDriver, QPQ, PQP, CovNeuralnet, STOUT_Layer,
    MDActionSet, MDForceGroup,
    Gradientflow, Heatbath, ILDG
import Gaugefields: gauge_configuration, measure_plaquette, gaussian_momenta!, gauge_momenta,
    md_action_workspace, md_potential, md_force!, md_hamiltonian, update_gaugefields!, update_momenta!,
    md_trajectory!, md_step!, md_step_size, gradient_flow
import Gaugefields.AbstractGaugefields_module: Gaugefields_4D, TA_Gaugefields_4D, calculate_Plaquette,
    substitute_U!, set_wing_U!, clear_U!, initialize_TA_Gaugefields, gauss_distribution!, topological_charge,
    topological_charge_density
import Gaugefields.GaugeAction_module: evaluate_GaugeAction, calc_dSdUμ!, evaluate_GaugeAction_untraced,
    evaluate_GaugeAction_untraced!
import Gaugefields.Gradientflow_module: flow!
import Gaugefields.Abstractsmearing_module: calc_smearedU, back_prop, forward!, layer_pullback!
import Gaugefields.heatbath_module: heatbath!, overrelaxation!
import Gaugefields.ILDG_format: load_gaugefield!, save_binarydata
import Wilsonloop: make_loops_fromname

export HIPBackend, Gaugefields_4D_HIP, TA_Gaugefields_4D_HIP, OmelyanPQP, StoutGaugeAction,
    stout_smear, stout_smear!, stout_pullback, force_field, backup, calculate_energy_density, energy_density,
    flow_measure!

const libgfhip = joinpath(@__DIR__, "..", "libgfhip.so")

const GF_F64 = Cint(0)
const GF_F32 = Cint(1)
const GF_F32_FORCE = Cint(2)  # FP64 links and momenta, every force evaluated in FP32
const GF_EINVAL = Cint(-1)
const GF_QPQ = Cint(0)
const GF_PQP = Cint(1)
const GF_OMELYAN2MN = Cint(2)
const GF_FUSED = Cint(0x100)
const GF_STOUT_STORE = Cint(0)
const GF_STOUT_RECOMPUTE = Cint(1)
const GF_STOUT_AUTO = Cint(2)
const GF_STOUT_MODES = Dict(:store => GF_STOUT_STORE, :recompute => GF_STOUT_RECOMPUTE, :auto => GF_STOUT_AUTO)
const GF_TOPO_METHODS = Dict(:plaquette => Cint(0), :clover => Cint(1), :improved => Cint(2))
const OMELYAN_LAMBDA = 0.1931833275037836

# ----------------------------------------------------------------------------- contexts

"""Select the MI355X backend in `gauge_configuration(lattice; backend=HIPBackend())`."""
struct HIPBackend <: AbstractGaugeBackend
    device::Int
end
HIPBackend() = HIPBackend(0)

"""One device configuration (gf_ctx): created on first use, destroyed by the finalizer."""
mutable struct HIPContext
    ptr::Ptr{Cvoid}
    dims::NTuple{4,Int}
    precision::Cint
    device::Int
    stout::Tuple{Int,Float64,Symbol}  # (levels, rho, mode) last passed to gf_stout_config
    function HIPContext(dims::NTuple{4,Int}, precision::Cint, device::Int)
        c = new(C_NULL, dims, precision, device, (0, 0.0, :auto))
        finalizer(c) do c
            c.ptr == C_NULL || ccall((:gf_destroy, libgfhip), Cint, (Ptr{Cvoid},), c.ptr)
            c.ptr = C_NULL
        end
        return c
    end
end

function _ctx(c::HIPContext)
    if c.ptr == C_NULL
        out = Ref{Ptr{Cvoid}}(C_NULL)
        rc = ccall((:gf_create, libgfhip), Cint, (Ptr{Cint}, Cint, Cint, Ref{Ptr{Cvoid}}),
                   Cint[c.dims...], c.precision, Cint(c.device), out)
        rc == 0 || (rc == GF_EINVAL ?
                    throw(ArgumentError("gf_create rejected lattice $(c.dims) (descriptor extent or arguments)")) :
                    error("gf_create failed with status $rc"))
        c.ptr = out[]
    end
    return c.ptr
end

function _check(c::HIPContext, rc::Cint)
    if rc != 0
        msg = unsafe_string(ccall((:gf_last_error, libgfhip), Cstring, (Ptr{Cvoid},), c.ptr))
        rc == GF_EINVAL ? throw(ArgumentError(msg)) : error("libgfhip status $rc: $msg")
    end
    return nothing
end

"""One link direction `μ` of a device-resident SU(3) configuration (halo-free SoA layout:
the stencils wrap periodically, so `NDW == 0` and set_wing_U! is a no-op)."""
struct Gaugefields_4D_HIP{NC} <: Gaugefields_4D{NC}
    ctx::HIPContext
    μ::Int
    NC::Int
    NX::Int
    NY::Int
    NZ::Int
    NT::Int
    NV::Int
    NDW::Int
end

"""One direction of the conjugate momenta of a configuration (`slot = :momenta`), or of the
force field the context's `md_force!` writes (`slot = :force`)."""
struct TA_Gaugefields_4D_HIP{NC} <: TA_Gaugefields_4D{NC}
    ctx::HIPContext
    μ::Int
    slot::Symbol
    NC::Int
    NX::Int
    NY::Int
    NZ::Int
    NT::Int
    NV::Int
    NumofBasis::Int
end

const HIPLinks = AbstractVector{<:Gaugefields_4D_HIP}
const HIPMomenta = AbstractVector{<:TA_Gaugefields_4D_HIP}

_links_of(c::HIPContext) = [Gaugefields_4D_HIP{3}(c, μ, 3, c.dims..., prod(c.dims), 0) for μ in 1:4]
_ta_of(c::HIPContext, slot::Symbol) =
    [TA_Gaugefields_4D_HIP{3}(c, μ, slot, 3, c.dims..., prod(c.dims), 8) for μ in 1:4]

function _context(U::HIPLinks)
    length(U) == 4 || throw(ArgumentError("the gauge configuration has length $(length(U)); expected 4"))
    c = U[1].ctx
    all(u -> u.ctx === c, U) && all(μ -> U[μ].μ == μ, 1:4) ||
        throw(ArgumentError("the four directions must be those of one HIP configuration"))
    return c
end
function _same(U::HIPLinks, P::HIPMomenta)
    length(U) == length(P) || throw(ArgumentError("U and P must have the same number of directions"))
    c = _context(U)
    all(p -> p.ctx === c && p.slot === :momenta, P) ||
        throw(ArgumentError("P must be the momenta of U's configuration (gauge_momenta(U))"))
    return c
end

# element type of the stored links: md_driver converts its step sizes to its real type
# (molecular_dynamics.jl:425-465)
# (an FP32-force context holds FP64 links: its element type is ComplexF64)
_elementtype(c::HIPContext) = c.precision == GF_F32 ? ComplexF32 : ComplexF64
Base.eltype(u::Gaugefields_4D_HIP) = _elementtype(u.ctx)

Base.size(u::Gaugefields_4D_HIP) = (u.NC, u.NC, u.NX, u.NY, u.NZ, u.NT)

# ----------------------------------------------------------------------------- creation

"""
    gauge_configuration(lattice, backend::HIPBackend; colors=3, halo=1, start=:cold, seed=nothing,
                        process_grid=nothing, comm=nothing, boundary=:periodic, eltype=ComplexF64,
                        force_eltype=eltype, kwargs...)

The HIP method behind `gauge_configuration(lattice; backend=HIPBackend(), ...)` (API.jl:178-251;
INTEGRATION.md gives the one-line forward the keyword method needs).  `start=:hot` draws the
hot start on the device from the Philox4x32-10 stream of `seed` (a random seed if `nothing`).
`halo` is accepted and ignored: the device layout has no halo.  `force_eltype=ComplexF32` with
`eltype=ComplexF64` (a HIP-only keyword) keeps the links and momenta in FP64 and evaluates every
force in FP32 (GF_F32_FORCE).
"""
function gauge_configuration(lattice::NTuple{4,<:Integer}, backend::HIPBackend; colors::Integer=3,
                             halo::Integer=1, start::Symbol=:cold, seed=nothing, process_grid=nothing,
                             comm=nothing, boundary=:periodic, eltype::Type=ComplexF64,
                             force_eltype::Type=eltype, kwargs...)
    all(>(0), lattice) || throw(ArgumentError("all lattice extents must be positive; got $lattice"))
    colors == 3 || throw(ArgumentError("HIPBackend implements SU(3); got colors=$colors"))
    start in (:cold, :hot) || throw(ArgumentError("start must be :cold or :hot; got $start"))
    halo >= 0 || throw(ArgumentError("halo must be nonnegative; got $halo"))
    (process_grid === nothing || process_grid === :auto) && comm === nothing ||
        throw(ArgumentError("HIPBackend runs one configuration on one GPU (no process_grid or comm)"))
    boundary === :periodic || all(==(1), boundary) ||
        throw(ArgumentError("HIPBackend supports periodic boundaries only"))
    prec = eltype == ComplexF64 ? GF_F64 : eltype == ComplexF32 ? GF_F32 :
           throw(ArgumentError("eltype must be ComplexF64 or ComplexF32; got $eltype"))
    force_eltype in (eltype, ComplexF32) ||
        throw(ArgumentError("force_eltype must be the links' eltype or ComplexF32; got $force_eltype"))
    prec == GF_F64 && force_eltype == ComplexF32 && (prec = GF_F32_FORCE)
    c = HIPContext(ntuple(i -> Int(lattice[i]), 4), prec, backend.device)
    p = _ctx(c)  # gf_create leaves the links at the identity (cold start)
    if start === :hot
        s = seed === nothing ? rand(UInt64) : UInt64(seed)
        _check(c, ccall((:gf_hot_start, libgfhip), Cint, (Ptr{Cvoid}, Culonglong), p, s))
    end
    return _links_of(c)
end

"""A field of the same lattice with a context of its own, created when first used."""
function Base.similar(u::Gaugefields_4D_HIP{NC}) where {NC}
    c = HIPContext(u.ctx.dims, u.ctx.precision, u.ctx.device)
    return Gaugefields_4D_HIP{NC}(c, u.μ, u.NC, u.NX, u.NY, u.NZ, u.NT, u.NV, 0)
end
Base.similar(U::Vector{<:Gaugefields_4D_HIP}) = _links_of(HIPContext(U[1].ctx.dims, U[1].ctx.precision, U[1].ctx.device))

set_wing_U!(::Gaugefields_4D_HIP) = nothing
set_wing_U!(::HIPLinks) = nothing

# ----------------------------------------------------------------------------- host transfers

# interior sites of a reference link field as Array{ComplexF64,6}(3, 3, NX, NY, NZ, NT): the
# halo'd LatticeMatrix storage (gaugefields_4D_MPILattice.jl:166-174) through its interior
# getindex, or a plain host array
function _host_links(u)
    A = Array{ComplexF64}(undef, 3, 3, u.NX, u.NY, u.NZ, u.NT)
    @inbounds for it in 1:u.NT, iz in 1:u.NZ, iy in 1:u.NY, ix in 1:u.NX, j in 1:3, i in 1:3
        A[i, j, ix, iy, iz, it] = u[i, j, ix, iy, iz, it]
    end
    return A
end
function _host_links(u::Gaugefields.AbstractGaugefields_module.Gaugefields_4D_MPILattice)
    w = u.NDW
    A = Array(u.U.A)  # device -> host for JACC device arrays
    return ComplexF64.(A[:, :, w+1:w+u.NX, w+1:w+u.NY, w+1:w+u.NZ, w+1:w+u.NT])
end
_host_links(A::AbstractArray{<:Complex,6}) = ComplexF64.(A)

function _upload_links!(c::HIPContext, host::AbstractVector)
    length(host) == 4 || throw(ArgumentError("a gauge configuration has 4 directions"))
    arrays = [_host_links(h) for h in host]
    all(a -> size(a) == (3, 3, c.dims...), arrays) ||
        throw(DimensionMismatch("links must be 3x3 on the lattice $(c.dims)"))
    buf = reduce(vcat, vec.(arrays))  # double[4][V][9][2], element (i,j) at i + 3j
    _check(c, ccall((:gf_upload_links, libgfhip), Cint, (Ptr{Cvoid}, Ptr{ComplexF64}), _ctx(c), buf))
end

function _download_links(c::HIPContext)
    V = prod(c.dims)
    buf = Vector{ComplexF64}(undef, 36V)
    _check(c, ccall((:gf_download_links, libgfhip), Cint, (Ptr{Cvoid}, Ptr{ComplexF64}), _ctx(c), buf))
    return [reshape(buf[(μ-1)*9V+1:μ*9V], 3, 3, c.dims...) for μ in 1:4]
end

"""substitute_U!(U, Uref): the device configuration takes the links of a reference
configuration (interior sites; halo stripped), of another HIP configuration, or of four
(3, 3, NX, NY, NZ, NT) host arrays.  Rejects links that are not SU(3) (ArgumentError)."""
function substitute_U!(U::HIPLinks, Uref::AbstractVector)
    c = _context(U)
    if eltype(Uref) <: Gaugefields_4D_HIP
        Uref[1].ctx === c && return U
        _upload_links!(c, _download_links(_context(Uref)))
    else
        _upload_links!(c, Uref)
    end
    return U
end
"""substitute_U!(Uhost, U): copy a device configuration back into reference fields or arrays."""
function substitute_U!(Uhost::AbstractVector, U::HIPLinks)
    links = _download_links(_context(U))
    for μ in 1:4
        h = Uhost[μ]
        if h isa AbstractArray
            h .= links[μ]
        else
            for it in 1:h.NT, iz in 1:h.NZ, iy in 1:h.NY, ix in 1:h.NX, j in 1:3, i in 1:3
                h[i, j, ix, iy, iz, it] = links[μ][i, j, ix, iy, iz, it]
            end
            set_wing_U!(h)
        end
    end
    return Uhost
end
Base.Array(U::HIPLinks) = _download_links(_context(U))

# ----------------------------------------------------------------------------- measurements

function calculate_Plaquette(U::HIPLinks)
    c = _context(U)
    s = Ref{Float64}(0)
    _check(c, ccall((:gf_plaquette, libgfhip), Cint, (Ptr{Cvoid}, Ref{Float64}), _ctx(c), s))
    return s[]
end
calculate_Plaquette(U::HIPLinks, temp1, temp2) = calculate_Plaquette(U)

function measure_plaquette(U::HIPLinks; normalize::Bool=true)
    value = calculate_Plaquette(U)
    return normalize ? value / (6 * U[1].NV * 3) : value
end

function topological_charge(U::HIPLinks; method=:plaquette)
    haskey(GF_TOPO_METHODS, method) ||
        throw(ArgumentError("supported topological_charge methods are :plaquette, :clover, and :improved"))
    c = _context(U)
    Q = Ref{Float64}(0)
    _check(c, ccall((:gf_topological_charge, libgfhip), Cint, (Ptr{Cvoid}, Cint, Ref{Float64}),
                    _ctx(c), GF_TOPO_METHODS[method], Q))
    return Q[]
end

"""topological_charge_density(U; method) (AbstractGaugefields.jl:1430-1446): q(x) as an
Array{Float64,4}(NX, NY, NZ, NT)."""
function topological_charge_density(U::HIPLinks; method=:plaquette)
    haskey(GF_TOPO_METHODS, method) ||
        throw(ArgumentError("supported topological_charge_density methods are :plaquette, :clover, and :improved"))
    c = _context(U)
    q = Vector{Float64}(undef, prod(c.dims))
    _check(c, ccall((:gf_topological_charge_density, libgfhip), Cint, (Ptr{Cvoid}, Cint, Ptr{Float64}),
                    _ctx(c), GF_TOPO_METHODS[method], q))
    return reshape(q, c.dims...)
end

"""The clover energy density of samples/measurements/energydensity.jl:32-76 (calculate_energy_density)."""
function calculate_energy_density(U::HIPLinks)
    c = _context(U)
    E = Ref{Float64}(0)
    _check(c, ccall((:gf_energy_density_clover, libgfhip), Cint, (Ptr{Cvoid}, Ref{Float64}), _ctx(c), E))
    return E[]
end

"""The plaquette energy density E = 36 (1 - plaquette) (Lüscher, JHEP 08 (2010) 071)."""
function energy_density(U::HIPLinks)
    c = _context(U)
    E = Ref{Float64}(0)
    _check(c, ccall((:gf_energy_density, libgfhip), Cint, (Ptr{Cvoid}, Ref{Float64}), _ctx(c), E))
    return E[]
end

"""flow_measure!(U, eps, nsteps, every): nsteps RK3 flow steps with, every `every` steps, the
measurements of samples/measurements/topologicalcharge.jl:236-245 and energydensity.jl:100-118:
a (nsteps ÷ every) × 7 matrix of rows (t, plaquette/(18V), E plaquette, E clover, Q plaquette,
Q clover, Q improved)."""
function flow_measure!(U::HIPLinks, eps::Real, nsteps::Integer, every::Integer)
    every > 0 || throw(ArgumentError("every must be positive; got $every"))
    c = _context(U)
    nrow = nsteps ÷ every
    out = Vector{Float64}(undef, 7 * nrow)
    _check(c, ccall((:gf_flow_measure, libgfhip), Cint, (Ptr{Cvoid}, Cdouble, Cint, Cint, Ptr{Float64}),
                    _ctx(c), Float64(eps), Cint(nsteps), Cint(every), out))
    return permutedims(reshape(out, 7, nrow))
end

# ----------------------------------------------------------------------------- Wilson action

# the device evaluates the loop set plaquette ∪ plaquette' only (md_driver's validation style,
# molecular_dynamics.jl:447-466): the summed coefficient of such terms, or ArgumentError
const _PLAQ = Ref{Vector{String}}()
function _plaquette_strings()
    if !isassigned(_PLAQ)
        plaq = make_loops_fromname("plaquette", Dim=4)
        _PLAQ[] = sort(string.(vcat(plaq, adjoint.(plaq))))
    end
    return _PLAQ[]
end
function _plaquette_coefficient(action::GaugeAction)
    action.hascovnet && throw(ArgumentError("use StoutGaugeAction for smeared actions on HIPBackend"))
    isempty(action.dataset) && throw(ArgumentError("the action has no terms"))
    coeff = zero(ComplexF64)
    for d in action.dataset
        sort(string.(d.closedloops)) == _plaquette_strings() || throw(ArgumentError(
            "HIPBackend evaluates exactly the loop set plaquette ∪ plaquette'; other Wilson loops are not on this path"))
        coeff += d.β
    end
    imag(coeff) == 0 || throw(ArgumentError("HIPBackend needs a real action coefficient"))
    return real(coeff)
end

# the smearing every evaluation of the context applies: none for a GaugeAction, the layers of
# a StoutGaugeAction (gf_stout_config is idempotent and allocates only on a change)
function _stout!(c::HIPContext, nlev::Int, rho::Float64, mode::Symbol)
    _check(c, ccall((:gf_stout_config, libgfhip), Cint, (Ptr{Cvoid}, Cint, Cdouble, Cint),
                    _ctx(c), Cint(nlev), rho, GF_STOUT_MODES[mode]))
    c.stout = (nlev, rho, mode)
    return nothing
end

"""evaluate_GaugeAction for push!(action, β/2, plaquette ∪ plaquette'): 2 coeff Σ Re tr P."""
function evaluate_GaugeAction(action::GaugeAction, U::HIPLinks)
    coeff = _plaquette_coefficient(action)
    c = _context(U)
    _stout!(c, 0, 0.0, :auto)
    s = Ref{Float64}(0)
    _check(c, ccall((:gf_action, libgfhip), Cint, (Ptr{Cvoid}, Cdouble, Ref{Float64}), _ctx(c), coeff, s))
    return s[]
end

"""calc_dSdUμ!(dSdUμ, action, μ, U) with a host (3, 3, NX, NY, NZ, NT) array as the output
(the device computes all four directions in one kernel)."""
"""evaluate_GaugeAction_untraced!(uout, action, U) (GaugeActions.jl:190-213) for the loop set
plaquette ∪ plaquette': the per-site loop sum coeff Σ_{μ<ν} (P_μν(x) + P_μν(x)†) into a host
array of the reference's layout (3, 3, NX, NY, NZ, NT)."""
function evaluate_GaugeAction_untraced!(uout::AbstractArray{ComplexF64,6}, action::GaugeAction, U::HIPLinks)
    c = _context(U)
    size(uout) == (3, 3, c.dims...) || throw(DimensionMismatch("uout must be (3, 3, $(join(c.dims, ", ")))"))
    _check(c, ccall((:gf_action_untraced, libgfhip), Cint, (Ptr{Cvoid}, Cdouble, Ptr{ComplexF64}),
                    _ctx(c), _plaquette_coefficient(action), uout))
    return nothing
end

"""evaluate_GaugeAction_untraced(action, U) (GaugeActions.jl:144-160): the same, allocated."""
function evaluate_GaugeAction_untraced(action::GaugeAction, U::HIPLinks)
    c = _context(U)
    uout = Array{ComplexF64,6}(undef, 3, 3, c.dims...)
    evaluate_GaugeAction_untraced!(uout, action, U)
    return uout
end

function calc_dSdUμ!(dSdUμ::AbstractArray{ComplexF64,6}, action::GaugeAction, μ, U::HIPLinks)
    1 <= μ <= 4 || throw(ArgumentError("μ must be in 1:4; got $μ"))
    coeff = _plaquette_coefficient(action)
    c = _context(U)
    _stout!(c, 0, 0.0, :auto)
    V = U[1].NV
    buf = Vector{ComplexF64}(undef, 36V)
    _check(c, ccall((:gf_dsdu, libgfhip), Cint, (Ptr{Cvoid}, Cdouble, Ptr{ComplexF64}), _ctx(c), coeff, buf))
    dSdUμ .= reshape(buf[(μ-1)*9V+1:μ*9V], size(dSdUμ))
    return nothing
end

# ----------------------------------------------------------------------------- momenta

"""initialize_TA_Gaugefields(U) / gauge_momenta(U): the momenta of U's configuration."""
initialize_TA_Gaugefields(U::HIPLinks) = _ta_of(_context(U), :momenta)
"""The context's force field (what md_force! writes on the device)."""
force_field(U::HIPLinks) = _ta_of(_context(U), :force)

function clear_U!(P::HIPMomenta)
    c = P[1].ctx
    P[1].slot === :momenta || throw(ArgumentError("only the momenta of a configuration can be cleared"))
    _check(c, ccall((:gf_clear_momenta, libgfhip), Cint, (Ptr{Cvoid},), _ctx(c)))
    return nothing
end

"""gauss_distribution!(p; σ, seed, sweep) (TA_gaugefields_4D_MPILattice.jl:157-193): every
coefficient ~ N(0, σ²) from the device's Philox4x32-10 stream (seed, sweep)."""
function gauss_distribution!(P::HIPMomenta; σ=1.0, seed=nothing, sweep::Integer=0, kwargs...)
    sweep >= 0 || throw(ArgumentError("sweep must be nonnegative; got $sweep"))
    c = P[1].ctx
    P[1].slot === :momenta || throw(ArgumentError("gauss_distribution! fills momenta"))
    s = seed === nothing ? rand(UInt64) : UInt64(seed)
    _check(c, ccall((:gf_gaussian_momenta, libgfhip), Cint, (Ptr{Cvoid}, Cdouble, Culonglong, Culonglong),
                    _ctx(c), Float64(σ), s, UInt64(sweep)))
    return P
end
gaussian_momenta!(P::HIPMomenta; sigma=1.0, seed=nothing, sweep::Integer=0, kwargs...) =
    gauss_distribution!(P; σ=sigma, seed, sweep)

"""p*p (TA_gaugefields_4D_MPILattice.jl:147-155): Σ p_a² over all links."""
function Base.:*(P::HIPMomenta, Q::HIPMomenta)
    P === Q || (P[1].ctx === Q[1].ctx && P[1].slot === Q[1].slot) ||
        throw(ArgumentError("p*p is the squared norm of one momentum field"))
    c = P[1].ctx
    s = Ref{Float64}(0)
    _check(c, ccall((:gf_momenta_dot, libgfhip), Cint, (Ptr{Cvoid}, Ref{Float64}), _ctx(c), s))
    return s[]
end

# host momenta: Vector of four Array{Float64,5}(8, NX, NY, NZ, NT), Gell-Mann order
# (TA_gaugefields_4D_serial.jl:243-260)
_host_momenta(p::AbstractArray{<:Real,5}) = Float64.(p)
_host_momenta(p::Gaugefields.AbstractGaugefields_module.TA_Gaugefields_4D_serial) = p.a
function substitute_U!(P::HIPMomenta, Pref::AbstractVector)
    c = P[1].ctx
    buf = reduce(vcat, vec.(_host_momenta.(Pref)))
    length(buf) == 32 * P[1].NV || throw(DimensionMismatch("momenta must be 8 x lattice per direction"))
    _check(c, ccall((:gf_upload_momenta, libgfhip), Cint, (Ptr{Cvoid}, Ptr{Float64}), _ctx(c), buf))
    return P
end
function Base.Array(P::HIPMomenta)
    c = P[1].ctx
    V = P[1].NV
    buf = Vector{Float64}(undef, 32V)
    sym = P[1].slot === :momenta
    rc = sym ? ccall((:gf_download_momenta, libgfhip), Cint, (Ptr{Cvoid}, Ptr{Float64}), _ctx(c), buf) :
               ccall((:gf_download_force, libgfhip), Cint, (Ptr{Cvoid}, Ptr{Float64}), _ctx(c), buf)
    _check(c, rc)
    return [reshape(buf[(μ-1)*8V+1:μ*8V], 8, c.dims...) for μ in 1:4]
end

# ----------------------------------------------------------------------------- MD provider

"""The Omelyan 2MN step P(λε) Q(ε/2) P((1-2λ)ε) Q(ε/2) P(λε) of docs/src/howtoimplement.md:94-112."""
struct OmelyanPQP
    λ::Float64
end
OmelyanPQP() = OmelyanPQP(OMELYAN_LAMBDA)

"""MD action provider: `action` on `nn`'s stout-smeared links (test/HMCstout_test_nowing.jl:36-118),
with the smearing forward/backward and the force on the device.  mode: :store, :recompute or
:auto (store each level's C when the device has room)."""
struct StoutGaugeAction
    action::GaugeAction
    levels::Int
    ρ::Float64
    mode::Symbol
end
function StoutGaugeAction(action::GaugeAction, nn::CovNeuralnet; mode::Symbol=:auto)
    levels, ρ = _stout_layers(nn)
    haskey(GF_STOUT_MODES, mode) || throw(ArgumentError("mode must be :store, :recompute or :auto"))
    return StoutGaugeAction(action, levels, ρ, mode)
end

# (coefficient, stout setting) of an action provider, applied to the context
_configure!(c::HIPContext, a::GaugeAction) = (_stout!(c, 0, 0.0, :auto); _plaquette_coefficient(a))
_configure!(c::HIPContext, a::StoutGaugeAction) =
    (_stout!(c, a.levels, a.ρ, a.mode); _plaquette_coefficient(a.action))

md_action_workspace(action::Union{GaugeAction,StoutGaugeAction}, U::HIPLinks) = nothing

function _hamiltonian(c::HIPContext, coeff)
    H = Ref{Float64}(0)
    _check(c, ccall((:gf_hamiltonian, libgfhip), Cint, (Ptr{Cvoid}, Cdouble, Ref{Float64}), _ctx(c), coeff, H))
    return H[]
end
function _momenta_dot(c::HIPContext)
    s = Ref{Float64}(0)
    _check(c, ccall((:gf_momenta_dot, libgfhip), Cint, (Ptr{Cvoid}, Ref{Float64}), _ctx(c), s))
    return s[]
end

# the potential of one provider: the Hamiltonian minus p*p/2 of the same context
function _provider_potential(c::HIPContext, action)
    coeff = _configure!(c, action)
    return _hamiltonian(c, coeff) - _momenta_dot(c) / 2
end

"""md_potential(action, U, workspace) = -real(evaluate_GaugeAction(action, U)) / NC, on the smeared
links for a StoutGaugeAction."""
function md_potential(action::Union{GaugeAction,StoutGaugeAction}, U::HIPLinks, workspace)
    return _provider_potential(_context(U), action)
end

"""md_force!(force, action, U, workspace): F = -(1/3) TA(U dS/dU) (through the stout chain for a
StoutGaugeAction) into the context's force field (`force_field(U)`)."""
function md_force!(force::HIPMomenta, action::Union{GaugeAction,StoutGaugeAction}, U::HIPLinks, workspace)
    length(force) == length(U) || throw(ArgumentError("force and U must have the same number of directions"))
    c = _context(U)
    all(f -> f.ctx === c && f.slot === :force, force) ||
        throw(ArgumentError("the HIP force is the context's force field: pass force_field(U)"))
    coeff = _configure!(c, action)
    _check(c, ccall((:gf_stout_force, libgfhip), Cint, (Ptr{Cvoid}, Cdouble), _ctx(c), coeff))
    return nothing
end

# ---- several time scales: an MDActionSet's members (GaugeAction / StoutGaugeAction providers)
# one after another on the device (molecular_dynamics.jl:45-236); the reference's own
# md_step!(::SextonWeingarten, ...) (:636-700) drives them through update_momenta!(..., group)

_members(a) = (a,)
_members(actions::MDActionSet) = values(actions.terms)
_members(actions::MDActionSet, ::MDForceGroup{Names}) where {Names} = map(n -> getproperty(actions.terms, n), Names)

md_action_workspace(actions::MDActionSet, U::HIPLinks) = nothing

"""md_potential of an MDActionSet (molecular_dynamics.jl:111-126): the sum of the members'."""
function md_potential(actions::MDActionSet, U::HIPLinks, workspace)
    c = _context(U)
    return sum(a -> _provider_potential(c, a), _members(actions))
end

# the first member's force written, every further one added (_md_write_named_forces!,
# molecular_dynamics.jl:128-183)
function _set_force!(force::HIPMomenta, U::HIPLinks, members)
    length(force) == length(U) || throw(ArgumentError("force and U must have the same number of directions"))
    c = _context(U)
    all(f -> f.ctx === c && f.slot === :force, force) ||
        throw(ArgumentError("the HIP force is the context's force field: pass force_field(U)"))
    for (k, a) in enumerate(members)
        coeff = _configure!(c, a)
        if k == 1
            _check(c, ccall((:gf_stout_force, libgfhip), Cint, (Ptr{Cvoid}, Cdouble), _ctx(c), coeff))
        else
            _check(c, ccall((:gf_stout_force_add, libgfhip), Cint, (Ptr{Cvoid}, Cdouble), _ctx(c), coeff))
        end
    end
    return nothing
end

"""md_force!(force, actions::MDActionSet, U, workspace) (molecular_dynamics.jl:201-214): the sum
of the members' forces in the context's force field."""
function md_force!(force::HIPMomenta, actions::MDActionSet, U::HIPLinks, workspace)
    return _set_force!(force, U, _members(actions))
end

"""md_force!(force, actions, U, workspace, group) (molecular_dynamics.jl:216-236): only the
members the MDForceGroup names."""
function md_force!(force::HIPMomenta, actions::MDActionSet, U::HIPLinks, workspace, group::MDForceGroup)
    return _set_force!(force, U, _members(actions, group))
end

# P += step F_m for each member m: force and kick in one library call each
function _kick!(c::HIPContext, members, step_size)
    isfinite(step_size) || throw(ArgumentError("the momentum step size must be finite; got $step_size"))
    for a in members
        coeff = _configure!(c, a)
        _check(c, ccall((:gf_stout_force_kick, libgfhip), Cint, (Ptr{Cvoid}, Cdouble, Cdouble),
                        _ctx(c), coeff, Float64(step_size)))
    end
    return nothing
end

"""md_hamiltonian(U, p, driver) (molecular_dynamics.jl:494-505): md_potential + p*p/2."""
function md_hamiltonian(U::HIPLinks, P::HIPMomenta, driver::MDDriver)
    c = _same(U, P)
    driver.action isa MDActionSet && return sum(a -> _provider_potential(c, a), _members(driver.action)) +
                                            _momenta_dot(c) / 2
    return _hamiltonian(c, _configure!(c, driver.action))
end

"""update_gaugefields!(U, P, step_size, driver) (molecular_dynamics.jl:513-531): U_μ ← exp(step P_μ) U_μ."""
function update_gaugefields!(U::HIPLinks, P::HIPMomenta, step_size, driver::MDDriver)
    c = _same(U, P)
    isfinite(step_size) || throw(ArgumentError("the gauge-field step size must be finite; got $step_size"))
    _check(c, ccall((:gf_exp_update, libgfhip), Cint, (Ptr{Cvoid}, Cdouble), _ctx(c), Float64(step_size)))
    return U
end

"""update_momenta!(P, U, step_size, driver) (molecular_dynamics.jl:539-551): P ← P + step F(U),
force and kick in one kernel (through the stout chain for a StoutGaugeAction; member by member
for an MDActionSet)."""
function update_momenta!(P::HIPMomenta, U::HIPLinks, step_size, driver::MDDriver)
    _kick!(_same(U, P), _members(driver.action), step_size)
    return P
end

"""update_momenta!(P, U, step_size, driver, group) (molecular_dynamics.jl:553-583): the kick of the
members of the driver's MDActionSet that the MDForceGroup names (Sexton-Weingarten's slow and
fast kicks)."""
function update_momenta!(P::HIPMomenta, U::HIPLinks, step_size, driver::MDDriver, group::MDForceGroup)
    driver.action isa MDActionSet ||
        throw(ArgumentError("a force group selects members of an MDActionSet; got $(typeof(driver.action))"))
    _kick!(_same(U, P), _members(driver.action, group), step_size)
    return P
end

function md_step!(integ::OmelyanPQP, U::HIPLinks, P::HIPMomenta, step_size, driver::MDDriver)
    λ = integ.λ
    update_momenta!(P, U, λ * step_size, driver)
    update_gaugefields!(U, P, step_size / 2, driver)
    update_momenta!(P, U, (1 - 2λ) * step_size, driver)
    update_gaugefields!(U, P, step_size / 2, driver)
    update_momenta!(P, U, λ * step_size, driver)
    return nothing
end

_integrator_code(::QPQ) = (GF_QPQ, OMELYAN_LAMBDA)
_integrator_code(::PQP) = (GF_PQP, OMELYAN_LAMBDA)
_integrator_code(i::OmelyanPQP) = (GF_OMELYAN2MN, i.λ)

"""md_trajectory!(U, P, driver; diagnostics=true, fused=false) (molecular_dynamics.jl:712-730).
QPQ, PQP and OmelyanPQP of a single action run as one library call per trajectory (`fused=true` merges adjacent
half steps into single force + kick + exp kernels); other integrators step through md_step!."""
function md_trajectory!(U::HIPLinks, P::HIPMomenta, driver::MDDriver; diagnostics::Bool=true,
                        fused::Bool=false)
    c = _same(U, P)
    integ = driver.integrator
    if integ isa Union{QPQ,PQP,OmelyanPQP} && !(driver.action isa MDActionSet)
        coeff = _configure!(c, driver.action)
        code, λ = _integrator_code(integ)
        H = zeros(Float64, 3)
        _check(c, ccall((:gf_md_trajectory, libgfhip), Cint,
                        (Ptr{Cvoid}, Cdouble, Cint, Cdouble, Cint, Cdouble, Ptr{Float64}),
                        _ctx(c), coeff, code | (fused ? GF_FUSED : Cint(0)), λ, Cint(driver.steps),
                        Float64(driver.trajectory_length), diagnostics ? H : C_NULL))
        diagnostics || return nothing
        return (initial_hamiltonian=H[1], final_hamiltonian=H[2], delta_hamiltonian=H[3])
    end
    h0 = diagnostics ? md_hamiltonian(U, P, driver) : nothing
    ε = md_step_size(driver)
    for _ in 1:driver.steps
        md_step!(integ, U, P, ε, driver)
    end
    diagnostics || return nothing
    h1 = md_hamiltonian(U, P, driver)
    return (initial_hamiltonian=h0, final_hamiltonian=h1, delta_hamiltonian=h1 - h0)
end

"""Metropolis backup: substitute_U!(Uold, U) / substitute_U!(U, Uold) between a configuration and
`backup(U)` stay on the device (test/HMCstout_test_nowing.jl:47-74)."""
struct HIPBackup
    ctx::HIPContext
end
backup(U::HIPLinks) = HIPBackup(_context(U))
function substitute_U!(b::HIPBackup, U::HIPLinks)
    _context(U) === b.ctx || throw(ArgumentError("a backup belongs to one configuration"))
    _check(b.ctx, ccall((:gf_save_links, libgfhip), Cint, (Ptr{Cvoid},), _ctx(b.ctx)))
    return b
end
function substitute_U!(U::HIPLinks, b::HIPBackup)
    _context(U) === b.ctx || throw(ArgumentError("a backup belongs to one configuration"))
    _check(b.ctx, ccall((:gf_restore_links, libgfhip), Cint, (Ptr{Cvoid},), _ctx(b.ctx)))
    return U
end

# ----------------------------------------------------------------------------- Wilson flow

"""flow!(U, g::Gradientflow) (gradientflow.jl:171-238): g.Nflow Lüscher RK3 steps of size g.eps
on the device (the temporaries of `g` are not used)."""
function flow!(U::HIPLinks, g::Gradientflow)
    c = _context(U)
    E = zeros(Float64, 1)
    _check(c, ccall((:gf_flow, libgfhip), Cint, (Ptr{Cvoid}, Cdouble, Cint, Cint, Ptr{Float64}),
                    _ctx(c), Float64(g.eps), Cint(g.Nflow), Cint(0), E))
    return nothing
end

"""flow! with the plaquette energy density E(t) = 36 (1 - plaquette) sampled every `every` steps."""
function flow!(U::HIPLinks, eps::Real, nsteps::Integer, every::Integer)
    every > 0 || throw(ArgumentError("every must be positive; got $every"))
    c = _context(U)
    E = zeros(Float64, nsteps ÷ every)
    _check(c, ccall((:gf_flow, libgfhip), Cint, (Ptr{Cvoid}, Cdouble, Cint, Cint, Ptr{Float64}),
                    _ctx(c), Float64(eps), Cint(nsteps), Cint(every), E))
    return E
end

# ----------------------------------------------------------------------------- stout smearing

# levels and rho of a CovNeuralnet of plaquette STOUT_Layers with one common rho
function _stout_layers(nn::CovNeuralnet)
    layers = nn.layers
    isempty(layers) && return (0, 0.0)
    all(l -> l isa STOUT_Layer, layers) || throw(ArgumentError("HIPBackend smears with STOUT_Layers only"))
    ρs = [l.ρs for l in layers]
    all(r -> length(r) == 1, ρs) && all(r -> r[1] == ρs[1][1], ρs) ||
        throw(ArgumentError("HIPBackend smears with plaquette layers of one common rho"))
    return (length(layers), Float64(real(ρs[1][1])))
end

# the device keeps every level; this view is level `level` (1..nlev) of the context's smearing
struct HIPSmearedLevel
    ctx::HIPContext
    level::Int
end

function _smear!(c::HIPContext, nn::CovNeuralnet, host::Union{Nothing,Vector{ComplexF64}})
    nlev, ρ = _stout_layers(nn)
    _stout!(c, nlev, ρ, c.stout[3])
    _check(c, ccall((:gf_stout_smear, libgfhip), Cint, (Ptr{Cvoid}, Ptr{ComplexF64}), _ctx(c),
                    host === nothing ? C_NULL : host))
    return nlev
end

"""calc_smearedU(U, nn) (Abstractsmearing.jl:237-272): the smeared configuration as host arrays
(four (3, 3, NX, NY, NZ, NT)), the level views the device keeps, and no dSdU temporaries."""
function calc_smearedU(Uin::HIPLinks, nn::CovNeuralnet; calcdSdU=false, temps=nothing)
    c = _context(Uin)
    V = Uin[1].NV
    host = Vector{ComplexF64}(undef, 36V)
    nlev = _smear!(c, nn, host)
    nlev == 0 && return Uin, nothing, nothing
    Uout = [reshape(host[(μ-1)*9V+1:μ*9V], 3, 3, c.dims...) for μ in 1:4]
    return Uout, [HIPSmearedLevel(c, l) for l in 1:nlev], nothing
end

# D (reference dS/dU convention: δS = 2 Re tr(D δU)) through the device's stout chain
function _stout_backward(c::HIPContext, D::AbstractVector)
    V = prod(c.dims)
    dout = reduce(vcat, [vec(ComplexF64.(d)) for d in D])
    length(dout) == 36V || throw(DimensionMismatch("dS/dU must be four 3x3 fields on $(c.dims)"))
    din = Vector{ComplexF64}(undef, 36V)
    _check(c, ccall((:gf_stout_backward, libgfhip), Cint, (Ptr{Cvoid}, Ptr{ComplexF64}, Ptr{ComplexF64}),
                    _ctx(c), dout, din))
    return [reshape(din[(μ-1)*9V+1:μ*9V], 3, 3, c.dims...) for μ in 1:4]
end

"""back_prop(δL, nn, Uout_multi, Uin) (Abstractsmearing.jl:316-411): dS/dU of the bare links
from dS/dU of the smeared ones, host arrays in the reference's convention."""
function back_prop(δL::AbstractVector, nn::CovNeuralnet, Uout_multi::Vector{HIPSmearedLevel}, Uin::HIPLinks)
    c = _context(Uin)
    all(l -> l.ctx === c, Uout_multi) || throw(ArgumentError("Uout_multi must come from calc_smearedU(Uin, nn)"))
    _smear!(c, nn, nothing)  # the levels of Uin (gf_stout_backward uses the last smear)
    return _stout_backward(c, δL)
end

"""forward!(layer, Uout, ρs, Uin) for one STOUT_Layer (stout_fast.jl:250-274): Uout (host
arrays) = exp(Q) Uin with the device's level-1 smearing at ρs[1]."""
function forward!(layer::STOUT_Layer, Uout::AbstractVector, ρs::Vector{<:Number}, Uin::HIPLinks)
    length(ρs) == 1 || throw(ArgumentError("HIPBackend smears with one plaquette rho per layer"))
    c = _context(Uin)
    V = Uin[1].NV
    _stout!(c, 1, Float64(real(ρs[1])), c.stout[3])
    host = Vector{ComplexF64}(undef, 36V)
    _check(c, ccall((:gf_stout_smear, libgfhip), Cint, (Ptr{Cvoid}, Ptr{ComplexF64}), _ctx(c), host))
    for μ in 1:4
        Uout[μ] .= reshape(host[(μ-1)*9V+1:μ*9V], 3, 3, c.dims...)
    end
    return nothing
end

"""layer_pullback!(δ_prev, δ_current, layer, Uprev, temps, tempf) (stout_fast.jl:223-247): the
cotangent through one layer at Uprev (host arrays)."""
function layer_pullback!(δ_prev::AbstractVector, δ_current::AbstractVector, layer::STOUT_Layer,
                         Uprev::HIPLinks, temps, tempf)
    c = _context(Uprev)
    length(layer.ρs) == 1 || throw(ArgumentError("HIPBackend smears with one plaquette rho per layer"))
    _stout!(c, 1, Float64(real(layer.ρs[1])), c.stout[3])
    _check(c, ccall((:gf_stout_smear, libgfhip), Cint, (Ptr{Cvoid}, Ptr{ComplexF64}), _ctx(c), C_NULL))
    din = _stout_backward(c, δ_current)
    for μ in 1:4
        δ_prev[μ] .= din[μ]
    end
    return nothing
end

"""Smearing map for AD (host arrays in and out): stout_smear(U, nn, links) sets the device
configuration to `links`, smears with nn's layers and returns the top level.  Its ChainRules
rrule and Enzyme reverse rule (ext/) pull the cotangent back through gf_stout_backward."""
function stout_smear(U::HIPLinks, nn::CovNeuralnet, links::AbstractVector)
    c = _context(U)
    _upload_links!(c, links)
    V = U[1].NV
    host = Vector{ComplexF64}(undef, 36V)
    _smear!(c, nn, host)
    return [reshape(host[(μ-1)*9V+1:μ*9V], 3, 3, c.dims...) for μ in 1:4]
end
function stout_smear!(out::AbstractVector, U::HIPLinks, nn::CovNeuralnet, links::AbstractVector)
    r = stout_smear(U, nn, links)
    for μ in 1:4
        out[μ] .= r[μ]
    end
    return nothing
end

"""Cotangent of a real loss through stout_smear(U, nn, links): Ū (∂L/∂Re U + i ∂L/∂Im U, the
ChainRules/Enzyme convention) -> D = Ū†/2 per site (δL = 2 Re tr(D δU), the library's
convention) -> gf_stout_backward -> 2 D_bare† per site.  The links are uploaded and smeared
again first, so the levels the backward pass reads are those of `links`."""
function stout_pullback(U::HIPLinks, nn::CovNeuralnet, links::AbstractVector, Ūout::AbstractVector)
    c = _context(U)
    _upload_links!(c, links)
    _smear!(c, nn, nothing)
    dag(A) = conj(permutedims(A, (2, 1, 3, 4, 5, 6)))
    D = [dag(ComplexF64.(Ū)) ./ 2 for Ū in Ūout]
    return [2 .* dag(d) for d in _stout_backward(c, D)]
end

# ----------------------------------------------------------------------------- heatbath

"""heatbath!(U, h::Heatbath): one plaquette-action heatbath sweep on the device (Cabibbo-Marinari
subgroups with Kennedy-Pendleton SU(2) steps), stream (h.seed, h.sweep); h.sweep advances."""
function heatbath!(U::HIPLinks, h::Heatbath)
    c = _context(U)
    failures = Ref{Clonglong}(0)
    _check(c, ccall((:gf_heatbath, libgfhip), Cint,
                    (Ptr{Cvoid}, Cdouble, Culonglong, Culonglong, Cint, Ref{Clonglong}),
                    _ctx(c), h.β, h.seed, h.sweep, Cint(min(h.ITERATION_MAX, 65534)), failures))
    h.sweep += 1
    return Int(failures[])
end

"""overrelaxation!(U, h::Heatbath): three SU(2) reflections per link, action unchanged."""
function overrelaxation!(U::HIPLinks, h::Heatbath)
    c = _context(U)
    _check(c, ccall((:gf_overrelaxation, libgfhip), Cint, (Ptr{Cvoid}, Culonglong, Culonglong),
                    _ctx(c), h.seed, h.overrelaxation_sweep))
    h.overrelaxation_sweep += 1
    return nothing
end

# ----------------------------------------------------------------------------- ILDG files

function _check_ildg(rc::Cint)
    rc == 0 || throw(ArgumentError(unsafe_string(ccall((:gf_ildg_last_error, libgfhip), Cstring, ()))))
    return nothing
end

"""load_gaugefield!(U, i, ildg, L, NC) (ildg_format.jl:902-969): configuration `i` (1) of the file."""
function load_gaugefield!(U::HIPLinks, i, ildg::ILDG, L, NC)
    NC == 3 || throw(ArgumentError("HIPBackend loads SU(3) configurations; got NC=$NC"))
    i == 1 || throw(ArgumentError("an ILDG file holds one configuration; got index $i"))
    c = _context(U)
    Tuple(L) == c.dims || throw(DimensionMismatch("file lattice $L, configuration $(c.dims)"))
    buf = Vector{ComplexF64}(undef, 36 * prod(c.dims))
    prec = Ref{Cint}(0)
    _check_ildg(ccall((:gf_ildg_read, libgfhip), Cint, (Cstring, Ptr{Cint}, Ptr{ComplexF64}, Ref{Cint}),
                      ildg.filename, Cint[c.dims...], buf, prec))
    _check(c, ccall((:gf_upload_links, libgfhip), Cint, (Ptr{Cvoid}, Ptr{ComplexF64}), _ctx(c), buf))
    return nothing
end

"""save_binarydata(U, filename; precision=64) (ildg_format.jl:797-833)."""
function save_binarydata(U::HIPLinks, filename; precision::Integer=64)
    precision in (32, 64) || throw(ArgumentError("precision must be 32 or 64; got $precision"))
    c = _context(U)
    buf = Vector{ComplexF64}(undef, 36 * prod(c.dims))
    _check(c, ccall((:gf_download_links, libgfhip), Cint, (Ptr{Cvoid}, Ptr{ComplexF64}), _ctx(c), buf))
    _check_ildg(ccall((:gf_ildg_write, libgfhip), Cint, (Cstring, Ptr{Cint}, Ptr{ComplexF64}, Cint),
                      filename, Cint[c.dims...], buf, Cint(precision)))
    return nothing
end

end # module
