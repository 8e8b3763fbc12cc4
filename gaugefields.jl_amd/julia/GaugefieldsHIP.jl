"""
    GaugefieldsHIP

Julia-side binding of libgfhip.so (include/gfhip.h) for Gaugefields.jl.  A maintainer adds
this module (or its methods) to Gaugefields.jl to route the SU(3) hot path of a
configuration created with `gauge_configuration(...; backend=HIPBackend())` to the MI355X
kernels.  The methods specialize the reference's own entry points:

  calculate_Plaquette        (AbstractGaugefields.jl:2684-2699)
  evaluate_GaugeAction        (GaugeActions.jl:132-142)
  md_potential / md_force!   (molecular_dynamics.jl:247-267)
  update_gaugefields! / update_momenta! / md_trajectory!  (molecular_dynamics.jl:513-730)
  flow!(U, ::Gradientflow)   (gradientflow.jl:171-238)
  back_prop / STOUT_Layer    (Abstractsmearing.jl:352-411, stout_fast.jl:223-274)
  gauss_distribution!, hot start, heatbath! / overrelaxation!, topological_charge,
  calculate_energy_density, ILDG load_gaugefield! / save_binarydata

This file is not executed in this repository (there is no Julia toolchain here); the same
C ABI is exercised by the Python ctypes mirror in ../api.py and the tests.
"""
module GaugefieldsHIP

using Libdl

const libgfhip = joinpath(@__DIR__, "..", "libgfhip.so")

const GF_F64 = Cint(0)
const GF_F32 = Cint(1)
const GF_QPQ = Cint(0)
const GF_PQP = Cint(1)
const GF_OMELYAN2MN = Cint(2)
const GF_FUSED = Cint(0x100)

"""Device-resident SU(3) configuration: the four link fields plus its momenta and work fields."""
mutable struct HIPConfiguration
    ctx::Ptr{Cvoid}
    dims::NTuple{4,Int}
    precision::Cint
    function HIPConfiguration(dims::NTuple{4,Int}; eltype::Type=ComplexF64, device::Integer=0)
        prec = eltype == ComplexF64 ? GF_F64 :
               eltype == ComplexF32 ? GF_F32 :
               throw(ArgumentError("eltype must be ComplexF64 or ComplexF32"))
        ctx = Ref{Ptr{Cvoid}}(C_NULL)
        cdims = Cint[dims...]
        rc = ccall((:gf_create, libgfhip), Cint, (Ptr{Cint}, Cint, Cint, Ref{Ptr{Cvoid}}),
                   cdims, prec, Cint(device), ctx)
        rc == 0 || error("gf_create failed with status $rc")
        U = new(ctx[], dims, prec)
        finalizer(U) do u
            u.ctx == C_NULL || ccall((:gf_destroy, libgfhip), Cint, (Ptr{Cvoid},), u.ctx)
            u.ctx = C_NULL
        end
        return U
    end
end

function _check(U::HIPConfiguration, rc::Cint)
    if rc != 0
        msg = unsafe_string(ccall((:gf_last_error, libgfhip), Cstring, (Ptr{Cvoid},), U.ctx))
        rc == -1 ? throw(ArgumentError(msg)) : error("libgfhip status $rc: $msg")
    end
    return nothing
end

# links: Vector of four Array{ComplexF64,6}(3,3,NX,NY,NZ,NT) (interior), concatenated
function upload_links!(U::HIPConfiguration, links::Vector{<:AbstractArray{ComplexF64,6}})
    buf = reduce(vcat, vec.(links))
    _check(U, ccall((:gf_upload_links, libgfhip), Cint, (Ptr{Cvoid}, Ptr{ComplexF64}), U.ctx, buf))
end

function download_links(U::HIPConfiguration)
    V = prod(U.dims)
    buf = Vector{ComplexF64}(undef, 4 * 9 * V)
    _check(U, ccall((:gf_download_links, libgfhip), Cint, (Ptr{Cvoid}, Ptr{ComplexF64}), U.ctx, buf))
    return [reshape(buf[(mu-1)*9V+1:mu*9V], 3, 3, U.dims...) for mu in 1:4]
end

# momenta: Vector of four Array{Float64,5}(8,NX,NY,NZ,NT) (TA_Gaugefields a[k,...])
function upload_momenta!(U::HIPConfiguration, p::Vector{<:AbstractArray{Float64,5}})
    buf = reduce(vcat, vec.(p))
    _check(U, ccall((:gf_upload_momenta, libgfhip), Cint, (Ptr{Cvoid}, Ptr{Float64}), U.ctx, buf))
end

function calculate_Plaquette(U::HIPConfiguration)
    s = Ref{Float64}(0)
    _check(U, ccall((:gf_plaquette, libgfhip), Cint, (Ptr{Cvoid}, Ref{Float64}), U.ctx, s))
    return s[]
end

"evaluate_GaugeAction for push!(action, coeff, plaquette ∪ plaquette')"
function evaluate_plaquette_action(U::HIPConfiguration, coeff::Real)
    s = Ref{Float64}(0)
    _check(U, ccall((:gf_action, libgfhip), Cint, (Ptr{Cvoid}, Cdouble, Ref{Float64}), U.ctx, coeff, s))
    return s[]
end

md_potential(U::HIPConfiguration, coeff::Real) = -evaluate_plaquette_action(U, coeff) / 3

update_momenta!(U::HIPConfiguration, coeff::Real, eps::Real) =
    _check(U, ccall((:gf_force_kick, libgfhip), Cint, (Ptr{Cvoid}, Cdouble, Cdouble), U.ctx, coeff, eps))

update_gaugefields!(U::HIPConfiguration, eps::Real) =
    _check(U, ccall((:gf_exp_update, libgfhip), Cint, (Ptr{Cvoid}, Cdouble), U.ctx, eps))

function md_trajectory!(U::HIPConfiguration, coeff::Real, steps::Integer, tau::Real;
                        integrator::Cint=GF_QPQ, lambda::Real=0.1931833275037836, fused::Bool=false)
    H = zeros(Float64, 3)
    code = integrator | (fused ? GF_FUSED : Cint(0))
    _check(U, ccall((:gf_md_trajectory, libgfhip), Cint,
                    (Ptr{Cvoid}, Cdouble, Cint, Cdouble, Cint, Cdouble, Ptr{Float64}),
                    U.ctx, coeff, code, lambda, Cint(steps), tau, H))
    return (initial_hamiltonian=H[1], final_hamiltonian=H[2], delta_hamiltonian=H[3])
end

function flow!(U::HIPConfiguration, eps::Real, nsteps::Integer; every::Integer=0)
    E = zeros(Float64, every > 0 ? nsteps ÷ every : 1)
    _check(U, ccall((:gf_flow, libgfhip), Cint, (Ptr{Cvoid}, Cdouble, Cint, Cint, Ptr{Float64}),
                    U.ctx, eps, Cint(nsteps), Cint(every), E))
    return every > 0 ? E : nothing
end

stout_config!(U::HIPConfiguration, nlev::Integer, rho::Real; recompute::Bool=false) =
    _check(U, ccall((:gf_stout_config, libgfhip), Cint, (Ptr{Cvoid}, Cint, Cdouble, Cint),
                    U.ctx, Cint(nlev), rho, Cint(recompute ? 1 : 0)))

# Seeded fields drawn on the device (Philox4x32-10; stream key = seed, sweep in the counter):
# gauss_distribution! / randomGaugefields (TA_gaugefields_4D_MPILattice.jl:157-193,
# gaugefields_4D_MPILattice.jl:430-472)
gaussian_momenta!(U::HIPConfiguration; sigma::Real=1.0, seed::Integer=0, sweep::Integer=0) =
    _check(U, ccall((:gf_gaussian_momenta, libgfhip), Cint, (Ptr{Cvoid}, Cdouble, Culonglong, Culonglong),
                    U.ctx, sigma, seed, sweep))

hot_start!(U::HIPConfiguration; seed::Integer=0) =
    _check(U, ccall((:gf_hot_start, libgfhip), Cint, (Ptr{Cvoid}, Culonglong), U.ctx, seed))

# heatbath! / overrelaxation! for the plaquette action (heatbath/heatbathmodule.jl:50-200,
# 1248-1321, 2119-2276); ITERATION_MAX above 65534 acts as 65534
function heatbath!(U::HIPConfiguration, beta::Real; seed::Integer=0, sweep::Integer=0,
                   ITERATION_MAX::Integer=65534)
    failures = Ref{Clonglong}(0)
    _check(U, ccall((:gf_heatbath, libgfhip), Cint,
                    (Ptr{Cvoid}, Cdouble, Culonglong, Culonglong, Cint, Ref{Clonglong}),
                    U.ctx, beta, seed, sweep, Cint(min(ITERATION_MAX, 65534)), failures))
    return Int(failures[])
end

overrelaxation!(U::HIPConfiguration; seed::Integer=0, sweep::Integer=0) =
    _check(U, ccall((:gf_overrelaxation, libgfhip), Cint, (Ptr{Cvoid}, Culonglong, Culonglong),
                    U.ctx, seed, sweep))

# topological_charge / calculate_energy_density (AbstractGaugefields.jl:1172-1471;
# samples/measurements/energydensity.jl:32-76)
const GF_TOPO_METHODS = Dict(:plaquette => Cint(0), :clover => Cint(1), :improved => Cint(2))

function topological_charge(U::HIPConfiguration; method::Symbol=:plaquette)
    haskey(GF_TOPO_METHODS, method) ||
        throw(ArgumentError("supported topological_charge methods are :plaquette, :clover, and :improved"))
    Q = Ref{Float64}(0)
    _check(U, ccall((:gf_topological_charge, libgfhip), Cint, (Ptr{Cvoid}, Cint, Ref{Float64}),
                    U.ctx, GF_TOPO_METHODS[method], Q))
    return Q[]
end

function calculate_energy_density(U::HIPConfiguration)
    E = Ref{Float64}(0)
    _check(U, ccall((:gf_energy_density_clover, libgfhip), Cint, (Ptr{Cvoid}, Ref{Float64}), U.ctx, E))
    return E[]
end

# ILDG files (ildg_format.jl:686-969): host I/O in the library, no CLIME executables
function _check_ildg(rc::Cint)
    rc == 0 || throw(ArgumentError(unsafe_string(ccall((:gf_ildg_last_error, libgfhip), Cstring, ()))))
    return nothing
end

function load_gaugefield!(U::HIPConfiguration, filename::AbstractString)
    buf = Vector{ComplexF64}(undef, 4 * 9 * prod(U.dims))
    prec = Ref{Cint}(0)
    _check_ildg(ccall((:gf_ildg_read, libgfhip), Cint, (Cstring, Ptr{Cint}, Ptr{ComplexF64}, Ref{Cint}),
                      filename, Cint[U.dims...], buf, prec))
    _check(U, ccall((:gf_upload_links, libgfhip), Cint, (Ptr{Cvoid}, Ptr{ComplexF64}), U.ctx, buf))
    return Int(prec[])
end

function save_binarydata(U::HIPConfiguration, filename::AbstractString; precision::Integer=64)
    buf = Vector{ComplexF64}(undef, 4 * 9 * prod(U.dims))
    _check(U, ccall((:gf_download_links, libgfhip), Cint, (Ptr{Cvoid}, Ptr{ComplexF64}), U.ctx, buf))
    _check_ildg(ccall((:gf_ildg_write, libgfhip), Cint, (Cstring, Ptr{Cint}, Ptr{ComplexF64}, Cint),
                      filename, Cint[U.dims...], buf, Cint(precision)))
end

end # module
