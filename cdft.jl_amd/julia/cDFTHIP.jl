# cDFTHIP.jl -- Julia binding of libcdft_hip.so for cDFT.jl.
#
# Adds a device type `HIPNative` for `DFTOptions` and a more specific
# `converge!(prob::DFTProblem, method::AASol, rho)` for PC-SAFT systems built with
# it (src/methods/converge.jl:220).  System construction, bulk thermodynamics
# (Clapeyron) and the external potential stay in cDFT.jl on the host; the functional
# evaluation and the Picard/Anderson loop run in the HIP library.
#
#   using Clapeyron, cDFT, cDFTHIP
#   system = DFTSystem(model, structure, surface, DFTOptions(HIPNative(0)))
#   rho = initialize_profiles(system)
#   converge!(system, rho; tol = 1e-8)          # rho overwritten, returns nothing
#   cDFTHIP.last_result()                        # (idid, errcode, iters, ...)
module cDFTHIP

using cDFT, Clapeyron
import KernelAbstractions

const libcdft = get(ENV, "CDFT_HIP_LIB", joinpath(@__DIR__, "..", "libcdft_hip.so"))

"""Device backend whose arrays live on the host; the library moves them."""
struct HIPNative <: KernelAbstractions.Backend
    device::Int
end
HIPNative() = HIPNative(0)
KernelAbstractions.allocate(::HIPNative, ::Type{T}, dims::Tuple) where {T} = Array{T}(undef, dims)
KernelAbstractions.synchronize(::HIPNative) = nothing
cDFT.DFTOptions(dev::HIPNative; precision = Float64) = cDFT.DFTOptions(dev, :reverse; precision)

# --- C structs (include/cdft_hip.h) -------------------------------------------
struct CdftSystemDesc
    nc::Cint; ndim::Cint; ngrid::NTuple{3,Cint}
    lb::NTuple{3,Cdouble}; ub::NTuple{3,Cdouble}
    T::Cdouble; p::Cdouble; L::Cdouble; psi::Cdouble
    m::Ptr{Cdouble}; d::Ptr{Cdouble}; sigma::Ptr{Cdouble}; eps::Ptr{Cdouble}
    rho_bulk::Ptr{Cdouble}; mu_res::Ptr{Cdouble}; vext::Ptr{Cdouble}
    precision::Cint
    assoc_nsite::Cint; assoc_npair::Cint
    assoc_site_comp::Ptr{Cint}; assoc_site_count::Ptr{Cdouble}; assoc_pair_site::Ptr{Cint}
    assoc_pair_eps::Ptr{Cdouble}; assoc_pair_kappa::Ptr{Cdouble}
    coord::Cint                    # 0 Cartesian, 1 spherical, 2 cylindrical
end

struct CdftSolverOpts
    maxit::Cint; beta::Cdouble; tol::Cdouble; anderson_start::Cdouble
    m::Cint; hist_cap::Cint; hist::Ptr{Cdouble}
    progress::Ptr{Cvoid}; progress_user::Ptr{Cvoid}; progress_every::Cint
end
CdftSolverOpts(maxit, beta, tol, anderson_start, m) =
    CdftSolverOpts(maxit, beta, tol, anderson_start, m, 0, C_NULL, C_NULL, C_NULL, 0)

struct CdftResult
    idid::Cint; errcode::Cint; iters::Cint; nevals::Cint; nhist::Cint
    resnorm::Cdouble; tol::Cdouble; nevals_fp32::Cint; reserved::Cint; device_ms::Cdouble
end

const CTX = Ref{Ptr{Cvoid}}(C_NULL)
const LAST = Ref{Any}(nothing)
last_result() = LAST[]

function _ctx(dev::Int)
    if CTX[] == C_NULL
        h = Ref{Ptr{Cvoid}}(C_NULL)
        rc = ccall((:cdft_ctx_create, libcdft), Cint, (Cint, Ref{Ptr{Cvoid}}), dev, h)
        rc == 0 || error("cdft_ctx_create failed: ", unsafe_string(ccall((:cdft_last_error, libcdft), Cstring, (Ptr{Cvoid},), C_NULL)))
        CTX[] = h[]
    end
    return CTX[]
end

_check(rc, ctx) = rc == 0 || error("cdft error $rc: ", unsafe_string(ccall((:cdft_last_error, libcdft), Cstring, (Ptr{Cvoid},), ctx)))

# Wertheim association data in the descriptor's layout: pack_assoc_params
# (src/models/DFT/assoc.jl:18-57) with 0-based global site indices
function _assoc(model)
    site_comp = Cint[]; site_count = Float64[]
    pair_site = Cint[]; pair_eps = Float64[]; pair_kappa = Float64[]
    if hasproperty(model.params, :bondvol) && !isempty(model.params.bondvol.values.values)
        sites = Clapeyron.getsites(model)
        nper = [length(sites.n_sites[i]) for i in 1:length(model)]
        off = cumsum([0; nper])
        for i in 1:length(model), a in 1:nper[i]
            push!(site_comp, i - 1)
            push!(site_count, sites.n_sites[i][a])
        end
        ev = model.params.epsilon_assoc.values
        kv = model.params.bondvol.values
        for idx in 1:length(ev.values)
            i, j = ev.outer_indices[idx]
            a, b = ev.inner_indices[idx]
            push!(pair_site, off[i] + a - 1, off[j] + b - 1)
            push!(pair_eps, ev.values[idx])
            push!(pair_kappa, kv.values[idx])
        end
    end
    return site_comp, site_count, pair_site, pair_eps, pair_kappa
end

function _system(system, ctx)
    model = system.model
    st = system.structure
    site_comp, site_count, pair_site, pair_eps, pair_kappa = _assoc(model)
    nd = cDFT.dimension(st)
    (p, T) = st.conditions
    ngrid = ntuple(i -> i <= nd ? Cint(st.ngrid[i]) : Cint(1), 3)
    lb = ntuple(i -> i <= nd ? st.bounds[i][1] : 0.0, 3)
    ub = ntuple(i -> i <= nd ? st.bounds[i][2] : 0.0, 3)
    m = Float64.(model.params.segment.values)
    d = Float64.(system.species.size)
    sigma = Matrix{Float64}(model.params.sigma.values)
    eps = Matrix{Float64}(model.params.epsilon.values)
    rhob = Float64.(system.species.bulk_density)
    mu = Float64.(system.species.chempot_res)
    vext = nothing
    if system.external_field !== nothing
        cache = cDFT.preallocate_external_potential(system, zeros(st.ngrid..., length(model)))
        vext = sum(c[end] for c in cache)          # beta V_ext (steele.jl:49-81)
    end
    prec = cDFT.fptype(system.options) == Float64 ? Cint(0) : Cint(1)
    coord = st isa cDFT.DFTStructByCoord{cDFT.Spherical} ? Cint(1) :
            st isa cDFT.DFTStructByCoord{cDFT.Cylindrical} ? Cint(2) : Cint(0)
    npair = length(pair_eps)
    GC.@preserve m d sigma eps rhob mu vext site_comp site_count pair_site pair_eps pair_kappa begin
        desc = CdftSystemDesc(length(model), nd, ngrid, lb, ub, T, p, cDFT.length_scale(model), 1.3862,
                              pointer(m), pointer(d), pointer(sigma), pointer(eps), pointer(rhob), pointer(mu),
                              vext === nothing ? Ptr{Cdouble}(C_NULL) : pointer(vext), prec,
                              npair > 0 ? length(site_comp) : 0, npair,
                              npair > 0 ? pointer(site_comp) : Ptr{Cint}(C_NULL),
                              npair > 0 ? pointer(site_count) : Ptr{Cdouble}(C_NULL),
                              npair > 0 ? pointer(pair_site) : Ptr{Cint}(C_NULL),
                              npair > 0 ? pointer(pair_eps) : Ptr{Cdouble}(C_NULL),
                              npair > 0 ? pointer(pair_kappa) : Ptr{Cdouble}(C_NULL), coord)
        h = Ref{Ptr{Cvoid}}(C_NULL)
        _check(ccall((:cdft_system_create, libcdft), Cint, (Ptr{Cvoid}, Ref{CdftSystemDesc}, Ref{Ptr{Cvoid}}), ctx, desc, h), ctx)
        return h[]
    end
end

const HIPSystem = cDFT.DFTSystem{<:Clapeyron.PCSAFTModel,<:Any,<:Any,<:Any,<:Any,<:Any,<:cDFT.DFTOptions{HIPNative}}

# Device systems kept across converge! calls of the same DFTSystem object (keyed by
# objectid; at most SYS_CACHE_MAX, the oldest destroyed first), so repeated solves do
# not rebuild the kernels.  A system changed in place needs clear_cache().
const SYS_CACHE_MAX = 16
const SYS_CACHE = Vector{Pair{UInt,Ptr{Cvoid}}}()

function _cached_system(system, ctx)
    key = objectid(system)
    for (k, h) in SYS_CACHE
        k == key && return h
    end
    h = _system(system, ctx)
    push!(SYS_CACHE, key => h)
    if length(SYS_CACHE) > SYS_CACHE_MAX
        _, old = popfirst!(SYS_CACHE)
        ccall((:cdft_system_destroy, libcdft), Cint, (Ptr{Cvoid},), old)
    end
    return h
end

"""Destroy the device systems kept for earlier converge! calls."""
function clear_cache()
    foreach(p -> ccall((:cdft_system_destroy, libcdft), Cint, (Ptr{Cvoid},), p.second), SYS_CACHE)
    empty!(SYS_CACHE)
    return nothing
end

# F_ideal on the host (cDFT's own, BasicIdeal.jl:10-34) + F_res from the library
function _free_energy(system, sys, ctx, r)
    fres = Ref{Cdouble}(0.0)
    out = similar(r, Float64)
    _check(ccall((:cdft_functional_eval, libcdft), Cint, (Ptr{Cvoid}, Ptr{Cdouble}, Ptr{Cdouble}, Ref{Cdouble}, Ptr{Cdouble}),
                 sys, r, out, fres, C_NULL), ctx)
    return cDFT.F_ideal(system, r) + fres[]
end

# converge!'s GFix_logger (src/methods/converge.jl:256-267), called by the library
# after every gcd(log_interval, save_interval)-th map evaluation with the profile
# that evaluation used
mutable struct _Progress
    prob::Any; sys::Ptr{Cvoid}; ctx::Ptr{Cvoid}; dims::Tuple; err::Any
end

function _progress(user::Ptr{Cvoid}, it::Cint, rho::Ptr{Cdouble}, n::Clong)::Cint
    st = unsafe_pointer_to_objref(user)::_Progress
    try
        r = reshape(copy(unsafe_wrap(Array, rho, Int(n))), st.dims)
        prob = st.prob
        if prob.log_interval > 0 && it % prob.log_interval == 0
            H = _free_energy(prob.system, st.sys, st.ctx, r)
            @info "DFT iter $(lpad(it, 5)) | F = $(round(H; sigdigits=6))"
        end
        if prob.save_interval > 0 && it % prob.save_interval == 0 && prob.save_callback !== nothing
            prob.save_callback(Int(it), r)
        end
        return Cint(0)
    catch e
        st.err = e                     # re-thrown once the library has returned
        return Cint(1)
    end
end

function cDFT.converge!(prob::cDFT.DFTProblem{<:HIPSystem}, method::cDFT.AASol, rho::AbstractArray)
    system = prob.system
    ctx = _ctx(system.options.device.device)
    sys = _cached_system(system, ctx)
    begin
        every = prob.log_interval > 0 && prob.save_interval > 0 ? gcd(prob.log_interval, prob.save_interval) :
                max(prob.log_interval, prob.save_interval)
        st = _Progress(prob, sys, ctx, size(rho), nothing)
        cb = @cfunction(_progress, Cint, (Ptr{Cvoid}, Cint, Ptr{Cdouble}, Clong))
        res = Ref{CdftResult}()
        buf = Array{Float64}(rho)
        GC.@preserve st begin
            opts = CdftSolverOpts(method.maxit, method.beta, method.rtol, method.picard_rtol, method.m, 0, C_NULL,
                                  every > 0 ? cb : C_NULL, every > 0 ? pointer_from_objref(st) : C_NULL, every)
            rc = ccall((:cdft_converge, libcdft), Cint, (Ptr{Cvoid}, Ptr{Cdouble}, Ref{CdftSolverOpts}, Ref{CdftResult}),
                       sys, buf, opts, res)
        end
        st.err === nothing || throw(st.err)
        _check(rc, ctx)
        copyto!(rho, buf)
        LAST[] = res[]
        if method.verbose                 # converge.jl:247-252
            r = res[]
            msg = r.idid == 1 ? "DFT converged" : "DFT did not converge"
            H = _free_energy(system, sys, ctx, buf)
            @info "$msg after $(r.nevals) iterations: err = $(round(r.resnorm; sigdigits=3)) | F = $(round(H; sigdigits=6))"
        end
    end
    return nothing
end

"""Batched solve of many HIPNative systems in one persistent launch (one-component,
mixed-precision systems of at most 512 points, anderson_m <= 5; others raise, solve them
with converge!)."""
function converge_batch!(systems::AbstractVector, rhos::AbstractVector; maxit = 10_000, beta = 1e-2, tol = 1e-4,
                         anderson_start = 1e-1, anderson_m = 5)
    ctx = _ctx(first(systems).options.device.device)
    hs = [_system(s, ctx) for s in systems]
    bufs = [Array{Float64}(r) for r in rhos]
    res = Vector{CdftResult}(undef, length(systems))
    opts = CdftSolverOpts(maxit, beta, tol, anderson_start, anderson_m)
    try
        GC.@preserve bufs begin
            ptrs = [pointer(b) for b in bufs]
            _check(ccall((:cdft_converge_batch, libcdft), Cint,
                         (Ptr{Cvoid}, Cint, Ptr{Ptr{Cvoid}}, Ptr{Ptr{Cdouble}}, Ref{CdftSolverOpts}, Ptr{CdftResult}),
                         ctx, length(hs), hs, ptrs, opts, res), ctx)
        end
        foreach(copyto!, rhos, bufs)
    finally
        foreach(h -> ccall((:cdft_system_destroy, libcdft), Cint, (Ptr{Cvoid},), h), hs)
    end
    return res
end

"""Pullback of sum_i cot_i Phi_i (cdft_phi_vjp): the rrule of F_res with cot = 1."""
function phi_vjp(system, rho::AbstractArray, cot::AbstractVector)
    ctx = _ctx(system.options.device.device)
    sys = _system(system, ctx)
    out = zeros(Float64, length(rho))
    try
        _check(ccall((:cdft_phi_vjp, libcdft), Cint, (Ptr{Cvoid}, Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}),
                     sys, Array{Float64}(vec(rho)), Array{Float64}(cot), out), ctx)
    finally
        ccall((:cdft_system_destroy, libcdft), Cint, (Ptr{Cvoid},), sys)
    end
    return reshape(out, size(rho))
end

"""Adjoint of the converged fixed point (cdft_adjoint_solve): lambda with (I - J') lambda = g."""
function adjoint_solve(system, rho::AbstractArray, g::AbstractArray; maxit = 10_000, beta = 1e-2, tol = 1e-10,
                       anderson_start = 1e-1, anderson_m = 5)
    ctx = _ctx(system.options.device.device)
    sys = _system(system, ctx)
    lam = zeros(Float64, length(rho))
    res = Ref{CdftResult}()
    opts = CdftSolverOpts(maxit, beta, tol, anderson_start, anderson_m)
    try
        _check(ccall((:cdft_adjoint_solve, libcdft), Cint,
                     (Ptr{Cvoid}, Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}, Ref{CdftSolverOpts}, Ref{CdftResult}),
                     sys, Array{Float64}(vec(rho)), Array{Float64}(vec(g)), lam, opts, res), ctx)
    finally
        ccall((:cdft_system_destroy, libcdft), Cint, (Ptr{Cvoid},), sys)
    end
    LAST[] = res[]
    return reshape(lam, size(rho))
end

"""Parameter sensitivities through the adjoint (cdft_adjoint_params): for objectives with
gradients g[:, k] = rho .* dJ_k/drho (one column each) at the converged rho of `system`,
and pairs (plus, minus) of the same system rebuilt with theta_p +- h[p], returns the
multipliers (one column per objective) and the implicit parts lambda_k . dG/dtheta_p as a
(objective, parameter) matrix; add the explicit partials dJ_k/dtheta_p at fixed rho."""
function adjoint_params(system, rho::AbstractArray, g::AbstractMatrix, pairs::AbstractVector, h::AbstractVector;
                        maxit = 10_000, beta = 1e-2, tol = 1e-10, anderson_start = 1e-1, anderson_m = 5)
    ctx = _ctx(system.options.device.device)
    sys = _system(system, ctx)
    hp = [_system(p, ctx) for (p, _) in pairs]
    hm = [_system(m, ctx) for (_, m) in pairs]
    nobj = size(g, 2)
    lam = zeros(Float64, length(rho), nobj)
    out = zeros(Float64, max(length(pairs), 1), nobj)           # out[p, k] = C's out[k nparam + p]
    res = Vector{CdftResult}(undef, nobj)
    opts = CdftSolverOpts(maxit, beta, tol, anderson_start, anderson_m)
    try
        _check(ccall((:cdft_adjoint_params, libcdft), Cint,
                     (Ptr{Cvoid}, Ptr{Cdouble}, Cint, Ptr{Cdouble}, Ptr{Cdouble}, Cint, Ptr{Ptr{Cvoid}}, Ptr{Ptr{Cvoid}},
                      Ptr{Cdouble}, Ref{CdftSolverOpts}, Ptr{Cdouble}, Ptr{CdftResult}),
                     sys, Array{Float64}(vec(rho)), nobj, Array{Float64}(g), lam, length(pairs), hp, hm,
                     Array{Float64}(h), opts, out, res), ctx)
    finally
        foreach(x -> ccall((:cdft_system_destroy, libcdft), Cint, (Ptr{Cvoid},), x), vcat(hp, hm, [sys]))
    end
    LAST[] = res
    return lam, permutedims(out[1:length(pairs), :])
end

export HIPNative, converge_batch!, phi_vjp, adjoint_solve, adjoint_params, clear_cache

end # module
