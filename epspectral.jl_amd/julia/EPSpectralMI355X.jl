# EPSpectral backend that runs the scGD0 and ladder loops in libepsx (gfx950) through its C ABI
# (include/epsx.h).  Loaded next to EPSpectral; a maintainer adds `backend::MI355XBackend`
# methods of `_loop_scgd0` / `_loop_ladder` (src/scgd0.jl:189, src/ladder.jl:303) that forward
# here.  Arrays are passed as they are laid out in Julia: column-major, 1-based indices,
# ComplexF64 interleaved.
module EPSpectralMI355X

const libepsx = get(ENV, "EPSX_LIB", joinpath(@__DIR__, "..", "epsx", "libepsx.so"))

const EPSX_OK = Cint(0)
const EPSX_ERR_ARG = Cint(1)
const EPSX_ERR_NO_MU_ROOT = Cint(2)

struct MI355XBackend
    ctx::Ptr{Cvoid}
end

function MI355XBackend(device::Integer = 0)
    h = Ref{Ptr{Cvoid}}(C_NULL)
    rc = ccall((:epsx_open, libepsx), Cint, (Cint, Ref{Ptr{Cvoid}}), device, h)
    rc == EPSX_OK || error("epsx_open failed with code $rc")
    b = MI355XBackend(h[])
    finalizer(x -> ccall((:epsx_close, libepsx), Cvoid, (Ptr{Cvoid},), x.ctx), Ref(b))
    b
end

function check(b::MI355XBackend, rc)
    rc == EPSX_OK && return nothing
    msg = unsafe_string(ccall((:epsx_last_error, libepsx), Cstring, (Ptr{Cvoid},), b.ctx))
    (rc == EPSX_ERR_ARG || rc == EPSX_ERR_NO_MU_ROOT) && throw(ArgumentError(msg))
    error("libepsx error $rc: $msg")
end

# epsx_scgd0_problem
struct Scgd0Problem
    struct_size::Int64
    nm::Int64; ni::Int64; nf::Int64
    nq::Int64; nki::Int64; nkf::Int64
    g2::Ptr{Float64}; wph::Ptr{Float64}; iq_kk::Ptr{Int32}; iks_i::Ptr{Int32}; iks_f::Ptr{Int32}
    e_i::Ptr{Float64}; w_i::Ptr{Float64}; e_f::Ptr{Float64}; w_f::Ptr{Float64}; f_to_i::Ptr{Int32}
    W::Float64; nstates_base::Float64
    nw::Int64; w_origin::Float64; dw_unit::Float64; w_offsets::Ptr{Int64}
    nwa::Int64; wa_origin::Float64; dwa_unit::Float64; wa_offsets::Ptr{Int64}
    n_mult::Int64; mult_ptr::Ptr{Int64}; mult_idx::Ptr{Int32}
end

struct Scgd0Params
    struct_size::Int64
    nT::Int64
    Tlist::Ptr{Float64}; nlist::Ptr{Float64}; mu_start::Ptr{Float64}; eta_min::Ptr{Float64}
    nelec::Float64; spin_degeneracy::Float64
    is_semiconductor::Int32; fix_mu::Int32
    eta_init::Float64; tol::Float64; w_acoustic::Float64
    maxiter::Int64; f_tile::Int64
    fm_method::Int32; precision::Int32; batched::Int32; verbosity::Int32
    Sigma_init::Ptr{ComplexF64}
    reuse_star::Int32; pad::Int32
end

mutable struct Scgd0Outputs
    struct_size::Int64
    Sigma::Ptr{ComplexF64}; mu::Ptr{Float64}; converged::Ptr{Int32}
    hist_err::Ptr{Float64}; hist_mu::Ptr{Float64}; hist_minA::Ptr{Float64}; wall::Ptr{Float64}
    iter_frozen::Ptr{Int64}
    niter::Int64
    fm_method_used::Int32; n_low_precision::Int32
end

"""
    loop_scgd0(b::MI355XBackend, setup, occ; η_min, η_init, maxiter, tol, Σ_init, fix_μ,
               ω_acoustic, gpu_tile = nothing, fm_method = 3, precision = 0, batched = true)

`_loop_scgd0` (src/scgd0.jl:189-346) on the device: the same `setup` fields and keywords, the
same `(; Σ, μ, converged, history)` result.
"""
function loop_scgd0(b::MI355XBackend, setup, occ; η_min, η_init, maxiter, tol, Σ_init = nothing,
        fix_μ = false, ω_acoustic, gpu_tile = nothing, fm_method = 3, precision = 0, batched = true,
        verbosity = 0)
    (; g2, ωph, iq_kk, iks_i, iks_f, ε_i, w_i, ε_f, w_f, f_to_i, W, nstates_base, multiplets, ωs,
       ωs_spectral, μ_start) = setup
    nm, ni, nf = size(g2)
    nT = length(occ.Tlist)
    ptrs = Int64[1]; idx = Int32[]
    for m in multiplets
        append!(idx, Int32.(m)); push!(ptrs, length(idx) + 1)
    end
    mult_ptr = ptrs .- 1                              # 0-based CSR
    arrays = (g2 = Array{Float64}(g2), wph = Array{Float64}(ωph), iq = Array{Int32}(iq_kk),
        iki = Array{Int32}(iks_i), ikf = Array{Int32}(iks_f), ei = Array{Float64}(ε_i),
        wi = Array{Float64}(w_i), ef = Array{Float64}(ε_f), wf = Array{Float64}(w_f),
        fi = Array{Int32}(f_to_i), wo = Array{Int64}(ωs.offsets), wao = Array{Int64}(ωs_spectral.offsets),
        mp = mult_ptr, mi = isempty(idx) ? Int32[0] : idx)
    GC.@preserve arrays begin
        P = Ref(Scgd0Problem(sizeof(Scgd0Problem), nm, ni, nf, size(ωph, 2), size(iq_kk)...,
            pointer(arrays.g2), pointer(arrays.wph), pointer(arrays.iq), pointer(arrays.iki),
            pointer(arrays.ikf), pointer(arrays.ei), pointer(arrays.wi), pointer(arrays.ef),
            pointer(arrays.wf), pointer(arrays.fi), W, nstates_base,
            length(ωs), ωs.ω_origin, ωs.dω_unit, pointer(arrays.wo),
            length(ωs_spectral), ωs_spectral.ω_origin, ωs_spectral.dω_unit, pointer(arrays.wao),
            length(multiplets), pointer(arrays.mp), pointer(arrays.mi)))
        check(b, ccall((:epsx_scgd0_set_problem, libepsx), Cint, (Ptr{Cvoid}, Ref{Scgd0Problem}), b.ctx, P))
    end
    T = collect(Float64, occ.Tlist); nl = collect(Float64, occ.nlist)
    μ0 = collect(Float64, μ_start); ηm = η_min isa Number ? fill(Float64(η_min), nT) : collect(Float64, η_min)
    Σi = Σ_init === nothing ? nothing : Array{ComplexF64}(Σ_init)
    Σ = Array{ComplexF64}(undef, length(ωs), ni, nT); μ = Vector{Float64}(undef, nT)
    conv = Vector{Int32}(undef, nT)
    herr, hμ, hA = (Matrix{Float64}(undef, nT, maxiter) for _ in 1:3)
    wall = Vector{Float64}(undef, maxiter); frozen = Vector{Int64}(undef, nT)
    GC.@preserve T nl μ0 ηm Σi Σ μ conv herr hμ hA wall frozen begin
        par = Ref(Scgd0Params(sizeof(Scgd0Params), nT, pointer(T), pointer(nl), pointer(μ0), pointer(ηm),
            occ.nelec, occ.spin_degeneracy, Int32(occ.type === :Semiconductor), Int32(fix_μ),
            η_init, tol, ω_acoustic, maxiter, gpu_tile === nothing ? 0 : gpu_tile,
            Int32(fm_method), Int32(precision), Int32(batched), Int32(verbosity),
            Σi === nothing ? Ptr{ComplexF64}(C_NULL) : pointer(Σi), Int32(0), Int32(0)))
        out = Scgd0Outputs(sizeof(Scgd0Outputs), pointer(Σ), pointer(μ), pointer(conv), pointer(herr),
            pointer(hμ), pointer(hA), pointer(wall), pointer(frozen), 0, 0, 0)
        check(b, ccall((:epsx_scgd0_run, libepsx), Cint,
            (Ptr{Cvoid}, Ref{Scgd0Params}, Ref{Scgd0Outputs}), b.ctx, par, out))
        n = out.niter
        converged = conv .!= 0
        any(!, converged) && maxiter > 1 &&
            @warn "scGD0 not converged after maxiter = $maxiter iterations at T = $(T[.!converged])"
        history = (; err = herr[:, 1:n], μ = hμ[:, 1:n], min_spectral_weight = hA[:, 1:n],
            wall = wall[1:n], iter_frozen = frozen)
        return (; Σ, μ, converged, history)
    end
end

end # module
