# EPSpectral backend that runs the scGD0 and ladder loops in libepsx (gfx950) through its C ABI
# (include/epsx.h).  Loaded next to EPSpectral.  Arrays are passed as they are laid out in Julia:
# column-major, 1-based indices, ComplexF64 interleaved.
#
# Entry points:
# * `MI355XBackend(device)`: a device context (one HIP stream); freed by its finalizer or `close!`.
# * `loop_scgd0(b, setup, occ; ...)`: `_loop_scgd0` (src/scgd0.jl:189-346) on a `setup` tuple.
# * `loop_ladder(b, ladder, Σ, μ, Tlist, Ωs; ...)`: `_loop_ladder` (src/ladder.jl:303-379) on the
#   tuple `_setup_ladder` returns.
# * `run_scgd0_mi355x(model, nk, window; occ, ωs, backend, ...)` and
#   `run_scgd0_ladder_mi355x(model, nk, window, scgd0; occ, backend, ...)`: `run_scgd0` /
#   `run_scgd0_ladder` (src/scgd0.jl:75-121, src/ladder.jl:65-105) with the extraction on the
#   host and the loops in libepsx; they return the reference's `SCGD0Result` / `SCGD0LadderResult`.
# * `net_create`, `net_forward`, `net_gradient`, `net_train`, `net_params`, `net_params!`: the
#   warm-start network (no reference counterpart; it fills `Σ_init`).
module EPSpectralMI355X

using EPSpectral
using EPSpectral: SCGD0Result, SCGD0LadderResult, FrequencyGrid
import ElectronPhonon

const libepsx = get(ENV, "EPSX_LIB", joinpath(@__DIR__, "..", "epsx", "libepsx.so"))

const EPSX_OK = Cint(0)
const EPSX_ERR_ARG = Cint(1)
const EPSX_ERR_NO_MU_ROOT = Cint(2)
const FM_AUTO = Int32(3)
const PREC = Dict(:fp64 => Int32(0), :fp32 => Int32(1), :bf16 => Int32(2))

"""
    MI355XBackend(device = 0)

A libepsx context on one GPU: the `backend` of this implementation.  Its device memory (the
resident coupling, pooled temporaries, pinned staging buffers) is released by `close!(b)` or, at
the latest, by the finalizer when `b` is garbage-collected.
"""
mutable struct MI355XBackend
    ctx::Ptr{Cvoid}
    device::Int
    function MI355XBackend(device::Integer = 0)
        h = Ref{Ptr{Cvoid}}(C_NULL)
        rc = ccall((:epsx_open, libepsx), Cint, (Cint, Ref{Ptr{Cvoid}}), device, h)
        rc == EPSX_OK || error("epsx_open(device = $device) failed with code $rc: no usable GPU")
        b = new(h[], Int(device))
        finalizer(close!, b)
        b
    end
end

"""Release the context (idempotent; the finalizer calls it too)."""
function close!(b::MI355XBackend)
    if b.ctx != C_NULL
        ccall((:epsx_close, libepsx), Cvoid, (Ptr{Cvoid},), b.ctx)
        b.ctx = C_NULL
    end
    nothing
end

function handle(b::MI355XBackend)
    b.ctx == C_NULL && error("the MI355XBackend context is closed")
    b.ctx
end

function check(b::MI355XBackend, rc)
    rc == EPSX_OK && return nothing
    msg = unsafe_string(ccall((:epsx_last_error, libepsx), Cstring, (Ptr{Cvoid},), handle(b)))
    (rc == EPSX_ERR_ARG || rc == EPSX_ERR_NO_MU_ROOT) && throw(ArgumentError(msg))
    error("libepsx error $rc: $msg")
end

# ---- struct mirrors of include/epsx.h ---------------------------------------------------------

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
    reuse_star::Int32; warm_start::Int32
end

mutable struct Scgd0Outputs
    struct_size::Int64
    Sigma::Ptr{ComplexF64}; mu::Ptr{Float64}; converged::Ptr{Int32}
    hist_err::Ptr{Float64}; hist_mu::Ptr{Float64}; hist_minA::Ptr{Float64}; wall::Ptr{Float64}
    iter_frozen::Ptr{Int64}
    niter::Int64
    fm_method_used::Int32; n_low_precision::Int32
    warm_kept::Ptr{Int32}
end

struct LadderProblem
    struct_size::Int64
    R::Ptr{Float64}
    v::Ptr{Float64}
end

struct LadderParams
    struct_size::Int64
    nT::Int64; nOmega::Int64
    Tlist::Ptr{Float64}; mu::Ptr{Float64}; Sigma::Ptr{ComplexF64}; Omegas::Ptr{Float64}
    solver::Int32; precision::Int32
    rtol::Float64
    maxiter::Int64
    w_acoustic::Float64
    store_vertex::Int32; verbosity::Int32
    batched::Int32; pad::Int32
end

struct LadderOutputs
    struct_size::Int64
    sigma_bubble::Ptr{Float64}; sigma_ladder::Ptr{Float64}
    converged::Ptr{Int32}; niter::Ptr{Int64}
    norms::Ptr{Float64}; sigma_iter::Ptr{Float64}
    J::Ptr{ComplexF64}
end

struct NetConfig
    struct_size::Int64
    n_layers::Int32; kernel::Int32; c_in::Int32; c_hidden::Int32; c_out::Int32; pad::Int32
    length::Int64
end

struct NetTrainParams
    struct_size::Int64
    epochs::Int64; batch::Int64
    lr::Float64; beta1::Float64; beta2::Float64; eps::Float64; weight_decay::Float64
    precision::Int32; verbosity::Int32
    seed::UInt64
end

# ---- scGD0 --------------------------------------------------------------------------------------

# Upload the `setup` fields (src/scgd0.jl:173-183) to the context; g2 stays resident.
function set_problem!(b::MI355XBackend, setup)
    (; g2, ωph, iq_kk, iks_i, iks_f, ε_i, w_i, ε_f, w_f, f_to_i, W, nstates_base, multiplets, ωs,
       ωs_spectral) = setup
    nm, ni, nf = size(g2)
    ptrs = Int64[0]; idx = Int32[]
    for m in multiplets
        append!(idx, Int32.(m)); push!(ptrs, length(idx))
    end
    arrays = (g2 = Array{Float64}(g2), wph = Array{Float64}(ωph), iq = Array{Int32}(iq_kk),
        iki = Array{Int32}(iks_i), ikf = Array{Int32}(iks_f), ei = Array{Float64}(ε_i),
        wi = Array{Float64}(w_i), ef = Array{Float64}(ε_f), wf = Array{Float64}(w_f),
        fi = Array{Int32}(f_to_i), wo = Array{Int64}(ωs.offsets), wao = Array{Int64}(ωs_spectral.offsets),
        mp = ptrs, mi = isempty(idx) ? Int32[0] : idx)
    GC.@preserve arrays begin
        P = Ref(Scgd0Problem(sizeof(Scgd0Problem), nm, ni, nf, size(ωph, 2), size(iq_kk)...,
            pointer(arrays.g2), pointer(arrays.wph), pointer(arrays.iq), pointer(arrays.iki),
            pointer(arrays.ikf), pointer(arrays.ei), pointer(arrays.wi), pointer(arrays.ef),
            pointer(arrays.wf), pointer(arrays.fi), W, nstates_base,
            length(ωs), ωs.ω_origin, ωs.dω_unit, pointer(arrays.wo),
            length(ωs_spectral), ωs_spectral.ω_origin, ωs_spectral.dω_unit, pointer(arrays.wao),
            length(multiplets), pointer(arrays.mp), pointer(arrays.mi)))
        check(b, ccall((:epsx_scgd0_set_problem, libepsx), Cint, (Ptr{Cvoid}, Ref{Scgd0Problem}),
            handle(b), P))
    end
    nothing
end

"""
    loop_scgd0(b::MI355XBackend, setup, occ; η_min, η_init, maxiter, tol, Σ_init = nothing,
               fix_μ = false, ω_acoustic, gpu_tile = nothing, verbosity = 0,
               precision = :fp64, batched = true) -> (; Σ, μ, converged, history)

`_loop_scgd0` (src/scgd0.jl:189-346) on the device: the same `setup` fields and keywords, the
same result.  `precision = :fp32` or `:bf16` runs the Fan-Migdal part in that precision until
its floor, then FP64 to the same `tol`.
"""
function loop_scgd0(b::MI355XBackend, setup, occ; η_min, η_init, maxiter, tol, Σ_init = nothing,
        fix_μ = false, ω_acoustic, gpu_tile = nothing, verbosity = 0, precision = :fp64,
        batched = true, upload = true, warm_start = nothing)
    upload && set_problem!(b, setup)
    nω, ni = length(setup.ωs), size(setup.g2, 2)
    nT = length(occ.Tlist)
    T = collect(Float64, occ.Tlist); nl = collect(Float64, occ.nlist)
    μ0 = collect(Float64, setup.μ_start)
    ηm = η_min isa Number ? fill(Float64(η_min), nT) : collect(Float64, η_min)
    Σi = Σ_init === nothing ? nothing : Array{ComplexF64}(Σ_init)
    Σi === nothing || size(Σi) == (nω, ni, nT) || throw(ArgumentError(
        "Σ_init must have size (nω, n_i, nT) = $((nω, ni, nT)), got $(size(Σi))"))
    Σ = Array{ComplexF64}(undef, nω, ni, nT); μ = Vector{Float64}(undef, nT)
    conv = Vector{Int32}(undef, nT)
    herr, hμ, hA = (Matrix{Float64}(undef, nT, maxiter) for _ in 1:3)
    wall = Vector{Float64}(undef, maxiter); frozen = Vector{Int64}(undef, nT)
    kept = zeros(Int32, nT)
    # warm_start: nothing, or the precision (:bf16 / :fp32) of the context's network (net_create /
    # net_train): Σ_init predicted on the device from the first iteration's G0D0 self-energy
    ws = warm_start === nothing ? Int32(0) : PREC[warm_start]
    out = Scgd0Outputs(sizeof(Scgd0Outputs), pointer(Σ), pointer(μ), pointer(conv), pointer(herr),
        pointer(hμ), pointer(hA), pointer(wall), pointer(frozen), 0, 0, 0, pointer(kept))
    GC.@preserve T nl μ0 ηm Σi Σ μ conv herr hμ hA wall frozen kept out begin
        par = Ref(Scgd0Params(sizeof(Scgd0Params), nT, pointer(T), pointer(nl), pointer(μ0), pointer(ηm),
            occ.nelec, occ.spin_degeneracy, Int32(occ.type === :Semiconductor), Int32(fix_μ),
            η_init, tol, ω_acoustic, maxiter, gpu_tile === nothing ? 0 : gpu_tile,
            FM_AUTO, PREC[precision], Int32(batched), Int32(verbosity),
            Σi === nothing ? Ptr{ComplexF64}(C_NULL) : pointer(Σi), Int32(0), ws))
        check(b, ccall((:epsx_scgd0_run, libepsx), Cint,
            (Ptr{Cvoid}, Ref{Scgd0Params}, Ref{Scgd0Outputs}), handle(b), par, out))
    end
    n = out.niter
    converged = conv .!= 0
    any(!, converged) && maxiter > 1 &&
        @warn "scGD0 not converged after maxiter = $maxiter iterations at T = $(T[.!converged])"
    history = (; err = herr[:, 1:n], μ = hμ[:, 1:n], min_spectral_weight = hA[:, 1:n],
        wall = wall[1:n], iter_frozen = frozen)
    (; Σ, μ, converged, history, warm_kept = kept .!= 0)
end

"""
    run_scgd0_mi355x(model, nk, window; occ, ωs, backend::MI355XBackend, kwargs...) -> SCGD0Result

`run_scgd0` (src/scgd0.jl:75-121) with the loop in libepsx: the same argument checks and defaults
(the same `ArgumentError` texts), the extraction `_setup_scgd0` on the host, then `loop_scgd0`.
"""
function run_scgd0_mi355x(model::ElectronPhonon.Model{FT}, nk, window; occ, ωs,
        backend::MI355XBackend, ωs_spectral = nothing, η_min = 0,
        η_init = 5 * ElectronPhonon.unit_to_aru(:meV), maxiter = 20,
        tol = 0.05 * ElectronPhonon.unit_to_aru(:meV), symmetry = model.symmetry, verbosity = 1,
        gpu_tile = nothing, Σ_init = nothing, fix_μ = false,
        ω_acoustic = ElectronPhonon.omega_acoustic,
        electron_degen_cutoff = ElectronPhonon.electron_degen_cutoff, precision = :fp64,
        ep_kwargs...) where {FT}
    nk isa NTuple{3, Integer} ||
        throw(ArgumentError("nk must be a grid size (n1, n2, n3), got $nk"))
    occ.occ_type === :FermiDirac || throw(ArgumentError(
        "run_scgd0 uses the Fermi-Dirac occupation only, got occ.occ_type = $(occ.occ_type)"))
    occ.type ∈ (:Metal, :Semiconductor) || throw(ArgumentError(
        "occ.type must be :Metal or :Semiconductor, got $(occ.type)"))
    ωs isa FrequencyGrid{FT} || throw(ArgumentError("ωs must be a FrequencyGrid{$FT}, got a " *
        "$(typeof(ωs)); build one with `graded_frequency_grid` or `FrequencyGrid(ωs::AbstractVector)`"))
    if ωs_spectral === nothing
        nω_spectral = length(range(extrema(ωs)...; step = ωs.dω_min))
        ωs_spectral = FrequencyGrid(first(ωs), ωs.dω_min, 0:nω_spectral - 1)
    else
        ωs_spectral isa FrequencyGrid{FT} || throw(ArgumentError("ωs_spectral must be a " *
            "FrequencyGrid{$FT} or nothing, got a $(typeof(ωs_spectral))"))
        (first(ωs_spectral) <= first(ωs) + ωs.dω_min &&
            last(ωs_spectral) >= last(ωs) - ωs.dω_min) || throw(ArgumentError(
            "ωs_spectral spans [$(first(ωs_spectral)), $(last(ωs_spectral))], which does not " *
            "cover ωs = [$(first(ωs)), $(last(ωs))] to within ωs.dω_min"))
        (first(ωs_spectral) < first(ωs) - ωs.dω_min || last(ωs_spectral) > last(ωs) + ωs.dω_min) &&
            @warn "ωs_spectral spans [$(first(ωs_spectral)), $(last(ωs_spectral))], past ωs = " *
                "[$(first(ωs)), $(last(ωs))]; Σ is held at its end values outside ωs"
    end
    fix_μ && !ElectronPhonon.chemical_potential_is_computed(occ) && throw(ArgumentError(
        "fix_μ = true keeps μ at occ.μlist, which is not given"))
    η_min = η_min isa Number ? fill(FT(η_min), length(occ)) : collect(FT, η_min)
    length(η_min) == length(occ) || throw(ArgumentError("η_min must be a scalar or have one " *
        "entry per temperature ($(length(occ))), got $(length(η_min))"))
    (gpu_tile === nothing || gpu_tile >= 1) || throw(ArgumentError(
        "gpu_tile must be nothing or a positive number of inner states, got $gpu_tile"))
    maxiter >= 1 || throw(ArgumentError("maxiter must be at least 1, got $maxiter"))

    # the extraction on the host: `setup` holds host arrays, which loop_scgd0 uploads
    setup = EPSpectral._setup_scgd0(model, nk, window, occ; ωs, ωs_spectral,
        backend = ElectronPhonon.CPUBackend(), symmetry, verbosity, electron_degen_cutoff, ep_kwargs...)
    (; Σ, μ, converged, history) = loop_scgd0(backend, setup, occ; η_min, η_init, maxiter, tol,
        Σ_init, fix_μ, ω_acoustic, gpu_tile, verbosity, precision)
    SCGD0Result(ωs, ωs_spectral, Σ, μ, setup.el_i, collect(FT, occ.Tlist), collect(FT, occ.nlist),
        converged, history)
end

# ---- ladder -------------------------------------------------------------------------------------

"""
    loop_ladder(b::MI355XBackend, ladder, Σ, μ, Tlist, Ωs; solver, rtol, maxiter, store_vertex,
                ω_acoustic, verbosity = 0, gpu_tile = nothing, precision = :fp64)
        -> (; σ_bubble, σ_ladder, converged, niter, history, J)

`_loop_ladder` (src/ladder.jl:303-379) on the device, for the tuple `_setup_ladder` returns
(src/ladder.jl:112-158; its scGD0 fields are uploaded, then R and v): σ unsymmetrised and without
`spin_degeneracy / volume`, `history[Ω, T] = (; norms, σ_iter)` as the reference fills it.
`gpu_tile` only reorders sums and is not needed on this backend.
"""
function loop_ladder(b::MI355XBackend, ladder, Σ, μ, Tlist, Ωs; solver, rtol, maxiter, store_vertex,
        ω_acoustic, verbosity = 0, gpu_tile = nothing, precision = :fp64, batched = false)
    solver ∈ (:gmres, :fixed_point) || throw(ArgumentError(
        "solver must be :gmres or :fixed_point, got $solver"))
    set_problem!(b, ladder)
    R = Array{Float64}(ladder.R); v = Array{Float64}(ladder.v)
    GC.@preserve R v begin
        LP = Ref(LadderProblem(sizeof(LadderProblem), pointer(R), pointer(v)))
        check(b, ccall((:epsx_ladder_set_problem, libepsx), Cint, (Ptr{Cvoid}, Ref{LadderProblem}),
            handle(b), LP))
    end
    nω, ni, nT = size(Σ)
    nΩ = length(Ωs)
    Σh = Array{ComplexF64}(Σ); μh = collect(Float64, μ); Th = collect(Float64, Tlist)
    Ωh = collect(Float64, Ωs)
    σb, σl = (Array{Float64}(undef, 3, 3, nΩ, nT) for _ in 1:2)
    conv = Matrix{Int32}(undef, nΩ, nT); nit = Matrix{Int64}(undef, nΩ, nT)
    norms = Array{Float64}(undef, maxiter + 1, nΩ, nT)
    σit = Array{Float64}(undef, 3, 3, maxiter, nΩ, nT)
    J = store_vertex ? Array{ComplexF64}(undef, nω, 4, 3, 2, ni, nΩ, nT) : nothing
    GC.@preserve Σh μh Th Ωh σb σl conv nit norms σit J begin
        par = Ref(LadderParams(sizeof(LadderParams), nT, nΩ, pointer(Th), pointer(μh), pointer(Σh),
            pointer(Ωh), Int32(solver === :gmres ? 0 : 1), PREC[precision], rtol, maxiter, ω_acoustic,
            Int32(store_vertex), Int32(verbosity), Int32(batched), Int32(0)))
        out = Ref(LadderOutputs(sizeof(LadderOutputs), pointer(σb), pointer(σl), pointer(conv),
            pointer(nit), pointer(norms), pointer(σit),
            J === nothing ? Ptr{ComplexF64}(C_NULL) : pointer(J)))
        check(b, ccall((:epsx_ladder_run, libepsx), Cint,
            (Ptr{Cvoid}, Ref{LadderParams}, Ref{LadderOutputs}), handle(b), par, out))
    end
    converged = conv .!= 0
    niter = Int.(nit)
    # the history as the reference keeps it: residual norms (GMRES) or σ per iteration (fixed point)
    history = [solver === :gmres ?
        (; norms = filter(!isnan, norms[:, iΩ, iT]), σ_iter = Matrix{Float64}[]) :
        (; norms = Float64[], σ_iter = [σit[:, :, k, iΩ, iT] for k in 1:niter[iΩ, iT]])
        for iΩ in 1:nΩ, iT in 1:nT]
    if !all(converged)
        bad = [(Ωs[iΩ], Tlist[iT]) for iΩ in 1:nΩ, iT in 1:nT if !converged[iΩ, iT]]
        @warn "ladder not converged (rtol = $rtol, maxiter = $maxiter) at (Ω, T) = $bad"
    end
    (; σ_bubble = σb, σ_ladder = σl, converged, niter, history, J)
end

"""
    run_scgd0_ladder_mi355x(model, nk, window, scgd0::SCGD0Result; occ, backend::MI355XBackend,
                            kwargs...) -> SCGD0LadderResult

`run_scgd0_ladder` (src/ladder.jl:65-105) with the loop in libepsx: the same checks, the
extraction and `_setup_ladder` on the host, `loop_ladder`, then the symmetrisation with
`spin_degeneracy / volume`.
"""
function run_scgd0_ladder_mi355x(model::ElectronPhonon.Model{FT}, nk, window, scgd0::SCGD0Result;
        occ, backend::MI355XBackend, Ωs = [zero(FT)], solver = :gmres, rtol = 1e-2, maxiter = 30,
        store_vertex = false, symmetry = model.symmetry, gpu_tile = nothing, verbosity = 1,
        ω_acoustic = ElectronPhonon.omega_acoustic,
        electron_degen_cutoff = ElectronPhonon.electron_degen_cutoff, precision = :fp64,
        ep_kwargs...) where {FT}
    solver ∈ (:gmres, :fixed_point) || throw(ArgumentError(
        "solver must be :gmres or :fixed_point, got $solver"))
    maxiter >= 1 || throw(ArgumentError("maxiter must be at least 1, got $maxiter"))
    (gpu_tile === nothing || gpu_tile >= 1) || throw(ArgumentError(
        "gpu_tile must be nothing or a positive number of inner states, got $gpu_tile"))
    isempty(Ωs) && throw(ArgumentError("Ωs must hold at least one probe frequency"))
    ω_acoustic > 0 || throw(ArgumentError("ω_acoustic must be positive, got $ω_acoustic: the " *
        "shift-grid coupling build has no row for a negative phonon frequency"))
    (length(occ.Tlist) == length(scgd0.Tlist) && occ.Tlist ≈ scgd0.Tlist) || throw(ArgumentError(
        "occ.Tlist = $(occ.Tlist) must equal the temperatures of scgd0, $(scgd0.Tlist)"))
    host = ElectronPhonon.CPUBackend()
    setup = EPSpectral._setup_scgd0(model, nk, window, occ; ωs = scgd0.ωs, scgd0.ωs_spectral,
        backend = host, symmetry, verbosity, electron_degen_cutoff, ep_kwargs...)
    el_new, el_old = setup.el_i, scgd0.el_i
    (el_new.iks == el_old.iks && el_new.ibands == el_old.ibands &&
     isapprox(el_new.es, el_old.es; rtol = 1e-10)) || throw(ArgumentError(
        "the states of the new extraction differ from scgd0.el_i: run_scgd0_ladder needs the " *
        "same model, nk, window and symmetry as the run that produced scgd0"))
    ladder = EPSpectral._setup_ladder(setup, scgd0, symmetry; backend = host)
    res = loop_ladder(backend, ladder, scgd0.Σ, scgd0.μ, scgd0.Tlist, collect(FT, Ωs); solver, rtol,
        maxiter, store_vertex, ω_acoustic, verbosity, precision)
    function to_symmetrized_σ(σ)
        out = similar(σ)
        for iT in axes(σ, 4), iΩ in axes(σ, 3)
            out[:, :, iΩ, iT] = occ.spin_degeneracy / occ.volume *
                ElectronPhonon.symmetrize(ElectronPhonon.Mat3{FT}(σ[:, :, iΩ, iT]), symmetry)
        end
        out
    end
    SCGD0LadderResult(collect(FT, Ωs), collect(FT, scgd0.Tlist), collect(FT, scgd0.μ),
        to_symmetrized_σ(res.σ_bubble), to_symmetrized_σ(res.σ_ladder), res.converged, res.niter,
        res.history, ladder.dω_unit, res.J)
end

# ---- the warm-start network (fills Σ_init; no reference counterpart) ---------------------------

"""
    net_create(b; length, n_layers = 6, kernel = 9, c_in = 5, c_hidden = 64, c_out = 2, seed = 0)

A new network in the context (seeded Kaiming-uniform weights): `epsx_net_create`.
"""
function net_create(b::MI355XBackend; length::Integer, n_layers = 6, kernel = 9, c_in = 5,
        c_hidden = 64, c_out = 2, seed = 0)
    cfg = Ref(NetConfig(sizeof(NetConfig), n_layers, kernel, c_in, c_hidden, c_out, 0, length))
    check(b, ccall((:epsx_net_create, libepsx), Cint, (Ptr{Cvoid}, Ref{NetConfig}, UInt64),
        handle(b), cfg, UInt64(seed)))
    nothing
end

function net_param_count(b::MI355XBackend)
    n = Ref{Int64}(0)
    check(b, ccall((:epsx_net_param_count, libepsx), Cint, (Ptr{Cvoid}, Ref{Int64}), handle(b), n))
    n[]
end

"""The network's FP32 parameters (per layer W [cout, cin, kernel] then b [cout])."""
function net_params(b::MI355XBackend)
    p = Vector{Float32}(undef, net_param_count(b))
    check(b, ccall((:epsx_net_get_params, libepsx), Cint, (Ptr{Cvoid}, Ptr{Float32}), handle(b), p))
    p
end

"""Set the parameters (and clear the optimiser state)."""
function net_params!(b::MI355XBackend, p::Vector{Float32})
    length(p) == net_param_count(b) || throw(ArgumentError("expected $(net_param_count(b)) parameters"))
    check(b, ccall((:epsx_net_set_params, libepsx), Cint, (Ptr{Cvoid}, Ptr{Float32}), handle(b), p))
    nothing
end

# Samples are float [n, channels, length] in C order: a Julia array of size (length, channels, n).
"""y = net(x) for x of size (length, c_in, n); returns (length, c_out, n)."""
function net_forward(b::MI355XBackend, x::Array{Float32, 3}; c_out = 2, precision = :bf16)
    L, _, n = size(x)
    y = Array{Float32}(undef, L, c_out, n)
    check(b, ccall((:epsx_net_forward, libepsx), Cint,
        (Ptr{Cvoid}, Ptr{Float32}, Int64, Cint, Ptr{Float32}), handle(b), x, n, PREC[precision], y))
    y
end

"""(loss, gradient) of the mean squared error against `target` (length, c_out, n)."""
function net_gradient(b::MI355XBackend, x::Array{Float32, 3}, target::Array{Float32, 3};
        precision = :bf16)
    loss = Ref{Float64}(0)
    g = Vector{Float32}(undef, net_param_count(b))
    check(b, ccall((:epsx_net_gradient, libepsx), Cint,
        (Ptr{Cvoid}, Ptr{Float32}, Ptr{Float32}, Int64, Cint, Ref{Float64}, Ptr{Float32}),
        handle(b), x, target, size(x, 3), PREC[precision], loss, g))
    loss[], g
end

"""Adam over shuffled mini-batches; returns the mean batch loss of each epoch."""
function net_train(b::MI355XBackend, x::Array{Float32, 3}, target::Array{Float32, 3}; epochs,
        batch = 64, lr = 1e-3, beta1 = 0.9, beta2 = 0.999, eps = 1e-8, weight_decay = 0.0,
        precision = :bf16, verbosity = 0, seed = 0)
    losses = Vector{Float64}(undef, epochs)
    par = Ref(NetTrainParams(sizeof(NetTrainParams), epochs, batch, lr, beta1, beta2, eps,
        weight_decay, PREC[precision], verbosity, UInt64(seed)))
    check(b, ccall((:epsx_net_train, libepsx), Cint,
        (Ptr{Cvoid}, Ptr{Float32}, Ptr{Float32}, Int64, Ref{NetTrainParams}, Ptr{Float64}),
        handle(b), x, target, size(x, 3), par, losses))
    losses
end

end # module
