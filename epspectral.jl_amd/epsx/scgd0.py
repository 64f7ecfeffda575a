"""The scGD0 loop on libepsx: the Python side of src/scgd0.jl.

`Scgd0Setup` holds the fields of the reference's `setup` NamedTuple (src/scgd0.jl:173-183; the
hand-built form of test/holstein_1d_reference.jl:62-77).  `loop_scgd0` mirrors `_loop_scgd0`
(src/scgd0.jl:189-346) and `run_scgd0` the argument checks and defaults of `run_scgd0`
(src/scgd0.jl:75-121) on a given setup (the e-ph extraction of `_setup_scgd0` is upstream of
this path).  The iteration itself runs in libepsx on the GPU.
"""
import dataclasses
import math
import warnings
from typing import Optional

import numpy as np

from . import _lib
from ._lib import ArgumentError, c128, f64, i32, i64, ptr
from .grids import FrequencyGrid, uniform_spectral_grid

# Modes below this frequency are skipped (caller units).  The reference takes
# ElectronPhonon.omega_acoustic; this implementation takes the value as a parameter.
OMEGA_ACOUSTIC = 1e-6
ELECTRON_DEGEN_CUTOFF = 1e-6


@dataclasses.dataclass
class Occupation:
    """The fields of ElectronOccupationParams that the loop reads."""
    Tlist: np.ndarray
    nlist: np.ndarray
    nelec: float
    spin_degeneracy: float = 2.0
    type: str = "Metal"                 # "Metal" or "Semiconductor"
    μlist: Optional[np.ndarray] = None
    volume: float = 1.0
    occ_type: str = "FermiDirac"

    def __post_init__(self):
        self.Tlist = np.atleast_1d(np.asarray(self.Tlist, dtype=np.float64))
        nl = np.atleast_1d(np.asarray(self.nlist, dtype=np.float64))
        self.nlist = np.broadcast_to(nl, self.Tlist.shape).copy() if nl.size == 1 else nl
        if self.μlist is not None:
            ml = np.atleast_1d(np.asarray(self.μlist, dtype=np.float64))
            self.μlist = np.broadcast_to(ml, self.Tlist.shape).copy() if ml.size == 1 else ml

    def __len__(self):
        return len(self.Tlist)


@dataclasses.dataclass
class Scgd0Setup:
    """The `setup` of `_loop_scgd0`.  Arrays use the reference's index order (column-major),
    state and q indices are 1-based."""
    g2: np.ndarray           # (nm, ni, nf)
    ωph: np.ndarray          # (nm, nq)
    iq_kk: np.ndarray        # (nki, nkf) int32, 1-based q
    iks_i: np.ndarray        # (ni,) 1-based rows of iq_kk
    iks_f: np.ndarray        # (nf,) 1-based columns of iq_kk
    ε_i: np.ndarray
    w_i: np.ndarray
    ε_f: np.ndarray
    w_f: np.ndarray
    f_to_i: np.ndarray       # (nf,) 1-based
    W: float
    nstates_base: float
    multiplets: list         # list of lists of 1-based outer states
    ωs: FrequencyGrid
    ωs_spectral: FrequencyGrid
    μ_start: np.ndarray

    @property
    def dims(self):
        nm, ni, nf = self.g2.shape
        return nm, ni, nf


@dataclasses.dataclass
class SCGD0Result:
    """`SCGD0Result` (src/scgd0.jl:25-35) without `el_i` (the states belong to the setup)."""
    ωs: FrequencyGrid
    ωs_spectral: FrequencyGrid
    Σ: np.ndarray            # (nω, ni, nT) complex
    μ: np.ndarray
    Tlist: np.ndarray
    nlist: np.ndarray
    converged: np.ndarray
    history: dict
    info: dict = dataclasses.field(default_factory=dict)


def build_multiplets(iks, es, cutoff=ELECTRON_DEGEN_CUTOFF):
    """Degenerate multiplets of the outer states (src/scgd0.jl:146-160): sorted by (k, ε),
    neighbours at one k chained by gaps below `cutoff`; groups of at least two, 1-based."""
    iks = np.asarray(iks)
    es = np.asarray(es, dtype=np.float64)
    order = np.lexsort((np.arange(len(es)), es, iks))
    out, group = [], [int(order[0])]
    for i in order[1:]:
        j = group[-1]
        if iks[i] == iks[j] and es[i] - es[j] < cutoff:
            group.append(int(i))
        else:
            if len(group) > 1:
                out.append([g + 1 for g in group])
            group = [int(i)]
    if len(group) > 1:
        out.append([g + 1 for g in group])
    return out


class _Problem:
    """The C structs of one setup, with the numpy arrays they point into kept alive."""

    def __init__(self, setup: Scgd0Setup):
        s = setup
        nm, ni, nf = s.dims
        self.keep = keep = {}
        keep["g2"] = f64(s.g2)
        keep["wph"] = f64(s.ωph)
        keep["iq_kk"] = i32(s.iq_kk)
        keep["iks_i"] = i32(s.iks_i)
        keep["iks_f"] = i32(s.iks_f)
        for k in ("ε_i", "w_i", "ε_f", "w_f"):
            keep[k] = f64(getattr(s, k))
        keep["f_to_i"] = i32(s.f_to_i)
        keep["woff"] = i64(s.ωs.offsets)
        keep["waoff"] = i64(s.ωs_spectral.offsets)
        mptr = [0]
        midx = []
        for m in s.multiplets:
            midx.extend(int(x) for x in m)
            mptr.append(len(midx))
        keep["mptr"] = i64(mptr)
        keep["midx"] = i32(midx if midx else [0])
        P = _lib.Scgd0Problem()
        P.struct_size = C_sizeof(_lib.Scgd0Problem)
        P.nm, P.ni, P.nf = nm, ni, nf
        P.nq = keep["wph"].shape[1]
        P.nki, P.nkf = keep["iq_kk"].shape
        P.g2, P.wph, P.iq_kk = ptr(keep["g2"]), ptr(keep["wph"]), ptr(keep["iq_kk"])
        P.iks_i, P.iks_f = ptr(keep["iks_i"]), ptr(keep["iks_f"])
        P.e_i, P.w_i = ptr(keep["ε_i"]), ptr(keep["w_i"])
        P.e_f, P.w_f = ptr(keep["ε_f"]), ptr(keep["w_f"])
        P.f_to_i = ptr(keep["f_to_i"])
        P.W, P.nstates_base = float(s.W), float(s.nstates_base)
        P.nw, P.w_origin, P.dw_unit, P.w_offsets = len(s.ωs), s.ωs.ω_origin, s.ωs.dω_unit, ptr(keep["woff"])
        P.nwa, P.wa_origin, P.dwa_unit = len(s.ωs_spectral), s.ωs_spectral.ω_origin, s.ωs_spectral.dω_unit
        P.wa_offsets = ptr(keep["waoff"])
        P.n_mult = len(s.multiplets)
        P.mult_ptr, P.mult_idx = ptr(keep["mptr"]), ptr(keep["midx"])
        self.P = P


def C_sizeof(t):
    import ctypes

    return ctypes.sizeof(t)


class Scgd0Device:
    """A setup uploaded to the device (g2 resident), reusable across runs."""

    def __init__(self, setup: Scgd0Setup, ctx=None):
        self.ctx = ctx or _lib.default_context()
        self.setup = setup
        prob = _Problem(setup)
        self.ctx.check(_lib.lib().epsx_scgd0_set_problem(self.ctx.handle, ctypes_ref(prob.P)))
        del prob  # the library copied everything

    def loop(self, occ: Occupation, *, η_min, η_init, maxiter, tol, gpu_tile=None, Σ_init=None,
             fix_μ=False, ω_acoustic=OMEGA_ACOUSTIC, verbosity=0, method="direct",
             precision="fp64", batched=False, μ_start=None, reuse_star=False, warm_start=None):
        """`_loop_scgd0` (src/scgd0.jl:189-346) on the device.

        warm_start: None, or the precision ("bf16" / "fp32") of the context's network (set up with
        epsx.learn.WarmStartNet), which predicts Σ_init from the G0D0 self-energy of the first
        iteration on the device; each temperature goes on from the predicted or the plain start,
        whichever has the smaller residual after one more iteration (info["warm_kept"])."""
        s = self.setup
        nm, ni, nf = s.dims
        nω, nT = len(s.ωs), len(occ)
        if Σ_init is not None:
            Σ_init = np.asarray(Σ_init)
            if Σ_init.shape != (nω, ni, nT):
                raise ArgumentError(f"Σ_init must have size (nω, n_i, nT) = {(nω, ni, nT)}, got "
                                    f"{Σ_init.shape}")
            Σ_init = c128(Σ_init)
        η_min = f64(np.broadcast_to(np.asarray(η_min, dtype=np.float64), (nT,)))
        T = f64(occ.Tlist)
        nl = f64(occ.nlist)
        mu0 = f64(s.μ_start if μ_start is None else μ_start)
        if mu0.shape != (nT,):
            raise ArgumentError(f"μ_start must have one entry per temperature ({nT})")
        par = _lib.Scgd0Params()
        par.struct_size = C_sizeof(_lib.Scgd0Params)
        par.nT = nT
        par.Tlist, par.nlist, par.mu_start, par.eta_min = ptr(T), ptr(nl), ptr(mu0), ptr(η_min)
        par.nelec, par.spin_degeneracy = float(occ.nelec), float(occ.spin_degeneracy)
        par.is_semiconductor = int(occ.type == "Semiconductor")
        par.fix_mu = int(bool(fix_μ))
        par.eta_init, par.tol, par.w_acoustic = float(η_init), float(tol), float(ω_acoustic)
        par.maxiter = int(maxiter)
        par.f_tile = 0 if gpu_tile is None else int(gpu_tile)
        par.fm_method = _lib.FM_METHODS[method]
        par.precision = _lib.PRECISIONS[precision]
        par.reuse_star = int(bool(reuse_star))
        par.batched = int(bool(batched))
        par.verbosity = int(verbosity)
        par.Sigma_init = ptr(Σ_init)
        par.warm_start = 0 if warm_start is None else _lib.PRECISIONS[warm_start]
        Σ = np.empty((nω, ni, nT), dtype=np.complex128, order="F")
        μ = np.empty(nT)
        conv = np.empty(nT, dtype=np.int32)
        herr = np.empty((nT, maxiter), order="F")
        hmu = np.empty((nT, maxiter), order="F")
        hmin = np.empty((nT, maxiter), order="F")
        wall = np.empty(maxiter)
        frozen = np.empty(nT, dtype=np.int64)
        out = _lib.Scgd0Outputs()
        out.struct_size = C_sizeof(_lib.Scgd0Outputs)
        out.Sigma, out.mu, out.converged = ptr(Σ), ptr(μ), ptr(conv)
        out.hist_err, out.hist_mu, out.hist_minA = ptr(herr), ptr(hmu), ptr(hmin)
        out.wall, out.iter_frozen = ptr(wall), ptr(frozen)
        kept = np.zeros(nT, dtype=np.int32)
        out.warm_kept = ptr(kept)
        self.ctx.check(_lib.lib().epsx_scgd0_run(self.ctx.handle, ctypes_ref(par), ctypes_ref(out)))
        n = int(out.niter)
        converged = conv.astype(bool)
        if not converged.all() and maxiter > 1:
            warnings.warn(f"scGD0 not converged after maxiter = {maxiter} iterations at T = "
                          f"{list(occ.Tlist[~converged])}")
        history = dict(err=herr[:, :n].copy(), μ=hmu[:, :n].copy(), min_spectral_weight=hmin[:, :n].copy(),
                       wall=wall[:n].copy(), iter_frozen=frozen.copy())
        info = dict(fm_method=int(out.fm_method_used), n_low_precision=int(out.n_low_precision),
                    warm_kept=kept.astype(bool), timings=self.ctx.timings())
        return dict(Σ=Σ, μ=μ, converged=converged, history=history, info=info)


def ctypes_ref(struct):
    import ctypes

    return ctypes.addressof(struct)


def loop_scgd0(setup: Scgd0Setup, occ: Occupation, *, η_min, η_init, maxiter, tol, gpu_tile=None,
               Σ_init=None, fix_μ=False, ω_acoustic=OMEGA_ACOUSTIC, verbosity=0, ctx=None,
               **kw):
    """`_loop_scgd0(setup, occ; ...)` (src/scgd0.jl:189-346): uploads the setup and runs."""
    dev = Scgd0Device(setup, ctx)
    return dev.loop(occ, η_min=η_min, η_init=η_init, maxiter=maxiter, tol=tol, gpu_tile=gpu_tile,
                    Σ_init=Σ_init, fix_μ=fix_μ, ω_acoustic=ω_acoustic, verbosity=verbosity, **kw)


def check_run_arguments(setup, occ, ωs, ωs_spectral, η_min, gpu_tile, maxiter, fix_μ):
    """The argument checks and defaults of run_scgd0 (src/scgd0.jl:81-113).  Returns
    (ωs_spectral, η_min as a vector)."""
    if occ.occ_type != "FermiDirac":
        raise ArgumentError("run_scgd0 uses the Fermi-Dirac occupation only, got occ.occ_type = "
                            f"{occ.occ_type}")
    if occ.type not in ("Metal", "Semiconductor"):
        raise ArgumentError(f"occ.type must be :Metal or :Semiconductor, got {occ.type}")
    if not isinstance(ωs, FrequencyGrid):
        raise ArgumentError(f"ωs must be a FrequencyGrid{{Float64}}, got a {type(ωs).__name__}; build "
                            "one with `graded_frequency_grid` or `FrequencyGrid(ωs::AbstractVector)`")
    if ωs_spectral is None:
        ωs_spectral = uniform_spectral_grid(ωs)
    else:
        if not isinstance(ωs_spectral, FrequencyGrid):
            raise ArgumentError("ωs_spectral must be a FrequencyGrid{Float64} or nothing, got a "
                                f"{type(ωs_spectral).__name__}")
        if not (ωs_spectral[0] <= ωs[0] + ωs.dω_min and ωs_spectral[-1] >= ωs[-1] - ωs.dω_min):
            raise ArgumentError(f"ωs_spectral spans [{ωs_spectral[0]}, {ωs_spectral[-1]}], which does "
                                f"not cover ωs = [{ωs[0]}, {ωs[-1]}] to within ωs.dω_min")
        if ωs_spectral[0] < ωs[0] - ωs.dω_min or ωs_spectral[-1] > ωs[-1] + ωs.dω_min:
            warnings.warn(f"ωs_spectral spans [{ωs_spectral[0]}, {ωs_spectral[-1]}], past ωs = "
                          f"[{ωs[0]}, {ωs[-1]}]; Σ is held at its end values outside ωs")
    if fix_μ and occ.μlist is None:
        raise ArgumentError("fix_μ = true keeps μ at occ.μlist, which is not given")
    η = np.atleast_1d(np.asarray(η_min, dtype=np.float64))
    if η.size == 1 and np.ndim(η_min) == 0:
        η = np.full(len(occ), float(η_min))
    if len(η) != len(occ):
        raise ArgumentError("η_min must be a scalar or have one entry per temperature "
                            f"({len(occ)}), got {len(η)}")
    if not (gpu_tile is None or gpu_tile >= 1):
        raise ArgumentError(f"gpu_tile must be nothing or a positive number of inner states, got {gpu_tile}")
    if maxiter < 1:
        raise ArgumentError(f"maxiter must be at least 1, got {maxiter}")
    return ωs_spectral, η


def run_scgd0(setup: Scgd0Setup, occ: Occupation, *, ωs_spectral=None, η_min=0.0, η_init=5e-3,
              maxiter=20, tol=5e-5, gpu_tile=None, Σ_init=None, fix_μ=False,
              ω_acoustic=OMEGA_ACOUSTIC, verbosity=0, ctx=None, **kw) -> SCGD0Result:
    """`run_scgd0` (src/scgd0.jl:75-121) on a setup: checks, the default spectral grid, the
    starting μ (occ.μlist if given, else the setup's bare μ), the loop, the result.  The
    defaults η_init = 5 meV and tol = 0.05 meV assume eV units."""
    ωs_spectral, η = check_run_arguments(setup, occ, setup.ωs, ωs_spectral, η_min, gpu_tile,
                                        maxiter, fix_μ)
    s = dataclasses.replace(setup, ωs_spectral=ωs_spectral)
    μ0 = occ.μlist if occ.μlist is not None else setup.μ_start
    s = dataclasses.replace(s, μ_start=np.asarray(μ0, dtype=np.float64))
    r = loop_scgd0(s, occ, η_min=η, η_init=η_init, maxiter=maxiter, tol=tol, gpu_tile=gpu_tile,
                   Σ_init=Σ_init, fix_μ=fix_μ, ω_acoustic=ω_acoustic, verbosity=verbosity, ctx=ctx, **kw)
    return SCGD0Result(setup.ωs, ωs_spectral, r["Σ"], r["μ"], occ.Tlist.copy(), occ.nlist.copy(),
                       r["converged"], r["history"], r["info"])
