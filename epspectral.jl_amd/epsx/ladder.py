"""The ladder (vertex-corrected Kubo conductivity) set-up: the Python side of src/ladder.jl.

`setup_ladder` mirrors `_setup_ladder` (src/ladder.jl:112-158): the rotations R[:, :, f] of
the vertex (identity without symmetry), the inner states grouped by representative, the
hat-basis shift tables on the grid's own step, the quadrature weights and the velocities.
The device loop lives in libepsx (`LadderDevice`).
"""
import dataclasses
import math

import numpy as np

from ._lib import ArgumentError
from .grids import FrequencyGrid, trapezoid_weights, vertex_shift_tables
from .scgd0 import Scgd0Setup


@dataclasses.dataclass
class LadderSetup:
    scgd0: Scgd0Setup
    R: np.ndarray                # (3, 3, nf) rotation of each inner state's vertex
    v: np.ndarray                # (3, ni) band velocities of the outer states
    inner_states_by_star: np.ndarray   # 1-based, sortperm(f_to_i)
    inner_star_ptr: np.ndarray         # 1-based, length ni + 1
    nshift: int
    dω_unit: float
    ωs_weights: np.ndarray
    klo: np.ndarray
    wlo: np.ndarray
    whi: np.ndarray

    # the reference's ladder NamedTuple fields, read through to the scGD0 setup
    def __getattr__(self, k):
        if k in ("scgd0",):
            raise AttributeError(k)
        return getattr(self.scgd0, k)


def setup_ladder(setup: Scgd0Setup, R=None, v=None):
    """`_setup_ladder` (src/ladder.jl:112-158) on a scGD0 setup."""
    nm, ni, nf = setup.dims
    ωs = setup.ωs
    if not isinstance(ωs, FrequencyGrid):
        raise ArgumentError("the ladder needs ωs as a FrequencyGrid")
    if R is None:
        R = np.zeros((3, 3, nf), order="F")
        for a in range(3):
            R[a, a, :] = 1.0
    R = np.asfortranarray(R, dtype=np.float64)
    if R.shape != (3, 3, nf):
        raise ArgumentError(f"R must have size (3, 3, nf) = {(3, 3, nf)}")
    v = np.zeros((3, ni), order="F") if v is None else np.asfortranarray(v, dtype=np.float64)
    f_to_i = np.asarray(setup.f_to_i)
    by_star = np.argsort(f_to_i, kind="stable") + 1
    sizes = np.bincount(f_to_i - 1, minlength=ni)
    ptr = np.concatenate([[1], 1 + np.cumsum(sizes)]).astype(np.int64)
    h = ωs.dω_unit
    nshift = int(math.ceil(np.max(setup.ωph) / h)) + 2
    klo, wlo, whi = vertex_shift_tables(ωs, nshift)
    return LadderSetup(scgd0=setup, R=R, v=v, inner_states_by_star=by_star, inner_star_ptr=ptr, nshift=nshift,
                       dω_unit=float(h), ωs_weights=trapezoid_weights(ωs), klo=klo, wlo=wlo, whi=whi)


@dataclasses.dataclass
class SCGD0LadderResult:
    """`SCGD0LadderResult` (src/ladder.jl:27-38)."""
    Ωs: np.ndarray
    Tlist: np.ndarray
    μ: np.ndarray
    σ_bubble: np.ndarray        # (3, 3, nΩ, nT), symmetrised, with spin_degeneracy / volume
    σ_ladder: np.ndarray
    converged: np.ndarray       # (nΩ, nT)
    niter: np.ndarray
    history: list               # [iΩ][iT] -> dict(norms, σ_iter)
    dω_unit: float
    J: object = None


class LadderDevice:
    """A ladder problem on the device, on top of an uploaded scGD0 setup (`Scgd0Device`)."""

    def __init__(self, dev, ladder: LadderSetup):
        import ctypes
        from . import _lib
        self.dev, self.ctx, self.ladder = dev, dev.ctx, ladder
        R = np.asfortranarray(ladder.R, dtype=np.float64)
        v = np.asfortranarray(ladder.v, dtype=np.float64)
        P = _lib.LadderProblem(ctypes.sizeof(_lib.LadderProblem), _lib.ptr(R), _lib.ptr(v))
        self.ctx.check(_lib.lib().epsx_ladder_set_problem(self.ctx.handle, ctypes.addressof(P)))

    def apply(self, J, Gm, Gp, T, ω_acoustic, precision="fp64"):
        """`apply_ladder_kernel!` (src/ladder.jl:245-296) on host arrays."""
        from . import _lib
        J = np.asfortranarray(J, dtype=np.complex128)
        Gm = np.asfortranarray(Gm, dtype=np.complex128)
        Gp = np.asfortranarray(Gp, dtype=np.complex128)
        LJ = np.empty_like(J, order="F")
        self.ctx.check(_lib.lib().epsx_apply_ladder_kernel(
            self.ctx.handle, _lib.ptr(J), _lib.ptr(Gm), _lib.ptr(Gp), float(T), float(ω_acoustic),
            _lib.PRECISIONS[precision], _lib.ptr(LJ)))
        return LJ

    def loop(self, Σ, μ, Tlist, Ωs, *, solver="gmres", rtol=1e-2, maxiter=30, store_vertex=False,
             ω_acoustic=None, precision="fp64", verbosity=0, batched=False):
        """`_loop_ladder` (src/ladder.jl:303-379): unsymmetrised σ without spin / volume.  With
        `batched` (GMRES only) the Ωs of each temperature are solved together as instances of one
        batched GMRES, each from J = 0 instead of the previous Ω's solution."""
        import ctypes
        import warnings
        from . import _lib
        from .scgd0 import OMEGA_ACOUSTIC
        ω_acoustic = OMEGA_ACOUSTIC if ω_acoustic is None else ω_acoustic
        if solver not in ("gmres", "fixed_point"):
            raise ArgumentError(f"solver must be :gmres or :fixed_point, got {solver}")
        Σ = np.asfortranarray(Σ, dtype=np.complex128)
        Tl = np.ascontiguousarray(np.atleast_1d(Tlist), dtype=np.float64)
        mu = np.ascontiguousarray(np.atleast_1d(μ), dtype=np.float64)
        Om = np.ascontiguousarray(np.atleast_1d(Ωs), dtype=np.float64)
        nT, nO = len(Tl), len(Om)
        nw, ni = Σ.shape[0], Σ.shape[1]
        sb = np.empty((3, 3, nO, nT), order="F")
        sl = np.empty((3, 3, nO, nT), order="F")
        conv = np.empty((nO, nT), dtype=np.int32, order="F")
        nit = np.empty((nO, nT), dtype=np.int64, order="F")
        norms = np.empty((maxiter + 1, nO, nT), order="F")
        sit = np.empty((3, 3, maxiter, nO, nT), order="F")
        J = np.empty((nw, 4, 3, 2, ni, nO, nT), dtype=np.complex128, order="F") if store_vertex else None
        par = _lib.LadderParams(ctypes.sizeof(_lib.LadderParams), nT, nO, _lib.ptr(Tl), _lib.ptr(mu), _lib.ptr(Σ),
                                _lib.ptr(Om), 0 if solver == "gmres" else 1, _lib.PRECISIONS[precision], float(rtol),
                                int(maxiter), float(ω_acoustic), int(store_vertex), int(verbosity), int(bool(batched)),
                                0)
        out = _lib.LadderOutputs(ctypes.sizeof(_lib.LadderOutputs), _lib.ptr(sb), _lib.ptr(sl), _lib.ptr(conv),
                                 _lib.ptr(nit), _lib.ptr(norms), _lib.ptr(sit), _lib.ptr(J))
        self.ctx.check(_lib.lib().epsx_ladder_run(self.ctx.handle, ctypes.addressof(par), ctypes.addressof(out)))
        converged = conv.astype(bool)
        history = [[dict(norms=norms[:nit[iO, iT] + 1, iO, iT].copy() if solver == "gmres" else np.zeros(0),
                         σ_iter=[sit[:, :, k, iO, iT].copy() for k in range(nit[iO, iT])] if solver != "gmres" else [])
                    for iT in range(nT)] for iO in range(nO)]
        if not converged.all():
            bad = [(Om[a], Tl[b]) for a in range(nO) for b in range(nT) if not converged[a, b]]
            warnings.warn(f"ladder not converged (rtol = {rtol}, maxiter = {maxiter}) at (Ω, T) = {bad}")
        return dict(σ_bubble=sb, σ_ladder=sl, converged=converged, niter=nit, history=history, J=J)


def symmetrize(σ, ops=None):
    """The 3x3 tensor averaged over the Cartesian point-group matrices `ops` (identity if None)."""
    if ops is None:
        return σ.copy()
    return sum(g @ σ @ g.T for g in ops) / len(ops)


def run_scgd0_ladder(dev, ladder: LadderSetup, scgd0, occ, *, Ωs=(0.0,), solver="gmres", rtol=1e-2, maxiter=30,
                     store_vertex=False, ω_acoustic=None, symmetry_ops=None, precision="fp64", verbosity=0):
    """`run_scgd0_ladder` (src/ladder.jl:65-105) on an uploaded scGD0 setup and its result: the
    argument checks, the loop and σ symmetrised with spin_degeneracy / volume."""
    from .scgd0 import OMEGA_ACOUSTIC
    ω_acoustic = OMEGA_ACOUSTIC if ω_acoustic is None else ω_acoustic
    if solver not in ("gmres", "fixed_point"):
        raise ArgumentError(f"solver must be :gmres or :fixed_point, got {solver}")
    if maxiter < 1:
        raise ArgumentError(f"maxiter must be at least 1, got {maxiter}")
    if len(Ωs) == 0:
        raise ArgumentError("Ωs must hold at least one probe frequency")
    if not ω_acoustic > 0:
        raise ArgumentError(f"ω_acoustic must be positive, got {ω_acoustic}: the shift-grid coupling build has no "
                            "row for a negative phonon frequency")
    if not (len(occ.Tlist) == len(scgd0.Tlist) and np.allclose(occ.Tlist, scgd0.Tlist)):
        raise ArgumentError(f"occ.Tlist = {list(occ.Tlist)} must equal the temperatures of scgd0, {list(scgd0.Tlist)}")
    ld = LadderDevice(dev, ladder)
    r = ld.loop(scgd0.Σ, scgd0.μ, scgd0.Tlist, np.asarray(Ωs, dtype=float), solver=solver, rtol=rtol,
                maxiter=maxiter, store_vertex=store_vertex, ω_acoustic=ω_acoustic, precision=precision,
                verbosity=verbosity)
    fac = occ.spin_degeneracy / occ.volume

    def sym(σ):
        out = np.empty_like(σ)
        for iT in range(σ.shape[3]):
            for iO in range(σ.shape[2]):
                out[:, :, iO, iT] = fac * symmetrize(σ[:, :, iO, iT], symmetry_ops)
        return out

    return SCGD0LadderResult(np.asarray(Ωs, dtype=float), np.asarray(scgd0.Tlist), np.asarray(scgd0.μ),
                             sym(r["σ_bubble"]), sym(r["σ_ladder"]), r["converged"], r["niter"], r["history"],
                             ladder.dω_unit, r["J"])
