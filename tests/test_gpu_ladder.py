"""GPU parity of the ladder: libepsx (through its C ABI) against the CPU oracle.

Tolerances: one kernel application 1e-12 relative (the device regroups the shift sums, see
epsx_ladder.hip, so the sum order differs); σ and J of full solves 1e-10 relative (the
reference's CPU-vs-GPU gate, test/test_ladder_gpu.jl:46-48) with equal iteration counts;
bitwise run to run.
"""
import os
import warnings

import numpy as np
import pytest

import epsx
import oracle
from epsx.ladder import LadderDevice, setup_ladder
from test_oracle_ladder import disordered_case, holstein_case, propagators
from ref_numpy import pseudorandom_complex

pytestmark = pytest.mark.gpu


def rel(a, b):
    return np.linalg.norm(np.ravel(a - b)) / np.linalg.norm(np.ravel(b))


def device(ctx, s, L):
    dev = epsx.Scgd0Device(s, ctx)
    return LadderDevice(dev, L)


@pytest.mark.parametrize("case", ["holstein", "forward"])
def test_apply_ladder_kernel(ctx, case):
    s, L, r = holstein_case(False) if case == "holstein" else disordered_case()
    ld = device(ctx, s, L)
    nw, ni = len(s.ωs), s.g2.shape[1]
    J = pseudorandom_complex(nw, 4, 3, 2, ni, seed=1.3)
    J[0] = 0
    J[-1] = 0
    for iT, T in enumerate((0.1, 0.2)):
        Gm, Gp = propagators(L, r, iT, 0.2, T)
        ref = oracle.apply_ladder_kernel(L, J, Gm, Gp, T, 1e-6)
        LJ = ld.apply(J, Gm, Gp, T, 1e-6)
        assert rel(LJ, ref) <= 1e-12
        assert np.array_equal(ld.apply(J, Gm, Gp, T, 1e-6), LJ)


@pytest.mark.parametrize("prec,tol", [("fp64", 1e-13), ("fp32", 1e-6), ("bf16", 1e-5)])
def test_shift_gemm_blas_matches_own_kernels(ctx, prec, tol):
    """The shift-table product in rocBLAS (the default) against the library's own MFMA kernels
    (EPSX_LADDER_BLAS=0): the same operands, so one application agrees to the accumulation order
    (FP64, or FP32 for the stored FP32 / BF16 operands); the default path is the one the other
    ladder tests and the oracle comparisons run."""
    s, L, r = disordered_case()
    ld = device(ctx, s, L)
    nw, ni = len(s.ωs), s.g2.shape[1]
    J = pseudorandom_complex(nw, 4, 3, 2, ni, seed=3.7)
    Gm, Gp = propagators(L, r, 0, 0.2, 0.1)
    blas = ld.apply(J, Gm, Gp, 0.1, 1e-6, precision=prec)
    os.environ["EPSX_LADDER_BLAS"] = "0"
    try:
        own = ld.apply(J, Gm, Gp, 0.1, 1e-6, precision=prec)
    finally:
        del os.environ["EPSX_LADDER_BLAS"]
    assert rel(blas, own) <= tol
    if prec == "fp64":
        assert rel(blas, oracle.apply_ladder_kernel(L, J, Gm, Gp, 0.1, 1e-6)) <= 1e-12


def oracle_loop(L, r, Ωs, T_idx, **kw):
    out = []
    for iT in T_idx:
        T = (0.1, 0.2)[iT]
        J = None
        for Ω in Ωs:
            o = oracle.ladder_solve(L, r["Σ"][:, :, iT], r["μ"][iT], T, Ω, J=J, **kw)
            J = o["J"]
            out.append(o)
    return out


def test_gmres_loop_against_oracle(ctx):
    s, L, r = disordered_case()
    ld = device(ctx, s, L)
    Ωs = [0.0, 0.1, 0.3]
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        dv = ld.loop(r["Σ"], r["μ"], [0.1, 0.2], Ωs, rtol=1e-8, maxiter=60, store_vertex=True)
    ref = oracle_loop(L, r, Ωs, (0, 1), rtol=1e-8, maxiter=60)
    k = 0
    for iT in range(2):
        for iO in range(len(Ωs)):
            o = ref[k]
            k += 1
            assert dv["niter"][iO, iT] == o["niter"] and dv["converged"][iO, iT] == o["converged"]
            assert rel(dv["σ_bubble"][:, :, iO, iT], o["σ_bubble"]) <= 1e-12
            assert rel(dv["σ_ladder"][:, :, iO, iT], o["σ_ladder"]) <= 1e-10
            assert rel(dv["J"][..., iO, iT], o["J"]) <= 1e-10
            h = dv["history"][iO][iT]["norms"]
            assert len(h) == len(o["norms"]) and np.allclose(h, o["norms"], rtol=1e-6, atol=1e-12 * o["norms"][0])
    # GMRES without restart reaches 1e-8 within 60 iterations only at the largest Ω here; the
    # others stop at maxiter, as the oracle's do
    assert dv["converged"][-1].all()
    assert np.all(dv["niter"][~dv["converged"]] == 60)
    # the correction is not negligible, so the agreement is not a trivial one
    assert np.linalg.norm(dv["σ_ladder"] - dv["σ_bubble"]) > 1e-2 * np.linalg.norm(dv["σ_bubble"])
    again = ld.loop(r["Σ"], r["μ"], [0.1, 0.2], Ωs, rtol=1e-8, maxiter=60, store_vertex=True)
    assert np.array_equal(again["J"], dv["J"]) and np.array_equal(again["σ_ladder"], dv["σ_ladder"])


def test_fixed_point_against_oracle(ctx):
    s, L, r = disordered_case()
    ld = device(ctx, s, L)
    dv = ld.loop(r["Σ"], r["μ"], [0.1], [0.3], solver="fixed_point", rtol=1e-6, maxiter=200)
    o = oracle.ladder_solve(L, r["Σ"][:, :, 0], r["μ"][0], 0.1, 0.3, solver="fixed_point", rtol=1e-6, maxiter=200)
    assert dv["niter"][0, 0] == o["niter"] and dv["converged"][0, 0] and o["converged"]
    assert rel(dv["σ_ladder"][:, :, 0, 0], o["σ_ladder"]) <= 1e-10


def test_holstein_ladder_equals_bubble(ctx):
    """test/test_ladder_holstein.jl:23-25: with the k -> -k symmetry the current correction
    vanishes, σ_ladder = σ_bubble, and run_scgd0_ladder symmetrises with spin / volume."""
    s, L, r = holstein_case(True)
    dev = epsx.Scgd0Device(s, ctx)
    res = epsx.SCGD0Result(s.ωs, s.ωs_spectral, r["Σ"], r["μ"], np.array([0.1, 0.2]), np.zeros(2),
                           r["converged"], r["history"])
    occ = epsx.Occupation(Tlist=[0.1, 0.2], nlist=0.0, nelec=1.0, spin_degeneracy=2.0, volume=1.0)
    out = epsx.run_scgd0_ladder(dev, L, res, occ, Ωs=[0.0, 0.2])
    assert out.converged.all()
    assert rel(out.σ_ladder, out.σ_bubble) <= 1e-12
    ref = oracle.ladder_solve(L, r["Σ"][:, :, 1], r["μ"][1], 0.2, 0.2, rtol=1e-2, maxiter=30)
    assert rel(out.σ_bubble[:, :, 1, 1], 2.0 * ref["σ_bubble"]) <= 1e-12
    with pytest.raises(epsx.ArgumentError, match="ω_acoustic"):
        epsx.run_scgd0_ladder(dev, L, res, occ, ω_acoustic=-1.0)
    with pytest.raises(epsx.ArgumentError, match="solver"):
        epsx.run_scgd0_ladder(dev, L, res, occ, solver="newton")


@pytest.mark.parametrize("prec,tol_apply", [("fp32", 1e-5), ("bf16", 2e-2)])
def test_low_precision_shift_gemm(ctx, prec, tol_apply):
    """The shift-table GEMMs in FP32 / BF16 (FP32 accumulation) against FP64: one application, and
    σ of a GMRES solve at rtol 1e-2 against the FP64 solve at rtol 1e-8 (C4's check)."""
    s, L, r = disordered_case()
    ld = device(ctx, s, L)
    nw, ni = len(s.ωs), s.g2.shape[1]
    J = pseudorandom_complex(nw, 4, 3, 2, ni, seed=2.1)
    Gm, Gp = propagators(L, r, 0, 0.2, 0.1)
    ref = ld.apply(J, Gm, Gp, 0.1, 1e-6)
    lp = ld.apply(J, Gm, Gp, 0.1, 1e-6, precision=prec)
    assert rel(lp, ref) <= tol_apply
    # from B̃ and the vertex stored in BF16 / FP32 (k-contiguous) against the product that rounds
    # the FP64 operands as it stages them: the same rounded operands, FP32 accumulation order apart
    os.environ["EPSX_LP_STAGED"] = "1"
    try:
        staged = ld.apply(J, Gm, Gp, 0.1, 1e-6, precision=prec)
    finally:
        del os.environ["EPSX_LP_STAGED"]
    assert rel(lp, staged) <= (1e-5 if prec == "bf16" else 1e-6)
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        tight = ld.loop(r["Σ"], r["μ"], [0.1], [0.3], rtol=1e-8, maxiter=60)
        loose64 = ld.loop(r["Σ"], r["μ"], [0.1], [0.3], rtol=1e-2, maxiter=30)
        loose = ld.loop(r["Σ"], r["μ"], [0.1], [0.3], rtol=1e-2, maxiter=30, precision=prec)
    assert np.array_equal(loose["σ_bubble"], tight["σ_bubble"])
    assert loose["converged"].all()
    # the precision effect (same rtol) is far below the solver-tolerance effect (rtol 1e-2 vs 1e-8)
    assert rel(loose["σ_ladder"], loose64["σ_ladder"]) <= (1e-6 if prec == "fp32" else 5e-3)
    assert rel(loose["σ_ladder"], tight["σ_ladder"]) <= 3e-2


def test_t2g_oh_gmres_against_oracle(ctx):
    """The t2g model at nk = 4 with its Oh vertex rotations (C4's model reduced; the Σ of the
    oracle's scGD0 run): GMRES at rtol 1e-8 for Ω = 0 and 20 meV at 100 K, warm-started from the
    previous Ω, against the oracle: equal iteration counts, σ and J at 1e-10."""
    from test_oracle_ladder import t2g_case
    s, L, r, meta = t2g_case()
    ld = device(ctx, s, L)
    T = meta["Tlist"][0]
    Ωs = [0.0, 0.02]
    dv = ld.loop(r["Σ"][:, :, :1], r["μ"][:1], [T], Ωs, rtol=1e-8, maxiter=60, store_vertex=True)
    J = None
    for iO, Ω in enumerate(Ωs):
        o = oracle.ladder_solve(L, r["Σ"][:, :, 0], r["μ"][0], T, Ω, J=J, rtol=1e-8, maxiter=60)
        J = o["J"]
        assert o["converged"] and dv["converged"][iO, 0]
        assert dv["niter"][iO, 0] == o["niter"]
        assert rel(dv["σ_bubble"][:, :, iO, 0], o["σ_bubble"]) <= 1e-12
        assert rel(dv["σ_ladder"][:, :, iO, 0], o["σ_ladder"]) <= 1e-10
        assert rel(dv["J"][..., iO, 0], o["J"]) <= 1e-10
    assert rel(dv["σ_ladder"], dv["σ_bubble"]) > 0.05


def test_c4_t2g16_gmres_solves_the_equation(ctx):
    """C4 at its size: the t2g model at 16³ (ni = 331, nf = 8448, Oh rotations) from its converged
    FP64 scGD0 Σ at 100 K; GMRES at rtol 1e-2 in FP64, FP32 and BF16 shift-table GEMMs for Ω = 0 and
    100 meV.  The FP64 solution satisfies J = J0 + L(J) to 10 rtol with the device's own L (J0 from
    the oracle's propagators, test/test_ladder_pb.jl:48-57), and σ_ladder of each precision lies
    within 3 rtol of the FP64 GMRES at rtol 1e-8."""
    from epsx.models import t2g_ladder_setup
    from test_oracle_ladder import propagators
    s, R, v, meta = t2g_ladder_setup(nk=16, Tlist_K=(100.0,))
    dev = epsx.Scgd0Device(s, ctx)
    occ = epsx.Occupation(Tlist=meta["Tlist"], nlist=0.0, nelec=1.0)
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        r = dev.loop(occ, η_min=0.0, η_init=5e-3, maxiter=200, tol=1e-6, ω_acoustic=1e-6, method="auto",
                     batched=True)
    assert r["converged"].all()
    L = setup_ladder(s, R, v)
    ld = LadderDevice(dev, L)
    T = float(meta["Tlist"][0])
    Ωs = [0.0, 0.1]
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        ref = ld.loop(r["Σ"], r["μ"], [T], Ωs, rtol=1e-8, maxiter=80)
        runs = {p: ld.loop(r["Σ"], r["μ"], [T], Ωs, rtol=1e-2, maxiter=30, precision=p, store_vertex=(p == "fp64"))
                for p in ("fp64", "fp32", "bf16")}
    g = runs["fp64"]
    assert g["converged"].all() and np.all(g["niter"] >= 1)
    for iO, Ω in enumerate(Ωs):
        Gm, Gp = propagators(L, r, 0, Ω, T)
        J0 = oracle.attach_greens_bare(Gm, L.v, Gp)
        J = g["J"][..., iO, 0]
        res = J - J0 - ld.apply(J, Gm, Gp, T, 1e-6)
        assert np.linalg.norm(res) <= 10 * 1e-2 * np.linalg.norm(J0)
    for p, x in runs.items():
        assert x["converged"].all(), p
        assert np.array_equal(x["σ_bubble"], ref["σ_bubble"]), p
        assert rel(x["σ_ladder"], ref["σ_ladder"]) <= 3e-2, p


def test_batched_gmres_against_oracle_cold_start(ctx):
    """The Ωs of each temperature as instances of one batched GMRES (stacked shift-table product,
    batched MGS, Givens rotations on the device), every instance from J = 0: against the oracle's
    GMRES from J = 0 for each (Ω, T) of the disordered case at rtol 1e-8, equal iteration counts and
    residual histories, σ and J at 1e-10, bitwise on repeat."""
    s, L, r = disordered_case()
    ld = device(ctx, s, L)
    Ωs = [0.0, 0.1, 0.3]
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        dv = ld.loop(r["Σ"], r["μ"], [0.1, 0.2], Ωs, rtol=1e-8, maxiter=60, store_vertex=True, batched=True)
        again = ld.loop(r["Σ"], r["μ"], [0.1, 0.2], Ωs, rtol=1e-8, maxiter=60, store_vertex=True, batched=True)
    for iT, T in enumerate((0.1, 0.2)):
        for iO, Ω in enumerate(Ωs):
            o = oracle.ladder_solve(L, r["Σ"][:, :, iT], r["μ"][iT], T, Ω, J=None, rtol=1e-8, maxiter=60)
            assert dv["niter"][iO, iT] == o["niter"] and dv["converged"][iO, iT] == o["converged"]
            assert rel(dv["σ_bubble"][:, :, iO, iT], o["σ_bubble"]) <= 1e-12
            assert rel(dv["σ_ladder"][:, :, iO, iT], o["σ_ladder"]) <= 1e-10
            assert rel(dv["J"][..., iO, iT], o["J"]) <= 1e-10
            h = dv["history"][iO][iT]["norms"]
            assert len(h) == len(o["norms"]) and np.allclose(h, o["norms"], rtol=1e-6, atol=1e-12 * o["norms"][0])
    assert len(set(dv["niter"].ravel().tolist())) > 1     # the instances stop at different steps
    assert np.array_equal(again["J"], dv["J"]) and np.array_equal(again["σ_ladder"], dv["σ_ladder"])


@pytest.mark.parametrize("prec", ["fp64", "bf16"])
def test_batched_gmres_t2g_matches_sequential(ctx, prec):
    """The t2g model with Oh rotations: the batched solve (cold starts) and the reference's sequential
    loop (warm starts from the previous Ω) reach the same σ_ladder to the solver tolerance; in BF16
    shift-table products as well."""
    from test_oracle_ladder import t2g_case
    s, L, r, meta = t2g_case()
    ld = device(ctx, s, L)
    T = meta["Tlist"][:1]
    Ωs = [0.0, 0.01, 0.02, 0.04]
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        ref = ld.loop(r["Σ"][:, :, :1], r["μ"][:1], T, Ωs, rtol=1e-8, maxiter=60)
        b = ld.loop(r["Σ"][:, :, :1], r["μ"][:1], T, Ωs, rtol=1e-3, maxiter=60, batched=True, precision=prec)
        sq = ld.loop(r["Σ"][:, :, :1], r["μ"][:1], T, Ωs, rtol=1e-3, maxiter=60, precision=prec)
    assert b["converged"].all() and sq["converged"].all()
    assert np.array_equal(b["σ_bubble"], ref["σ_bubble"])
    tol = 3e-2 if prec == "bf16" else 3e-3   # BF16 products: their rounding, not rtol, sets the floor
    assert rel(b["σ_ladder"], ref["σ_ladder"]) <= tol and rel(sq["σ_ladder"], ref["σ_ladder"]) <= tol
