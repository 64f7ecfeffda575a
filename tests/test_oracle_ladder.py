"""The CPU oracle's ladder pinned against the reference's ladder test oracles (CPU only)."""
import math

import numpy as np
import pytest

import oracle
import epsx
from epsx.ladder import setup_ladder
from epsx.models import holstein_ladder_setup
from ref_numpy import (hat_value, naive_apply_ladder, pseudorandom, pseudorandom_complex, spectral_bubble)


def rel(a, b):
    return np.linalg.norm(np.ravel(a - b)) / np.linalg.norm(np.ravel(b))


def test_contour_identities():
    """test/test_ladder_bubble.jl:44-60."""
    f = lambda x, T: 1 / (math.exp(x / T) + 1)
    for ω, ε, Σ, μ, T in ((0.3, -0.1, -0.02 - 0.05j, 0.05, 0.1), (-0.7, 0.4, 0.03 - 0.2j, -0.1, 0.25),
                          (1.1, 0.2, -0.1 - 0.01j, 0.3, 0.05)):
        GR = 1 / (ω - ε - Σ)
        GA = GR.conjugate()
        G11, G21, G12, G22 = (oracle.contour_green(ω, ε, Σ, μ, T, c) for c in range(1, 5))
        fω = f(ω - μ, T)
        tol = 1e-14 * abs(GR - GA)
        assert abs(G11 - G12 - GR) <= tol and abs(G21 - G22 - GR) <= tol and abs(G11 - G21 - GA) <= tol
        assert abs(G21 - (1 - fω) * (GR - GA)) <= tol and abs(G12 + fω * (GR - GA)) <= tol


def test_attach_against_2x2():
    """test/test_ladder_bubble.jl:71-99."""
    nw, ni = 5, 3
    Gm = pseudorandom_complex(nw, 4, ni, seed=1.1)
    Gp = pseudorandom_complex(nw, 4, ni, seed=2.3)
    X = pseudorandom_complex(nw, 4, 3, 2, ni, seed=3.7)
    J = oracle.attach_greens(Gm, X, Gp)
    J_ref = np.zeros_like(J)
    for i in range(ni):
        for c0 in range(2):
            for a in range(3):
                for w in range(nw):
                    J_ref[w, :, a, c0, i] = (Gm[w, :, i].reshape(2, 2, order="F") @ X[w, :, a, c0, i].reshape(2, 2, order="F")
                                             @ Gp[w, :, i].reshape(2, 2, order="F")).reshape(4, order="F")
    assert rel(J, J_ref) <= 1e-14
    v = pseudorandom(3, ni, seed=4.1)
    V = np.zeros((nw, 4, 3, 2, ni), dtype=complex)
    for c0 in range(2):
        V[:, c0 + 2 * c0, :, c0, :] = v[None, :, :]
    J0 = oracle.attach_greens_bare(Gm, v, Gp)
    assert rel(J0, oracle.attach_greens(Gm, V, Gp)) <= 1e-14 and np.linalg.norm(J0) > 0


def holstein_case(symmetry, ω0=0.173, nk=16, ωs=None):
    s, R, v = holstein_ladder_setup(symmetry=symmetry, ω0=ω0, nk=nk, ωs=ωs)
    r = oracle.loop_scgd0(s, [0.1, 0.2], 0.0, 1.0, η_min=0.0, η_init=0.05, tol=1e-6, maxiter=100, ω_acoustic=1e-6)
    assert r["converged"].all()
    return s, setup_ladder(s, R, v), r


def propagators(L, r, iT, Ω, T):
    ws = np.asarray(L.ωs)
    jm, tm = oracle.linear_interpolation_table(ws, ws - Ω / 2)
    jp, tp = oracle.linear_interpolation_table(ws, ws + Ω / 2)
    Sm = oracle.apply_interpolation_table(r["Σ"][:, :, iT], jm, tm)
    Sp = oracle.apply_interpolation_table(r["Σ"][:, :, iT], jp, tp)
    return oracle.contour_greens(ws, L.ε_i, Sm, Sp, Ω, r["μ"][iT], T)


@pytest.mark.parametrize("Ω", [0.0, 0.3])
def test_bubble_identity(Ω):
    """test/test_ladder_bubble.jl:101-122: the contour bubble equals the spectral bubble."""
    s, L, r = holstein_case(True, ω0=0.2, nk=64, ωs=epsx.FrequencyGrid.from_nodes(np.arange(-200, 201) * 0.02))
    for iT, T in enumerate((0.1, 0.2)):
        Gm, Gp = propagators(L, r, iT, Ω, T)
        J0 = oracle.attach_greens_bare(Gm, L.v, Gp)
        σ = oracle.conductivity_from_vertex(J0, L.v, s.w_i, oracle.trapezoid_weights(np.asarray(s.ωs)), Ω, T)
        σ_ref = spectral_bubble(np.asarray(s.ωs), r["Σ"][:, :, iT], s.ε_i, s.w_i, L.v, r["μ"][iT], T, Ω)
        assert σ_ref[0, 0] > 0
        assert rel(σ, σ_ref) <= 1e-12


def test_shift_regrouping_and_tables():
    """test/test_ladder_kernel.jl:56-81, and the product's tables bitwise equal to the oracle's."""
    nodes = np.concatenate([np.arange(-10, -3) * 0.1, np.arange(-7, 8) * 0.05, np.arange(4, 11) * 0.1])
    ωs = epsx.FrequencyGrid.from_nodes(np.round(nodes, 12))
    h, nshift = ωs.dω_unit, 13
    assert math.isclose(h, 0.05)
    klo, wlo, whi = oracle.vertex_shift_tables(ωs.offsets, nshift)
    kp, lp, hp = epsx.vertex_shift_tables(ωs, nshift)
    assert np.array_equal(klo, kp) and np.array_equal(wlo, lp) and np.array_equal(whi, hp)
    ws = np.asarray(ωs)
    y0 = pseudorandom_complex(len(ws), seed=0.7)
    y = y0.copy()
    y[0] = y[-1] = 0
    for s, sign in ((0, 1), (1, -1)):
        for ωq in (0.0123, 0.2237, 0.4999, 0.55):
            m = math.floor(ωq / h)
            r = ωq / h - m
            sh = lambda yy, mm: wlo[:, mm, s] * yy[klo[:, mm, s] - 1] + whi[:, mm, s] * yy[klo[:, mm, s]]
            reg = (1 - r) * sh(y, m) + r * sh(y, m + 1)
            direct = hat_value(y, ws, ws + sign * ωq)
            assert np.linalg.norm(reg - direct) <= 1e-14 * np.linalg.norm(direct)
            reg_e = (1 - r) * sh(y0, m) + r * sh(y0, m + 1)
            dir_e = hat_value(y0, ws, ws + sign * ωq)
            inter = (ws > ws[0] + ωq + h) & (ws < ws[-1] - ωq - h)
            assert np.linalg.norm((reg_e - dir_e)[inter]) <= 1e-14 * np.linalg.norm(dir_e[inter])


def disordered_case():
    """A case with a large vertex correction: the Holstein chain without symmetry, two modes (one
    spanning several shift steps, one below the step) and a forward-scattering coupling that
    peaks at small |k_f - k_i|, with a pseudorandom modulation."""
    s, R, v = holstein_ladder_setup(symmetry=False)
    nk = 16
    s.ωph = np.asfortranarray(np.vstack([np.full(nk, 0.173), np.full(nk, 0.015)]))
    j = np.arange(nk)
    d = np.minimum(np.abs(j[:, None] - j[None, :]), nk - np.abs(j[:, None] - j[None, :]))
    fw = 0.1 * np.exp(-(d / 1.5) ** 2)
    s.g2 = np.asfortranarray(np.stack([0.02 + fw, 0.01 + 0.3 * fw]) * (1 + 0.2 * pseudorandom(2, nk, nk, seed=2.7)))
    r = oracle.loop_scgd0(s, [0.1, 0.2], 0.0, 1.0, η_min=0.0, η_init=0.05, tol=1e-6, maxiter=100, ω_acoustic=1e-6)
    assert r["converged"].all()
    return s, setup_ladder(s, R, v), r


@pytest.mark.parametrize("case", ["holstein", "disordered"])
def test_apply_ladder_against_naive(case):
    """test/test_ladder_kernel.jl:134-172: one application against the naive loop with exact
    hat interpolation, J pseudorandom with zero end values, both temperatures, two chunkings."""
    if case == "holstein":
        s, L, r = holstein_case(False)
    else:
        s, L, r = disordered_case()
    nw, ni = len(s.ωs), s.g2.shape[1]
    J = pseudorandom_complex(nw, 4, 3, 2, ni, seed=1.3)
    J[0] = 0
    J[-1] = 0
    KK = oracle.kramers_kronig_matrix(np.asarray(s.ωs))
    for iT, T in enumerate((0.1, 0.2)):
        Gm, Gp = propagators(L, r, iT, 0.2, T)
        ref = naive_apply_ladder(s.g2, s.ωph, s.iq_kk, s.iks_i, s.iks_f, s.w_f, s.f_to_i, L.R, s.ωs, KK, J, Gm, Gp, T,
                                 1e-6)
        assert np.linalg.norm(ref) > 1e-3 * np.linalg.norm(J)
        for tile in (None, 7):
            LJ = oracle.apply_ladder_kernel(L, J, Gm, Gp, T, 1e-6, gpu_tile=tile)
            assert rel(LJ, ref) <= 1e-12


def test_holstein_current_correction_vanishes():
    """test/test_ladder_holstein.jl:5-26."""
    s, L, r = holstein_case(True)
    for iT, T in enumerate((0.1, 0.2)):
        for Ω in (0.0, 0.2):
            Gm, Gp = propagators(L, r, iT, Ω, T)
            J0 = oracle.attach_greens_bare(Gm, L.v, Gp)
            LJ = oracle.apply_ladder_kernel(L, J0, Gm, Gp, T, 1e-6)
            assert np.max(np.abs(LJ)) <= 1e-12 * np.max(np.abs(J0))
            import dataclasses
            wrong = dataclasses.replace(L, R=np.abs(L.R))
            assert np.max(np.abs(oracle.apply_ladder_kernel(wrong, J0, Gm, Gp, T, 1e-6))) > 1e-3 * np.max(np.abs(J0))
    out = oracle.ladder_solve(L, r["Σ"][:, :, 0], r["μ"][0], 0.1, 0.0, rtol=1e-2, maxiter=30)
    assert out["converged"] and rel(out["σ_ladder"], out["σ_bubble"]) <= 1e-12


def test_solvers_satisfy_the_equation():
    """test/test_ladder_pb.jl:11-22, :48-57 on the disordered case: a converged GMRES solve
    satisfies J = J0 + L(J) to 10 rtol, the fixed point agrees, the correction is not small."""
    s, L, r = disordered_case()
    T, Ω = 0.1, 0.3
    g = oracle.ladder_solve(L, r["Σ"][:, :, 0], r["μ"][0], T, Ω, rtol=1e-8, maxiter=60)
    assert g["converged"]
    Gm, Gp = propagators(L, r, 0, Ω, T)
    J0 = oracle.attach_greens_bare(Gm, L.v, Gp)
    res = g["J"] - J0 - oracle.apply_ladder_kernel(L, g["J"], Gm, Gp, T, 1e-6)
    assert np.linalg.norm(res) <= 10 * 1e-8 * np.linalg.norm(J0)
    assert np.linalg.norm(g["σ_ladder"] - g["σ_bubble"]) > 1e-2 * np.linalg.norm(g["σ_bubble"])
    assert len(g["norms"]) == g["niter"] + 1 and np.all(np.diff(g["norms"]) <= 1e-12 * g["norms"][0])
    fp = oracle.ladder_solve(L, r["Σ"][:, :, 0], r["μ"][0], T, Ω, solver="fixed_point", rtol=1e-8, maxiter=200)
    assert fp["converged"]
    assert np.linalg.norm(fp["σ_ladder"] - g["σ_ladder"]) <= 10 * 1e-6 * np.linalg.norm(g["σ_ladder"])
    # warm start from the solution: converged at once
    w = oracle.ladder_solve(L, r["Σ"][:, :, 0], r["μ"][0], T, Ω, rtol=1e-8, maxiter=60, J=g["J"])
    assert w["niter"] <= 1 and w["converged"]


_T2G = {}


def t2g_case():
    """The t2g model at nk = 4 with its Oh vertex rotations R_f and band velocities (the C4 model
    reduced), 3 modes, a coarse graded grid; Σ from 200 oracle iterations (converged or not, it is
    the input of the ladder)."""
    if not _T2G:
        from epsx.models import t2g_ladder_setup
        s, R, v, meta = t2g_ladder_setup(nk=4, Tlist_K=(100.0, 300.0), nmodes=3,
                                         grid_windows=((0.3, 8e-3), (1.2, 32e-3), (3.0, 128e-3)))
        r = oracle.loop_scgd0(s, meta["Tlist"], 0.0, 1.0, η_min=0.0, η_init=5e-2, tol=1e-7, maxiter=200,
                              ω_acoustic=1e-6)
        _T2G.update(s=s, L=setup_ladder(s, R, v), r=r, meta=meta)
    return _T2G["s"], _T2G["L"], _T2G["r"], _T2G["meta"]


def test_apply_ladder_against_naive_t2g_oh():
    """One application on the t2g model with the Oh rotations of the vertex (R_f != identity, 3x3
    mixing of the Cartesian components) against the naive loop (test/test_ladder_kernel.jl:134-172)."""
    s, L, r, meta = t2g_case()
    assert np.count_nonzero(np.abs(L.R[0, 1, :]) + np.abs(L.R[0, 2, :])) > 0
    nw, ni = len(s.ωs), s.g2.shape[1]
    J = pseudorandom_complex(nw, 4, 3, 2, ni, seed=1.3)
    J[0] = 0
    J[-1] = 0
    T = meta["Tlist"][0]
    Gm, Gp = propagators(L, r, 0, 0.02, T)
    KK = oracle.kramers_kronig_matrix(np.asarray(s.ωs))
    ref = naive_apply_ladder(s.g2, s.ωph, s.iq_kk, s.iks_i, s.iks_f, s.w_f, s.f_to_i, L.R, s.ωs, KK, J, Gm, Gp, T,
                             1e-6)
    LJ = oracle.apply_ladder_kernel(L, J, Gm, Gp, T, 1e-6)
    assert np.linalg.norm(ref) > 1e-2 * np.linalg.norm(J)
    assert rel(LJ, ref) <= 1e-12


def test_t2g_sigma_symmetrised_over_oh():
    """run_scgd0_ladder symmetrises the irreducible-state sum over the point group
    (src/ladder.jl:94-101).  Over the 48 operations of Oh a rank-2 tensor becomes tr(σ)/3 times the
    identity: the t2g model's unsymmetrised σ_bubble and σ_ladder (from the oracle's GMRES, a vertex
    correction of tens of percent) are anisotropic, their symmetrised forms isotropic with the same
    trace, and a wrong group (the identity only, or a single mirror) leaves them anisotropic."""
    from epsx.ladder import symmetrize
    from epsx.models import oh_operations
    s, L, r, meta = t2g_case()
    T = meta["Tlist"][0]
    g = oracle.ladder_solve(L, r["Σ"][:, :, 0], r["μ"][0], T, 0.02, rtol=1e-8, maxiter=60)
    assert g["converged"] and rel(g["σ_ladder"], g["σ_bubble"]) > 0.05
    ops = oh_operations()
    assert len(ops) == 48 and all(np.allclose(o @ o.T, np.eye(3)) for o in ops)
    for σ in (g["σ_bubble"], g["σ_ladder"]):
        aniso = np.linalg.norm(σ - np.trace(σ) / 3 * np.eye(3))
        assert aniso > 1e-2 * abs(np.trace(σ))
        sym = symmetrize(σ, ops)
        assert np.linalg.norm(sym - np.trace(σ) / 3 * np.eye(3)) <= 1e-14 * abs(np.trace(σ))
        assert np.linalg.norm(symmetrize(σ, [np.diag([-1.0, 1, 1])]) - np.trace(σ) / 3 * np.eye(3)) > 1e-3 * abs(np.trace(σ))
