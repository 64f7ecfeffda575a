"""The CPU oracle's scGD0 path pinned against the reference's test oracles (CPU only).

Each check restates a reference test with an independent NumPy computation (tests/ref_numpy.py):
a wrong sign, index or dropped term in the oracle fails at least one of them.
"""
import dataclasses
import json
import math
import os

import numpy as np
import pytest

import oracle
from epsx.grids import FrequencyGrid
from epsx.models import holstein_1d_setup
from epsx.scgd0 import build_multiplets
from ref_numpy import (holstein_bare_mu, holstein_local_scgd0, holstein_oneshot_imsigma, interp_flat,
                       n_old, naive_fan_migdal, pseudorandom, scgd0_v2_reference, trapezoid)

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "reference_known_answers.json")))
NK, T0, OMEGA0, G2 = 64, 0.1, 0.2, 0.2


def rel(a, b):
    return np.linalg.norm(np.ravel(a - b)) / np.linalg.norm(np.ravel(b))


def grid():
    return FrequencyGrid.from_nodes(np.arange(-200, 201) * 0.02)


def setup(μ_start, nk=NK, ωs=None):
    return holstein_1d_setup(nk, 1.0, OMEGA0, G2, grid() if ωs is None else ωs, μ_start)


def loop(s, T=(T0,), nlist=0.0, **kw):
    kw.setdefault("η_min", 0.0)
    kw.setdefault("η_init", 0.05)
    kw.setdefault("tol", 1e-6)
    kw.setdefault("maxiter", 30)
    kw.setdefault("ω_acoustic", 1e-6)
    return oracle.loop_scgd0(s, list(T), nlist, 1.0, **kw)


def test_fan_migdal_against_naive_loop():
    """test/test_scgd0_kernel.jl:6-44: graded grid, per-pair ωq, an acoustic mode with Inf g2
    (skipped, so no NaN), a random f -> i map and a complex Σin."""
    from test_gpu_scgd0 import kernel_case, run_fm_oracle
    c = kernel_case()
    ref = run_fm_oracle(c)
    naive = naive_fan_migdal(np.full_like(ref, 0.5), np.asarray(c["ωs"]), c["Sin"], c["g2"], c["wq"], c["e_f"],
                             c["f_to_i"], c["w_f"], c["mu"], c["T"], c["w_ac"])
    assert np.all(np.isfinite(ref))
    assert rel(ref, naive) <= 1e-13
    # a single term against the closed form with a constant Σin = -iη (no interpolation error)
    ws = np.asarray(c["ωs"])
    Sc = np.full((len(ws), 1), -0.05j)
    table = oracle.cell_search_table(ws, c["ωs"].dω_min)
    w, q, ef, T, mu = 0.1, 0.13, -0.2, 0.03, 0.07
    got = oracle.lib().orc_fm_term(w, 2.0, q, ef, 1, Sc.ctypes.data, len(ws), ws.ctypes.data, table[0].ctypes.data,
                                   table[1], table[2], mu, T, 0.5, 1e-3)
    nB = 1 / math.expm1(q / T)
    f = lambda x: 1 / (math.exp(x / T) + 1)
    la = -0.05 / ((w + q - ef) ** 2 + 0.05 ** 2)
    le = -0.05 / ((w - q - ef) ** 2 + 0.05 ** 2)
    assert math.isclose(got, 0.5 * 2.0 * (la * (nB + f(w + q - mu)) + le * (nB + 1 - f(w - q - mu))), rel_tol=1e-14)


@pytest.mark.parametrize("nlist,μ_fixed", [(0.0, 0.0), (-0.3, 0.3)])
def test_one_shot_closed_form(nlist, μ_fixed):
    """test/test_scgd0_holstein.jl:21-32 on the hand-built setup: maxiter = 1, fix_μ."""
    ωs = grid()
    r = loop(setup([μ_fixed]), nlist=nlist, maxiter=1, fix_μ=True)
    ImS = holstein_oneshot_imsigma(np.asarray(ωs), nk=NK, t=1.0, ω0=OMEGA0, g2=G2, T=T0, μ=μ_fixed, η=0.05)
    S_ref = oracle.kramers_kronig_matrix(np.asarray(ωs)) @ ImS + 1j * ImS
    assert r["Σ"].shape == (len(ωs), NK // 2 + 1, 1)
    for i in range(r["Σ"].shape[1]):
        assert np.linalg.norm(r["Σ"][:, i, 0] - S_ref) <= 1e-12 * np.linalg.norm(S_ref)
    assert r["μ"][0] == μ_fixed and not r["converged"][0]


@pytest.mark.parametrize("nlist", [0.0, -0.3])
def test_self_consistent_against_local_iteration(nlist):
    """test/test_scgd0_holstein.jl:34-50: Σ at 1e-10, μ at 1e-9, k-independence, chunking."""
    n_target = 1 + nlist
    μ0 = holstein_bare_mu(nk=NK, t=1.0, T=T0, n_target=n_target, spin_degeneracy=2)
    s = setup([μ0])
    r = loop(s, nlist=nlist)
    niter = r["history"]["err"].shape[1]
    assert r["converged"][0]
    assert np.max(np.abs(r["Σ"] - r["Σ"][:, :1, :])) < 1e-12
    ws = np.asarray(s.ωs)
    S_ref, μ_ref = holstein_local_scgd0(ws, np.asarray(s.ωs_spectral), oracle.kramers_kronig_matrix(ws), nk=NK,
                                        t=1.0, ω0=OMEGA0, g2=G2, T=T0, η_init=0.05, μ_start=μ0, niter=niter,
                                        fix_μ=False, n_target=n_target, spin_degeneracy=2)
    assert np.linalg.norm(r["Σ"][:, 0, 0] - S_ref) <= 1e-10 * np.linalg.norm(S_ref)
    assert abs(r["μ"][0] - μ_ref) < 1e-9
    if nlist != 0:
        assert abs(r["μ"][0] - μ0) > 1e-6
    rt = loop(s, nlist=nlist, gpu_tile=7)
    assert rel(rt["Σ"], r["Σ"]) <= 1e-12 and abs(rt["μ"][0] - r["μ"][0]) < 1e-12


def test_restart_two_temperatures_maxiter_no_root():
    μs = [holstein_bare_mu(nk=NK, t=1.0, T=T, n_target=1.0, spin_degeneracy=2) for T in (T0, 0.2)]
    s1 = setup([μs[0]])
    r = loop(s1)
    # restart from the converged result: one iteration (test/test_scgd0_holstein.jl:70-75)
    rr = loop(s1, Σ_init=r["Σ"], μ_start=r["μ"])
    assert rr["history"]["err"].shape[1] == 1 and rr["converged"][0]
    assert np.max(np.abs(rr["Σ"] - r["Σ"])) < 1e-6 and abs(rr["μ"][0] - r["μ"][0]) < 1e-6
    # two temperatures = each alone, each freezing at its own iteration (:77-85)
    r2 = loop(setup([μs[1]]), T=(0.2,))
    rb = loop(setup(μs), T=(T0, 0.2))
    for iT, one in enumerate((r, r2)):
        assert np.linalg.norm(rb["Σ"][:, :, iT] - one["Σ"][:, :, 0]) <= 1e-12 * np.linalg.norm(one["Σ"])
        assert abs(rb["μ"][iT] - one["μ"][0]) < 1e-12
    assert list(rb["history"]["iter_frozen"]) == [one["history"]["err"].shape[1] for one in (r, r2)]
    # maxiter without convergence (test/test_scgd0_loop.jl:12-18)
    rm = loop(setup([0.0]), maxiter=2)
    assert not rm["converged"][0] and rm["history"]["iter_frozen"][0] == 0 and rm["history"]["err"].shape[1] == 2
    # 2.5 electrons per cell do not fit in one spin-degenerate band (:20-23)
    with pytest.raises(oracle.OracleError) as e:
        loop(setup([0.0]), nlist=1.5)
    assert e.value.args[0] == 2


def test_multiplet_average():
    """test/test_scgd0_loop.jl:25-37."""
    s = setup([0.0])
    ni = NK // 2 + 1
    g2v = np.array([[0.2 * (1 + 0.1 * i)] * NK for i in range(1, ni + 1)]).reshape(1, ni, NK)
    sv = dataclasses.replace(s, g2=np.asfortranarray(g2v))
    kw = dict(maxiter=1, fix_μ=True)
    ra = loop(sv, **kw)
    rb = loop(dataclasses.replace(sv, multiplets=[[2, 3]]), **kw)
    assert np.linalg.norm(ra["Σ"][:, 1, 0] - ra["Σ"][:, 2, 0]) > 1e-3 * np.linalg.norm(ra["Σ"][:, 1, 0])
    assert np.array_equal(rb["Σ"][:, 1, 0], rb["Σ"][:, 2, 0])
    assert rel(rb["Σ"][:, 1, 0], (ra["Σ"][:, 1, 0] + ra["Σ"][:, 2, 0]) / 2) <= 1e-14
    keep = [0] + list(range(3, ni))
    assert np.array_equal(rb["Σ"][:, keep, :], ra["Σ"][:, keep, :])


def test_build_multiplets():
    """src/scgd0.jl:146-160: sort by (k, ε), chain gaps below the cutoff, groups >= 2."""
    iks = np.array([2, 1, 1, 2, 1, 2, 3], dtype=np.int32)
    es = np.array([0.5, 0.1, 0.1 + 5e-7, 0.5 + 9e-7, 0.3, 0.5 + 1.8e-6, 0.0])
    want = [[2, 3], [1, 4, 6]]
    assert build_multiplets(iks, es, 1e-6) == want
    assert [[int(x) for x in m] for m in oracle.build_multiplets(iks, es, 1e-6)] == want
    assert build_multiplets(iks, es, 1e-7) == []


def test_carrier_count_formula():
    """test/test_scgd0_mu.jl:33-51 against the old per-state formula at rtol 1e-12."""
    wsa = np.arange(-500, 501) / 500.0
    ni = 9
    ε_i = 0.7 * pseudorandom(ni, seed=12)
    w_i = 0.05 + 0.1 * np.abs(pseudorandom(ni, seed=13))
    Ssp = 0.05 * pseudorandom(len(wsa), ni, seed=14) + 1j * (-0.01 - 0.03 * np.abs(pseudorandom(len(wsa), ni, seed=15)))
    A = -np.imag(1 / (wsa[:, None] - ε_i[None, :] - Ssp)) / math.pi
    D = A @ w_i * trapezoid(wsa)
    W, nsb = w_i.sum(), 0.37
    for μ in (0.1, 0.9, -0.95):
        for T in (0.01, 0.05):
            for holes in (False, True):
                n = oracle.ncarrier_spectral(μ, T, wsa, D, W, holes, nsb, 2.0)
                assert math.isclose(n, n_old(μ, T, wsa, Ssp, ε_i, w_i, nsb, 2.0, holes), rel_tol=1e-12)


def test_mu_branch_per_type():
    """test/test_scgd0_mu.jl:53-74: the μ of one iteration satisfies the selected count only."""
    s = setup([0.0])
    ws, wsa = np.asarray(s.ωs), np.asarray(s.ωs_spectral)
    for semi, nlist, holes in ((False, -0.3, False), (True, 0.3, False), (True, -0.3, True)):
        r = oracle.loop_scgd0(s, [T0], nlist, 1.0, η_min=0.0, η_init=0.05, tol=1e-6, maxiter=1, ω_acoustic=1e-6,
                              semiconductor=semi)
        Ssp = np.stack([interp_flat(wsa, ws, r["Σ"][:, i, 0]) for i in range(r["Σ"].shape[1])], axis=1)
        n_of = lambda h: n_old(r["μ"][0], T0, wsa, Ssp, s.ε_i, s.w_i, s.nstates_base, 2.0, h)
        assert math.isclose(n_of(holes), 1 + nlist, rel_tol=1e-12)
        assert abs(n_of(not holes) - (1 + nlist)) > 1e-5


def test_run_record():
    """examples/scgd0/run_scgd0_holstein_1d.jl:33-36 (CPU, 16 threads; an earlier
    ElectronPhonon version): 11 iterations; the error sequence and μ to the record's 2-3 digits
    within the drift of that version (err within 30% per iteration, μ within 1e-5)."""
    rr = GOLD["holstein_run_record"]
    ωs = FrequencyGrid.from_nodes(np.arange(-200, 201) * rr["grid_step"])
    μ0 = holstein_bare_mu(nk=rr["nk"], t=rr["t"], T=rr["T"], n_target=1.0, spin_degeneracy=2)
    s = holstein_1d_setup(rr["nk"], rr["t"], rr["omega0"], rr["g2"], ωs, [μ0])
    r = oracle.loop_scgd0(s, [rr["T"]], 0.0, 1.0, η_min=0.0, η_init=rr["eta_init"], tol=rr["tol"], maxiter=20,
                          ω_acoustic=1e-6)
    err = r["history"]["err"][0]
    assert len(err) == rr["niter"] and r["converged"][0]
    assert np.all(np.abs(np.log(err / np.array(rr["err"]))) < math.log(1.3))
    assert abs(r["μ"][0] - rr["mu"]) < 1e-5
    assert abs(r["history"]["min_spectral_weight"][0, -1] - rr["min_spectral_weight"]) < 1e-4


def _v2_case(identity=True, nki=5, nkf=6):
    from toys import dispersive_setup, pair_frequencies
    s = dispersive_setup(identity=identity, nki=nki, nkf=nkf)
    return s, pair_frequencies(s)


def _v2_reference(s, ωq, Tlist, η_min, niter=2):
    ws = np.asarray(s.ωs)
    return scgd0_v2_reference(s.ε_i, s.w_i, s.nstates_base, s.g2, ωq, ws, np.asarray(s.ωs_spectral),
                              oracle.kramers_kronig_matrix(ws), Tlist, [0.0] * len(Tlist), 1.0, 2.0,
                              [s.μ_start[0]] * len(Tlist), η_min, 0.05, niter, 1e-6)


V2_T, V2_ETA = [0.03, 0.06], [0.0, 0.03]


def _oracle_v2(s, niter=2):
    return oracle.loop_scgd0(s, V2_T, 0.0, 1.0, η_min=V2_ETA, η_init=0.05, maxiter=niter, tol=0.0,
                             ω_acoustic=1e-6, μ_start=[s.μ_start[0]] * 2)


def test_dispersive_gather_and_eta_clamp_against_v2_reference():
    """test/test_scgd0_pb.jl:70-104 restated on a toy: phonon frequencies varying with q (one below
    ω_acoustic), a non-square, non-symmetric iq_kk, iks_i != iks_f, per-temperature η_min = [0, 0.03]
    (test/pb_fixture.jl:16), two iterations with the μ update.  The oracle's loop against the v2
    iteration over the per-pair ωq[ν, i, f] built by a plain loop: Σ at 1e-10, μ at 1e-9."""
    s, ωq = _v2_case()
    assert len(np.unique(ωq)) > 10 and np.any(ωq < 1e-6)
    r = _oracle_v2(s)
    Σ_ref, μ_ref = _v2_reference(s, ωq, V2_T, V2_ETA)
    assert rel(r["Σ"], Σ_ref) <= 1e-10
    assert np.max(np.abs(r["μ"] - μ_ref)) <= 1e-9
    # the clamp acts: the η_min = 0.03 temperature differs from the same run without it
    Σ_0, _ = _v2_reference(s, ωq, V2_T, [0.0, 0.0])
    assert rel(Σ_0[:, :, 1], Σ_ref[:, :, 1]) > 1e-3


@pytest.mark.parametrize("plant", ["transposed iq_kk", "swapped k maps", "no clamp"])
def test_dispersive_check_catches_planted_mistakes(plant):
    """Plausible mistakes planted in the oracle's inputs, each of which the v2 check must catch: the
    q map read transposed (iq_kk[iks_f[f], iks_i[i]]), the outer and inner k maps swapped, and the
    η_min clamp dropped.  Square iq_kk so that the transposed read stays in range."""
    import dataclasses
    s, ωq = _v2_case(nki=6, nkf=6)
    assert not np.array_equal(s.iq_kk, s.iq_kk.T) and not np.array_equal(s.iks_i, s.iks_f)
    Σ_ref, μ_ref = _v2_reference(s, ωq, V2_T, V2_ETA)
    assert rel(_oracle_v2(s)["Σ"], Σ_ref) <= 1e-10
    if plant == "transposed iq_kk":
        bad = dataclasses.replace(s, iq_kk=np.asfortranarray(s.iq_kk.T))
        r = _oracle_v2(bad)
    elif plant == "swapped k maps":
        bad = dataclasses.replace(s, iks_i=s.iks_f.copy(), iks_f=s.iks_i.copy())
        r = _oracle_v2(bad)
    else:
        r = oracle.loop_scgd0(s, V2_T, 0.0, 1.0, η_min=[0.0, 0.0], η_init=0.05, maxiter=2, tol=0.0,
                              ω_acoustic=1e-6, μ_start=[s.μ_start[0]] * 2)
    assert rel(r["Σ"], Σ_ref) > 1e-4
