"""Synthetic scGD0 setups: the inputs that `_setup_scgd0` would extract, built directly.

* `holstein_1d_setup`: the 1D Holstein chain of test/holstein_1d_reference.jl:62-77 (outer
  states on the k -> -k wedge, one dispersionless mode, constant coupling).
* `t2g_setup`: a cubic three-band t2g tight-binding model on a Γ-centred nk³ grid reduced by
  the 48 operations of Oh, with Einstein modes and seeded log-normal couplings.

Energies at symmetry-related k are computed from one canonical k, so ε_f equals ε of its IBZ
representative bitwise and the degenerate multiplets are exact.
"""
import math

import numpy as np

from .grids import FrequencyGrid, graded_frequency_grid, uniform_spectral_grid
from .scgd0 import Scgd0Setup, build_multiplets, ELECTRON_DEGEN_CUTOFF

KB_EV = 8.617333262e-5   # Boltzmann constant in eV/K (CODATA 2018)


def fermi(x, T):
    with np.errstate(over="ignore"):
        return 1.0 / (np.exp(x / T) + 1.0)


def bare_mu(es, ws, T, n_target, spin_degeneracy=2.0, nstates_base=0.0):
    """The μ at which spin * (Σ w f(ε - μ) + nstates_base) = n_target (the bare-band μ of the
    window states), by bisection to adjacent floats."""
    es = np.asarray(es, dtype=np.float64)
    ws = np.asarray(ws, dtype=np.float64)

    def g(mu):
        return spin_degeneracy * (np.sum(ws * fermi(es - mu, T)) + nstates_base) - n_target

    a, b = es.min() - 1.0, es.max() + 1.0
    fa = g(a)
    for _ in range(200):
        m = 0.5 * (a + b)
        if m == a or m == b:
            break
        fm = g(m)
        if (fm > 0) == (fa > 0):
            a, fa = m, fm
        else:
            b = m
    return a


def holstein_energies(nk, t):
    """ε_j = -2t cos(2πj/nk), evaluated at min(j, nk - j) so that k and -k agree bitwise."""
    j = np.arange(nk)
    jj = np.minimum(j, nk - j)
    return -2.0 * t * np.cos(2.0 * np.pi * jj / nk)


def holstein_1d_setup(nk, t, ω0, g2, ωs, μ_start, ωs_spectral=None):
    """The hand-built Holstein input of test/holstein_1d_reference.jl:62-77."""
    ni = nk // 2 + 1
    εs = holstein_energies(nk, t)
    ji = np.arange(ni)
    jf = np.arange(nk)
    iq_kk = (np.mod(jf[None, :] - ji[:, None], nk) + 1).astype(np.int32)
    return Scgd0Setup(
        g2=np.full((1, ni, nk), float(g2), order="F"),
        ωph=np.full((1, nk), float(ω0), order="F"),
        iq_kk=np.asfortranarray(iq_kk),
        iks_i=np.arange(1, ni + 1, dtype=np.int32),
        iks_f=np.arange(1, nk + 1, dtype=np.int32),
        ε_i=εs[:ni].copy(),
        w_i=np.array([(1 if (j == 0 or 2 * j == nk) else 2) / nk for j in range(ni)]),
        ε_f=εs.copy(),
        w_f=np.full(nk, 1.0 / nk),
        f_to_i=np.array([min(j, nk - j) + 1 for j in range(nk)], dtype=np.int32),
        W=1.0, nstates_base=0.0, multiplets=[],
        ωs=ωs, ωs_spectral=ωs_spectral if ωs_spectral is not None else uniform_spectral_grid(ωs),
        μ_start=np.atleast_1d(np.asarray(μ_start, dtype=np.float64)),
    )


def t2g_bands(cx, cy, cz, t, tp):
    """The three t2g bands (xy, yz, zx) at cos k = (cx, cy, cz), sorted ascending."""
    exy = -2 * t * (cx + cy) - 2 * tp * cz
    eyz = -2 * t * (cy + cz) - 2 * tp * cx
    ezx = -2 * t * (cz + cx) - 2 * tp * cy
    return np.sort(np.stack([exy, eyz, ezx], axis=-1), axis=-1)


def oh_ibz(nk):
    """Γ-centred nk³ grid reduced by Oh (sign flips and permutations of the axes).

    Returns (canon index of each full-grid k [nk³], list of canonical triples (a<=b<=c, each in
    0..nk//2), star sizes)."""
    n = np.arange(nk)
    fold = np.minimum(n, nk - n)
    k1, k2, k3 = np.meshgrid(fold, fold, fold, indexing="ij")
    trip = np.sort(np.stack([k1.ravel(order="F"), k2.ravel(order="F"), k3.ravel(order="F")], axis=1), axis=1)
    h = nk // 2 + 1
    key = trip[:, 0] * h * h + trip[:, 1] * h + trip[:, 2]
    uniq, inv, counts = np.unique(key, return_inverse=True, return_counts=True)
    canon = np.stack([uniq // (h * h), (uniq // h) % h, uniq % h], axis=1)
    return inv, canon, counts


def t2g_setup(nk=24, t=0.26, tp=0.03, nmodes=9, wmin=0.010, wmax=0.080, g2_median=2e-3,
              g2_sigma_ln=1.0, window=0.8, nelec=1.0, Tlist_K=(100.0, 300.0), nlist=0.0, seed=1234,
              grid_windows=((0.3, 1e-3), (1.2, 4e-3), (3.0, 16e-3)), spectral_step=None,
              degen_cutoff=ELECTRON_DEGEN_CUTOFF, make_g2=True, uniform_grid=None, forward_q=None):
    """The synthetic cubic t2g model of the benchmark, in eV.

    * bands ε_xy = -2t(cos kx + cos ky) - 2t' cos kz and cyclic, sorted per k;
    * outer states: Oh-irreducible k x bands within ±`window` of the bare μ at the first
      temperature; inner states: the full grid in the same window;
    * `nmodes` Einstein modes ω_ν ~ U(wmin, wmax); g2[ν,i,f] = |g|²/(2ω_ν) with
      |g|² ~ LogNormal(ln g2_median, g2_sigma_ln), i.i.d., seeded;
    * ωs graded around the bare μ: 1 meV to ±0.3 eV, 4 meV to ±1.2 eV, 16 meV to ±3 eV, or, with
      `uniform_grid = (L, half_width)`, L uniform nodes over bare μ ± half_width;
      ωs_spectral uniform at ωs.dω_min over the span of ωs (the run_scgd0 default);
    * with `forward_q = q0` (grid steps) the couplings are also weighted by the momentum transfer,
      0.25 + 3 exp(-|k_f - k_i|² / (2 q0²)) with |k_f - k_i| the minimal-image distance between the
      inner state's k and the outer state's representative k: forward scattering, which gives the
      ladder a vertex correction (i.i.d. couplings average it out).
    """
    rng = np.random.default_rng(seed)
    inv, canon, counts = oh_ibz(nk)
    nkf = nk ** 3
    ctab = np.cos(2 * np.pi * np.arange(nk // 2 + 1) / nk)
    e_ibz = t2g_bands(ctab[canon[:, 0]], ctab[canon[:, 1]], ctab[canon[:, 2]], t, tp)  # (nkibz, 3)
    e_full = e_ibz[inv]                                                                 # (nkf, 3)
    Tlist = np.asarray(Tlist_K, dtype=np.float64) * KB_EV
    mu_all = bare_mu(e_full.ravel(), np.full(e_full.size, 1.0 / nkf), Tlist[0], nelec)
    lo, hi = mu_all - window, mu_all + window
    nstates_base = float(np.sum(e_full < lo)) / nkf
    # outer states: (IBZ k, band) in the window, ordered by k then band
    sel_i = [(k, b) for k in range(len(canon)) for b in range(3) if lo <= e_ibz[k, b] <= hi]
    index_i = {kb: n for n, kb in enumerate(sel_i)}
    ε_i = np.array([e_ibz[k, b] for k, b in sel_i])
    w_i = np.array([counts[k] / nkf for k, b in sel_i])
    iks_i = np.array([k + 1 for k, b in sel_i], dtype=np.int32)
    # inner states: (full k, band) in the window
    sel_f = [(k, b) for k in range(nkf) for b in range(3) if lo <= e_full[k, b] <= hi]
    ε_f = np.array([e_full[k, b] for k, b in sel_f])
    f_to_i = np.array([index_i[(inv[k], b)] + 1 for k, b in sel_f], dtype=np.int32)
    ni, nf = len(sel_i), len(sel_f)
    ω_ν = np.sort(rng.uniform(wmin, wmax, size=nmodes))
    if make_g2:
        g2 = np.empty((nmodes, ni, nf), order="F")
        chunk = max(1, (1 << 25) // (nmodes * ni))
        for f0 in range(0, nf, chunk):
            f1 = min(nf, f0 + chunk)
            absg2 = rng.lognormal(mean=math.log(g2_median), sigma=g2_sigma_ln, size=(nmodes, ni, f1 - f0))
            if forward_q is not None:
                ki = np.array([canon[k] for k, b in sel_i], dtype=np.int64)             # (ni, 3)
                kf = np.array([k for k, b in sel_f[f0:f1]], dtype=np.int64)
                kfv = np.stack([kf % nk, (kf // nk) % nk, kf // (nk * nk)], axis=1)    # (nc, 3)
                d = np.abs(kfv[None, :, :] - ki[:, None, :]) % nk
                d2 = np.sum(np.minimum(d, nk - d) ** 2, axis=2).astype(np.float64)    # (ni, nc)
                absg2 *= (0.25 + 3.0 * np.exp(-d2 / (2.0 * forward_q ** 2)))[None, :, :]
            g2[:, :, f0:f1] = absg2 / (2.0 * ω_ν[:, None, None])
    else:
        g2 = np.empty((nmodes, ni, 0), order="F")
    μ_start = np.array([bare_mu(ε_i, w_i, T, nelec + nl, 2.0, nstates_base)
                        for T, nl in zip(Tlist, np.broadcast_to(nlist, Tlist.shape))])
    if uniform_grid is None:
        wins = [(mu_all - w, mu_all + w, dw) for (w, dw) in grid_windows]
        ωs = graded_frequency_grid(wins)
    else:
        L, hw = uniform_grid
        ωs = FrequencyGrid(mu_all - hw, 2.0 * hw / (L - 1), np.arange(L, dtype=np.int64))
    ωs_spectral = (uniform_spectral_grid(ωs) if spectral_step is None else
                   FrequencyGrid(ωs[0], spectral_step, np.arange(int(round((ωs[-1] - ωs[0]) / spectral_step)) + 1)))
    setup = Scgd0Setup(
        g2=g2, ωph=np.asfortranarray(ω_ν.reshape(nmodes, 1)),
        iq_kk=np.ones((1, 1), dtype=np.int32, order="F"),
        iks_i=np.ones(ni, dtype=np.int32), iks_f=np.ones(nf, dtype=np.int32),
        ε_i=ε_i, w_i=w_i, ε_f=ε_f, w_f=np.full(nf, 1.0 / nkf), f_to_i=f_to_i,
        W=float(w_i.sum()), nstates_base=nstates_base,
        multiplets=build_multiplets(iks_i, ε_i, degen_cutoff),
        ωs=ωs, ωs_spectral=ωs_spectral, μ_start=μ_start)
    meta = dict(nk=nk, t=t, tp=tp, ω_ν=ω_ν, Tlist=Tlist, Tlist_K=np.asarray(Tlist_K, dtype=float),
                mu_bare=mu_all, window=(lo, hi), ni=ni, nf=nf, seed=seed, canon=canon, inv=inv,
                counts=counts, sel_i=sel_i, sel_f=sel_f)
    return setup, meta


def holstein_ladder_grid():
    """The ladder fixture grid of test/ladder_fixtures.jl:23: 0.02 cells inside ±1, 0.04 outside."""
    nodes = np.concatenate([np.arange(-100, -24) * 0.04, np.arange(-49, 50) * 0.02, np.arange(25, 101) * 0.04])
    return FrequencyGrid.from_nodes(np.round(nodes, 12))


def holstein_ladder_setup(nk=16, ω0=0.173, t=1.0, λ=0.5, symmetry=False, ωs=None, μ_start=(0.0, 0.0)):
    """The 1D Holstein ladder case of test/ladder_fixtures.jl:22-31 as (scGD0 setup, R, v).

    symmetry=False: outer states are all nk states (f_to_i the identity, R = 1).  symmetry=True:
    the k -> -k wedge, with R_f = -1 for the partner -k of a representative (the velocity is
    odd under k -> -k).  g2 = 2 λ ω0 t (one dimension); v_x = 2t sin k."""
    from .scgd0 import Scgd0Setup
    ωs = holstein_ladder_grid() if ωs is None else ωs
    g2 = 2.0 * λ * ω0 * t
    if symmetry:
        s = holstein_1d_setup(nk, t, ω0, g2, ωs, list(μ_start))
        ni = nk // 2 + 1
        R = np.zeros((3, 3, nk), order="F")
        for f in range(nk):
            sgn = 1.0 if f <= nk // 2 else -1.0
            for a in range(3):
                R[a, a, f] = sgn
        js = np.arange(ni)
    else:
        εs = holstein_energies(nk, t)
        j = np.arange(nk)
        iq_kk = (np.mod(j[None, :] - j[:, None], nk) + 1).astype(np.int32)
        s = Scgd0Setup(
            g2=np.full((1, nk, nk), g2, order="F"), ωph=np.full((1, nk), ω0, order="F"),
            iq_kk=np.asfortranarray(iq_kk), iks_i=np.arange(1, nk + 1, dtype=np.int32),
            iks_f=np.arange(1, nk + 1, dtype=np.int32), ε_i=εs.copy(), w_i=np.full(nk, 1.0 / nk), ε_f=εs.copy(),
            w_f=np.full(nk, 1.0 / nk), f_to_i=np.arange(1, nk + 1, dtype=np.int32), W=1.0, nstates_base=0.0,
            multiplets=[], ωs=ωs, ωs_spectral=uniform_spectral_grid(ωs),
            μ_start=np.atleast_1d(np.asarray(μ_start, dtype=np.float64)))
        R = np.zeros((3, 3, nk), order="F")
        for a in range(3):
            R[a, a, :] = 1.0
        js = np.arange(nk)
    v = np.zeros((3, len(js)), order="F")
    v[0] = 2.0 * t * np.sin(2.0 * np.pi * js / nk)
    return s, R, v


def oh_operations():
    """The 48 signed permutation matrices of Oh acting on Cartesian k (and on vectors)."""
    import itertools
    ops = []
    for perm in itertools.permutations(range(3)):
        for signs in itertools.product((1, -1), repeat=3):
            S = np.zeros((3, 3))
            for a in range(3):
                S[a, perm[a]] = signs[a]
            ops.append(S)
    return np.array(ops)


def t2g_velocities(kvec, t, tp):
    """Band energies and velocities ∇ε (eV times the lattice constant) of the sorted t2g bands at
    the reduced k points kvec [n, 3] (radians).  Returns (ε [n, 3], v [n, 3 bands, 3 cart])."""
    s = np.sin(kvec)
    c = np.cos(kvec)
    cx, cy, cz = c[:, 0], c[:, 1], c[:, 2]
    e = np.stack([-2 * t * (cx + cy) - 2 * tp * cz,     # xy
                  -2 * t * (cy + cz) - 2 * tp * cx,     # yz
                  -2 * t * (cz + cx) - 2 * tp * cy],    # zx
                 axis=1)
    coef = np.array([[t, t, tp], [tp, t, t], [t, tp, t]])  # orbital x cart hopping of ∂/∂k
    v = 2.0 * coef[None, :, :] * s[:, None, :]
    order = np.argsort(e, axis=1, kind="stable")
    idx = np.arange(len(kvec))[:, None]
    return e[idx, order], v[idx, order]


def t2g_ladder_setup(nk=16, t=0.26, tp=0.03, **kw):
    """The t2g model of `t2g_setup` with what the ladder adds (src/ladder.jl:112-158):

    * v[:, i]: the band velocity of outer state i at its IBZ representative;
    * R[:, :, f]: the averaged vector unfolding of the representative to inner state f, the mean
      of the Oh operations S with S k_rep = k_f (mod the grid), so that v_f = R_f v_{i_f} where
      the band is non-degenerate (a vector field odd under k -> -k, as inversion is in Oh).

    Returns (scGD0 setup, R, v, meta)."""
    setup, meta = t2g_setup(nk=nk, t=t, tp=tp, **kw)
    canon, sel_i, sel_f = meta["canon"], meta["sel_i"], meta["sel_f"]
    inv = meta["inv"]
    two_pi_n = 2.0 * np.pi / nk
    # outer-state velocities at the canonical k of each representative
    rep_k = np.array([canon[k] for k, b in sel_i], dtype=np.float64) * two_pi_n
    _, v_rep = t2g_velocities(rep_k, t, tp)
    bands_i = np.array([b for k, b in sel_i])
    v = np.asfortranarray(v_rep[np.arange(len(sel_i)), bands_i].T)      # (3, ni)
    # inner states: full-grid triples (n1 fastest, as the grid is enumerated)
    kf = np.array([k for k, b in sel_f])
    nvec = np.stack([kf % nk, (kf // nk) % nk, kf // (nk * nk)], axis=1)
    rep = canon[inv[kf]]                                                   # (nf, 3) canonical triple
    ops = oh_operations()                                                  # (48, 3, 3)
    mapped = np.einsum("sab,fb->sfa", ops, rep.astype(np.float64))
    hit = np.all(np.mod(np.rint(mapped).astype(np.int64), nk) == nvec[None, :, :], axis=2)   # (48, nf)
    count = hit.sum(axis=0)
    if not np.all(count > 0):
        raise RuntimeError("an inner state has no Oh operation from its representative")
    R = np.einsum("sf,sab->abf", hit.astype(np.float64), ops) / count[None, None, :]
    meta = dict(meta, v_rep=v_rep, bands_i=bands_i, kf=kf, nvec=nvec)
    return setup, np.asfortranarray(R), v, meta
