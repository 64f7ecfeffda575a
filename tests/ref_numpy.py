"""Independent NumPy restatements of the reference's *test* oracles (not of the package code).

They pin the C oracle (oracle/epsx_oracle.c) without sharing its implementation:

* `pseudorandom`, `pseudorandom_complex`: deterministic test data, test/common.jl:11-13.
* `interp_flat`: `extrapolate(interpolate((ωs,), y, Gridded(Linear())), Flat())` = numpy.interp.
* `holstein_energies`, `holstein_oneshot_imsigma`, `holstein_local_scgd0`, `holstein_bare_mu`:
  test/holstein_1d_reference.jl:6-57.
* `naive_fan_migdal`: the plain 4-loop of test/test_scgd0_kernel.jl:32-42.
* `simpson_pv_kk`: the composite-Simpson subtracted PV integral of
  test/test_frequency_grids.jl:84-107.
* `n_old`: the per-state carrier count of test/test_scgd0_mu.jl:7-25.
* `scgd0_v2_reference`: the v2 scFM iteration as plain loops over a full per-pair ωq[ν, i, f]
  (test/test_scgd0_pb.jl:12-67): f -> i the identity, per-temperature η clamp, μ by a root
  finder on the bracket of the spectral grid ± 0.1.
"""
import math

import numpy as np
from scipy.optimize import brentq


def pseudorandom(*dims, seed):
    n = int(np.prod(dims))
    return np.sin(seed * np.arange(1, n + 1, dtype=np.float64) ** 1.5).reshape(dims, order="F")


def pseudorandom_complex(*dims, seed):
    return pseudorandom(*dims, seed=seed) + 1j * pseudorandom(*dims, seed=seed + 0.5)


def interp_flat(x, xp, fp):
    fp = np.asarray(fp)
    if np.iscomplexobj(fp):
        return np.interp(x, xp, fp.real) + 1j * np.interp(x, xp, fp.imag)
    return np.interp(x, xp, fp)


def fermi(x, T):
    with np.errstate(over="ignore"):
        return 1.0 / (np.exp(x / T) + 1.0)


def trapezoid(ws):
    ws = np.asarray(ws)
    n = len(ws)
    return np.array([(ws[min(r + 1, n - 1)] - ws[max(r - 1, 0)]) / 2 for r in range(n)])


def holstein_energies(nk, t):
    return np.array([-2 * t * math.cos(2 * math.pi * j / nk) for j in range(nk)])


def holstein_oneshot_imsigma(ws, *, nk, t, ω0, g2, T, μ, η):
    nB = 1 / math.expm1(ω0 / T)
    εs = holstein_energies(nk, t)
    ws = np.asarray(ws)
    out = np.zeros(len(ws))
    for k, ω in enumerate(ws):
        la = η / ((ω + ω0 - εs) ** 2 + η ** 2) * (nB + fermi(ω + ω0 - μ, T))
        le = η / ((ω - ω0 - εs) ** 2 + η ** 2) * (nB + 1 - fermi(ω - ω0 - μ, T))
        out[k] = -g2 / nk * np.sum(la + le)
    return out


def holstein_bare_mu(*, nk, t, T, n_target, spin_degeneracy):
    εs = holstein_energies(nk, t)
    f = lambda μ: spin_degeneracy * np.sum(fermi(εs - μ, T)) / nk - n_target
    return brentq(f, εs.min() - 1, εs.max() + 1, xtol=1e-15, rtol=4 * np.finfo(float).eps)


def holstein_local_scgd0(ws, wsa, KK, *, nk, t, ω0, g2, T, η_init, μ_start, niter, fix_μ, n_target,
                         spin_degeneracy, η_min=0.0):
    """test/holstein_1d_reference.jl:24-50: Σ is k-independent; one local G per iteration."""
    εs = holstein_energies(nk, t)
    nB = 1 / math.expm1(ω0 / T)
    ws, wsa = np.asarray(ws), np.asarray(wsa)
    Ww = trapezoid(wsa)
    Σ = np.full(len(ws), -1j * η_init)
    μ = μ_start
    for _ in range(niter):
        Σin = Σ.real + 1j * np.minimum(Σ.imag, -η_min)

        def ImG_loc(x):
            S = interp_flat(x, ws, Σin)
            return np.imag(np.sum(1 / (x[:, None] - εs[None, :] - S[:, None]), axis=1) / nk)

        ImΣ = g2 * (ImG_loc(ws + ω0) * (nB + fermi(ws + ω0 - μ, T))
                    + ImG_loc(ws - ω0) * (nB + 1 - fermi(ws - ω0 - μ, T)))
        Σ = KK @ ImΣ + 1j * ImΣ
        if fix_μ:
            continue
        Ssp = interp_flat(wsa, ws, Σ)
        D = np.sum(-np.imag(1 / (wsa[:, None] - εs[None, :] - Ssp[:, None])) / math.pi, axis=1) / nk * Ww
        g = lambda m: spin_degeneracy * np.sum(D * fermi(wsa - m, T)) - n_target
        μ = brentq(g, wsa[0] - 0.1, wsa[-1] + 0.1, xtol=1e-15, rtol=4 * np.finfo(float).eps, maxiter=500)
    return Σ, μ


def naive_fan_migdal(ImS, ws, Sin, g2, wq, e_f, f_to_i, w_f, mu, T, w_acoustic):
    """test/test_scgd0_kernel.jl:32-42 with numpy.interp lookups; wq is [ν, i, f]."""
    ImS = ImS.copy()
    nm, ni, nf = g2.shape
    ws = np.asarray(ws)
    for f in range(nf):
        col = f_to_i[f] - 1
        for i in range(ni):
            for nu in range(nm):
                q = wq[nu, i, f]
                if q < w_acoustic:
                    continue
                n = 1 / math.expm1(q / T)
                Sp = interp_flat(ws + q, ws, Sin[:, col])
                Sm = interp_flat(ws - q, ws, Sin[:, col])
                ImS[:, i] += w_f[f] * g2[nu, i, f] * (
                    np.imag(1 / (ws + q - e_f[f] - Sp)) * (n + fermi(ws + q - mu, T))
                    + np.imag(1 / (ws - q - e_f[f] - Sm)) * (n + 1 - fermi(ws - q - mu, T)))
    return ImS


def simpson_pv_kk(xg, zg, m=4000):
    """P∫ h(y)/(y-x) dy / π of the hat interpolant with one end cell of decay per side, as the
    subtracted integral plus h(x) log|(R-x)/(L-x)|, composite Simpson per cell."""
    xg = np.asarray(xg, dtype=float)
    xe = np.concatenate([[2 * xg[0] - xg[1]], xg, [2 * xg[-1] - xg[-2]]])
    ze = np.concatenate([[0.0], zg, [0.0]])
    h = lambda y: np.interp(y, xe, ze)
    out = []
    for x in xg:
        tot = h(x) * math.log(abs((xe[-1] - x) / (xe[0] - x)))
        for c in range(len(xe) - 1):
            if xe[c] <= x <= xe[c + 1]:
                tot += (ze[c + 1] - ze[c]) / (xe[c + 1] - xe[c]) * (xe[c + 1] - xe[c])
            else:
                ys = np.linspace(xe[c], xe[c + 1], m + 1)
                fv = (h(ys) - h(x)) / (ys - x)
                dy = ys[1] - ys[0]
                tot += dy / 3 * (fv[0] + fv[-1] + 4 * fv[1:-1:2].sum() + 2 * fv[2:-1:2].sum())
        out.append(tot / math.pi)
    return np.array(out)


def n_old(μ, T, wsa, Ssp, ε_i, w_i, nstates_base, spin_degeneracy, count_holes):
    Ww = trapezoid(wsa)
    fer = fermi(μ - wsa, T) if count_holes else fermi(wsa - μ, T)
    nocc = nbase = 0.0
    for i in range(len(ε_i)):
        A = -np.imag(1 / (wsa - ε_i[i] - Ssp[:, i])) / math.pi
        so = np.sum(A * fer * Ww)
        if count_holes:
            nocc -= so * w_i[i]
            nbase += w_i[i]
        else:
            nocc += so * w_i[i]
    return (nocc + nbase + nstates_base) * spin_degeneracy


def hat_value(y, ws, x):
    """Linear interpolation of y on ws, zero outside (test/test_ladder_kernel.jl:6-11); vector x."""
    ws = np.asarray(ws)
    x = np.asarray(x, dtype=float)
    out = np.zeros(x.shape, dtype=np.result_type(y, np.float64))
    inside = (x >= ws[0]) & (x <= ws[-1])
    k = np.minimum(np.searchsorted(ws, x[inside], side="right") - 1, len(ws) - 2)
    t = (x[inside] - ws[k]) / (ws[k + 1] - ws[k])
    out[inside] = (1 - t) * y[k] + t * y[k + 1]
    return out


def naive_apply_ladder(g2, wph, iq_kk, iks_i, iks_f, w_f, f_to_i, R, ws, KK, J, Gm, Gp, T, w_acoustic):
    """L(J) = -Gm (s ⊙ K[J]) Gp by loops, with the channel sums of apply_phonon_exchange_kernel!
    (test/test_ladder_kernel.jl:14-54): the shifted vertex by exact hat interpolation at
    ωs[j] ± ωq, no shift grid."""
    nm, ni, nf = g2.shape
    ws = np.asarray(ws)
    nw = len(ws)
    F = np.zeros((nw, 4, 3, 2, nf), dtype=complex)
    for f in range(nf):
        for c0 in range(2):
            for a in range(3):
                for c12 in range(4):
                    F[:, c12, a, c0, f] = sum(R[a, b, f] * J[:, c12, b, c0, f_to_i[f] - 1] for b in range(3))
    Up, Um, Vp, Vm = (np.zeros((nw, 4, 3, 2, ni), dtype=complex) for _ in range(4))
    for i in range(ni):
        for f in range(nf):
            q = iq_kk[iks_i[i] - 1, iks_f[f] - 1] - 1
            for nu in range(nm):
                om = wph[nu, q]
                if om < w_acoustic:
                    continue
                w = w_f[f] * g2[nu, i, f]
                n = 1 / math.expm1(om / T)
                for c0 in range(2):
                    for a in range(3):
                        for c12 in range(4):
                            y = F[:, c12, a, c0, f]
                            up = w * hat_value(y, ws, ws + om)
                            um = w * hat_value(y, ws, ws - om)
                            Up[:, c12, a, c0, i] += up
                            Um[:, c12, a, c0, i] += um
                            Vp[:, c12, a, c0, i] += n * up
                            Vm[:, c12, a, c0, i] += n * um
    KY = (KK @ ((Up - Um) / 2).reshape(nw, -1, order="F")).reshape(Up.shape, order="F")
    LJ = np.zeros((nw, 4, 3, 2, ni), dtype=complex)
    s = np.array([[1, -1], [-1, 1]])
    for i in range(ni):
        for c0 in range(2):
            for a in range(3):
                k11 = -1j * KY[:, 0, a, c0, i] - (Up[:, 0, a, c0, i] + Um[:, 0, a, c0, i]) / 2 - (Vp[:, 0, a, c0, i] + Vm[:, 0, a, c0, i])
                k22 = 1j * KY[:, 3, a, c0, i] - (Up[:, 3, a, c0, i] + Um[:, 3, a, c0, i]) / 2 - (Vp[:, 3, a, c0, i] + Vm[:, 3, a, c0, i])
                k21 = -(Vp[:, 1, a, c0, i] + Um[:, 1, a, c0, i] + Vm[:, 1, a, c0, i])
                k12 = -(Up[:, 2, a, c0, i] + Vp[:, 2, a, c0, i] + Vm[:, 2, a, c0, i])
                for j in range(nw):
                    Kab = s * np.array([[k11[j], k12[j]], [k21[j], k22[j]]])
                    gm = Gm[j, :, i].reshape(2, 2, order="F")
                    gp = Gp[j, :, i].reshape(2, 2, order="F")
                    LJ[j, :, a, c0, i] = -(gm @ Kab @ gp).reshape(4, order="F")
    return LJ


def spectral_bubble(ws, Σ, e_i, w_i, v, mu, T, Ω):
    """σ_ab = π Σ_i w_i v_a v_b Σ_ε Wω A_i(ε-Ω/2) A_i(ε+Ω/2) (f(ε-Ω/2-μ) - f(ε+Ω/2-μ))/Ω, with
    f(1-f)/T at Ω = 0 (test/test_ladder_bubble.jl:10-26); Σ [nω, ni] read by numpy.interp."""
    ws = np.asarray(ws)
    Ww = trapezoid(ws)
    σ = np.zeros((3, 3))
    for i in range(len(e_i)):
        A = lambda x: -np.imag(1 / (x - e_i[i] - interp_flat(x, ws, Σ[:, i]))) / math.pi
        if Ω == 0:
            fac = fermi(ws - mu, T) * (1 - fermi(ws - mu, T)) / T
        else:
            fac = (fermi(ws - Ω / 2 - mu, T) - fermi(ws + Ω / 2 - mu, T)) / Ω
        win = np.sum(Ww * A(ws - Ω / 2) * A(ws + Ω / 2) * fac)
        σ += math.pi * w_i[i] * win * np.outer(v[:, i], v[:, i])
    return σ


def scgd0_v2_reference(es, weights, nstates_base, g2, ωq, ws, wsa, KK, Tlist, nlist, nelec, spin_degeneracy,
                       μlist, ηlist, η_init, niter, ω_acoustic):
    """test/test_scgd0_pb.jl:12-67.  The outer and inner states are the same n states (f -> i the
    identity); ωq[ν, i, f] is the per-pair phonon frequency.  Returns (Σ [nω, n, nT], μ [nT])."""
    ws, wsa = np.asarray(ws, dtype=float), np.asarray(wsa, dtype=float)
    nω, n, nT, nm = len(ws), len(es), len(Tlist), g2.shape[0]
    Σs = np.full((nω, n, nT), complex(0.0, -η_init))
    μlist = [float(m) for m in μlist]
    Ww = trapezoid(wsa)
    for _ in range(niter):
        Σs_new = np.zeros((nω, n, nT), dtype=complex)
        for iT in range(nT):
            T, μ = Tlist[iT], μlist[iT]
            for f in range(n):
                Σ_f = Σs[:, f, iT]
                Σ_clamped = Σ_f.real + 1j * np.minimum(Σ_f.imag, -ηlist[iT])
                ekq, wkq = es[f], weights[f]
                for i in range(n):
                    for nu in range(nm):
                        q = ωq[nu, i, f]
                        if q < ω_acoustic:
                            continue
                        nq = 1 / math.expm1(q / T)
                        A1 = np.imag(1 / (ws + q - ekq - interp_flat(ws + q, ws, Σ_clamped))) * (nq + fermi(ws + q - μ, T))
                        A2 = np.imag(1 / (ws - q - ekq - interp_flat(ws - q, ws, Σ_clamped))) * (nq + 1 - fermi(ws - q - μ, T))
                        Σs_new[:, i, iT] += (A1 + A2) * g2[nu, i, f] * wkq * 1j
        for iT in range(nT):
            for i in range(n):
                Σs_new[:, i, iT] += KK @ Σs_new[:, i, iT].imag
        Σs = Σs_new
        for iT in range(nT):
            T = Tlist[iT]
            Sp = np.stack([interp_flat(wsa, ws, Σs[:, i, iT]) for i in range(n)], axis=1)

            def nocc(μ_):
                fer = fermi(wsa - μ_, T)
                tot = 0.0
                for i in range(n):
                    As = -np.imag(1 / (wsa - es[i] - Sp[:, i])) / math.pi
                    tot += np.sum(As * fer * Ww) * weights[i]
                return (tot + nstates_base) * spin_degeneracy

            n_target = nlist[iT] + nelec
            μlist[iT] = brentq(lambda m: nocc(m) - n_target, wsa[0] - 0.1, wsa[-1] + 0.1, xtol=1e-15,
                               rtol=4 * np.finfo(float).eps, maxiter=500)
    return Σs, np.array(μlist)
